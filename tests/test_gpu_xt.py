"""XT maps (pic_sim_set_xt / pic_sim_read_xt) against the oracle's
restatement of eduPIC's XT diagnostics (oracle/xt.py): one measured step
on an ion step, fed with the uploaded particles and the field and densities
of the step's solve; and the bookkeeping of several steps over two bins."""
import numpy as np
import pytest

from helpers import synthetic_state
from oracle import xt as ox

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


def _oracle_step(orc, p, t, E, I, grid, bin_steps=20):
    pot, ef, ne, ni = grid
    n_xt = p.n_t // bin_steps
    raw, cnt = ox.new_raw(p.n_g, n_xt)
    tb = t // bin_steps
    ox.add_grid(raw, cnt, tb, pot, ef, ne, ni)
    cs, _, _ = orc.cross_sections(p.gas_density)
    sigma_ion = cs[2 * orc.CS_RANGES:3 * orc.CS_RANGES]
    ox.add_particles(raw, tb, *E, ef, p.dx, p.dt_e, True, sigma_ion, p.gas_density)
    if t % p.n_sub == 0:
        ox.add_particles(raw, tb, *I, ef, p.dx, p.dt_i, False)
    return ox.normalise(raw, cnt, p.weight, p.area, p.dx, p.dt_e)


@pytest.mark.parametrize("t0", [20, 21])
def test_one_measured_step_matches_the_oracle(ca, orc, t0):
    p = ca.params(weight=2.0e3, precision=2)
    E, I = synthetic_state(60_000, 63_000, seed=91, te_ev=6.0)
    with ca.Sim(p, capacity=150_000) as sim:
        sim.set_xt(20)
        sim.set_time(t0)
        sim.upload(0, *E)
        sim.upload(1, *I)
        sim.run(1, measure=True)
        got = sim.xt()
        grid = sim.grid()
    want = _oracle_step(orc, p, t0, E, I, grid)
    tb = t0 // 20
    for name in ox.NAMES:
        g, w = got[name], want[name]
        assert g.shape == (p.n_g, p.n_t // 20)
        others = np.delete(g, tb, axis=1)
        assert np.all(others == 0.0), name
        scale = np.abs(w[:, tb]).max()
        if t0 % 20 and name in ("ji", "poweri", "meanei"):
            assert scale == 0.0 and np.all(g[:, tb] == 0.0), name  # no ion step, no ion maps
            continue
        assert scale > 0.0, name
        # f32 LDS sums on the GPU against f64 sums of the same particles; the
        # mean velocities cancel thermal spread, and the ionisation rate may
        # take the neighbouring table entry at a rounding boundary
        err = np.abs(g[:, tb] - w[:, tb]).max() / scale
        tol = 1e-4 if name in ("je", "ji", "powere", "poweri", "ioniz") else 2e-5
        assert err < tol, (name, err)


def test_bins_and_reset(ca):
    p = ca.params(weight=2.0e3, precision=2)
    E, I = synthetic_state(30_000, 31_500, seed=92)
    with ca.Sim(p, capacity=80_000) as sim:
        sim.upload(0, *E)
        sim.upload(1, *I)
        with pytest.raises(Exception):
            sim.xt()                       # not enabled
        with pytest.raises(Exception):
            sim.set_xt(7)                  # does not divide n_t
        sim.set_xt(40)
        sim.run(80, measure=True)          # bins 0 and 1 of 100
        m = sim.xt()
        assert m["pot"].shape == (p.n_g, 100)
        assert np.all(m["ne"][:, :2].sum(axis=0) > 0) and np.all(m["ne"][:, 2:] == 0)
        assert np.all(m["meanee"][100:300, :2] > 0.1)
        sim.run(80)                        # not measured: nothing changes
        np.testing.assert_array_equal(sim.xt()["ne"], m["ne"])
        sim.reset_measurements()
        assert not np.any(sim.xt()["ne"])
        sim.set_xt(0)
        with pytest.raises(Exception):
            sim.xt()


def test_xt_is_a_fused_path_measurement(ca):
    p = ca.params(weight=2.0e3)
    with ca.Sim(p, capacity=10_000, path=ca.PIC_PATH_REFERENCE) as sim:
        with pytest.raises(Exception):
            sim.set_xt(20)



def test_xt_of_an_empty_discharge(ca):
    """No particles: the grid maps hold the vacuum potential and field, the
    particle maps stay zero."""
    p = ca.params(weight=2.0e3, precision=2)
    with ca.Sim(p, capacity=10_000) as sim:
        sim.set_xt(20)
        sim.upload(0, *[np.zeros(0, np.float32)] * 4)
        sim.upload(1, *[np.zeros(0, np.float32)] * 4)
        sim.run(40, measure=True)
        m = sim.xt()
        for name in ("ne", "ni", "je", "ji", "meanee", "meanei", "ioniz", "powere", "poweri"):
            assert not np.any(m[name]), name
        assert np.abs(m["pot"][0, :2]).max() > 0 and np.isfinite(m["efield"]).all()


def test_xt_on_the_finest_grid(ca):
    """N_G = 4000 (128 KB of LDS accumulators per workgroup) still measures."""
    p = ca.params(weight=2.0e3, n_g=4000, precision=2)
    E, I = synthetic_state(30_000, 31_500, seed=93)
    with ca.Sim(p, capacity=80_000) as sim:
        sim.set_xt(20)
        sim.upload(0, *E)
        sim.upload(1, *I)
        sim.run(40, measure=True)
        m = sim.xt()
        assert m["ne"].shape == (4000, 200)
        for name, a in m.items():
            assert np.isfinite(a).all(), name
        assert np.all(m["ne"][1000:3000, :2] > 0) and np.all(m["meanee"][1000:3000, :2] > 0)
        assert not np.any(m["ne"][:, 2:])
