"""Ensembles of independent discharges in one launch per stage: each
instance's step equals the single-discharge step of its own parameters
(checked stage by stage against the oracle as in test_gpu_sim.py), and
whole runs reproduce the reference's init-cycle statistics per instance."""
import numpy as np
import pytest

from helpers import SEED_E, SEED_I, synthetic_state
from test_gpu_sim import _push_ref, _replay_species, _split_ions

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


def test_ensemble_step_per_instance_replay(ca, orc):
    specs = [(250.0, 10.0, 2.0e3, 60_000), (120.0, 4.0, 5.0e3, 30_000), (380.0, 25.0, 1.0e3, 90_000)]
    ps = [ca.params(v0=v, pressure=pr, weight=w, precision=2) for v, pr, w, _ in specs]
    caps = [int(n * 1.3) for *_, n in specs]
    ens = ca.Ensemble(ps, caps, use_graph=False)
    states = []
    for i, (v, pr, w, n) in enumerate(specs):
        E, I = synthetic_state(n, int(n * 1.05), seed=80 + i, te_ev=6.0)
        ens.upload(i, 0, *E)
        ens.upload(i, 1, *I)
        states.append((E, I))
    rng0 = [(ens.rng_download(i, 0, caps[i]), ens.rng_download(i, 1, caps[i])) for i in range(3)]
    ens.run(1, measure=True)
    ne, ni = ens.counts()
    born_total = 0
    for i, (v, pr, w, n) in enumerate(specs):
        po = orc.params(v0=v, pressure=pr, weight=w)
        E, I = states[i]
        pot, ef = ens.grid(i)
        ce, ci, cnt = ens.measurements(i)
        # densities of the initial state (cumul after one measured step)
        for got, parts in ((ce, E), (ci, I)):
            want = orc.deposit(parts[0], po)
            assert np.abs(got - want).max() <= 2e-6 * want.max()
        pw, ew = orc.poisson_scan_f64(ce, ci, v, po)
        assert np.abs(pot - pw).max() <= 1e-5 * np.abs(pw).max()
        assert np.abs(ef - ew).max() <= 1e-5 * np.abs(ew).max()
        cs, se, si = orc.cross_sections(po.gas_density)
        ge, gi = ens.download(i, 0), ens.download(i, 1)
        cnt_e = _push_ref(orc, po, ef, E, "e")[0]
        cnt_i = _push_ref(orc, po, ef, I, "i")[0]
        re1 = ens.rng_download(i, 0, caps[i])
        ri1 = ens.rng_download(i, 1, caps[i])
        # survivors fill the instance's region from its start, newborn
        # electrons wait in the staging array (no slot)
        slots_e = np.concatenate([np.arange(cnt_e), -np.ones(len(ge[0]) - cnt_e, np.int64)])
        born_x, ions = _replay_species(orc, po, se, cs, E, ef, ge, slots_e, rng0[i][0], re1, "e")
        _, old_ion, n_amb, _ = _split_ions(gi, _push_ref(orc, po, ef, I, "i")[1][0])
        # survivors fill the first cnt_i slots in order
        assert not old_ion[cnt_i:].any()
        _replay_species(orc, po, si, cs, I, ef, [a[old_ion] for a in gi],
                        np.arange(len(gi[0]))[old_ion], rng0[i][1], ri1, "i", missing=n_amb)
        # (slots are assigned in atomic order, so the RNG stream a particle
        # meets and the number of ionisations vary from run to run)
        assert ni[i] - cnt_i == ne[i] - cnt_e == len(born_x)
        born_total += len(born_x)
        assert cnt[6] == 1 and cnt[7] == 1 and cnt[4] > 0
    assert born_total > 0
    ens.check()
    ens.close()


def test_ensemble_init_cycle_statistics(ca):
    """Four copies of the reference case from 1000 e + 1000 i at rest end the
    init cycle near eduPIC's published 888 e / 3345 i; a fifth instance at
    100 V / 3 Pa runs alongside without disturbing them."""
    p_ref = ca.params()
    ps = [p_ref] * 4 + [ca.params(v0=100.0, pressure=3.0)]
    ens = ca.Ensemble(ps, [20_000] * 5, use_graph=True)
    rng = np.random.default_rng(9)
    z = np.zeros(1000, np.float32)
    for i in range(5):
        ens.upload(i, 0, (rng.random(1000) * 0.025).astype(np.float32), z, z, z)
        ens.upload(i, 1, (rng.random(1000) * 0.025).astype(np.float32), z, z, z)
    ens.run(4000)
    ens.check()
    ne, ni = ens.counts()
    assert abs(ne[:4].mean() - 888) < 0.15 * 888, ne
    assert abs(ni[:4].mean() - 3345) < 0.15 * 3345, ni
    assert ne[4] != ne[0] and ne[4] > 0
    ens.close()


def test_ensemble_rejects_mismatched_grids(ca):
    with pytest.raises(ca.PicError):
        ca.Ensemble([ca.params(), ca.params(n_g=200)], [1000, 1000])
