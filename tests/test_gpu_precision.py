"""Precision stress (N_G = 4000, ~3.7 M electrons + ~3.9 M ions, W = 2e3):
the FP32 scan with the f32 density flush (the reference), the compensated
(float-float) scan with the f32 flush, and the FP64 scan with the FP64
flush, each against an all-FP64 oracle -- FP64 CIC weights of the same
particles and eduPIC's FP64 Thomas solve -- on ion steps of a running
discharge (the fused path), and the 20-cycle averaged densities of each
mode against the FP64 mode, beside the spread between two FP64 runs with
different RNG seeds (the statistical error of the average)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
N_G = 4000
MODES = [(0, "fp32"), (1, "comp"), (2, "fp64")]
E_CHARGE = 1.60217662e-19


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


@pytest.fixture(scope="module")
def state():
    from cudapic_amd import synthetic
    return synthetic.state(3_740_000, 3_930_000, seed=13)


def _sim(ca, prec, state, seed=None):
    p = ca.params(n_g=N_G, weight=2.0e3, precision=prec)
    kw = {} if seed is None else {"seed_e": (seed, 2, 3, 4), "seed_i": (seed, 6, 7, 8)}
    s = ca.Sim(p, capacity=8_000_000, **kw)
    s.upload(0, *state[0])
    s.upload(1, *state[1])
    return p, s


@pytest.mark.parametrize("prec,name", MODES)
def test_field_error_per_step_against_fp64_oracle(ca, orc, state, prec, name):
    p, sim = _sim(ca, prec, state)
    po = orc.params(n_g=N_G, weight=2.0e3)
    sim.run(400)
    worst = {"n": 0.0, "phi": 0.0, "E": 0.0}
    for _ in range(3):
        t = sim.time()[0]
        assert t % 20 == 0  # an ion step: both densities are those of the current positions
        E = sim.download(0)
        I = sim.download(1)
        sim.run(1)
        pot, ef, ed, idn = sim.grid()
        ne = orc.deposit_f64(E[0], po)
        ni = orc.deposit_f64(I[0], po)
        pot0 = 250.0 * np.cos(2.0 * np.pi * t / 4000.0)
        pw, ew = orc.poisson_thomas(E_CHARGE * (ni - ne), pot0, po)
        worst["n"] = max(worst["n"], np.abs(ed - ne).max() / ne.max(), np.abs(idn - ni).max() / ni.max())
        worst["phi"] = max(worst["phi"], np.abs(pot - pw).max() / np.abs(pw).max())
        worst["E"] = max(worst["E"], np.abs(ef - ew).max() / np.abs(ew).max())
        sim.run(19)
    sim.check()
    sim.close()
    print(f"N_G={N_G} {name}: max rel error density {worst['n']:.3e} phi {worst['phi']:.3e} E {worst['E']:.3e}")
    # measured on the MI355X: fp32 phi 4.8e-4, E 1.4e-1 (the f32 scan loses
    # digits over 4000 nodes and the central difference divides them by
    # 2 dx); comp 2.0e-7 / 1.3e-7 and fp64 6.3e-8 / 7.0e-8 (the f32 pot and
    # efield outputs); densities 7e-6 in every mode (f32 CIC weights)
    assert worst["n"] < 2e-5
    if name == "fp32":
        assert 1e-5 < worst["phi"] < 2e-3 and worst["E"] < 0.5
    else:
        assert worst["phi"] < 1e-6 and worst["E"] < 1e-6


def test_averaged_densities_against_fp64(ca, state):
    cycles = 20
    avg = {}
    for prec, name, seed in [(2, "fp64", None), (2, "fp64_seed2", 0x9e3779b9), (1, "comp", None), (0, "fp32", None)]:
        p, sim = _sim(ca, prec, state, seed)
        sim.run(400)
        sim.reset_measurements()
        sim.run(4000 * cycles, measure=True)
        sim.check()
        m = sim.measurements()
        c = m["counters"]
        avg[name] = (m["cumul_e"] / c["e_meas_steps"], m["cumul_i"] / c["i_meas_steps"])
        sim.close()
    ref_e, ref_i = avg["fp64"]

    def l2(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)
    noise = max(l2(avg["fp64_seed2"][0], ref_e), l2(avg["fp64_seed2"][1], ref_i))
    res = {}
    for name in ("comp", "fp32"):
        res[name] = (l2(avg[name][0], ref_e), l2(avg[name][1], ref_i))
        print(f"20-cycle averages, {name} vs fp64: L2 n_e {res[name][0]:.3e} n_i {res[name][1]:.3e} "
              f"(fp64 seed spread {noise:.3e})")
    # the compensated scan is as good as FP64 within the statistical error;
    # the FP32 scan's field error shows in the averages (measured 2-4x the
    # seed spread at N_G = 4000), but stays at the percent level
    assert max(res["comp"]) < 3 * noise
    assert max(res["fp32"]) < 0.05
