"""The benchmark configurations at their full sizes, through properties that
do not depend on the size: the plain method at scale (3.8 M electrons +
4.0 M ions, N_G = 400, the converged reference case repeated 35 times at
W = 2e3) and the 64-discharge ensemble with its default cache-sized groups.

  * particle balance from the device counters: every ionisation adds one
    electron and one ion, every absorption removes one, so
    N_e(end) - N_e(start) + e absorbed == N_i(end) - N_i(start) + i absorbed
    exactly (both sides count the ionisations);
  * the charge invariant of the deposited densities:
    sum of the interior nodes + (n_0 + n_{N-1}) / 2 == N * weight factor;
  * one step after the run (an ion step) against the CPU oracle fed with
    the downloaded particles: densities (the oracle's CIC deposit) and the
    field (its FP64 scan solve).
The 4 M run crosses a compaction (every 800 steps)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


def _charge(dens, n, wf):
    return (dens[1:-1].sum() + 0.5 * (dens[0] + dens[-1])) / (n * wf)


def test_plain_method_at_scale_880_steps(ca, orc):
    from cudapic_amd import synthetic
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    E, I = synthetic.converged_state(35)
    n_e0, n_i0 = len(E[0]), len(I[0])
    print(f"config 2: {n_e0} electrons + {n_i0} ions, 880 steps", flush=True)
    assert n_e0 > 3_500_000 and n_i0 > 3_700_000
    with ca.Sim(p, capacity=8_000_000) as sim:
        sim.upload(0, *E)
        sim.upload(1, *I)
        del E, I
        sim.run(880)
        sim.check()
        c = sim.measurements()["counters"]
        n_e1, n_i1 = sim.counts()
        ion_e = n_e1 - n_e0 + c["e_lost_powered"] + c["e_lost_grounded"]
        ion_i = n_i1 - n_i0 + c["i_lost_powered"] + c["i_lost_grounded"]
        print(f"ionisations from the electron balance {ion_e}, from the ion balance {ion_i}; "
              f"absorbed e {c['e_lost_powered'] + c['e_lost_grounded']}, "
              f"i {c['i_lost_powered'] + c['i_lost_grounded']}", flush=True)
        assert ion_e == ion_i and ion_e > 0
        assert c["e_lost_powered"] + c["e_lost_grounded"] > 0 and c["i_lost_powered"] + c["i_lost_grounded"] > 0
        # step 880 (an ion step) against the oracle
        assert sim.time()[0] == 880
        Ed, Id = sim.download(0), sim.download(1)
        sim.run(1)
        sim.check()
        pot, ef, ed, idn = sim.grid()
    for got, parts in ((ed, Ed), (idn, Id)):
        want = orc.deposit(parts[0], po)
        assert np.abs(got - want).max() <= 2e-6 * want.max()
        assert abs(_charge(got, len(parts[0]), po.weight_factor) - 1.0) < 1e-6
    pw, ew = orc.poisson_scan_f64(ed, idn, 250.0 * np.cos(2.0 * np.pi * 880 / 4000.0), po)
    assert np.abs(pot - pw).max() <= 1e-5 * np.abs(pw).max()
    assert np.abs(ef - ew).max() <= 1e-5 * np.abs(ew).max()


def test_ensemble_of_64_discharges_default_groups(ca, orc):
    """The batched configuration as cudapic_amd.ens_bench builds it (V0 in
    [100, 400] V, p in [3, 30] Pa, 50 k - 2 M particles per species, ~32 M
    electrons in all), default group cuts: 400 steps, then step 400 of the
    largest instance and step 480 of the smallest against the oracle."""
    from cudapic_amd import ens_bench, synthetic
    v0, pres, cnt = ens_bench.instances(64)
    wts = [7.0e3 * 1.2e6 / float(c) for c in cnt]
    ps = [ca.params(v0=float(v), pressure=float(q), weight=w, precision=2) for v, q, w in zip(v0, pres, wts)]
    caps = [int(c * 1.5) + 4096 for c in cnt]
    states = [synthetic.state(int(c), int(c * 1.05), seed=100 + i) for i, c in enumerate(cnt)]
    print(f"config 4: 64 discharges, {int(cnt.sum())} electrons", flush=True)
    ens = ca.Ensemble(ps, caps)
    try:
        for i, (E, I) in enumerate(states):
            ens.upload(i, 0, *E)
            ens.upload(i, 1, *I)
        ens.run(400)
        ens.check()
        n_e1, n_i1 = ens.counts()
        for i, (E, I) in enumerate(states):
            c = ens.measurements(i)[2]
            ion_e = int(n_e1[i]) - len(E[0]) + int(c[0] + c[1])
            ion_i = int(n_i1[i]) - len(I[0]) + int(c[2] + c[3])
            assert ion_e == ion_i, (i, ion_e, ion_i)
        t = 400
        for k in (int(np.argmax(cnt)), int(np.argmin(cnt))):
            Ed, Id = ens.download(k, 0), ens.download(k, 1)
            po = orc.params(v0=float(v0[k]), pressure=float(pres[k]), weight=wts[k])
            want_e, want_i = orc.deposit(Ed[0], po), orc.deposit(Id[0], po)
            ens.run(1)
            pot, ef = ens.grid(k)
            pw, ew = orc.poisson_scan_f64(want_e, want_i, float(v0[k]) * np.cos(2.0 * np.pi * t / 4000.0), po)
            print(f"instance {k}: {len(Ed[0])} e, field max |dE| / max |E| = "
                  f"{np.abs(ef - ew).max() / np.abs(ew).max():.2e}", flush=True)
            assert np.abs(ef - ew).max() <= 2e-5 * np.abs(ew).max(), k
            assert np.abs(pot - pw).max() <= 2e-5 * np.abs(pw).max(), k
            ens.run(79)    # to the next ion step (not a compacting one) for the next instance
            t += 80
    finally:
        ens.close()
