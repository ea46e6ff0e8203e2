"""Sim.report (eduPIC's run report, C/eduPIC.cc:926-978) on one measured
RF cycle of the converged reference case: the counters it divides are the
ones eduPIC divides.  The electron collision frequency equals the
population average of v sigma_tot_e(E) taken from the particles themselves
(every electron step counts); the ion one is of the order of g sigma_tot_i
over the ions (counted on ion steps only: counting on every electron step
would be 20 times larger); the centre mean energy implied by the Debye
length is that of the electrons in the EEPF window."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
E_CHARGE, E_MASS, AR_MASS = 1.60217662e-19, 9.10938356e-31, 6.63352090e-26


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


def _nu(v2, mass, table):
    energy = 0.5 * mass * v2 / E_CHARGE
    idx = np.minimum((energy / 1e-3).astype(np.int64), len(table) - 1)
    return np.sqrt(v2) * table[idx].astype(np.float64)


def test_report_of_a_measured_cycle(ca):
    from cudapic_amd import synthetic
    p = ca.params()
    E, I = synthetic.converged_state()
    _, se, si = ca.kernels.cross_sections(p.gas_density)
    with ca.Sim(p, capacity=400_000) as sim:
        sim.upload(0, *E)
        sim.upload(1, *I)
        sim.run(p.n_t, measure=True)
        r = sim.report()
        e_end = sim.download(0)
        i_end = sim.download(1)
        assert sim.time()[1] == p.n_t
    assert r["conditions_ok"]
    assert 0.0 < r["wp_dt_e"] < 0.2 and 0.0 < r["dx_over_debye"] < 1.0
    assert f"{r['nu_max_e_dt_e']:.3f}" == "0.013" and f"{r['nu_max_i_dt_i']:.3f}" == "0.020"
    # electrons: <v sigma_tot_e> over the population, start and end of the cycle
    nu_e = np.mean([_nu(vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2 + vz.astype(np.float64) ** 2,
                        E_MASS, se).mean() for _, vx, vy, vz in (E, e_end)])
    assert abs(r["coll_freq_e"] / nu_e - 1.0) < 0.1, (r["coll_freq_e"], nu_e)
    # ions: the ion speed stands in for the speed relative to the gas atoms
    nu_i = np.mean([_nu(vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2 + vz.astype(np.float64) ** 2,
                        AR_MASS / 2.0, si).mean() for _, vx, vy, vz in (I, i_end)])
    assert 0.5 < r["coll_freq_i"] / nu_i < 2.0, (r["coll_freq_i"], nu_i)
    # centre mean energy: kT = lambda_D^2 n e^2 / eps0 and mean = 3/2 kT
    mean_e = 1.5 * r["debye_length"] ** 2 * r["density_center"] * E_CHARGE / 8.85418781e-12
    x, vx, vy, vz = (a.astype(np.float64) for a in e_end)
    band = (x > 0.45 * p.gap) & (x < 0.55 * p.gap)
    want = (0.5 * E_MASS * (vx[band] ** 2 + vy[band] ** 2 + vz[band] ** 2) / E_CHARGE).mean()
    assert abs(mean_e / want - 1.0) < 0.15, (mean_e, want)
