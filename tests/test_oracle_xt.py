"""Pins of the XT-map restatement (oracle/xt.py) by closed forms (CPU only):
a particle on a node and one between nodes, a uniform drift (j_e = -e n u
and j_i = e n u), the half-step correction of a uniform field for both
species, the ionisation threshold, and the per-bin normalisation."""
import numpy as np

from oracle import xt as ox

N_G, N_XT = 11, 4
DX = 1.0e-3
DT_E = 1.0e-11
DT_I = 2.0e-10
W, AREA = 7.0e4, 1.0e-4
N_GAS = 2.0e21


def _sigma():
    # a step cross section: 0 below 15.8 eV, 1e-20 m^2 above
    s = np.zeros(1_000_000)
    s[15800:] = 1.0e-20
    return s


def _speed(ev, mass):
    return np.sqrt(2.0 * ev * ox.E_CHARGE / mass)


def test_particle_on_a_node_and_between_nodes():
    raw, cnt = ox.new_raw(N_G, N_XT)
    u = _speed(20.0, ox.E_MASS)
    ox.add_particles(raw, 2, [3 * DX, 5.25 * DX], [u, u], [0.0, 0.0], [0.0, 0.0], np.zeros(N_G), DX, DT_E, True,
                     _sigma(), N_GAS)
    ox.add_grid(raw, cnt, 2, np.zeros(N_G), np.zeros(N_G), np.zeros(N_G), np.zeros(N_G))
    w = raw[ox.CE][:, 2]
    assert w[3] == 1.0 and abs(w[5] - 0.75) < 1e-12 and abs(w[6] - 0.25) < 1e-12 and w.sum() == 2.0
    m = ox.normalise(raw, cnt, W, AREA, DX, DT_E)
    for k in (3, 5, 6):
        assert abs(m["meanee"][k, 2] - 20.0) < 1e-9
    # ionisation rate of one superparticle at one node: sigma v dt n_gas W / (A dx) / dt
    want = 1.0e-20 * u * DT_E * N_GAS * W / (AREA * DX) / DT_E
    assert abs(m["ioniz"][3, 2] - want) < 1e-9 * want
    assert np.all(m["meanee"][:, [0, 1, 3]] == 0.0)


def test_uniform_drift_gives_the_current_density():
    """Electrons (and ions) spread uniformly, all with velocity u: every
    node's mean velocity is u, so with the grid density n the current
    density is j = -e n u (ions +e n u) and the power j E."""
    n = 200_000
    rng = np.random.default_rng(3)
    x = rng.uniform(0.0, (N_G - 1) * DX, n)
    ue, ui = 3.0e5, -800.0
    dens = np.full(N_G, 1.0e15)
    for electrons, u, dt in ((True, ue, DT_E), (False, ui, DT_I)):
        raw, cnt = ox.new_raw(N_G, N_XT)
        ox.add_particles(raw, 0, x, np.full(n, u), np.zeros(n), np.zeros(n), np.zeros(N_G), DX, dt, electrons,
                         _sigma(), N_GAS)
        ox.add_grid(raw, cnt, 0, np.zeros(N_G), np.full(N_G, 50.0), dens, dens)
        m = ox.normalise(raw, cnt, W, AREA, DX, DT_E)
        j = m["je"] if electrons else m["ji"]
        sign = -1.0 if electrons else 1.0
        np.testing.assert_allclose(j[1:-1, 0], sign * ox.E_CHARGE * 1.0e15 * u, rtol=1e-12)
        pw = m["powere"] if electrons else m["poweri"]
        np.testing.assert_allclose(pw[1:-1, 0], j[1:-1, 0] * 50.0, rtol=1e-12)


def test_half_step_correction_of_a_uniform_field():
    """vx is half a step ahead of the mid-step velocity: electrons
    u = vx - 0.5 E dt e/m, ions u = vx + 0.5 E dt e/M."""
    e0 = 2.0e3
    for electrons, mass, dt in ((True, ox.E_MASS, DT_E), (False, ox.AR_MASS, DT_I)):
        raw, cnt = ox.new_raw(N_G, N_XT)
        vx = 1.0e4
        ox.add_particles(raw, 1, [4 * DX], [vx], [0.0], [0.0], np.full(N_G, e0), DX, dt, electrons, _sigma(),
                         N_GAS)
        ox.add_grid(raw, cnt, 1, np.zeros(N_G), np.full(N_G, e0), np.ones(N_G), np.ones(N_G))
        kick = 0.5 * e0 * dt * ox.E_CHARGE / mass
        u = vx - kick if electrons else vx + kick
        m = ox.normalise(raw, cnt, W, AREA, DX, DT_E)
        en = m["meanee" if electrons else "meanei"][4, 1]
        assert abs(en - 0.5 * mass * u * u / ox.E_CHARGE) < 1e-9 * en
        j = m["je" if electrons else "ji"][4, 1]
        assert abs(j - (-1.0 if electrons else 1.0) * u * ox.E_CHARGE) < 1e-9 * abs(j)


def test_ionisation_threshold_and_bin_averages():
    raw, cnt = ox.new_raw(N_G, N_XT)
    u_lo, u_hi = _speed(15.0, ox.E_MASS), _speed(16.0, ox.E_MASS)
    ox.add_particles(raw, 3, [2 * DX, 7 * DX], [u_lo, u_hi], [0.0, 0.0], [0.0, 0.0], np.zeros(N_G), DX, DT_E, True,
                     _sigma(), N_GAS)
    # two measured steps in bin 3: grid averages, particle sums per step
    for pot in (1.0, 3.0):
        ox.add_grid(raw, cnt, 3, np.full(N_G, pot), np.zeros(N_G), np.zeros(N_G), np.zeros(N_G))
    m = ox.normalise(raw, cnt, W, AREA, DX, DT_E)
    assert m["ioniz"][2, 3] == 0.0 and m["ioniz"][7, 3] > 0.0
    np.testing.assert_allclose(m["pot"][:, 3], 2.0)
    assert np.all(m["pot"][:, :3] == 0.0)
    # the rate per superparticle and measured step: two steps halve it
    want = 1.0e-20 * u_hi * DT_E * N_GAS * W / (AREA * DX) / (2 * DT_E)
    assert abs(m["ioniz"][7, 3] - want) < 1e-9 * want
