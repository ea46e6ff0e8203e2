"""Pins of the oracle's EEPF accumulator (CPU only).

The accumulator follows main.rs:1300-1334 (and eduPIC.cc:570-582): an
electron strictly inside the central band (0.45 L, 0.55 L) is binned by the
energy of its mid-step velocity, vx - 0.5 E(x) dt e / m with E interpolated
linearly between the two nodes around x, in bins of 0.05 eV below 2000 bins.
Each case here builds electrons whose expected bin is known by
construction (energies at bin centres, fields chosen to move the energy by
whole bins), so that a wrong band test, a wrong sign or factor of the
half-step correction, swapped interpolation weights or a missing overflow
test each change the histogram.
"""
import numpy as np

E_CHARGE = 1.60217662e-19
E_MASS = 9.10938356e-31
DE = 0.05
N_BINS = 2000


def _speed(energy_ev):
    return np.sqrt(2.0 * E_CHARGE * energy_ev / E_MASS)


def _centre(k):
    return (k + 0.5) * DE


def _hist(orc, p, x, vx, vy, vz, efield):
    return orc.accumulate_eepf(np.asarray(x, np.float32), np.asarray(vx, np.float32), np.asarray(vy, np.float32),
                               np.asarray(vz, np.float32), np.asarray(efield, np.float32), p)


def _expect(bins):
    h = np.zeros(N_BINS, np.uint32)
    for b in bins:
        h[b] += 1
    return h


def test_bins_of_known_energies_in_zero_field(orc):
    p = orc.params()
    xc = 0.5 * p.gap
    ks = [0, 1, 7, 100, 1234, 1999]
    x, vx, vy, vz = [], [], [], []
    for k in ks:
        u = _speed(_centre(k))
        # along x, and split over the three components
        x += [xc, xc]
        vx += [u, u / np.sqrt(3.0)]
        vy += [0.0, -u / np.sqrt(3.0)]
        vz += [0.0, u / np.sqrt(3.0)]
    h = _hist(orc, p, x, vx, vy, vz, np.zeros(p.n_g))
    np.testing.assert_array_equal(h, _expect(ks + ks))


def test_energies_beyond_the_last_bin_are_dropped(orc):
    p = orc.params()
    xc = 0.5 * p.gap
    e = [_centre(N_BINS - 1), _centre(N_BINS), _centre(N_BINS + 40)]
    h = _hist(orc, p, [xc] * 3, [_speed(v) for v in e], [0.0] * 3, [0.0] * 3, np.zeros(p.n_g))
    np.testing.assert_array_equal(h, _expect([N_BINS - 1]))


def test_band_is_open_at_both_ends(orc):
    p = orc.params()
    lo = np.float32(0.45 * p.gap)
    hi = np.float32(0.55 * p.gap)
    xs = [lo, hi, np.nextafter(lo, np.float32(1.0)), np.nextafter(hi, np.float32(0.0)),
          np.float32(0.3 * p.gap), np.float32(0.7 * p.gap)]
    u = _speed(_centre(42))
    h = _hist(orc, p, xs, [u] * len(xs), [0.0] * len(xs), [0.0] * len(xs), np.zeros(p.n_g))
    # only the two positions just inside the edges count
    np.testing.assert_array_equal(h, _expect([42, 42]))


def test_half_step_correction_of_a_uniform_field(orc):
    """vx is the velocity half a step after the mid-step velocity the bin
    is taken of: vx = u + 0.5 E dt e / m puts the electron at u's bin.  The
    field is chosen to shift the energy by about five bins, so that the
    opposite sign or a dropped factor 0.5 lands elsewhere."""
    p = orc.params()
    kick = p.dt_e * E_CHARGE / E_MASS  # m/s per V/m over one step
    k = 200  # 10.025 eV
    u = _speed(_centre(k))
    du = 5.0 * DE * E_CHARGE / (E_MASS * u)  # five bins of energy
    field = 2.0 * du / kick
    vx = u + 0.5 * field * kick
    h = _hist(orc, p, [0.5 * p.gap], [vx], [0.0], [0.0], np.full(p.n_g, field))
    np.testing.assert_array_equal(h, _expect([k]))
    # the same electron in zero field sits about five bins higher
    h0 = _hist(orc, p, [0.5 * p.gap], [vx], [0.0], [0.0], np.zeros(p.n_g))
    assert np.argmax(h0) in (k + 5, k + 4, k + 6)


def test_field_interpolated_between_the_nodes_around_x(orc):
    """E at x a quarter of a cell above node q is 0.75 E_q + 0.25 E_(q+1);
    the two node fields differ enough that swapped weights move the bin."""
    p = orc.params()
    kick = p.dt_e * E_CHARGE / E_MASS
    dx = p.gap / (p.n_g - 1)
    q = int(0.5 * p.gap / dx)
    x = (q + 0.25) * dx
    assert 0.45 * p.gap < x < 0.55 * p.gap
    k = 300
    u = _speed(_centre(k))
    du = DE * E_CHARGE / (E_MASS * u)  # one bin
    ef = np.zeros(p.n_g)
    ef[q] = 8.0 * du / kick           # weight 0.75: 6 du of the half step ...
    ef[q + 1] = -8.0 * du / kick      # ... weight 0.25: -2 du; E = 4 du / kick
    e_x = 0.75 * ef[q] + 0.25 * ef[q + 1]
    vx = u + 0.5 * e_x * kick         # mid-step velocity u
    h = _hist(orc, p, [x], [vx], [0.0], [0.0], ef)
    np.testing.assert_array_equal(h, _expect([k]))
    # with the weights swapped the field would be -4 du / kick: the
    # mid-step velocity u + 4 du, about four bins up
    e_sw = 0.25 * ef[q] + 0.75 * ef[q + 1]
    assert abs(0.5 * (e_x - e_sw) * kick) > 3.5 * du
