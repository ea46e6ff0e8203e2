"""Shared test utilities: synthetic discharge states and comparison helpers."""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
E_CHARGE = 1.60217662e-19
E_MASS = 9.10938356e-31
AR_MASS = 6.63352090e-26
K_B = 1.38064852e-23
SEED_E = (0x12345678, 0x11112222, 0x22223333, 0x33334444)   # Rust-CUDA/src/main.rs:1667
SEED_I = (0x44445555, 0x55556666, 0x66667777, 0x77778888)   # main.rs:1671


def refcase_density():
    """Published time-averaged densities of the eduPIC reference case
    (C/ReferenceCase_RawData/Figure10_density.dat): x [m], n_e, n_i [m^-3]."""
    d = np.loadtxt(os.path.join(GOLDEN, "refcase_density.dat"))
    return d[:, 0], d[:, 1], d[:, 2]


def sample_profile(rng, n, gap, profile, n_g=400):
    """Positions by inverse-CDF sampling of a nodal density profile (linear
    interpolation between the nodes of the published 400-node grid)."""
    xs = np.linspace(0.0, gap, len(profile))
    w = np.maximum(profile, 0.0)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (w[1:] + w[:-1]) * np.diff(xs))])
    cdf /= cdf[-1]
    u = rng.random(n)
    x = np.interp(u, cdf, xs)
    return np.clip(x, 0.0, gap).astype(np.float32)


def synthetic_state(n_e, n_i, gap=0.025, seed=0, te_ev=2.1, ion_beam_frac=0.1):
    """Plasma-like state: x from the published profiles, Maxwellian electrons
    (kT_e), thermal ions at 350 K with a fraction carrying sheath energies."""
    rng = np.random.default_rng(seed)
    _, ne, ni = refcase_density()
    ex = sample_profile(rng, n_e, gap, ne)
    ix = sample_profile(rng, n_i, gap, ni)
    sig_e = np.sqrt(te_ev * E_CHARGE / E_MASS)
    ev = (rng.standard_normal((3, n_e)) * sig_e).astype(np.float32)
    sig_i = np.sqrt(K_B * 350.0 / AR_MASS)
    iv = (rng.standard_normal((3, n_i)) * sig_i).astype(np.float32)
    nb = int(ion_beam_frac * n_i)
    if nb:
        e_b = rng.uniform(1.0, 60.0, nb)
        vb = np.sqrt(2 * e_b * E_CHARGE / AR_MASS) * np.sign(ix[:nb] - 0.5 * gap)
        iv[0, :nb] = vb.astype(np.float32)
    return (ex, ev[0], ev[1], ev[2]), (ix, iv[0], iv[1], iv[2])


def sort_rows(*cols):
    """Rows (x, vx, vy, vz) in lexicographic order: a canonical form of a
    multiset of particles."""
    a = np.stack([np.asarray(c, np.float32) for c in cols], axis=1)
    if len(a) == 0:
        return a
    order = np.lexsort(tuple(a[:, k] for k in range(a.shape[1] - 1, -1, -1)))
    return a[order]


def assert_same_multiset(got, want):
    g = sort_rows(*got)
    w = sort_rows(*want)
    assert g.shape == w.shape, (g.shape, w.shape)
    np.testing.assert_array_equal(g.view(np.uint32), w.view(np.uint32))
