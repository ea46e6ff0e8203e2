"""Plasma-like starting states for benchmarks and the CLI.

Positions are drawn from the published time-averaged electron and ion
density profiles of the eduPIC reference case (C/ReferenceCase_RawData/
Figure10_density.dat, shipped as data/refcase_density.dat); electrons get a
Maxwellian of temperature te_ev (2.1 eV ~ 2/3 of the 3.2 eV mean energy of
the bulk), ions the 350 K gas temperature with a fraction accelerated
towards the nearer electrode with 1-60 eV (the sheath energies of the
published ion flux-energy distribution).
"""
from __future__ import annotations

import os

import numpy as np

E_CHARGE = 1.60217662e-19
E_MASS = 9.10938356e-31
AR_MASS = 6.63352090e-26
K_B = 1.38064852e-23
PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "refcase_density.dat")


def refcase_density():
    """x [m], n_e, n_i [m^-3] of the published reference case."""
    d = np.loadtxt(PROFILE)
    return d[:, 0], d[:, 1], d[:, 2]


def sample_profile(rng, n, gap, profile):
    """Inverse-CDF sampling of a nodal profile (piecewise linear)."""
    xs = np.linspace(0.0, gap, len(profile))
    w = np.maximum(np.asarray(profile, np.float64), 0.0)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (w[1:] + w[:-1]) * np.diff(xs))])
    cdf /= cdf[-1]
    x = np.interp(rng.random(n), cdf, xs)
    return np.clip(x, 0.0, gap).astype(np.float32)


def state(n_e, n_i, gap=0.025, seed=0, te_ev=2.1, t_gas=350.0, ion_beam_frac=0.1):
    """((x, vx, vy, vz) electrons, (x, vx, vy, vz) ions) as float32 arrays."""
    rng = np.random.default_rng(seed)
    _, ne, ni = refcase_density()
    ex = sample_profile(rng, n_e, gap, ne)
    ix = sample_profile(rng, n_i, gap, ni)
    sig_e = np.sqrt(te_ev * E_CHARGE / E_MASS)
    ev = (rng.standard_normal((3, n_e)) * sig_e).astype(np.float32)
    sig_i = np.sqrt(K_B * t_gas / AR_MASS)
    iv = (rng.standard_normal((3, n_i)) * sig_i).astype(np.float32)
    nb = int(ion_beam_frac * n_i)
    if nb:
        e_b = rng.uniform(1.0, 60.0, nb)
        vb = np.sqrt(2 * e_b * E_CHARGE / AR_MASS) * np.sign(ix[:nb] - 0.5 * gap)
        iv[0, :nb] = vb.astype(np.float32)
    return (ex, ev[0], ev[1], ev[2]), (ix, iv[0], iv[1], iv[2])
