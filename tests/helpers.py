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


def synthetic_state(n_e, n_i, gap=0.025, seed=0, te_ev=2.1, ion_beam_frac=0.1):
    """Plasma-like state drawn from the published profiles (the generator the
    benchmark uses, cudapic_amd/synthetic.py)."""
    from cudapic_amd import synthetic
    return synthetic.state(n_e, n_i, gap=gap, seed=seed, te_ev=te_ev, ion_beam_frac=ion_beam_frac)


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


def no_ionisation_tables(orc, po):
    """The reference case's tables with the ionisation cross section set to
    zero (and the electron totals summed again as the reference sums them,
    f64 sum of the f32 entries times the gas density): electrons collide
    elastically or excite, so no particle is created and a run's slot
    contents do not depend on the order of the newborns' appends."""
    cs, _, si = orc.cross_sections(po.gas_density)
    cs = cs.copy()
    r = len(si)
    cs[2 * r:3 * r] = 0.0
    se = ((cs[:r].astype(np.float64) + cs[r:2 * r].astype(np.float64)) * po.gas_density).astype(np.float32)
    return cs, se, si
