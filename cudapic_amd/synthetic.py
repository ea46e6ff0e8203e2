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


# ---------------------------------------------------------------- converged state
# A converged state of the reference case (250 V, 10 Pa, 350 K, W = 7e4,
# N_G = 400): the fused path's particles after 1500 RF cycles from
# state(107000, 112500, seed=11) (eduPIC's converged counts,
# Figure9_conv.dat), made by `python -m cudapic_amd.synthetic converged`.
# Stored compactly: positions as uint16 fractions of the gap, velocities as
# float16 (electrons in km/s, ions in m/s).
CONVERGED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "refcase_state.npz")
GAP = 0.025


def pack_state(electrons, ions):
    out = {}
    for name, (x, vx, vy, vz), scale in (("e", electrons, 1e3), ("i", ions, 1.0)):
        out[name + "x"] = np.round(np.clip(np.asarray(x, np.float64) / GAP, 0.0, 1.0) * 65535.0).astype(np.uint16)
        out[name + "v"] = (np.stack([vx, vy, vz]) / scale).astype(np.float16)
    return out


def unpack_state(d):
    res = []
    for name, scale in (("e", 1e3), ("i", 1.0)):
        x = (d[name + "x"].astype(np.float64) / 65535.0 * GAP).astype(np.float32)
        v = d[name + "v"].astype(np.float32) * np.float32(scale)
        res.append((x, v[0], v[1], v[2]))
    return res[0], res[1]


def converged_state(replicas=1):
    """((x, vx, vy, vz) electrons, (...) ions) of the converged reference
    case, each particle repeated `replicas` times: at a weight 7e4 /
    replicas this is the same discharge (the same densities and
    distributions) with `replicas` times the particles; the copies part
    ways at their first collisions, which draw on different slots' streams."""
    E, I = unpack_state(np.load(CONVERGED))
    if replicas > 1:
        E = tuple(np.tile(a, replicas) for a in E)
        I = tuple(np.tile(a, replicas) for a in I)
    return E, I


def _make_converged(cycles=1500):
    import cudapic_amd as ca
    E, I = state(107_000, 112_500, seed=11)
    with ca.Sim(ca.params(), capacity=400_000) as sim:
        sim.upload(0, *E)
        sim.upload(1, *I)
        for c in range(0, cycles, 100):
            sim.run(4000 * min(100, cycles - c))
            sim.synchronize()
            print("cycle", c + min(100, cycles - c), sim.counts(), flush=True)
        sim.check()
        packed = pack_state(sim.download(0), sim.download(1))
    np.savez_compressed(CONVERGED, **packed)


if __name__ == "__main__":
    import sys
    if len(sys.argv) > 1 and sys.argv[1] == "converged":
        _make_converged(int(sys.argv[2]) if len(sys.argv) > 2 else 1500)
    else:
        print("usage: python -m cudapic_amd.synthetic converged [cycles]  (needs a GPU)", file=sys.stderr)
        sys.exit(2)
