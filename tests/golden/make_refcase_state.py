"""Writes refcase_state.npz: a converged state of the eduPIC reference case
(250 V, 10 Pa, 350 K, W = 7e4, N_G = 400) for the full-run statistics test.

Run on a GPU from the repository root:
    python tests/golden/make_refcase_state.py [cycles]

The state is the fused path's after `cycles` RF cycles (default 1500) from a
plasma-like synthetic state of 107000 electrons and 112500 ions (eduPIC's
converged counts, Figure9_conv.dat) drawn from the published density
profiles (seed 11).  It is stored compactly: positions as uint16 fractions of
the gap, velocities as float16 (electrons in km/s, ions in m/s).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
GAP = 0.025


def pack(electrons, ions):
    out = {}
    for name, (x, vx, vy, vz), scale in (("e", electrons, 1e3), ("i", ions, 1.0)):
        out[name + "x"] = np.round(np.clip(np.asarray(x, np.float64) / GAP, 0.0, 1.0) * 65535.0).astype(np.uint16)
        out[name + "v"] = (np.stack([vx, vy, vz]) / scale).astype(np.float16)
    return out


def unpack(d):
    """(electrons, ions) as (x, vx, vy, vz) float32 arrays."""
    res = []
    for name, scale in (("e", 1e3), ("i", 1.0)):
        x = (d[name + "x"].astype(np.float64) / 65535.0 * GAP).astype(np.float32)
        v = d[name + "v"].astype(np.float32) * np.float32(scale)
        res.append((x, v[0], v[1], v[2]))
    return res[0], res[1]


def main():
    import cudapic_amd as ca
    from cudapic_amd import synthetic
    cycles = int(sys.argv[1]) if len(sys.argv) > 1 else 1500
    E, I = synthetic.state(107_000, 112_500, seed=11)
    with ca.Sim(ca.params(), capacity=400_000) as sim:
        sim.upload(0, *E)
        sim.upload(1, *I)
        for c in range(0, cycles, 100):
            sim.run(4000 * min(100, cycles - c))
            sim.synchronize()
            print("cycle", c + min(100, cycles - c), sim.counts(), flush=True)
        sim.check()
        state = pack(sim.download(0), sim.download(1))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "refcase_state.npz"), **state)


if __name__ == "__main__":
    main()
