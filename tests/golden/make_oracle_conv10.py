"""Particle counts of the eduPIC reference case's first 10 RF cycles from the
CPU oracle (oracle/pic_oracle.c, the C restatement of the reference step,
main.rs:1725-1808): 1000 electrons + 1000 ions at rest at uniform random
positions (the eduPIC start, C/eduPIC.cc:283-294), N_SEEDS independent
seeds (positions and xoshiro128+ masters), counts after every cycle; the
fixture oracle_conv10.json holds their per-cycle mean and standard error.

    python tests/golden/make_oracle_conv10.py [--seeds 16] [--jobs 8]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
CYCLES = 10


def one_seed(seed):
    from oracle import oracle as orc
    p = orc.params()
    rng = np.random.default_rng(900 + seed)
    z = np.zeros(1000, np.float32)
    ex = (rng.random(1000) * p.gap).astype(np.float32)
    ix = (rng.random(1000) * p.gap).astype(np.float32)
    cap = 40_000
    se = [0x12345678 + seed, 0x11112222, 0x22223333, 0x33334444]
    si = [0x44445555 + seed, 0x55556666, 0x66667777, 0x77778888]
    sim = orc.Sim(p, cap, ex, z, z, z, ix, z, z, z, orc.seed_streams(se, cap), orc.seed_streams(si, cap))
    rows = []
    for _ in range(CYCLES):
        for t in range(p.n_t):
            sim.step(t)
        rows.append((sim.n_e, sim.n_i))
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=16)
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    with ProcessPoolExecutor(a.jobs) as ex:
        runs = np.array(list(ex.map(one_seed, range(a.seeds))), np.float64)   # [seed][cycle][species]
    out = {"source": "oracle/pic_oracle.c or_step, collisions on, eduPIC start at rest",
           "seeds": a.seeds, "cycles": CYCLES,
           "mean_e": runs[:, :, 0].mean(axis=0).tolist(), "mean_i": runs[:, :, 1].mean(axis=0).tolist(),
           "sd_e": runs[:, :, 0].std(axis=0, ddof=1).tolist(), "sd_i": runs[:, :, 1].std(axis=0, ddof=1).tolist()}
    path = os.path.join(HERE, "oracle_conv10.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
