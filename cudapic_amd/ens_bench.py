"""Ensemble throughput against sequential single-discharge runs.

    python -m cudapic_amd.ens_bench [--instances 64] [--steps 400] [--warmup 80]

Workload (the batched-ensemble configuration): independent argon discharges
on N_G = 400 with V0 uniform in [100, 400] V and p log-uniform in [3, 30] Pa
drawn from a seed, and live counts log-uniform in [50k, 2M] particles per
species per instance (plasma-like synthetic states drawn from the reference
case's published density profiles).  The superparticle weight is 7e3 for an
instance of 1.2M particles (the reference case's steady state at that
weight) and scales inversely with the count, so that every instance starts
at the reference case's physical density: a plasma of 50k particles at
W = 7e3 cannot shield a few hundred volts and runs away into an ionisation
avalanche within a few hundred steps.  The same instances, from
the same states, are run once as one ensemble (one Poisson launch and one
particle launch per species per step for each cache-sized group of
instances, each group a block of hipGraph periods at a time) and
once one after the other as single-discharge simulations: on the fused fast
path with its own hipGraph, and on the reference's kernel sequence (push,
compaction, collisions and deposit as separate launches, main.rs:1725-1808).  Both runs advance every instance by the same
number of steps, so their statistical errors are the same; the figure of
merit is total particle-steps per second.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np


def instances(n, seed=2024):
    rng = np.random.default_rng(seed)
    v0 = rng.uniform(100.0, 400.0, n)
    p = np.exp(rng.uniform(np.log(3.0), np.log(30.0), n))
    cnt = np.exp(rng.uniform(np.log(5.0e4), np.log(2.0e6), n)).astype(np.int64)
    return v0, p, cnt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--steps", type=int, default=400, help="timed steps (a multiple of 4*N_SUB)")
    ap.add_argument("--warmup", type=int, default=80)
    ap.add_argument("--weight", type=float, default=7.0e3, help="weight of a 1.2M-particle instance")
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--sequential-path", choices=["fused", "reference", "both"], default="both",
                    help="single-discharge path of the sequential runs: the fused fast path, the "
                         "reference's kernel sequence, or both")
    ap.add_argument("--grid", type=int, default=0, help="particle-kernel workgroups (0 = 4 per compute unit)")
    ap.add_argument("--ppt", type=int, default=0, help="particles per thread per tile (0 = 2)")
    ap.add_argument("--replicas", type=int, default=0, help="global density replicas (0 = 8)")
    ap.add_argument("--group-slots", type=int, default=0,
                    help="electron capacity per group of instances run a block at a time (0 = 9e6, "
                         "4294967295 = one group)")
    ap.add_argument("--group-periods", type=int, default=0, help="4*N_SUB-step periods per block (0 = 5)")
    ap.add_argument("--group-streams", type=int, default=0, help="streams the groups are dealt to (0 = 1)")
    ap.add_argument("--cand-masks", type=int, default=0, choices=[-1, 0, 1],
                    help="collision-candidate masks of the deferred electron runs (-1 off, 0 or 1 on)")
    ap.add_argument("--one-group", action="store_true",
                    help="also time the ensemble as one group of all instances (one launch per stage "
                         "and step over every instance, no temporal blocking)")
    args = ap.parse_args()
    import cudapic_amd as ca
    from cudapic_amd import synthetic

    v0, pres, cnt = instances(args.instances)
    ps = [ca.params(v0=float(v), pressure=float(p), weight=args.weight * 1.2e6 / float(c),
                    precision=ca.PIC_PREC_FP64) for v, p, c in zip(v0, pres, cnt)]
    caps = [int(c * 1.5) + 4096 for c in cnt]
    states = [synthetic.state(int(c), int(c * 1.05), seed=100 + i) for i, c in enumerate(cnt)]
    print(f"ens_bench: {args.instances} instances, {int(cnt.sum())} electrons in total "
          f"({int(cnt.min())}..{int(cnt.max())})", file=sys.stderr, flush=True)

    def ensemble(group_slots):
        ens = ca.Ensemble(ps, caps, grid_blocks=args.grid, ppt=args.ppt, replicas=args.replicas,
                          group_slots=group_slots, group_periods=args.group_periods,
                          group_streams=args.group_streams, cand_masks=args.cand_masks)
        for i, (E, I) in enumerate(states):
            ens.upload(i, 0, *E)
            ens.upload(i, 1, *I)
        ens.run(args.warmup)
        ens.synchronize()
        ens.check()
        t = ens.time_steps(args.steps)
        ens.check()
        ens.close()
        return t["particle_steps"] / (t["ms_total"] * 1e-3), t["ms_total"]

    ens_rate, ens_ms = ensemble(args.group_slots)
    print(f"ens_bench: ensemble {args.steps} steps in {ens_ms:.1f} ms", file=sys.stderr, flush=True)
    one = None
    if args.one_group:
        r1, ms1 = ensemble(0xFFFFFFFF)
        one = {"value": round(r1, 1), "ms_total": round(ms1, 2)}
        print(f"ens_bench: ensemble as one group {args.steps} steps in {ms1:.1f} ms", file=sys.stderr, flush=True)

    def sequential(path):
        tot_ps, tot_ms = 0, 0.0
        for i, (E, I) in enumerate(states):
            with ca.Sim(ps[i], capacity=caps[i], path=path) as sim:
                sim.upload(0, *E)
                sim.upload(1, *I)
                sim.run(args.warmup)
                sim.synchronize()
                r = sim.time_steps(args.steps)
                sim.check()
            tot_ps += r["particle_steps"]
            tot_ms += r["ms_total"]
        print(f"ens_bench: sequential ({'fused' if path == ca.PIC_PATH_FUSED else 'reference'} path) "
              f"{args.steps} steps x {args.instances} in {tot_ms:.1f} ms", file=sys.stderr, flush=True)
        return {"value": round(tot_ps / (tot_ms * 1e-3), 1), "ms_total": round(tot_ms, 2)}

    seq, seq_ref = None, None
    if not args.no_sequential:
        if args.sequential_path in ("fused", "both"):
            seq = sequential(ca.PIC_PATH_FUSED)
        if args.sequential_path in ("reference", "both"):
            seq_ref = sequential(ca.PIC_PATH_REFERENCE)
    out = {
        "metric": "ensemble particle-steps/s (batched vs sequential single-discharge runs)",
        "unit": "particle-steps/s",
        "instances": args.instances,
        "steps": args.steps,
        "electrons_total": int(cnt.sum()),
        "ensemble": {"value": round(ens_rate, 1), "ms_total": round(ens_ms, 2)},
        "sequential": seq,
        "speedup": round(ens_rate / seq["value"], 3) if seq else None,
        "one_group": one,
        "one_group_speedup": round(one["value"] / seq["value"], 3) if (one and seq) else None,
        "sequential_reference_path": seq_ref,
        "speedup_vs_reference_path": round(ens_rate / seq_ref["value"], 3) if seq_ref else None,
        "config": {"v0": "U[100,400] V", "pressure": "logU[3,30] Pa",
                   "weight": f"{args.weight} x 1.2e6 / count", "n_g": 400,
                   "counts": "logU[5e4,2e6] per species per instance", "precision": "fp64 Poisson"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
