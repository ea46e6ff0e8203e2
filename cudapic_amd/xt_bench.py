"""Cost of measuring eduPIC's XT maps on the fused path.

    python -m cudapic_amd.xt_bench [--steps 800] [--bin 20]

Workload: the 4 M-particle discharge of the plain method at scale (W = 2e3,
the converged reference case repeated 35 times: ~3.8 M electrons + ~4.0 M
ions, FP64 Poisson).  Times windows of `steps` steps, measuring without the
XT maps, measuring with them (bins of `bin` steps) and not measuring, each
with its graphs captured before the window; prints one JSON line with the
microseconds per step of each.
"""
from __future__ import annotations

import argparse
import json
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=800, help="steps per timed window (a multiple of 80)")
    ap.add_argument("--bin", type=int, default=20, help="steps per XT time bin")
    args = ap.parse_args()
    import cudapic_amd as ca
    from cudapic_amd import synthetic

    E, I = synthetic.converged_state(35)
    p = ca.params(weight=2.0e3, precision=ca.PIC_PREC_FP64)
    out = {"metric": "us per step: measured without / with XT maps, unmeasured", "steps": args.steps,
           "bin_steps": args.bin, "electrons": len(E[0]), "ions": len(I[0])}
    with ca.Sim(p, capacity=8_000_000) as sim:
        sim.upload(0, *E)
        sim.upload(1, *I)
        sim.run(400)
        sim.synchronize()
        for key, xt, measure in (("measured", 0, True), ("measured_xt", args.bin, True), ("unmeasured", 0, False)):
            sim.set_xt(xt)
            sim.run(80, measure=measure)
            sim.synchronize()
            r = sim.time_steps(args.steps, measure=measure)
            out[key + "_us_per_step"] = round(r["ms_total"] / args.steps * 1e3, 2)
            print(f"xt_bench: {key} {out[key + '_us_per_step']} us/step", file=sys.stderr, flush=True)
        sim.check()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
