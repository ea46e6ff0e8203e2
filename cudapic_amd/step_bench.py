"""Throughput of the RF step on caller-owned buffers (pic_step_graph over the
reference's GpuSimState arrays) at the benchmark's 4 M-particle workload.

    python -m cudapic_amd.step_bench [--steps 4000] [--warmup 400]

Workload as bench.py: the converged reference case repeated 35 times at
W = 2e3 (3.78 M electrons + 3.97 M ions, MAX = 8 M per species, N_G = 400,
FP64 Poisson).  The arrays are device buffers the caller owns
(cudapic_amd.StepState); one call of pic_step_graph runs the timed steps,
whole 80-step periods from captured graphs.  Timed with the host clock
around the call and a stream synchronisation.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time


def measure(steps=4000, warmup=400, capacity=8_000_000, weight=2.0e3, rng_defer=0):
    import cudapic_amd as ca
    from cudapic_amd import synthetic
    p = ca.params(weight=weight, precision=ca.PIC_PREC_FP64)
    E, I = synthetic.converged_state(int(round(7.0e4 / weight)))
    with ca.StepState(p, capacity, rng_defer=rng_defer) as st:
        st.upload(0, *E)
        st.upload(1, *I)
        del E, I
        st.run(0, warmup, reset=True)
        st.synchronize()
        t = warmup % p.n_t
        n0 = (len(st.download(0)[0][0]), len(st.download(1)[0][0]))
        t0 = time.perf_counter()
        st.run(t, steps)
        st.synchronize()
        wall = time.perf_counter() - t0
        n1 = (len(st.download(0)[0][0]), len(st.download(1)[0][0]))
        _, _, overflow = st.diagnostics()
    ion_steps = sum(1 for k in range(steps) if (t + k) % p.n_sub == 0)
    ps = 0.5 * (n0[0] + n1[0]) * steps + 0.5 * (n0[1] + n1[1]) * ion_steps
    return {"value": round(ps / wall, 1), "unit": "particle-steps/s", "steps": steps,
            "ms_per_step": round(wall * 1e3 / steps, 5), "n_electrons": n1[0], "n_ions": n1[1],
            "overflow": overflow, "rng_defer": rng_defer or 8,
            "entry_point": "pic_step_graph on caller-owned buffers"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--rng-defer", type=int, default=0,
                    help="steps per run of deferred electron RNG write-backs (0 = the default, 8)")
    a = ap.parse_args()
    r = measure(a.steps, a.warmup, rng_defer=a.rng_defer)
    print(f"step_bench: {a.steps} steps, {r['ms_per_step']} ms per step", file=sys.stderr, flush=True)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
