"""Cost of the adjoint against the forward pass at scale.

    python -m cudapic_amd.adjoint_bench [--steps 8000] [--checkpoint 200] [--n-e 3740000]

Workload (the differentiation configuration): the 4 M-particle discharge of
the plain method at scale (W = 2e3, ~3.74 M electrons + ~3.93 M ions from
the plasma-like synthetic state), a window of `steps` electron steps (8000 =
two RF cycles) driven by V(t) = sum_{k=1..16} a_k sin(2 pi k f t) with
a_1 = 250 V and a_2..a_16 ~ N(0, 10 V) from a seed, objective the L2 misfit
of the window-averaged ion density against a seeded target profile (the
published reference-case ion density times 1 + 0.1 * a seeded smooth
perturbation).  Timed with the stream synchronised around each pass:
  record    the forward pass that draws the collisions and tapes the events
  forward   one frozen-event forward pass (what every gradient step needs)
  backward  checkpoint restores, recomputation and the adjoint steps
Prints one JSON line with ms per step of each and the ratios
backward / forward and (forward + backward) / forward.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def waveform(seed=5):
    wc = np.zeros(16)
    ws = np.zeros(16)
    ws[0] = 250.0
    ws[1:] = np.random.default_rng(seed).normal(0.0, 10.0, 15)
    return wc, ws


def target_profile(n_g, seed=6):
    from .synthetic import refcase_density
    x, _, ni = refcase_density()
    xs = np.linspace(0.0, x[-1], n_g)
    base = np.interp(xs, x, ni)
    rng = np.random.default_rng(seed)
    k = np.arange(1, 5)
    pert = (rng.normal(0.0, 1.0, 4)[:, None] * np.sin(np.pi * k[:, None] * xs[None, :] / x[-1])).sum(0)
    return base * (1.0 + 0.1 * pert / max(1e-12, np.abs(pert).max()))


def measure(n_e=3_740_000, n_i=3_930_000, capacity=8_000_000, weight=2.0e3, steps=8000, checkpoint=200,
            tape_capacity=0, log=None):
    import cudapic_amd as ca
    from cudapic_amd import synthetic
    p = ca.params(weight=weight, precision=ca.PIC_PREC_FP64)
    wc, ws = waveform()
    for h in range(16):
        p.wave_cos[h], p.wave_sin[h] = wc[h], ws[h]
    E, I = synthetic.state(n_e, n_i, seed=1)
    if not tape_capacity:
        # electrons collide ~0.5 % per step, ions ~3 % per ion step
        tape_capacity = int(min(4.0e9, max(1 << 20, 0.01 * capacity * steps)))
    with ca.Adjoint(p, capacity, capacity, steps, checkpoint=checkpoint, tape_capacity=tape_capacity) as adj:
        adj.upload(0, *E)
        adj.upload(1, *I)
        del E, I
        adj.set_target(target_profile(p.n_g))
        lib = ca.lib()

        def timed(fn):
            lib.pic_device_synchronize()
            t0 = time.perf_counter()
            r = fn()
            lib.pic_device_synchronize()
            return r, time.perf_counter() - t0

        j_rec, t_rec = timed(adj.record)
        if log:
            log(f"record {t_rec:.2f} s, J = {j_rec:.4e}")
        j_fwd, t_fwd = timed(lambda: adj.forward(wc, ws))
        if log:
            log(f"forward {t_fwd:.2f} s, J = {j_fwd:.4e}")
        (gc, gs, _), t_bwd = timed(adj.backward)
        if log:
            log(f"backward {t_bwd:.2f} s")
        sizes = adj.sizes()
    return {
        "steps": steps, "checkpoint": checkpoint, "electrons": n_e, "ions": n_i,
        "record_ms_per_step": round(1e3 * t_rec / steps, 4),
        "forward_ms_per_step": round(1e3 * t_fwd / steps, 4),
        "backward_ms_per_step": round(1e3 * t_bwd / steps, 4),
        "backward_over_forward": round(t_bwd / t_fwd, 3),
        "gradient_over_forward": round((t_fwd + t_bwd) / t_fwd, 3),
        "J": j_fwd, "replay_equals_record": j_fwd == j_rec,
        "grad_sin_a1": gs[0], "grad_sin_a2": gs[1],
        "tape_events": sizes,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--checkpoint", type=int, default=200)
    ap.add_argument("--n-e", type=int, default=3_740_000)
    ap.add_argument("--n-i", type=int, default=3_930_000)
    ap.add_argument("--capacity", type=int, default=8_000_000)
    ap.add_argument("--weight", type=float, default=2.0e3)
    args = ap.parse_args()

    def log(msg):
        print(f"adjoint_bench: {msg}", file=sys.stderr, flush=True)

    r = measure(args.n_e, args.n_i, args.capacity, args.weight, args.steps, args.checkpoint, log=log)
    print(json.dumps({"metric": "adjoint cost against the forward pass (frozen-event model)", **r}), flush=True)


if __name__ == "__main__":
    main()
