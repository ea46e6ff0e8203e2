"""Launch-shape autotuner for the particle kernels.

    python -m cudapic_amd.autotune sim [--n-e 3740000 --n-i 3930000 --capacity 8000000 --weight 2e3]
    python -m cudapic_amd.autotune ensemble [--instances 64]

The fused path's speed depends on a handful of launch parameters whose best
values move with the device, the particle count and the grid size:

  block_size    lanes per workgroup of the particle kernels (128, 256, 512)
  lanes_per_cu  persistent lanes per compute unit (512, 1024, 2048; the grid
                is lanes_per_cu / block_size workgroups per compute unit)
  ppt_inplace   particles per lane per tile of the in-place kernel (1, 2, 4)
  ppt           particles per lane per tile of the compacting kernel (2, 4)
  lds_replicas  LDS density histograms per workgroup (1, 2, 4; the waves of
                a workgroup share them round-robin)
  replicas      global density replicas the workgroup flushes spread over
                (8, 16, 32)
  segments      segments per species (1, 8, 32; the single-discharge path)
  rng_defer     steps per run of deferred electron RNG write-backs (8, 4,
                16: fewer stores against more catch-up draws per load)

The tuner measures real steps: for each candidate it builds the simulation,
uploads the same starting state, runs a warm-up (which captures the step
graph) and times whole 4*N_SUB-step graph periods with HIP events.  The
search is coordinate descent from the defaults: one parameter at a time is
swept with the others held at their best values so far, and the sweep
repeats until a full pass changes nothing (at most `passes` passes).  A
candidate must beat the incumbent by more than `min_gain` (relative) to
replace it, so that timing noise does not walk the configuration around.

Results are cached in a JSON file (CUDAPIC_AUTOTUNE_CACHE, default
~/.cache/cudapic_amd/autotune.json) under a key made of the device, the
kind of run and the workload's shape (grid size, capacity, particle count
to the nearest power of two, precision).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys

SIM_SPACE = {
    "block_size": [256, 128, 512],
    "lanes_per_cu": [1024, 512, 2048],
    "ppt_inplace": [2, 1, 4],
    "lds_replicas": [0, 2, 1],
    "replicas": [8, 16, 32],
    "segments": [0, 8, 32],
    "ppt": [4, 2],
    "rng_defer": [8, 4, 16],
}
ENSEMBLE_SPACE = {
    "block_size": [256, 128, 512],
    "lanes_per_cu": [1024, 512, 2048],
    "ppt": [2, 1, 4],
    "lds_replicas": [0, 2, 1],
    "replicas": [8, 16, 32],
    # electron slots per group of instances run a block at a time
    # (0xFFFFFFFF: one group of all)
    "group_slots": [9_000_000, 6_000_000, 12_000_000, 0xFFFFFFFF],
    "rng_defer": [8, 4, 16],
}


def coordinate_descent(evaluate, space, passes=2, min_gain=0.01, log=None):
    """Maximise evaluate(config) over the product space by sweeping one
    parameter at a time.  The first value of each list is the starting
    point.  Returns (best_config, best_value, evaluations) where evaluations
    lists every (config, value) measured (each distinct config once)."""
    best = {k: v[0] for k, v in space.items()}
    seen = {}

    def score(cfg):
        key = tuple(sorted(cfg.items()))
        if key not in seen:
            seen[key] = evaluate(dict(cfg))
            if log:
                log(cfg, seen[key])
        return seen[key]

    best_val = score(best)
    for _ in range(passes):
        changed = False
        for name, values in space.items():
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best)
                cand[name] = v
                val = score(cand)
                if val > best_val * (1.0 + min_gain):
                    best, best_val, changed = cand, val, True
        if not changed:
            break
    return best, best_val, [(dict(k), v) for k, v in seen.items()]


def cache_path():
    return os.environ.get("CUDAPIC_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "cudapic_amd", "autotune.json"))


def _load_cache():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key, entry):
    data = _load_cache()
    data[key] = entry
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _device_key():
    from .device import device_info
    cus, _, arch = device_info()
    return f"{arch}x{cus}"


def workload_key(kind, n_g, capacity, n_particles, precision):
    return (f"{_device_key()}:{kind}:ng{n_g}:cap{capacity}:"
            f"n2^{round(math.log2(max(1, n_particles)))}:prec{precision}")


def _rate(timing):
    return timing["particle_steps"] / (timing["ms_total"] * 1e-3)


def tune_sim(p, capacity, electrons, ions, steps=80, warmup=80, passes=2, use_cache=True, log=None):
    """Best fused-path launch configuration for this workload.  electrons and
    ions are (x, vx, vy, vz) host arrays of the starting state.  Returns
    {"config": Sim keyword arguments, "rate": particle-steps/s, "key": ...}."""
    from . import Sim
    from .device import device_info
    key = workload_key("sim", p.n_g, capacity, len(electrons[0]), p.precision)
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            return hit
    cus = device_info()[0]

    def evaluate(cfg):
        kw = sim_kwargs(cfg, cus)
        with Sim(p, capacity=capacity, **kw) as sim:
            sim.upload(0, *electrons)
            sim.upload(1, *ions)
            sim.run(warmup)
            sim.synchronize()
            r = _rate(sim.time_steps(steps))
            sim.check()
        return r

    best, val, evals = coordinate_descent(evaluate, SIM_SPACE, passes=passes, log=log)
    entry = {"config": sim_kwargs(best, cus), "rate": val, "key": key, "evaluations": len(evals)}
    if use_cache:
        _store_cache(key, entry)
    return entry


def _grid(cfg, cus):
    return max(1, cfg["lanes_per_cu"] // cfg["block_size"]) * cus


def sim_kwargs(cfg, cus):
    return {"block_size": cfg["block_size"], "grid_blocks": _grid(cfg, cus), "ppt_inplace": cfg["ppt_inplace"],
            "lds_replicas": cfg["lds_replicas"], "replicas": cfg["replicas"], "segments": cfg["segments"],
            "ppt": cfg["ppt"], "rng_defer": cfg.get("rng_defer", 0)}


def ensemble_kwargs(cfg, cus):
    return {"block_size": cfg["block_size"], "grid_blocks": _grid(cfg, cus), "ppt": cfg["ppt"],
            "lds_replicas": cfg["lds_replicas"], "replicas": cfg["replicas"], "group_slots": cfg["group_slots"],
            "rng_defer": cfg.get("rng_defer", 0)}


def tune_ensemble(params_list, capacities, states, steps=80, warmup=80, passes=2, use_cache=True, log=None):
    """Best ensemble launch configuration; states[i] = (electrons, ions) of
    instance i.  Returns {"config": Ensemble keyword arguments, ...}."""
    from . import Ensemble
    from .device import device_info
    n_tot = sum(len(s[0][0]) for s in states)
    key = workload_key(f"ensemble{len(params_list)}", params_list[0].n_g, sum(capacities), n_tot,
                       params_list[0].precision)
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            return hit
    cus = device_info()[0]

    def evaluate(cfg):
        ens = Ensemble(params_list, capacities, **ensemble_kwargs(cfg, cus))
        try:
            for i, (E, I) in enumerate(states):
                ens.upload(i, 0, *E)
                ens.upload(i, 1, *I)
            ens.run(warmup)
            ens.synchronize()
            r = _rate(ens.time_steps(steps))
            ens.check()
        finally:
            ens.close()
        return r

    best, val, evals = coordinate_descent(evaluate, ENSEMBLE_SPACE, passes=passes, log=log)
    entry = {"config": ensemble_kwargs(best, cus), "rate": val, "key": key, "evaluations": len(evals)}
    if use_cache:
        _store_cache(key, entry)
    return entry


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=["sim", "ensemble"])
    ap.add_argument("--n-e", type=int, default=3_740_000)
    ap.add_argument("--n-i", type=int, default=3_930_000)
    ap.add_argument("--capacity", type=int, default=8_000_000)
    ap.add_argument("--weight", type=float, default=2.0e3)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--steps", type=int, default=80)
    ap.add_argument("--no-cache", action="store_true", help="measure even if a cached result exists")
    args = ap.parse_args()
    import cudapic_amd as ca
    from cudapic_amd import synthetic

    def log(cfg, val):
        print(f"autotune: {cfg} -> {val:.4e} particle-steps/s", file=sys.stderr, flush=True)

    if args.kind == "sim":
        p = ca.params(weight=args.weight, precision=ca.PIC_PREC_FP64)
        E, I = synthetic.state(args.n_e, args.n_i, seed=3)
        res = tune_sim(p, args.capacity, E, I, steps=args.steps, use_cache=not args.no_cache, log=log)
    else:
        from .ens_bench import instances
        v0, pres, cnt = instances(args.instances)
        ps = [ca.params(v0=float(v), pressure=float(pp), weight=7.0e3 * 1.2e6 / float(c),
                        precision=ca.PIC_PREC_FP64) for v, pp, c in zip(v0, pres, cnt)]
        caps = [int(c * 1.5) + 4096 for c in cnt]
        states = [synthetic.state(int(c), int(c * 1.05), seed=100 + i) for i, c in enumerate(cnt)]
        res = tune_ensemble(ps, caps, states, steps=args.steps, use_cache=not args.no_cache, log=log)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
