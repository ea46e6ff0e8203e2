#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the RF step on one MI355X.

Workload (the plain method at scale): the eduPIC argon CCP (25 mm, 10 Pa,
350 K, 13.56 MHz, 250 V, N_G = 400, N_T = 4000, N_SUB = 20) with the
superparticle weight lowered to 2e3 so that ~3.8 M electrons and ~4.0 M ions
are live (MAX_PARTICLES = 8 M per species, FP32 SoA), started from a
converged state: the engine's converged reference case (W = 7e4, 1500 RF
cycles, shipped as cudapic_amd/data/refcase_state.npz) with every particle
repeated 35 times, which is the same discharge at W = 2e3.

A step is one RF time step (electrons every step, ions every N_SUB steps);
a particle-step is one particle advanced through one of its own steps.
The timed region runs on the fast path (fused particle kernels, captured
into a hipGraph) with every input resident in HBM.  The dominant kernel (the
fused electron kernel) is then timed live with HIP events on the sim stream
over a second window, for the roofline.  The CPU baseline is the oracle's
serial C restatement of the reference step on a 40-step sample of the same
4 M-particle workload; beside it, the reference's own eduPIC (compiled from
source into oracle/_ref) is timed on one core over one RF cycle of the
reference case it is compiled for (~1e5 particles per species).

The metric also names the batched and the adjoint figures: after the main
measurement the line gains an "ensemble" object (64 discharges in one
launch per stage against the same 64 run one after the other on the fused
path, cudapic_amd.ens_bench's workload) and an "adjoint" object (backward
pass / frozen-event forward pass of the 4 M-particle discharge,
cudapic_amd.adjoint_bench).  The roofline gains an "hbm_resident" object
(the same step at 4x the particles, beyond the Infinity Cache) and the copy
cross-checks; "drop_in" gives the same workload through the two entry points
a reference host binds (the whole step on its own buffers, and the
reference's stage kernels).  --no-extras skips them all.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
ALG_BYTES = 80                 # per particle-step: deposit 4 + push/compact 32 + collisions 44
METRIC = "particle-steps/s at 4M particles, % HBM roofline; batched and adjoint/fwd ratio"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000, help="timed RF time steps (4000 = one RF cycle)")
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--kernel-steps", type=int, default=400, help="steps of the event-timed window")
    ap.add_argument("--n-e", type=int, default=3_740_000)
    ap.add_argument("--n-i", type=int, default=3_930_000)
    ap.add_argument("--capacity", type=int, default=8_000_000)
    ap.add_argument("--weight", type=float, default=2.0e3)
    ap.add_argument("--synthetic-start", action="store_true",
                    help="start from the plasma-like synthetic state of --n-e / --n-i particles instead of "
                         "the converged reference case replicated to the weight")
    ap.add_argument("--path", choices=["fused", "reference"], default="fused")
    ap.add_argument("--precision", choices=["fp32", "comp", "fp64"], default="fp64")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--block-size", type=int, default=0, help="lanes per workgroup of the fused kernels (128/256/512)")
    ap.add_argument("--ppt", type=int, default=0, help="particles per thread, compacting steps")
    ap.add_argument("--ppt-inplace", type=int, default=0, help="particles per thread, in-place steps")
    ap.add_argument("--replicas", type=int, default=0)
    ap.add_argument("--segments", type=int, default=0)
    ap.add_argument("--lds-replicas", type=int, default=0)
    ap.add_argument("--rng-defer", type=int, default=0,
                    help="steps per run of deferred electron RNG write-backs (0 = 8; 1 = every step)")
    ap.add_argument("--cand-masks", type=int, default=0, choices=(-1, 0, 1),
                    help="collision-candidate masks of the deferred runs (0 or 1 = on, -1 = off)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick grid, particles per thread and replication with cudapic_amd.autotune "
                         "(cached per device and workload; the search runs before the warm-up)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-ensemble and adjoint figures of the metric")
    ap.add_argument("--adjoint-steps", type=int, default=800,
                    help="window of the adjoint cost figure (the differentiation configuration "
                         "uses 8000: python -m cudapic_amd.adjoint_bench)")
    ap.add_argument("--cpu-sample", type=int, default=100_000, help="particles per species of the CPU sample")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the streaming-copy calibration kernel on MIB MiB, 20 times")
    ap.add_argument("--no-copy", action="store_true", help="skip the copy-bandwidth cross-checks")
    ap.add_argument("--long-cycles", type=int, default=50,
                    help="RF cycles of the sustained run after the headline window (the configuration's "
                         "50-cycle timing; 0 skips it; skipped with --no-extras)")
    return ap.parse_args()


def copy_bandwidth(mib: int, reps: int = 20) -> dict:
    """GB/s over mib MiB of three streams: pic_stream_copy (read + write, the
    calibration kernel of the traffic counters), the runtime's device-to-
    device hipMemcpyAsync (read + write) and the read-only pic_stream_read:
    what this device reaches from HBM, measured three ways."""
    import cudapic_amd as ca
    from cudapic_amd._abi import call
    n = mib * (1 << 20)
    a = ca.DeviceBuffer.zeroed(n // 4, np.uint32)
    b = ca.DeviceBuffer.zeroed(n // 4, np.uint32)
    acc = ca.DeviceBuffer.zeroed(1, np.uint64)

    def rate(launch, bytes_moved):
        launch()
        ca.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            launch()
        ca.synchronize()
        return round(bytes_moved * reps / (time.perf_counter() - t0) / 1e9, 1)

    out = {"copy_kernel": rate(lambda: call("pic_stream_copy", b.ptr, a.ptr, n, None), 2.0 * n),
           "hip_memcpy_d2d": rate(lambda: call("pic_memcpy_dtod_async", b.ptr, a.ptr, n, None), 2.0 * n),
           "read_kernel": rate(lambda: call("pic_stream_read", a.ptr, n, acc.ptr, None), 1.0 * n)}
    for buf in (a, b, acc):
        buf.free()
    return out


def traffic_from_profiles():
    """HBM bytes per launch of the fused electron kernel, from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration), or None if it has not been measured."""
    path = os.path.join(ROOT, "profiles", "fused_electron_4M_traffic.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(n_sample: int):
    """eduPIC (C/eduPIC.cc compiled into oracle/_ref/edupic) on one core: one
    RF cycle from n_sample electrons + ions of the reference case (weight 7e4
    as compiled, i.e. the same physical densities as the 4 M-particle run)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "edupic")
    if not os.path.exists(exe):
        return None
    from cudapic_amd import synthetic
    E, I = synthetic.state(n_sample, int(n_sample * 1.05), seed=7)
    tmp = tempfile.mkdtemp(prefix="edupic_")
    try:
        from cudapic_amd.edupic_io import write_picdata
        write_picdata(os.path.join(tmp, "picdata.bin"), 0.0, 1, E, I)
        t0 = time.perf_counter()
        r = subprocess.run([exe, "1"], cwd=tmp, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            return None
        conv = open(os.path.join(tmp, "conv.dat")).read().split()
        ne1, ni1 = int(conv[-2]), int(conv[-1])
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    ne0, ni0 = len(E[0]), len(I[0])
    ps = 0.5 * (ne0 + ne1) * 4000 + 0.5 * (ni0 + ni1) * 200
    host = os.cpu_count()
    return {"value": ps / wall, "unit": "particle-steps/s", "cores": 1, "host_cores": host, "kind": "reference",
            "sample": f"eduPIC C (g++ -O3 -march=x86-64-v3, serial: 1 of {host} host cores), 1 RF cycle "
                      f"(4000 steps) from {ne0} e + {ni0} i of the reference case at W=7e4 ({wall:.1f} s wall)"}


def port_baseline(weight=2.0e3, steps=40):
    """The same workload on the CPU: the oracle's serial C restatement of the
    reference step (oracle/pic_oracle.c: FP32 particles, FP64 Poisson, the
    reference's stage order) over a bounded sample of `steps` steps (two ion
    steps at 40) of the benchmark's starting state, one core.  The seeding of
    the per-slot RNG streams is outside the timed region."""
    from cudapic_amd import synthetic
    from oracle import oracle as orc
    p = orc.params(weight=weight)
    E, I = synthetic.converged_state(int(round(7.0e4 / weight)))
    cap = int(max(len(E[0]), len(I[0])) * 1.05)
    er = orc.seed_streams([0x12345678, 0x11112222, 0x22223333, 0x33334444], cap)
    ir = orc.seed_streams([0x44445555, 0x55556666, 0x66667777, 0x77778888], cap)
    sim = orc.Sim(p, cap, *E, *I, er, ir)
    n_e, n_i = len(E[0]), len(I[0])
    del E, I
    ps = 0.0
    t0 = time.perf_counter()
    for t in range(steps):
        ps += sim.n_e + (sim.n_i if t % 20 == 0 else 0)
        sim.step(t)
    wall = time.perf_counter() - t0
    host = os.cpu_count()
    return {"value": ps / wall, "unit": "particle-steps/s", "cores": 1, "host_cores": host, "kind": "port",
            "sample": f"CPU oracle (oracle/pic_oracle.c, the reference step restated in C, FP32 particles, "
                      f"FP64 Poisson), {steps} steps of the benchmark's start ({n_e} e + {n_i} i at W={weight:g}), "
                      f"serial: 1 of {host} host cores ({wall:.1f} s wall)"}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the RF step on a single MI355X "
              "(multi-GPU sharding is a design note, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import cudapic_amd as ca
    from cudapic_amd import synthetic

    if args.copy_probe:
        print(json.dumps({"copy_probe_MiB": args.copy_probe, "GB/s": copy_bandwidth(args.copy_probe)}))
        return
    # the ensemble figure and the HBM-resident roofline first, in child
    # processes, before this process allocates anything on the GPU; a failing
    # extra is reported in its own field and does not take the headline line
    # with it
    ens_fig = None if args.no_extras else _extra(ensemble_figure)
    hbm_fig = None if (args.no_extras or args.path != "fused") else _extra(hbm_resident_figure)
    drop_in = None if args.no_extras else {"caller_owned_step": _extra(caller_owned_step_figure),
                                           "reference_stage_kernels": _extra(reference_path_figure)}
    prec = {"fp32": ca.PIC_PREC_FP32, "comp": ca.PIC_PREC_COMP, "fp64": ca.PIC_PREC_FP64}[args.precision]
    path = ca.PIC_PATH_FUSED if args.path == "fused" else ca.PIC_PATH_REFERENCE
    p = ca.params(weight=args.weight, precision=prec)
    cu, ws, arch = ca.device_info()
    print(f"bench: device {arch}, {cu} CUs, wave{ws}; building state", file=sys.stderr, flush=True)
    reps = 7.0e4 / args.weight
    if not args.synthetic_start and abs(reps - round(reps)) < 1e-9 and round(reps) >= 1:
        # the converged reference case (W = 7e4), each particle repeated
        # 7e4 / W times: the same discharge at this weight
        E, I = synthetic.converged_state(int(round(reps)))
        start = f"converged reference case x{int(round(reps))}"
    else:
        E, I = synthetic.state(args.n_e, args.n_i, seed=1)
        start = "synthetic plasma-like state"
    print(f"bench: start: {start}, {len(E[0])} e + {len(I[0])} i", file=sys.stderr, flush=True)
    shape = dict(grid_blocks=args.grid, ppt=args.ppt, replicas=args.replicas, segments=args.segments,
                 ppt_inplace=args.ppt_inplace, lds_replicas=args.lds_replicas, block_size=args.block_size,
                 rng_defer=args.rng_defer, cand_masks=args.cand_masks)
    tuned = None
    if args.autotune and path == ca.PIC_PATH_FUSED:
        from cudapic_amd import autotune
        tuned = autotune.tune_sim(p, args.capacity, E, I)
        shape = tuned["config"]
        print(f"bench: autotuned {shape}", file=sys.stderr, flush=True)
    sim = ca.Sim(p, capacity=args.capacity, path=path, use_graph=not args.no_graph, **shape)
    sim.upload(ca.PIC_ELECTRONS, *E)
    sim.upload(ca.PIC_IONS, *I)
    del E, I
    sim_thr = sim.cand_threshold()
    sim.run(args.warmup)
    sim.synchronize()
    sim.check()
    print("bench: warmup done", file=sys.stderr, flush=True)
    # the timed window replays captured graphs only: whole 4*N_SUB-step
    # periods and the partial periods it starts or ends with, all captured
    # by a dry pass before the window opens (fused path with graphs, n_t a
    # multiple of 4*N_SUB); otherwise every step launches eagerly
    graph_steps = 0
    if not args.no_graph and path == ca.PIC_PATH_FUSED and p.n_t % (4 * p.n_sub) == 0:
        graph_steps = args.steps
    # timed region (one GPU: the barrier is the stream synchronisation)
    sim.synchronize()
    t = sim.time_steps(args.steps)
    sim.synchronize()
    sim.check()
    n_e, n_i = sim.counts()
    value = t["particle_steps"] / (t["ms_total"] * 1e-3)
    ms_per_step = t["ms_total"] / args.steps
    print(f"bench: timed {args.steps} steps in {t['ms_total']:.1f} ms", file=sys.stderr, flush=True)
    roofline = None
    if path == ca.PIC_PATH_FUSED and args.kernel_steps > 0:
        ne0, _ = sim.counts()
        k = sim.time_steps(args.kernel_steps, kernel_events=True)
        ne1, _ = sim.counts()
        avg_ms = k["ms_kernel"] / max(1, k["kernel_launches"])
        units = 0.5 * (ne0 + ne1)
        achieved = ALG_BYTES * units / (avg_ms * 1e-3) / 1e9
        traffic = traffic_from_profiles()
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": "electron particle kernel: k_inplace / k_fused compaction (every 800 steps); "
                              "push+collide+deposit, timed with HIP events on the sim stream",
                    "kernel_avg_us": round(avg_ms * 1e3, 2), "electrons_per_launch": int(units),
                    "bytes_per_unit": ALG_BYTES}
        # what the in-place electron kernel must move per electron: read x
        # and v, write x and vx (24 B); the RNG words read and written once
        # per run of deferred write-backs (32 B / R) and, with the
        # collision-candidate masks, a mask byte read on the run's other
        # steps, written once, and the words read only by the candidates
        # (a share thr of those steps); without masks the words are read
        # every step (40 + 16 / R B)
        defer = shape.get("rng_defer") or 8
        thr, masks_on = sim_thr
        if masks_on and defer > 1:
            compulsory = 24.0 + 33.0 / defer + (defer - 1) / defer * (1.0 + 16.0 * thr)
        else:
            compulsory = 40.0 + 16.0 / defer
        roofline["rng_defer"] = defer
        roofline["cand_masks"] = masks_on
        roofline["cand_threshold"] = round(thr, 5)
        roofline["compulsory_bytes_per_unit"] = round(compulsory, 2)
        roofline["compulsory_frac"] = round(compulsory * units / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    # the same discharge continued for long_cycles RF cycles in one timed
    # window (graphs): sustained throughput and how far the counts drift
    long_run = None
    if not args.no_extras and args.long_cycles > 0:
        ne0, ni0 = sim.counts()
        n_long = args.long_cycles * p.n_t
        lt = sim.time_steps(n_long)
        sim.check()
        ne1, ni1 = sim.counts()
        long_run = {"cycles": args.long_cycles, "steps": n_long,
                    "value": round(lt["particle_steps"] / (lt["ms_total"] * 1e-3), 1),
                    "unit": "particle-steps/s", "ms_per_step": round(lt["ms_total"] / n_long, 5),
                    "wall_s": round(lt["ms_total"] * 1e-3, 2), "n_electrons": [ne0, ne1], "n_ions": [ni0, ni1]}
        print(f"bench: {args.long_cycles} RF cycles in {lt['ms_total'] * 1e-3:.1f} s", file=sys.stderr, flush=True)
    sim.close()
    if roofline is not None and not args.no_copy:
        bw = copy_bandwidth(1024)
        roofline["copy_kernel_GBs"] = bw["copy_kernel"]
        roofline["hip_memcpy_d2d_GBs"] = bw["hip_memcpy_d2d"]
        roofline["read_kernel_GBs"] = bw["read_kernel"]
    if roofline is not None and hbm_fig is not None:
        roofline["hbm_resident"] = hbm_fig
    cpu = None if args.no_cpu_baseline else _extra(port_baseline, args.weight)
    cpu_edupic = None if args.no_cpu_baseline else cpu_baseline(args.cpu_sample)
    step_frac = value * ALG_BYTES / (HBM_PEAK_GBS * 1e9)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"argon CCP at scale: W={args.weight:g}, {n_e / 1e6:.2f}M e + {n_i / 1e6:.2f}M i, "
                               f"MAX={args.capacity / 1e6:g}M/species, N_G=400, N_T=4000, N_SUB=20 "
                               "(plain method at scale)", "start": start,
                   "n_electrons_end": n_e, "n_ions_end": n_i, "path": args.path,
                   "graph": graph_steps > 0, "graph_steps": graph_steps,
                   "eager_steps": args.steps - graph_steps, "poisson_and_density_flush": args.precision,
                   "step_hbm_frac_80B": round(step_frac, 4),
                   **({"autotuned": tuned["config"]} if tuned else {})},
        "roofline": roofline,
        "cpu_baseline": cpu,
        # eduPIC's own code cannot run the benchmark's workload (its weight is
        # compiled in: the 4M-particle state would be a 35x denser plasma),
        # so it is timed on the reference case it was built for
        "cpu_reference_edupic": cpu_edupic,
    }
    if not args.no_extras:
        out["long_run"] = long_run
        out["ensemble"] = ens_fig
        out["adjoint"] = _extra(adjoint_figure, args.adjoint_steps)
        out["drop_in"] = drop_in
    print(json.dumps(out), flush=True)


def _extra(fn, *a):
    try:
        return fn(*a)
    except Exception as e:  # reported in the JSON line instead
        print(f"bench: {fn.__name__} failed: {e}", file=sys.stderr, flush=True)
        return {"error": f"{type(e).__name__}: {e}"[:500]}


def ensemble_figure():
    """The batched half of the metric: 64 discharges (V0 ~ U[100,400] V,
    p ~ logU[3,30] Pa, 50 k - 2 M particles each) in one launch per stage
    per cache-sized group of instances, against the same 64 run one after
    the other on the fused path.  Measured by cudapic_amd.ens_bench in a
    child process started before this process touches the GPU: memory
    allocated earlier in a process (and still held by a live parent)
    changes where the ensemble's arrays land, which alone moves the
    cache-blocked ensemble by 3-5 %."""
    print("bench: ensemble of 64 discharges (child process)", file=sys.stderr, flush=True)
    r = subprocess.run([sys.executable, "-m", "cudapic_amd.ens_bench", "--sequential-path", "fused", "--one-group"],
                       cwd=ROOT, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"ens_bench failed ({r.returncode}):\n{r.stderr[-2000:]}")
    d = json.loads(r.stdout.strip().splitlines()[-1])
    return {"instances": d["instances"], "steps": d["steps"], "value": d["ensemble"]["value"],
            "sequential_fused": d["sequential"]["value"], "speedup": d["speedup"], "unit": d["unit"],
            "one_group_value": d["one_group"]["value"], "one_group_speedup": d["one_group_speedup"],
            "grouping": "cache-sized groups of instances, each a block of 5 graph periods at a time; "
                        "one_group: every instance in each launch, no temporal blocking",
            "process": "child (cudapic_amd.ens_bench)"}


def hbm_resident_figure():
    """The same step at 4x the particles (W = 500: the converged reference
    case repeated 140 times, ~15 M electrons + ~16 M ions, 2.6 GB of
    electron state), whose working set does not fit the 256 MB Infinity
    Cache: the electron kernel's roofline fraction streaming from HBM.
    Measured in a child process before this process touches the GPU."""
    print("bench: HBM-resident roofline at 4x the particles (child process)", file=sys.stderr, flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--capacity", "32000000", "--weight", "500",
                        "--steps", "400", "--warmup", "80", "--kernel-steps", "400", "--no-extras",
                        "--no-cpu-baseline", "--no-copy"], cwd=ROOT, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"HBM-resident run failed ({r.returncode}):\n{r.stderr[-2000:]}")
    d = json.loads(r.stdout.strip().splitlines()[-1])
    rf = d["roofline"]
    return {"value": d["value"], "electrons_per_launch": rf["electrons_per_launch"],
            "kernel_avg_us": rf["kernel_avg_us"], "achieved": rf["achieved"], "frac": rf["frac"],
            "workload": d["config"]["workload"], "process": "child (bench.py --capacity 32000000 --weight 500)"}


def caller_owned_step_figure():
    """The same workload through the entry point a reference host binds with
    its own GpuSimState buffers (pic_step_graph, cudapic_amd.step_bench),
    in a child process."""
    print("bench: caller-owned step (child process)", file=sys.stderr, flush=True)
    r = subprocess.run([sys.executable, "-m", "cudapic_amd.step_bench", "--steps", "4000", "--warmup", "400"],
                       cwd=ROOT, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"step_bench failed ({r.returncode}):\n{r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def reference_path_figure():
    """The same workload on the reference's stage order, one kernel per stage
    (the per-kernel drop-ins of main.rs:352-1358), in a child process."""
    print("bench: reference stage order (child process)", file=sys.stderr, flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--path", "reference", "--steps", "400",
                        "--warmup", "40", "--kernel-steps", "0", "--no-extras", "--no-cpu-baseline", "--no-copy"],
                       cwd=ROOT, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"reference path failed ({r.returncode}):\n{r.stderr[-2000:]}")
    d = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": d["value"], "unit": d["unit"], "steps": d["steps"], "ms_per_step": d["ms_per_step"],
            "entry_points": "pic_get_density, pic_solve_poisson, pic_move_and_compact, pic_check_collisions_e/_i"}


def adjoint_figure(steps):
    """The adjoint half of the metric: backward pass / frozen-event forward
    pass of the 4 M-particle discharge (cudapic_amd.adjoint_bench)."""
    from cudapic_amd import adjoint_bench
    print(f"bench: adjoint, {steps}-step window", file=sys.stderr, flush=True)
    r = adjoint_bench.measure(steps=steps)
    out = {k: r[k] for k in ("steps", "checkpoint", "forward_ms_per_step", "backward_ms_per_step",
                             "backward_over_forward", "gradient_over_forward", "replay_equals_record")}
    # the gradient checks run by the GPU tests (not measured here)
    out["gradient_check"] = {
        "where": "tests/test_gpu_adjoint.py",
        "particles": "200k e + 210k i (and 20k + 21k)", "window_steps": 400, "checkpoint": 200,
        "against": "FP64 CPU adjoint on the GPU's event tape; FP64 central differences of the frozen-event model",
        "h_V": [1e-3, 1e-2],
        "tolerance": "h = 1e-3 V: 1 % relative per coefficient; h = 1e-2 V: 1 % for the four coefficients "
                     "as a vector, 3 % per coefficient",
        "measured": "200k: 0.02-0.1 % at h = 1e-3 V; at h = 1e-2 V 0.4-1.5 % per coefficient (0.5 % as a "
                    "vector) on the current tape, 0.01-0.09 % on an earlier one; 20k: within 0.06 % at "
                    "h = 1e-3 V, up to 1.3 % off at h = 1e-2 V",
        "config_target_met": "no: the configuration's 1 % per coefficient at h = 1e-2 V fails on some tapes "
                             "(sin_1 1.5 % off); the frozen-event model is only piecewise smooth at that step",
        "not_checked": "the configuration's 4M-particle 8000-step window: beyond ~500 steps the frozen-event "
                       "model is ill-conditioned, so that window is measured for cost only"}
    return out


if __name__ == "__main__":
    main()
