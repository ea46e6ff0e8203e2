"""Electrode fluxes and mean ion energies of the eduPIC reference case, from
eduPIC's own CPU code (C/eduPIC.cc, built by `make -C oracle ref` into
oracle/_ref/edupic), as the fixture refcase_flux.json.

No published file holds these numbers: eduPIC prints them to info.txt
(C/eduPIC.cc:1003-1008) from its absorption counters (:625-667), and the
reference does not ship an info.txt.  This script makes them the way the GPU
test measures them: every chain starts from the converged state of
cudapic_amd/data/refcase_state.npz (written as eduPIC's picdata.bin, f64
headers, C/eduPIC.cc:726-781), relaxes RELAX cycles (`edupic RELAX`), then
runs BATCHES measurement runs of CYCLES cycles each (`edupic CYCLES m`);
eduPIC resets its counters in every process, so each run's info.txt is one
batch.  eduPIC seeds its Mersenne twister from std::random_device
(C/eduPIC.cc:164-165), so the chains are independent.

    python tests/golden/make_refcase_flux.py [--chains 6] [--workdir DIR] [--append]

Writes tests/golden/refcase_flux.json: per quantity the batch values (run
by run: the first BATCHES belong to the first chain), and their mean and
standard error over all batches.  Takes ~45 min on 6 cores of a 2.7 GHz
Xeon.  --append adds the new chains to the chains already in the output file
(the fixture holds 14: 6 made first, 8 appended).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

RELAX, CYCLES, BATCHES = 20, 25, 4

# info.txt labels (C/eduPIC.cc:1003-1008, 931) -> fixture keys
LABELS = {
    "Ion flux at powered electrode": "ion_flux_powered",
    "Ion flux at grounded electrode": "ion_flux_grounded",
    "Mean ion energy at powered electrode": "mean_ion_energy_powered",
    "Mean ion energy at grounded electrode": "mean_ion_energy_grounded",
    "Electron flux at powered electrode": "electron_flux_powered",
    "Electron flux at grounded electrode": "electron_flux_grounded",
    "Electron density @ center": "electron_density_center",
}


def parse_info(text):
    out = {}
    for line in text.splitlines():
        m = re.match(r"\s*(.+?)\s*=\s*([-+0-9.eE]+)", line)
        if m and m.group(1) in LABELS:
            out[LABELS[m.group(1)]] = float(m.group(2))
    missing = set(LABELS.values()) - set(out)
    if missing:
        raise RuntimeError(f"info.txt lacks {sorted(missing)} (a violated stability condition?)")
    return out


def run_chain(binary, wd, results, idx, log):
    from cudapic_amd import edupic_io, synthetic
    E, I = synthetic.converged_state()
    os.makedirs(wd, exist_ok=True)
    edupic_io.write_picdata(os.path.join(wd, "picdata.bin"), 0.0, 0, E, I)
    subprocess.run([binary, str(RELAX)], cwd=wd, check=True, stdout=subprocess.DEVNULL)
    for b in range(BATCHES):
        subprocess.run([binary, str(CYCLES), "m"], cwd=wd, check=True, stdout=subprocess.DEVNULL)
        with open(os.path.join(wd, "info.txt")) as f:
            results[idx].append(parse_info(f.read()))
        log(f"chain {idx}: batch {b + 1}/{BATCHES} done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--chains", type=int, default=6)
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--out", default=os.path.join(HERE, "refcase_flux.json"))
    ap.add_argument("--append", action="store_true", help="add the chains to those already in --out")
    a = ap.parse_args()
    binary = os.path.join(ROOT, "oracle", "_ref", "edupic")
    if not os.path.exists(binary):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    wd0 = a.workdir or tempfile.mkdtemp(prefix="edupic_flux_")
    lock = threading.Lock()

    def log(msg):
        with lock:
            print(msg, flush=True)

    results = [[] for _ in range(a.chains)]
    threads = [threading.Thread(target=run_chain, args=(binary, os.path.join(wd0, f"chain{c}"), results, c, log))
               for c in range(a.chains)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    batches = [b for r in results for b in r]
    if len(batches) != a.chains * BATCHES:
        raise SystemExit(f"only {len(batches)} of {a.chains * BATCHES} batches finished")
    chains = a.chains
    if a.append:
        with open(a.out) as f:
            old = json.load(f)
        if old["relax_cycles"] != RELAX or old["cycles_per_batch"] != CYCLES or \
                old["batches"] != old["chains"] * BATCHES:
            raise SystemExit(f"{a.out} was made with other run lengths")
        prev = [dict(zip(LABELS.values(), vals))
                for vals in zip(*(old["quantities"][k]["batches"] for k in LABELS.values()))]
        batches = prev + batches
        chains += old["chains"]
    out = {
        "source": "eduPIC 1.0 (C/eduPIC.cc) built with g++ -O3 -march=x86-64-v3, info.txt of each run",
        "start": "cudapic_amd/data/refcase_state.npz as picdata.bin",
        "relax_cycles": RELAX, "cycles_per_batch": CYCLES, "batches": len(batches), "chains": chains,
        "quantities": {},
    }
    for key in LABELS.values():
        v = np.array([b[key] for b in batches])
        out["quantities"][key] = {"mean": float(v.mean()), "stderr": float(v.std(ddof=1) / np.sqrt(len(v))),
                                  "batches": [float(x) for x in v]}
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {a.out}")


if __name__ == "__main__":
    main()
