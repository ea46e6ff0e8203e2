"""eduPIC's particle data file, picdata.bin (C/eduPIC.cc:726-781): Time,
cycles_done and N_e as f64, then x, vx, vy, vz of the electrons as f64,
N_i and the ions likewise.  The C++ eduPIC (Cpp/eduPIC.cpp:728-746) writes
cycles_done, N_e and N_i as int32 instead (int_headers=True).  The `cudapic`
command line imports and exports the same format (--import-edupic,
--export-edupic)."""
from __future__ import annotations

import numpy as np


def write_picdata(path, time, cycles, electrons, ions, int_headers=False):
    """electrons, ions: (x, vx, vy, vz) arrays of equal length."""
    with open(path, "wb") as f:
        np.array([time], np.float64).tofile(f)
        hdr = np.int32 if int_headers else np.float64
        np.array([cycles], hdr).tofile(f)
        for sp in (electrons, ions):
            n = len(sp[0])
            assert all(len(a) == n for a in sp)
            np.array([n], hdr).tofile(f)
            for a in sp:
                np.asarray(a, np.float64).tofile(f)


def read_picdata(path, int_headers=False):
    """{"time", "cycles", "electrons": (x, vx, vy, vz), "ions": (...)} as f64."""
    hdr = np.int32 if int_headers else np.float64
    with open(path, "rb") as f:
        time = float(np.fromfile(f, np.float64, 1)[0])
        cycles = int(np.fromfile(f, hdr, 1)[0])
        out = {"time": time, "cycles": cycles}
        for name in ("electrons", "ions"):
            n = int(np.fromfile(f, hdr, 1)[0])
            arrs = tuple(np.fromfile(f, np.float64, n) for _ in range(4))
            if any(len(a) != n for a in arrs):
                raise ValueError(f"{path}: truncated {name} block")
            out[name] = arrs
    return out
