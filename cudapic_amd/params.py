"""Discharge parameters (pic_params_t) with the reference case as default
(Rust-CUDA/src/main.rs:39-49)."""
from __future__ import annotations

import ctypes as C

from ._abi import PicParams, call, lib


def params(n_g=400, n_t=4000, n_sub=20, freq=13.56e6, v0=250.0, gap=0.025, pressure=10.0,
           t_gas=350.0, weight=7.0e4, area=1.0e-4, precision=0, wave_cos=None,
           wave_sin=None) -> PicParams:
    """Build and derive a pic_params_t.  ``wave_cos``/``wave_sin`` are the
    harmonic amplitudes of pot0(t) (default: V0 cos(2 pi t / N_T))."""
    p = PicParams()
    lib().pic_params_default(C.byref(p))
    p.n_g, p.n_t, p.n_sub, p.precision = n_g, n_t, n_sub, precision
    p.freq, p.v0, p.gap, p.pressure, p.t_gas = freq, v0, gap, pressure, t_gas
    p.weight, p.area = weight, area
    for h in range(16):
        p.wave_cos[h] = 0.0
        p.wave_sin[h] = 0.0
    for h, a in enumerate(wave_cos or []):
        p.wave_cos[h] = a
    for h, b in enumerate(wave_sin or []):
        p.wave_sin[h] = b
    call("pic_params_derive", C.byref(p))
    return p
