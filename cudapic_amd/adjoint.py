"""Gradients of the window-averaged ion density misfit with respect to the
driving waveform (pic_adj_* of the C ABI): a record pass tapes the discrete
events of a window, frozen-event forward passes replay them, and the
backward pass runs the adjoint push, deposit and field-solve kernels over
checkpointed segments."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import PicAdjConfig, call, lib

N_HARM = 16


class Adjoint:
    def __init__(self, p, capacity_e, capacity_i, steps, checkpoint=0, t0=0, tape_capacity=0,
                 seed_e=(0x12345678, 0x11112222, 0x22223333, 0x33334444),
                 seed_i=(0x44445555, 0x55556666, 0x66667777, 0x77778888)):
        cfg = PicAdjConfig()
        cfg.params = p
        cfg.capacity_e, cfg.capacity_i = int(capacity_e), int(capacity_i)
        cfg.steps, cfg.checkpoint, cfg.t0, cfg.tape_capacity = int(steps), int(checkpoint), int(t0), int(tape_capacity)
        cfg.seed_e[:] = list(seed_e)
        cfg.seed_i[:] = list(seed_i)
        self.p = p
        self.steps = int(steps)
        h = C.c_void_p()
        call("pic_adj_create", C.byref(cfg), C.byref(h))
        self.h = h.value

    def upload(self, species, x, vx, vy, vz):
        a = [np.ascontiguousarray(v, np.float32) for v in (x, vx, vy, vz)]
        call("pic_adj_upload", self.h, species, *(v.ctypes.data for v in a), len(a[0]))

    def set_target(self, n_i):
        t = np.ascontiguousarray(n_i, np.float64)
        assert len(t) == self.p.n_g
        call("pic_adj_set_target", self.h, t.ctypes.data)

    def record(self):
        j = C.c_double()
        call("pic_adj_record", self.h, C.byref(j))
        return j.value

    def forward(self, wave_cos=None, wave_sin=None, nbar=False):
        """J (and the window-averaged ion density) of a frozen-event pass."""
        wc = None if wave_cos is None else np.ascontiguousarray(_pad(wave_cos), np.float64)
        ws = None if wave_sin is None else np.ascontiguousarray(_pad(wave_sin), np.float64)
        j = C.c_double()
        nb = np.zeros(self.p.n_g, np.float64)
        call("pic_adj_forward", self.h, None if wc is None else wc.ctypes.data,
             None if ws is None else ws.ctypes.data, C.byref(j), nb.ctypes.data)
        return (j.value, nb) if nbar else j.value

    def backward(self):
        """(dJ/dwave_cos[16], dJ/dwave_sin[16], dJ/dpot0(t)[steps])."""
        gc = np.zeros(N_HARM, np.float64)
        gs = np.zeros(N_HARM, np.float64)
        dv = np.zeros(self.steps, np.float64)
        call("pic_adj_backward", self.h, gc.ctypes.data, gs.ctypes.data, dv.ctypes.data)
        return gc, gs, dv

    def sizes(self):
        """Event counts of the tape and the slots in use at the window's end."""
        v = [C.c_uint32() for _ in range(5)]
        call("pic_adj_tape_sizes", self.h, *(C.byref(x) for x in v))
        return dict(zip(("electron_collisions", "ion_collisions", "ionisations", "slots_e", "slots_i"),
                        (x.value for x in v)))

    def tape(self):
        """The recorded events as numpy arrays (see pic_adj_download_tape)."""
        ne, ni, nion, se, si = (C.c_uint32() for _ in range(5))
        call("pic_adj_tape_sizes", self.h, *(C.byref(v) for v in (ne, ni, nion, se, si)))
        T1 = self.steps + 1
        out = {
            "off_e": np.zeros(T1, np.uint32), "slot_e": np.zeros(max(1, ne.value), np.uint32),
            "v_e": np.zeros(max(1, ne.value), np.float32),
            "off_i": np.zeros(T1, np.uint32), "slot_i": np.zeros(max(1, ni.value), np.uint32),
            "v_i": np.zeros(max(1, ni.value), np.float32),
            "off_ion": np.zeros(T1, np.uint32), "parent": np.zeros(max(1, nion.value), np.uint32),
            "new_e": np.zeros(max(1, nion.value), np.uint32), "new_i": np.zeros(max(1, nion.value), np.uint32),
            "v_new_e": np.zeros(max(1, nion.value), np.float32), "v_new_i": np.zeros(max(1, nion.value), np.float32),
            "slots_e": np.zeros(T1, np.uint32), "slots_i": np.zeros(T1, np.uint32),
            "death_e": np.zeros(max(1, se.value), np.uint32), "death_i": np.zeros(max(1, si.value), np.uint32),
        }
        call("pic_adj_download_tape", self.h, *(a.ctypes.data for a in out.values()))
        n = {"slot_e": ne.value, "v_e": ne.value, "slot_i": ni.value, "v_i": ni.value, "parent": nion.value,
             "new_e": nion.value, "new_i": nion.value, "v_new_e": nion.value, "v_new_i": nion.value,
             "death_e": se.value, "death_i": si.value}
        return {k: (v[:n[k]] if k in n else v) for k, v in out.items()}

    def close(self):
        if getattr(self, "h", None):
            lib().pic_adj_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _pad(a):
    a = np.zeros(N_HARM) if a is None else np.asarray(a, np.float64)
    out = np.zeros(N_HARM)
    out[:len(a)] = a
    return out
