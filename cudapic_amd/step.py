"""The RF step on caller-owned buffers (pic_init / pic_step / pic_step_graph):
the reference's GpuSimState (Rust-CUDA/src/main.rs:132-185) as device
buffers this object owns the way the reference's host owns them, and a step
context (the fused step's workspace) from the library.

    st = StepState(params, capacity=400_000)
    st.upload(0, x, vx, vy, vz); st.upload(1, ...)
    st.step(t, reset=True)            # first step deposits the particles
    st.run(t + 1, 79)                 # graphs for whole 80-step periods
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import (PIC_CS_RANGES, PIC_N_CS, PIC_N_EEPF, PIC_N_IFED, PIC_STEP_MEASURE, PIC_STEP_RESET,
                   PicState, PicStepConfig, call)
from .device import DeviceBuffer

SEED_E = (0x12345678, 0x11112222, 0x22223333, 0x33334444)   # main.rs:1667
SEED_I = (0x44445555, 0x55556666, 0x66667777, 0x77778888)   # main.rs:1671

_E = ("e_x", "e_vx", "e_vy", "e_vz")
_I = ("i_x", "i_vx", "i_vy", "i_vz")
_TE = ("tmp_x", "tmp_vx", "tmp_vy", "tmp_vz")
_TI = ("tmp_i_x", "tmp_i_vx", "tmp_i_vy", "tmp_i_vz")


class StepState:
    def __init__(self, p, capacity, seed_e=SEED_E, seed_i=SEED_I, tables=None, stream=True,
                 compact_every=0, grid_blocks=0, ppt=0, ppt_inplace=0, lds_replicas=0, rng_defer=0, context=True):
        """tables: (cs, stot_e, stot_i) host arrays (default: pic_cross_sections
        of the discharge).  stream: run on a stream of its own (needed for
        graphs) or on the null stream.  context: create the pic_init context
        (False: only the buffers, for the no-allocation entry point)."""
        from . import kernels
        self.p = p
        self.capacity = cap = int(capacity)
        self._bufs = {}
        st = PicState()

        def buf(name, n, dtype):
            b = DeviceBuffer.zeroed(n, dtype)
            self._bufs[name] = b
            setattr(st, name, b.ptr)
            return b

        for names in (_E, _I, _TE, _TI):
            for n in names:
                buf(n, cap, np.float32)
        for n in ("rng_e0", "rng_e1", "rng_e2", "rng_e3", "rng_i0", "rng_i1", "rng_i2", "rng_i3"):
            buf(n, cap, np.uint32)
        buf("n_electrons", 1, np.uint32)
        buf("n_ions", 1, np.uint32)
        buf("efield", p.n_g, np.float32)
        buf("pot", p.n_g, np.float32)
        buf("e_density", p.n_g, np.float64)
        buf("i_density", p.n_g, np.float64)
        buf("cumul_e_density", p.n_g, np.float64)
        buf("cumul_i_density", p.n_g, np.float64)
        buf("eepf_counts", PIC_N_EEPF, np.uint32)
        cs = buf("cs", PIC_N_CS * PIC_CS_RANGES, np.float32)
        se = buf("sigma_tot_e", PIC_CS_RANGES, np.float32)
        si = buf("sigma_tot_i", PIC_CS_RANGES, np.float32)
        if tables is None:
            tables = kernels.cross_sections(p.gas_density)
        for b, a in zip((cs, se, si), tables):
            b.copy_from_host(np.asarray(a, np.float32))
        st.capacity = cap
        self.st = st
        # main.rs:1666-1675: slot k of a species gets its master jumped k times
        for seed, names in ((seed_e, ("rng_e0", "rng_e1", "rng_e2", "rng_e3")),
                            (seed_i, ("rng_i0", "rng_i1", "rng_i2", "rng_i3"))):
            m = (C.c_uint32 * 4)(*[int(v) for v in seed])
            call("pic_seed_streams", m, cap, *(self._bufs[n].ptr for n in names), None)
        call("pic_device_synchronize")
        self.stream = None
        if stream:
            h = C.c_void_p()
            call("pic_stream_create", C.byref(h))
            self.stream = h.value
        self.ctx = None
        if context:
            cfg = PicStepConfig(compact_every=compact_every, grid_blocks=grid_blocks, ppt=ppt,
                                ppt_inplace=ppt_inplace, lds_replicas=lds_replicas, rng_defer=rng_defer)
            h = C.c_void_p()
            call("pic_init", C.byref(p), C.byref(cfg), C.byref(h))
            self.ctx = h.value

    # ------------------------------------------------------------ buffers
    def _ptrs(self, species, spare=False):
        names = (_TE if spare else _E) if species == 0 else (_TI if spare else _I)
        return [getattr(self.st, n) for n in names]

    def _count_ptr(self, species):
        return self.st.n_electrons if species == 0 else self.st.n_ions

    def upload(self, species, x, vx, vy, vz):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (x, vx, vy, vz)]
        n = len(arrs[0])
        assert n <= self.capacity
        # the copies run on the null stream: finish the steps in flight first
        self.synchronize()
        for ptr, a in zip(self._ptrs(species), arrs):
            if n:
                call("pic_memcpy_htod", ptr, a.ctypes.data, a.nbytes)
        c = np.array([n], np.uint32)
        call("pic_memcpy_htod", self._count_ptr(species), c.ctypes.data, 4)

    def count(self, species):
        """The watermark *n_electrons / *n_ions (slots in use, holes included)."""
        self.synchronize()
        c = np.zeros(1, np.uint32)
        call("pic_memcpy_dtoh", c.ctypes.data, self._count_ptr(species), 4)
        return int(c[0])

    def slots_in_use(self, species):
        """x, vx, vy, vz of the slots below the watermark (holes: x < 0)."""
        n = min(self.count(species), self.capacity)
        out = []
        for ptr in self._ptrs(species):
            a = np.empty(max(n, 1), np.float32)
            if n:
                call("pic_memcpy_dtoh", a.ctypes.data, ptr, 4 * n)
            out.append(a[:n])
        return out

    def download(self, species):
        """The live particles (holes skipped) and their slot indices."""
        a = self.slots_in_use(species)
        live = a[0] >= 0.0
        return [v[live] for v in a], np.nonzero(live)[0]

    def rng_download(self, species, n=None):
        n = self.capacity if n is None else int(n)
        names = ("rng_e0", "rng_e1", "rng_e2", "rng_e3") if species == 0 else ("rng_i0", "rng_i1", "rng_i2", "rng_i3")
        self.synchronize()
        return [self._bufs[k].to_host_vec(n) for k in names]

    def grid(self):
        """pot, efield (f32) and the f64 node densities of the last solve."""
        self.synchronize()
        return (self._bufs["pot"].to_host_vec(), self._bufs["efield"].to_host_vec(),
                self._bufs["e_density"].to_host_vec(), self._bufs["i_density"].to_host_vec())

    def measurements(self):
        self.synchronize()
        return {"cumul_e": self._bufs["cumul_e_density"].to_host_vec(),
                "cumul_i": self._bufs["cumul_i_density"].to_host_vec(),
                "eepf": self._bufs["eepf_counts"].to_host_vec()}

    def diagnostics(self):
        """counters[8] (pic_sim_read_measurements order), IFED [2][200], overflow flag."""
        c = (C.c_ulonglong * 8)()
        ifed = np.zeros(2 * PIC_N_IFED, np.uint32)
        ov = C.c_uint32()
        call("pic_step_diagnostics", self.ctx, c, ifed.ctypes.data, C.byref(ov), self.stream)
        return list(c), ifed.reshape(2, PIC_N_IFED), ov.value

    # ------------------------------------------------------------ steps
    @staticmethod
    def _flags(measure, reset):
        return (PIC_STEP_MEASURE if measure else 0) | (PIC_STEP_RESET if reset else 0)

    def step(self, t, measure=False, reset=False):
        call("pic_step", self.ctx, C.byref(self.st), int(t), self._flags(measure, reset), self.stream)

    def step_noalloc(self, t, measure=False, reset=False):
        """pic_step_noalloc: the same step on the library's fixed workspace."""
        call("pic_step_noalloc", C.byref(self.p), C.byref(self.st), int(t), self._flags(measure, reset),
             self.stream)

    def step_lagged(self, t, lag, lag_out, measure=False, reset=False):
        """pic_step_lagged: one step whose live electron slots' RNG words are
        `lag` draws behind on entry and are left `lag_out` behind."""
        call("pic_step_lagged", self.ctx, C.byref(self.st), int(t), self._flags(measure, reset), int(lag),
             int(lag_out), self.stream)

    def run(self, t, n_steps, measure=False, reset=False):
        call("pic_step_graph", self.ctx, C.byref(self.st), int(t), int(n_steps), self._flags(measure, reset),
             self.stream)

    def synchronize(self):
        if self.stream:
            call("pic_stream_synchronize", self.stream)
        else:
            call("pic_device_synchronize")

    def close(self):
        if getattr(self, "ctx", None):
            self.synchronize()
            call("pic_destroy", self.ctx)
            self.ctx = None
        if getattr(self, "stream", None):
            call("pic_stream_destroy", self.stream)
            self.stream = None
        for b in self._bufs.values():
            b.free()
        self._bufs = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
