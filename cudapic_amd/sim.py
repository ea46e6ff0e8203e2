"""Simulation object: the reference's GpuSimState + RF loop
(Rust-CUDA/src/main.rs:132-350, 1611-1878) behind pic_sim_* of the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import (PIC_ELECTRONS, PIC_IONS, PIC_N_EEPF, PIC_N_IFED, PIC_PATH_FUSED, PicSimConfig,
                   call, lib)


class Sim:
    def __init__(self, p, capacity=120_000, path=PIC_PATH_FUSED, use_graph=True, seed_e=None,
                 seed_i=None, segments=0, block_size=0, grid_blocks=0, ppt=0, replicas=0,
                 ppt_inplace=0, lds_replicas=0, compact_every=0, rng_defer=0, cand_masks=0,
                 cand_threshold=0.0):
        cfg = PicSimConfig()
        lib().pic_sim_config_default(C.byref(cfg))
        cfg.params = p
        cfg.capacity = int(capacity)
        cfg.segments = int(segments)
        if seed_e is not None:
            cfg.seed_e[:] = [int(v) for v in seed_e]
        if seed_i is not None:
            cfg.seed_i[:] = [int(v) for v in seed_i]
        cfg.path = int(path)
        cfg.use_graph = int(bool(use_graph))
        cfg.block_size, cfg.grid_blocks, cfg.ppt, cfg.replicas = block_size, grid_blocks, ppt, replicas
        cfg.ppt_inplace = ppt_inplace
        cfg.lds_replicas = lds_replicas
        cfg.compact_every = int(compact_every)
        cfg.rng_defer = int(rng_defer)
        cfg.cand_masks = int(cand_masks)
        cfg.cand_threshold = float(cand_threshold)
        self.p = p
        self.cfg = cfg
        self.capacity = int(capacity)
        h = C.c_void_p()
        call("pic_sim_create", C.byref(cfg), C.byref(h))
        self.h = h.value

    # -------------------------------------------------------------- state
    def upload(self, species, x, vx, vy, vz):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (x, vx, vy, vz)]
        call("pic_sim_upload", self.h, species, *(a.ctypes.data for a in arrs), len(arrs[0]))

    def download(self, species):
        n = C.c_uint32()
        call("pic_sim_download", self.h, species, None, None, None, None, 0, C.byref(n))
        out = [np.zeros(max(1, n.value), np.float32) for _ in range(4)]
        call("pic_sim_download", self.h, species, *(a.ctypes.data for a in out), n.value, C.byref(n))
        return [a[:n.value] for a in out]

    def segments(self, species):
        """(n_seg, seg_size, counts[n_seg]): particle j of segment q is in slot
        q * seg_size + j; download() returns the segments in order."""
        ns, sz = C.c_uint32(), C.c_uint32()
        cnt = np.zeros(64, np.uint32)
        call("pic_sim_segments", self.h, species, C.byref(ns), C.byref(sz), cnt.ctypes.data)
        return ns.value, sz.value, cnt[:ns.value].copy()

    def slots(self, species):
        """Slot index of every particle in download() order."""
        n = C.c_uint32()
        call("pic_sim_download_slots", self.h, species, None, 0, C.byref(n))
        out = np.zeros(max(1, n.value), np.uint32)
        call("pic_sim_download_slots", self.h, species, out.ctypes.data, n.value, C.byref(n))
        return out[:n.value].astype(np.int64)

    def rng_upload(self, species, words):
        w = [np.ascontiguousarray(a, np.uint32) for a in words]
        call("pic_sim_rng_upload", self.h, species, *(a.ctypes.data for a in w), len(w[0]))

    def rng_download(self, species, n=None):
        if n is None:
            ns, sz, _ = self.segments(species)
            n = ns * sz
        out = [np.zeros(max(1, n), np.uint32) for _ in range(4)]
        call("pic_sim_rng_download", self.h, species, *(a.ctypes.data for a in out), n)
        return [a[:n] for a in out]

    def set_tables(self, cs, stot_e, stot_i):
        a = [np.ascontiguousarray(v, np.float32) for v in (cs, stot_e, stot_i)]
        call("pic_sim_set_tables", self.h, *(v.ctypes.data for v in a))

    def cand_threshold(self):
        """(threshold, masks on) of the collision-candidate masks."""
        t, on = C.c_float(), C.c_int()
        call("pic_sim_cand_threshold", self.h, C.byref(t), C.byref(on))
        return t.value, bool(on.value)

    def counts(self):
        ne, ni = C.c_uint32(), C.c_uint32()
        call("pic_sim_counts", self.h, C.byref(ne), C.byref(ni))
        return ne.value, ni.value

    def time(self):
        t, s = C.c_uint32(), C.c_ulonglong()
        call("pic_sim_time", self.h, C.byref(t), C.byref(s))
        return t.value, s.value

    def set_time(self, t):
        call("pic_sim_set_time", self.h, int(t))

    # -------------------------------------------------------------- stepping
    def run(self, n_steps, measure=False):
        call("pic_sim_run", self.h, int(n_steps), int(bool(measure)))

    def synchronize(self):
        call("pic_sim_synchronize", self.h)

    def check(self):
        call("pic_sim_check", self.h)

    def overflowed(self) -> bool:
        return lib().pic_sim_check(self.h) != 0

    def time_steps(self, n_steps, measure=False, kernel_events=False):
        tot, ker = C.c_float(), C.c_float()
        nl, ps = C.c_uint32(), C.c_ulonglong()
        call("pic_sim_time_steps", self.h, int(n_steps), int(bool(measure)), C.byref(tot),
             C.byref(ker) if kernel_events else None, C.byref(nl), C.byref(ps))
        return {"ms_total": tot.value, "ms_kernel": ker.value if kernel_events else None,
                "kernel_launches": nl.value, "particle_steps": ps.value}

    # -------------------------------------------------------------- readback
    def grid(self):
        n = self.p.n_g
        pot = np.zeros(n, np.float32)
        ef = np.zeros(n, np.float32)
        ed = np.zeros(n, np.float64)
        idn = np.zeros(n, np.float64)
        call("pic_sim_read_grid", self.h, pot.ctypes.data, ef.ctypes.data, ed.ctypes.data, idn.ctypes.data)
        return pot, ef, ed, idn

    def measurements(self):
        n = self.p.n_g
        ce = np.zeros(n, np.float64)
        ci = np.zeros(n, np.float64)
        eepf = np.zeros(PIC_N_EEPF, np.uint32)
        ifed = np.zeros(2 * PIC_N_IFED, np.uint32)
        cnt = np.zeros(8, np.uint64)
        call("pic_sim_read_measurements", self.h, ce.ctypes.data, ci.ctypes.data, eepf.ctypes.data,
             ifed.ctypes.data, cnt.ctypes.data)
        keys = ("e_lost_powered", "e_lost_grounded", "i_lost_powered", "i_lost_grounded",
                "e_collisions", "i_collisions", "e_meas_steps", "i_meas_steps")
        return {"cumul_e": ce, "cumul_i": ci, "eepf": eepf, "ifed": ifed.reshape(2, PIC_N_IFED),
                "counters": dict(zip(keys, (int(v) for v in cnt)))}

    def reset_measurements(self):
        call("pic_sim_reset_measurements", self.h)

    def report(self, steps=None):
        """eduPIC's run report (C/eduPIC.cc:926-978) from this simulation's
        accumulators: the centre electron density averaged over the measured
        steps, the centre mean energy from the EEPF counts at their bin
        centres, and collision frequencies over `steps` steps (default: every
        step run; pass the steps since reset_measurements() after a reset,
        which also clears the collision counters).  A dict as
        kernels.report returns; the collision maxima use the tables of the
        discharge's gas density."""
        from . import kernels
        m = self.measurements()
        c = m["counters"]
        n_meas = c["e_meas_steps"]
        density = m["cumul_e"][self.p.n_g // 2] / n_meas if n_meas else 0.0
        h = m["eepf"].astype(np.float64)
        # an empty EEPF window: NaN, as eduPIC's 0/0 (which passes its conditions)
        mean_energy = float(h @ ((np.arange(PIC_N_EEPF) + 0.5) * 0.05) / h.sum()) if h.sum() else float("nan")
        if steps is None:
            steps = self.time()[1]
        n_e, n_i = self.counts()
        return kernels.report(self.p, density, mean_energy, steps * self.p.dt_e, c["e_collisions"],
                              c["i_collisions"], n_e, n_i)

    # eduPIC's XT file names (C/eduPIC.cc:905-919), in pic_sim_read_xt's order
    XT_NAMES = ("pot", "efield", "ne", "ni", "je", "ji", "powere", "poweri", "meanee", "meanei", "ioniz")

    def set_xt(self, bin_steps=20):
        """Measure the spatio-temporal maps in time bins of bin_steps steps
        (0 switches them off)."""
        call("pic_sim_set_xt", self.h, int(bin_steps))

    def xt(self):
        """The normalised XT maps: {name: array[n_g, n_t / bin_steps]}."""
        n = C.c_uint32()
        call("pic_sim_read_xt", self.h, None, C.byref(n))
        out = np.zeros((len(self.XT_NAMES), self.p.n_g, n.value), np.float64)
        call("pic_sim_read_xt", self.h, out.ctypes.data, C.byref(n))
        return dict(zip(self.XT_NAMES, out))

    def save(self, path):
        call("pic_sim_save", self.h, str(path).encode())

    def load(self, path):
        call("pic_sim_load", self.h, str(path).encode())

    def close(self):
        if getattr(self, "h", None):
            lib().pic_sim_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


ELECTRONS, IONS = PIC_ELECTRONS, PIC_IONS
