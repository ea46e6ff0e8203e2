"""Ensembles of independent discharges (pic_ens_* of the C ABI): every
stage of every instance in one launch, the 4*N_SUB-step pattern replayed
from a hipGraph."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import PicEnsembleConfig, PicParams, call, lib


class Ensemble:
    def __init__(self, params_list, capacities, seed_e=(0x12345678, 0x11112222, 0x22223333, 0x33334444),
                 seed_i=(0x44445555, 0x55556666, 0x66667777, 0x77778888), use_graph=True, grid_blocks=0,
                 ppt=0, replicas=0, lds_replicas=0, compact_every=0, group_slots=0, group_periods=0,
                 group_streams=0, block_size=0, rng_defer=0, cand_masks=0):
        n = len(params_list)
        self._params = (PicParams * n)(*params_list)
        self._caps = (C.c_uint32 * n)(*[int(c) for c in capacities])
        cfg = PicEnsembleConfig()
        cfg.n_instances = n
        cfg.params = self._params
        cfg.capacity = self._caps
        cfg.seed_e[:] = list(seed_e)
        cfg.seed_i[:] = list(seed_i)
        cfg.use_graph, cfg.grid_blocks, cfg.ppt, cfg.replicas = int(use_graph), grid_blocks, ppt, replicas
        cfg.lds_replicas = lds_replicas
        cfg.compact_every = int(compact_every)
        cfg.group_slots = int(group_slots)
        cfg.group_periods = int(group_periods)
        cfg.group_streams = int(group_streams)
        cfg.block_size = int(block_size)
        cfg.rng_defer = int(rng_defer)
        cfg.cand_masks = int(cand_masks)
        self.n = n
        self.n_g = params_list[0].n_g
        h = C.c_void_p()
        call("pic_ens_create", C.byref(cfg), C.byref(h))
        self.h = h.value

    def upload(self, inst, species, x, vx, vy, vz):
        a = [np.ascontiguousarray(v, np.float32) for v in (x, vx, vy, vz)]
        call("pic_ens_upload", self.h, inst, species, *(v.ctypes.data for v in a), len(a[0]))

    def download(self, inst, species):
        n = C.c_uint32()
        call("pic_ens_download", self.h, inst, species, None, None, None, None, 0, C.byref(n))
        out = [np.zeros(max(1, n.value), np.float32) for _ in range(4)]
        call("pic_ens_download", self.h, inst, species, *(a.ctypes.data for a in out), n.value, C.byref(n))
        return [a[:n.value] for a in out]

    def slots(self, inst, species):
        """Slot (within the instance's region) of each particle download() returns."""
        n = C.c_uint32()
        call("pic_ens_download_slots", self.h, inst, species, None, 0, C.byref(n))
        out = np.zeros(max(1, n.value), np.uint32)
        call("pic_ens_download_slots", self.h, inst, species, out.ctypes.data, n.value, C.byref(n))
        return out[:n.value].astype(np.int64)

    def set_time(self, t):
        call("pic_ens_set_time", self.h, int(t))

    def rng_download(self, inst, species, n):
        out = [np.zeros(max(1, n), np.uint32) for _ in range(4)]
        call("pic_ens_rng_download", self.h, inst, species, *(a.ctypes.data for a in out), int(n))
        return [a[:n] for a in out]

    def counts(self):
        ne = np.zeros(self.n, np.uint32)
        ni = np.zeros(self.n, np.uint32)
        call("pic_ens_counts", self.h, ne.ctypes.data, ni.ctypes.data)
        return ne, ni

    def run(self, n_steps, measure=False):
        call("pic_ens_run", self.h, int(n_steps), int(bool(measure)))

    def synchronize(self):
        call("pic_ens_synchronize", self.h)

    def check(self):
        call("pic_ens_check", self.h)

    def time_steps(self, n_steps, measure=False):
        ms, ps = C.c_float(), C.c_ulonglong()
        call("pic_ens_time_steps", self.h, int(n_steps), int(bool(measure)), C.byref(ms), C.byref(ps))
        return {"ms_total": ms.value, "particle_steps": ps.value}

    def grid(self, inst):
        pot = np.zeros(self.n_g, np.float32)
        ef = np.zeros(self.n_g, np.float32)
        call("pic_ens_read_grid", self.h, inst, pot.ctypes.data, ef.ctypes.data)
        return pot, ef

    def measurements(self, inst):
        ce = np.zeros(self.n_g, np.float64)
        ci = np.zeros(self.n_g, np.float64)
        cnt = np.zeros(8, np.uint64)
        call("pic_ens_read_measurements", self.h, inst, ce.ctypes.data, ci.ctypes.data, cnt.ctypes.data)
        return ce, ci, cnt

    def close(self):
        if getattr(self, "h", None):
            lib().pic_ens_destroy(self.h)
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
