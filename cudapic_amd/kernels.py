"""Stage entry points, named after the kernels of the reference's
``#[cuda_module] mod kernels`` (Rust-CUDA/src/main.rs:352-1358).

Each call enqueues HIP work through the C ABI (include/cudapic_amd.h) on a
stream (default: the null stream) and returns without synchronising.
Buffers are ``DeviceBuffer`` objects or raw device addresses (int).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import PIC_CS_RANGES, PIC_N_CS, PicSpecies, call
from .device import DeviceBuffer


def _p(b):
    if b is None:
        return None
    return b.ptr if isinstance(b, DeviceBuffer) else int(b)


class Species:
    """Device SoA of one species plus its per-slot RNG words
    (e_x/e_vx/e_vy/e_vz + rng_e0..3, main.rs:134-170)."""

    def __init__(self, capacity: int):
        self.capacity = int(capacity)
        self.x, self.vx, self.vy, self.vz = (DeviceBuffer.zeroed(capacity, np.float32) for _ in range(4))
        self.rng = [DeviceBuffer.zeroed(capacity, np.uint32) for _ in range(4)]

    def abi(self) -> PicSpecies:
        s = PicSpecies()
        s.x, s.vx, s.vy, s.vz = (b.ptr for b in (self.x, self.vx, self.vy, self.vz))
        s.rng0, s.rng1, s.rng2, s.rng3 = (b.ptr for b in self.rng)
        s.capacity = self.capacity
        return s

    def upload(self, x, vx, vy, vz, rng=None):
        for b, a in zip((self.x, self.vx, self.vy, self.vz), (x, vx, vy, vz)):
            b.copy_from_host(np.asarray(a, np.float32))
        if rng is not None:
            for b, a in zip(self.rng, rng):
                b.copy_from_host(np.asarray(a, np.uint32))

    def download(self, n: int):
        return [b.to_host_vec(n) for b in (self.x, self.vx, self.vy, self.vz)]

    def rng_host(self, n: int):
        return [b.to_host_vec(n) for b in self.rng]


def cross_sections(gas_density: float):
    """Host tables (main.rs:1378-1444): cs[5*1e6], sigma_tot_e, sigma_tot_i."""
    cs = np.zeros(PIC_N_CS * PIC_CS_RANGES, np.float32)
    se = np.zeros(PIC_CS_RANGES, np.float32)
    si = np.zeros(PIC_CS_RANGES, np.float32)
    call("pic_cross_sections", gas_density, cs.ctypes.data, se.ctypes.data, si.ctypes.data)
    return cs, se, si


def report(p, density_center, mean_energy_center, sim_time, n_e_coll, n_i_coll, n_e, n_i, tables=None):
    """eduPIC's run report (C/eduPIC.cc:926-978): plasma characteristics at
    the centre and the stability and accuracy conditions, as a dict.
    `tables` = (sigma_tot_e, sigma_tot_i) of cross_sections(), computed if
    omitted.  Host only."""
    from ._abi import PicReport
    if tables is None:
        _, se, si = cross_sections(p.gas_density)
    else:
        se, si = (np.ascontiguousarray(t, np.float32) for t in tables)
    r = PicReport()
    call("pic_report", C.byref(p), se.ctypes.data, si.ctypes.data, float(density_center),
         float(mean_energy_center), float(sim_time), int(n_e_coll), int(n_i_coll), int(n_e), int(n_i),
         C.byref(r))
    out = {n: getattr(r, n) for n, _ in PicReport._fields_ if n != "pad"}
    out["conditions_ok"] = bool(out["conditions_ok"])
    return out


def get_density(x, density, n_active, p, density_f64=False, stream=None):
    call("pic_get_density", _p(x), _p(n_active), _p(density), int(density_f64), C.byref(p), stream)


def solve_poisson_scan(e_density, i_density, pot, efield, pot0, p, precision=None, stream=None):
    prec = p.precision if precision is None else precision
    call("pic_solve_poisson", _p(e_density), _p(i_density), float(pot0), int(prec), _p(pot),
         _p(efield), C.byref(p), stream)


def solve_poisson_dps_flexible(source, pot, n, psi_left, psi_right, dx, precision=0, stream=None):
    call("pic_solve_poisson_flexible", _p(source), int(n), psi_left, psi_right, dx, int(precision),
         _p(pot), stream)


def move_and_compact(efield, src, dst, alive_counter, n_active, factor, dt, p, half_m_over_e=0.0,
                     abs_counts=None, ifed=None, stream=None):
    call("pic_move_and_compact", _p(efield), *(_p(b) for b in src), *(_p(b) for b in dst),
         _p(alive_counter), _p(n_active), factor, dt, half_m_over_e, _p(abs_counts), _p(ifed),
         C.byref(p), stream)


def check_collisions_e(total_cs_e, cs, active_e, electrons: Species, ions: Species, alive_e, alive_i,
                       overflow, n_coll, p, stream=None):
    e, i = electrons.abi(), ions.abi()
    call("pic_check_collisions_e", _p(total_cs_e), _p(cs), _p(active_e), C.byref(e), C.byref(i),
         _p(alive_e), _p(alive_i), _p(overflow), _p(n_coll), C.byref(p), stream)


def check_collisions_i(total_cs_i, cs, active_i, ions: Species, n_coll, p, stream=None):
    i = ions.abi()
    call("pic_check_collisions_i", _p(total_cs_i), _p(cs), _p(active_i), C.byref(i), _p(n_coll),
         C.byref(p), stream)


def accumulate_density(density, cumul, n_g, density_f64=False, stream=None):
    call("pic_accumulate_density", _p(density), int(density_f64), _p(cumul), int(n_g), stream)


def accumulate_eepf(x, vx, vy, vz, efield, active_e, eepf, p, stream=None):
    call("pic_accumulate_eepf", _p(x), _p(vx), _p(vy), _p(vz), _p(efield), _p(active_e), _p(eepf),
         C.byref(p), stream)


def xoshiro128p_next_n(state, n, out, state_out=None, stream=None):
    st = (C.c_uint32 * 4)(*[int(v) for v in state])
    call("pic_xoshiro_next_n", st, int(n), _p(out), _p(state_out), stream)


def seed_streams(master, n, rng, stream=None):
    m = (C.c_uint32 * 4)(*[int(v) for v in master])
    call("pic_seed_streams", m, int(n), *(_p(b) for b in rng), stream)
