"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  The product
library (``cudapic_amd``) never imports it.

The C sources are ``oracle/pic_oracle.c`` (restatement of sekulas/cudaPIC
``Rust-CUDA/src/main.rs``) and the reference's own CPU code compiled into
``oracle/_ref`` by ``oracle/Makefile``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PIC_ORACLE_LIB selects another build of the same sources (the sanitizer
# build of `make -C oracle sanitize`, for tests/test_sanitizers.py)
LIB_PATH = os.environ.get("PIC_ORACLE_LIB") or os.path.join(HERE, "libpic_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")
CS_RANGES = 1_000_000
N_EEPF = 2000
N_IFED = 200

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


class OrParams(C.Structure):
    _fields_ = (
        [("n_g", C.c_int32), ("n_t", C.c_int32), ("n_sub", C.c_int32)]
        + [(n, C.c_double) for n in ("freq", "v0", "gap", "pressure", "t_gas", "weight", "area",
                                      "dt_e", "dt_i", "dx", "inv_dx", "gas_density")]
        + [(n, C.c_float) for n in (
            "inv_dx_f", "factor_e", "factor_i", "weight_factor", "dt_e_f", "dt_i_f", "gap_f",
            "normal_range", "f1", "f2", "half_me_over_e", "half_mu_over_e",
            "e_charge_f", "alpha_f", "half_dx_over_eps_f", "half_inv_dx_f",
            "min_x_f", "max_x_f", "de_eepf_f", "de_cs_f", "me_f", "exc_th_f", "ion_th_f")]
    )


class OrState(C.Structure):
    _fields_ = [
        ("cap", C.c_uint32),
        *[(n, C.POINTER(C.c_float)) for n in ("ex", "evx", "evy", "evz", "ix", "ivx", "ivy", "ivz",
                                                "tx", "tvx", "tvy", "tvz")],
        ("er", C.POINTER(C.c_uint32) * 4),
        ("ir", C.POINTER(C.c_uint32) * 4),
        ("n_e", C.c_uint32), ("n_i", C.c_uint32),
        ("e_dens", C.POINTER(C.c_double)), ("i_dens", C.POINTER(C.c_double)),
        ("pot", C.POINTER(C.c_float)), ("efield", C.POINTER(C.c_float)),
        ("cs", C.POINTER(C.c_float)), ("stot_e", C.POINTER(C.c_float)),
        ("stot_i", C.POINTER(C.c_float)),
        ("abs_e", C.c_uint64 * 2), ("abs_i", C.c_uint64 * 2),
        ("coll_e", C.c_uint64), ("coll_i", C.c_uint64),
        ("ifed", C.POINTER(C.c_uint32)),
    ]


def build() -> None:
    """Compile the C restatement (and, where the reference is mounted, _ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "libpic_oracle.so"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER(OrParams)
    sig = {
        "or_params_derive": (None, [P, C.c_int, C.c_int, C.c_int] + [C.c_double] * 7),
        "or_xo_next": (C.c_uint32, [_u32p]),
        "or_xo_jump": (None, [_u32p]),
        "or_xo_skip_soa": (None, [_u32p, _u32p, _u32p, _u32p, C.c_uint32, _u32p]),
        "or_u32_to_real": (C.c_float, [C.c_uint32]),
        "or_seed_streams": (None, [_u32p, C.c_uint32, _u32p, _u32p, _u32p, _u32p]),
        "or_three_normal": (None, [_u32p, C.c_float, _f32p]),
        "or_cross_sections": (None, [C.c_double, _f32p, _f32p, _f32p]),
        "or_deposit": (None, [_f32p, C.c_uint32, P, _f64p]),
        "or_deposit_f64": (None, [_f64p, C.c_uint32, P, _f64p]),
        "or_poisson_thomas": (None, [_f64p, C.c_double, P, _f64p, _f64p]),
        "or_poisson_scan_f64": (None, [_f64p, _f64p, C.c_double, P, _f64p, _f64p]),
        "or_poisson_scan_f32": (None, [_f32p, _f32p, C.c_float, P, _f32p, _f32p]),
        "or_poisson_flexible_f64": (None, [_f64p, C.c_int, C.c_double, C.c_double, C.c_double, _f64p]),
        "or_move_and_compact": (C.c_uint32, [_f32p, _f32p, _f32p, _f32p, _f32p, C.c_uint32,
                                             C.c_float, C.c_float, C.c_float, P,
                                             _f32p, _f32p, _f32p, _f32p, _u64p, C.c_void_p]),
        "or_collide_e": (None, [_f32p, _f32p, C.c_uint32, _f32p, _f32p, _f32p, _f32p,
                                _u32p, _u32p, _u32p, _u32p, _f32p, _f32p, _f32p, _f32p,
                                _u32p, _u32p, C.c_uint32, C.c_uint32, P, _u64p]),
        "or_collide_e_one": (C.c_int, [_f32p, _f32p, _f32p, _u32p, _f32p, _f32p, P]),
        "or_collide_i": (None, [_f32p, _f32p, C.c_uint32, _f32p, _f32p, _f32p,
                                _u32p, _u32p, _u32p, _u32p, P, _u64p]),
        "or_collide_i_one": (C.c_int, [_f32p, _f32p, _f32p, _u32p, P]),
        "or_accumulate_eepf": (None, [_f32p, _f32p, _f32p, _f32p, _f32p, C.c_uint32, P, _u32p]),
        "or_step": (None, [C.POINTER(OrState), P, C.c_int, C.c_int]),
        "or_bin_division_mismatches": (C.c_uint64, [C.c_float, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def params(n_g=400, n_t=4000, n_sub=20, freq=13.56e6, v0=250.0, gap=0.025, pressure=10.0,
           t_gas=350.0, weight=7.0e4, area=1.0e-4) -> OrParams:
    p = OrParams()
    lib().or_params_derive(C.byref(p), n_g, n_t, n_sub, freq, v0, gap, pressure, t_gas,
                           weight, area)
    return p


# ---------------------------------------------------------------- RNG
def xo_next(state: np.ndarray) -> int:
    return int(lib().or_xo_next(state))


def xo_jump(state: np.ndarray) -> None:
    lib().or_xo_jump(state)


def xo_skip_soa(words, counts):
    """The SoA streams (four uint32 arrays) advanced by counts[j] draws each
    (new arrays)."""
    out = [np.ascontiguousarray(w, np.uint32).copy() for w in words]
    k = np.ascontiguousarray(counts, np.uint32)
    assert all(len(w) == len(k) for w in out)
    lib().or_xo_skip_soa(*out, len(k), k)
    return out


def u32_to_real(x: int) -> float:
    return float(lib().or_u32_to_real(x))


def seed_streams(master, n: int):
    m = np.asarray(master, dtype=np.uint32)
    out = [np.zeros(n, np.uint32) for _ in range(4)]
    lib().or_seed_streams(m, n, *out)
    return out


def three_normal(state: np.ndarray, sigma: float) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().or_three_normal(state, sigma, out)
    return out


_CS_CACHE: dict = {}


def cross_sections(gas_density: float):
    key = float(gas_density)
    if key not in _CS_CACHE:
        cs = np.zeros(5 * CS_RANGES, np.float32)
        se = np.zeros(CS_RANGES, np.float32)
        si = np.zeros(CS_RANGES, np.float32)
        lib().or_cross_sections(gas_density, cs, se, si)
        _CS_CACHE.clear()
        _CS_CACHE[key] = (cs, se, si)
    return _CS_CACHE[key]


# ---------------------------------------------------------------- grid
def deposit(x: np.ndarray, p: OrParams) -> np.ndarray:
    d = np.zeros(p.n_g, np.float64)
    x = np.ascontiguousarray(x, np.float32)
    lib().or_deposit(x, len(x), C.byref(p), d)
    return d


def deposit_f64(x: np.ndarray, p: OrParams) -> np.ndarray:
    d = np.zeros(p.n_g, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    lib().or_deposit_f64(x, len(x), C.byref(p), d)
    return d


def poisson_thomas(rho: np.ndarray, pot0: float, p: OrParams):
    pot = np.zeros(p.n_g, np.float64)
    ef = np.zeros(p.n_g, np.float64)
    lib().or_poisson_thomas(np.ascontiguousarray(rho, np.float64), pot0, C.byref(p), pot, ef)
    return pot, ef


def poisson_scan_f64(e_d, i_d, pot0, p):
    pot = np.zeros(p.n_g, np.float64)
    ef = np.zeros(p.n_g, np.float64)
    lib().or_poisson_scan_f64(np.ascontiguousarray(e_d, np.float64),
                              np.ascontiguousarray(i_d, np.float64), pot0, C.byref(p), pot, ef)
    return pot, ef


def poisson_scan_f32(e_d, i_d, pot0, p):
    pot = np.zeros(p.n_g, np.float32)
    ef = np.zeros(p.n_g, np.float32)
    lib().or_poisson_scan_f32(np.ascontiguousarray(e_d, np.float32),
                              np.ascontiguousarray(i_d, np.float32), pot0, C.byref(p), pot, ef)
    return pot, ef


def poisson_flexible_f64(src, psi_l, psi_r, dx):
    src = np.ascontiguousarray(src, np.float64)
    pot = np.zeros(len(src), np.float64)
    lib().or_poisson_flexible_f64(src, len(src), psi_l, psi_r, dx, pot)
    return pot


# ---------------------------------------------------------------- particles
def move_and_compact(efield, x, vx, vy, vz, factor, dt, p, mass_over_2e=0.0, ifed=None):
    n = len(x)
    out = [np.zeros(max(n, 1), np.float32) for _ in range(4)]
    absc = np.zeros(2, np.uint64)
    ifed_ptr = None if ifed is None else ifed.ctypes.data_as(C.c_void_p)
    cnt = lib().or_move_and_compact(
        np.ascontiguousarray(efield, np.float32), *(np.ascontiguousarray(a, np.float32) for a in (x, vx, vy, vz)),
        n, factor, dt, mass_over_2e, C.byref(p), *out, absc, ifed_ptr)
    return cnt, [o[:cnt] for o in out], absc


def collide_e(stot_e, cs, x, vx, vy, vz, rng, ix, ivx, ivy, ivz, n_active, n_ions, p):
    """In-place electron collisions over slots [0, n_active); returns
    (n_electrons_after, n_ions_after, n_collisions)."""
    alive_e = np.array([n_active], np.uint32)
    alive_i = np.array([n_ions], np.uint32)
    ncoll = np.zeros(1, np.uint64)
    lib().or_collide_e(stot_e, cs, n_active, x, vx, vy, vz, *rng, ix, ivx, ivy, ivz,
                       alive_e, alive_i, len(x), len(ix), C.byref(p), ncoll)
    return int(alive_e[0]), int(alive_i[0]), int(ncoll[0])


def collide_e_one(stot_e, cs, v, s, p):
    v = np.array(v, np.float32)
    s = np.array(s, np.uint32)
    se = np.zeros(3, np.float32)
    si = np.zeros(3, np.float32)
    kind = lib().or_collide_e_one(stot_e, cs, v, s, se, si, C.byref(p))
    return kind, v, s, se, si


def collide_i(stot_i, cs, vx, vy, vz, rng, n_active, p):
    ncoll = np.zeros(1, np.uint64)
    lib().or_collide_i(stot_i, cs, n_active, vx, vy, vz, *rng, C.byref(p), ncoll)
    return int(ncoll[0])


def collide_i_one(stot_i, cs, v, s, p):
    v = np.array(v, np.float32)
    s = np.array(s, np.uint32)
    kind = lib().or_collide_i_one(stot_i, cs, v, s, C.byref(p))
    return kind, v, s


def accumulate_eepf(x, vx, vy, vz, efield, p, eepf=None):
    if eepf is None:
        eepf = np.zeros(N_EEPF, np.uint32)
    lib().or_accumulate_eepf(*(np.ascontiguousarray(a, np.float32) for a in (x, vx, vy, vz, efield)),
                             len(x), C.byref(p), eepf)
    return eepf


class Sim:
    """Whole-step oracle (main.rs:1725-1808) over numpy-owned SoA arrays."""

    def __init__(self, p: OrParams, cap: int, ex, evx, evy, evz, ix, ivx, ivy, ivz,
                 er, ir, with_ifed=False):
        self.p = p
        self.cap = cap
        f = lambda a: np.zeros(cap, np.float32) if a is None else a
        self.arrs = {}
        for name, src in (("ex", ex), ("evx", evx), ("evy", evy), ("evz", evz),
                          ("ix", ix), ("ivx", ivx), ("ivy", ivy), ("ivz", ivz)):
            a = np.zeros(cap, np.float32)
            a[:len(src)] = src
            self.arrs[name] = a
        for name in ("tx", "tvx", "tvy", "tvz"):
            self.arrs[name] = np.zeros(cap, np.float32)
        self.er = [np.ascontiguousarray(r[:cap], np.uint32).copy() for r in er]
        self.ir = [np.ascontiguousarray(r[:cap], np.uint32).copy() for r in ir]
        self.e_dens = np.zeros(p.n_g, np.float64)
        self.i_dens = np.zeros(p.n_g, np.float64)
        self.pot = np.zeros(p.n_g, np.float32)
        self.efield = np.zeros(p.n_g, np.float32)
        self.cs, self.stot_e, self.stot_i = cross_sections(p.gas_density)
        self.ifed = np.zeros(2 * N_IFED, np.uint32) if with_ifed else None
        st = OrState()
        st.cap = cap
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        up = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
        for name, a in self.arrs.items():
            setattr(st, name, fp(a))
        for k in range(4):
            st.er[k] = up(self.er[k])
            st.ir[k] = up(self.ir[k])
        st.n_e = len(ex)
        st.n_i = len(ix)
        st.e_dens = self.e_dens.ctypes.data_as(C.POINTER(C.c_double))
        st.i_dens = self.i_dens.ctypes.data_as(C.POINTER(C.c_double))
        st.pot = fp(self.pot)
        st.efield = fp(self.efield)
        st.cs = fp(self.cs)
        st.stot_e = fp(self.stot_e)
        st.stot_i = fp(self.stot_i)
        if self.ifed is not None:
            st.ifed = up(self.ifed)
        self.st = st

    def step(self, t: int, collisions: bool = True) -> None:
        lib().or_step(C.byref(self.st), C.byref(self.p), t, int(collisions))

    @property
    def n_e(self):
        return self.st.n_e

    @property
    def n_i(self):
        return self.st.n_i

    def electrons(self):
        n = self.n_e
        return [self.arrs[k][:n].copy() for k in ("ex", "evx", "evy", "evz")]

    def ions(self):
        n = self.n_i
        return [self.arrs[k][:n].copy() for k in ("ix", "ivx", "ivy", "ivz")]


# ---------------------------------------------------------------- _ref
def ref_available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libedupic.so"))


def ref_edupic():
    """eduPIC compiled from the reference's C/eduPIC.cc with main renamed;
    exposes its cross-section tables and Thomas Poisson solver."""
    return C.CDLL(os.path.join(REF_DIR, "libedupic.so"))


def ref_xoshiro():
    """xoshiro128p.c compiled from the reference; ``next`` advances a static
    state that starts at {1,1,1,1}."""
    lib_ = C.CDLL(os.path.join(REF_DIR, "libxoshiro128p.so"))
    lib_.next.restype = C.c_uint32
    return lib_
