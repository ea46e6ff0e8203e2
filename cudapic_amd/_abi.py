"""ctypes declaration of the C ABI in include/cudapic_amd.h.

The shared library is the product: if it is missing this module raises
instead of falling back to anything else.  Build it with
``python -m cudapic_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcudapic_amd.so")

PIC_OK, PIC_ERR_HIP, PIC_ERR_ARG, PIC_ERR_CAPACITY, PIC_ERR_UNSUPPORTED, PIC_ERR_IO = range(6)
PIC_PREC_FP32, PIC_PREC_COMP, PIC_PREC_FP64 = 0, 1, 2
PIC_ELECTRONS, PIC_IONS = 0, 1
PIC_PATH_REFERENCE, PIC_PATH_FUSED = 0, 1
PIC_CS_RANGES = 1_000_000
PIC_N_CS = 5
PIC_N_EEPF = 2000
PIC_N_IFED = 200
PIC_MAX_HARMONICS = 16


class PicParams(C.Structure):
    _fields_ = (
        [(n, C.c_int32) for n in ("n_g", "n_t", "n_sub", "precision")]
        + [(n, C.c_double) for n in ("freq", "v0", "gap", "pressure", "t_gas", "weight", "area",
                                      "dt_e", "dt_i", "dx", "inv_dx", "gas_density")]
        + [(n, C.c_float) for n in (
            "inv_dx_f", "factor_e", "factor_i", "weight_factor", "dt_e_f", "dt_i_f", "gap_f",
            "normal_range", "f1", "f2", "half_me_over_e", "half_mu_over_e", "half_mi_over_e",
            "e_charge_f", "alpha_f", "half_dx_over_eps_f", "half_inv_dx_f",
            "min_x_f", "max_x_f", "de_eepf_f", "de_cs_f", "me_f", "exc_th_f", "ion_th_f",
            "eepf_kick_f")]
        + [("wave_cos", C.c_double * PIC_MAX_HARMONICS), ("wave_sin", C.c_double * PIC_MAX_HARMONICS)]
    )


class PicSpecies(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("x", "vx", "vy", "vz", "rng0", "rng1", "rng2", "rng3")] + [
        ("capacity", C.c_uint32)]


class PicSimConfig(C.Structure):
    _fields_ = [
        ("params", PicParams),
        ("capacity", C.c_uint32), ("segments", C.c_uint32),
        ("seed_e", C.c_uint32 * 4), ("seed_i", C.c_uint32 * 4),
        *[(n, C.c_int32) for n in ("path", "use_graph", "block_size", "grid_blocks", "ppt",
                                    "replicas", "measure", "ppt_inplace", "lds_replicas")],
        ("compact_every", C.c_uint32), ("rng_defer", C.c_uint32), ("cand_masks", C.c_int32),
        ("cand_threshold", C.c_float),
    ]


class PicEnsembleConfig(C.Structure):
    _fields_ = [
        ("n_instances", C.c_int32),
        ("params", C.POINTER(PicParams)),
        ("capacity", C.POINTER(C.c_uint32)),
        ("seed_e", C.c_uint32 * 4), ("seed_i", C.c_uint32 * 4),
        *[(n, C.c_int32) for n in ("use_graph", "grid_blocks", "ppt", "replicas", "lds_replicas")],
        ("compact_every", C.c_uint32),
        ("group_slots", C.c_uint32), ("group_periods", C.c_uint32), ("group_streams", C.c_uint32),
        ("block_size", C.c_int32), ("rng_defer", C.c_uint32), ("cand_masks", C.c_int32),
    ]


class PicAdjConfig(C.Structure):
    _fields_ = [
        ("params", PicParams),
        *[(n, C.c_uint32) for n in ("capacity_e", "capacity_i", "steps", "checkpoint", "t0", "tape_capacity")],
        ("seed_e", C.c_uint32 * 4), ("seed_i", C.c_uint32 * 4),
    ]


class PicState(C.Structure):
    """pic_state_t: the reference's GpuSimState buffers (device pointers)."""
    _fields_ = (
        [(n, C.c_void_p) for n in ("e_x", "e_vx", "e_vy", "e_vz", "i_x", "i_vx", "i_vy", "i_vz",
                                    "n_electrons", "n_ions", "efield", "pot", "e_density", "i_density",
                                    "cs", "sigma_tot_e", "sigma_tot_i",
                                    "rng_e0", "rng_e1", "rng_e2", "rng_e3", "rng_i0", "rng_i1", "rng_i2", "rng_i3",
                                    "tmp_x", "tmp_vx", "tmp_vy", "tmp_vz", "tmp_i_x", "tmp_i_vx", "tmp_i_vy",
                                    "tmp_i_vz", "cumul_e_density", "cumul_i_density", "eepf_counts")]
        + [("capacity", C.c_uint32), ("pad", C.c_uint32)]
    )


class PicStepConfig(C.Structure):
    _fields_ = [("compact_every", C.c_uint32),
                *[(n, C.c_int32) for n in ("grid_blocks", "ppt", "ppt_inplace", "lds_replicas")],
                ("rng_defer", C.c_uint32)]


PIC_STEP_MEASURE, PIC_STEP_RESET = 1, 2

P = C.c_void_p
U32 = C.c_uint32
U64 = C.c_uint64
I32 = C.c_int
F32 = C.c_float
F64 = C.c_double
PP = C.POINTER(PicParams)
PS = C.POINTER(PicSpecies)


class PicReport(C.Structure):
    _fields_ = (
        [(n, C.c_double) for n in ("density_center", "plasma_freq", "debye_length", "coll_freq_e", "coll_freq_i",
                                    "wp_dt_e", "dx_over_debye", "nu_max_e_dt_e", "nu_max_i_dt_i", "cfl_energy")]
        + [("conditions_ok", C.c_int32), ("pad", C.c_int32)]
    )

# name -> (restype, argtypes); every function in include/cudapic_amd.h
SIGNATURES = {
    "pic_params_default": (None, [PP]),
    "pic_params_derive": (I32, [PP]),
    "pic_cross_sections": (I32, [F64, P, P, P]),
    "pic_report": (I32, [PP, P, P, F64, F64, F64, U64, U64, U32, U32, C.POINTER(PicReport)]),
    "pic_get_density": (I32, [P, P, P, I32, PP, P]),
    "pic_solve_poisson": (I32, [P, P, F64, I32, P, P, PP, P]),
    "pic_solve_poisson_flexible": (I32, [P, I32, F32, F32, F32, I32, P, P]),
    "pic_move_and_compact": (I32, [P, P, P, P, P, P, P, P, P, P, P, F32, F32, F32, P, P, PP, P]),
    "pic_check_collisions_e": (I32, [P, P, P, PS, PS, P, P, P, P, PP, P]),
    "pic_check_collisions_i": (I32, [P, P, P, PS, P, PP, P]),
    "pic_accumulate_density": (I32, [P, I32, P, I32, P]),
    "pic_accumulate_eepf": (I32, [P, P, P, P, P, P, P, PP, P]),
    "pic_xoshiro_next_n": (I32, [C.POINTER(U32), U32, P, P, P]),
    "pic_seed_streams": (I32, [C.POINTER(U32), U32, P, P, P, P, P]),
    "pic_sim_config_default": (None, [C.POINTER(PicSimConfig)]),
    "pic_sim_create": (I32, [C.POINTER(PicSimConfig), C.POINTER(P)]),
    "pic_sim_destroy": (I32, [P]),
    "pic_sim_upload": (I32, [P, I32, P, P, P, P, U32]),
    "pic_sim_download": (I32, [P, I32, P, P, P, P, U32, C.POINTER(U32)]),
    "pic_sim_segments": (I32, [P, I32, C.POINTER(U32), C.POINTER(U32), P]),
    "pic_sim_download_slots": (I32, [P, I32, P, U32, C.POINTER(U32)]),
    "pic_sim_rng_upload": (I32, [P, I32, P, P, P, P, U32]),
    "pic_sim_rng_download": (I32, [P, I32, P, P, P, P, U32]),
    "pic_sim_set_tables": (I32, [P, P, P, P]),
    "pic_sim_cand_threshold": (I32, [P, P, P]),
    "pic_sim_counts": (I32, [P, C.POINTER(U32), C.POINTER(U32)]),
    "pic_sim_time": (I32, [P, C.POINTER(U32), C.POINTER(C.c_ulonglong)]),
    "pic_sim_set_time": (I32, [P, U32]),
    "pic_sim_run": (I32, [P, U32, I32]),
    "pic_sim_synchronize": (I32, [P]),
    "pic_sim_check": (I32, [P]),
    "pic_sim_read_grid": (I32, [P, P, P, P, P]),
    "pic_sim_read_measurements": (I32, [P, P, P, P, P, P]),
    "pic_sim_reset_measurements": (I32, [P]),
    "pic_sim_set_xt": (I32, [P, U32]),
    "pic_sim_read_xt": (I32, [P, P, C.POINTER(U32)]),
    "pic_sim_time_steps": (I32, [P, U32, I32, C.POINTER(F32), C.POINTER(F32), C.POINTER(U32),
                                 C.POINTER(C.c_ulonglong)]),
    "pic_sim_save": (I32, [P, C.c_char_p]),
    "pic_sim_load": (I32, [P, C.c_char_p]),
    "pic_ens_create": (I32, [C.POINTER(PicEnsembleConfig), C.POINTER(P)]),
    "pic_ens_destroy": (I32, [P]),
    "pic_ens_upload": (I32, [P, I32, I32, P, P, P, P, U32]),
    "pic_ens_download": (I32, [P, I32, I32, P, P, P, P, U32, C.POINTER(U32)]),
    "pic_ens_download_slots": (I32, [P, I32, I32, P, U32, C.POINTER(U32)]),
    "pic_ens_set_time": (I32, [P, U32]),
    "pic_ens_counts": (I32, [P, P, P]),
    "pic_ens_rng_download": (I32, [P, I32, I32, P, P, P, P, U32]),
    "pic_ens_run": (I32, [P, U32, I32]),
    "pic_ens_synchronize": (I32, [P]),
    "pic_ens_check": (I32, [P]),
    "pic_ens_time_steps": (I32, [P, U32, I32, C.POINTER(F32), C.POINTER(C.c_ulonglong)]),
    "pic_ens_read_grid": (I32, [P, I32, P, P]),
    "pic_ens_read_measurements": (I32, [P, I32, P, P, P]),
    "pic_adj_create": (I32, [C.POINTER(PicAdjConfig), C.POINTER(P)]),
    "pic_adj_destroy": (I32, [P]),
    "pic_adj_upload": (I32, [P, I32, P, P, P, P, U32]),
    "pic_adj_set_target": (I32, [P, P]),
    "pic_adj_record": (I32, [P, C.POINTER(C.c_double)]),
    "pic_adj_forward": (I32, [P, P, P, C.POINTER(C.c_double), P]),
    "pic_adj_backward": (I32, [P, P, P, P]),
    "pic_adj_tape_sizes": (I32, [P, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32), C.POINTER(U32),
                                 C.POINTER(U32)]),
    "pic_adj_download_tape": (I32, [P] + [P] * 16),
    "pic_init": (I32, [PP, C.POINTER(PicStepConfig), C.POINTER(P)]),
    "pic_destroy": (I32, [P]),
    "pic_step": (I32, [P, C.POINTER(PicState), U32, I32, P]),
    "pic_step_graph": (I32, [P, C.POINTER(PicState), U32, U32, I32, P]),
    "pic_step_diagnostics": (I32, [P, P, P, P, P]),
    "pic_step_noalloc": (I32, [PP, C.POINTER(PicState), U32, I32, P]),
    "pic_step_lagged": (I32, [P, C.POINTER(PicState), U32, I32, U32, U32, P]),
    "pic_malloc": (I32, [C.POINTER(P), U64]),
    "pic_free": (I32, [P]),
    "pic_memcpy_htod": (I32, [P, P, U64]),
    "pic_memcpy_dtoh": (I32, [P, P, U64]),
    "pic_memset": (I32, [P, I32, U64]),
    "pic_device_synchronize": (I32, []),
    "pic_stream_create": (I32, [C.POINTER(P)]),
    "pic_stream_destroy": (I32, [P]),
    "pic_stream_synchronize": (I32, [P]),
    "pic_stream_copy": (I32, [P, P, U64, P]),
    "pic_stream_read": (I32, [P, U64, P, P]),
    "pic_memcpy_dtod_async": (I32, [P, P, U64, P]),
    "pic_last_error": (C.c_char_p, []),
    "pic_device_info": (I32, [C.POINTER(I32), C.POINTER(I32), C.c_char_p, I32]),
    "pic_abi_layout": (I32, [C.POINTER(U64)]),
}


class PicError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


_LIB = None


def lib():
    """Load the engine library (once).  Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: the HIP engine must be built first "
                "(python -m cudapic_amd.build)")
        lib_ = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib_, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib_
    return _LIB


def check(rc: int, where: str) -> None:
    if rc != PIC_OK:
        msg = lib().pic_last_error()
        raise PicError(rc, where, msg.decode() if msg else "")


def call(name: str, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)
    return rc


def layout_matches() -> bool:
    """True if the ctypes structures have the C sizes the library reports."""
    out = (U64 * 6)()
    lib().pic_abi_layout(out)
    return tuple(out) == (C.sizeof(PicParams), C.sizeof(PicSpecies), C.sizeof(PicSimConfig),
                          C.sizeof(PicEnsembleConfig), C.sizeof(PicState), C.sizeof(PicStepConfig))
