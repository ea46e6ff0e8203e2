"""CPU checks of the C ABI: the library loads, exports every function that
include/cudapic_amd.h declares, its struct layout matches the binding, and
its host-side tables and derived parameters equal the oracle's.  No call
here touches a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cudapic_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(pic_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_stage_entry_points():
    names = declared_functions()
    for n in ("pic_get_density", "pic_solve_poisson", "pic_solve_poisson_flexible",
              "pic_move_and_compact", "pic_check_collisions_e", "pic_check_collisions_i",
              "pic_accumulate_density", "pic_accumulate_eepf", "pic_xoshiro_next_n",
              "pic_seed_streams", "pic_sim_create", "pic_sim_run"):
        assert n in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    from cudapic_amd import _abi
    lib = C.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding declares a signature for each of them
    assert set(declared_functions()) <= set(_abi.SIGNATURES), \
        set(declared_functions()) - set(_abi.SIGNATURES)


def test_struct_layout_matches_binding():
    from cudapic_amd import _abi
    assert _abi.layout_matches()


def test_params_derive_matches_oracle(orc):
    import cudapic_amd as ca
    for kw in ({}, {"weight": 2.0e3}, {"n_g": 4000, "weight": 2.0e3}, {"v0": 137.0, "pressure": 3.0}):
        p = ca.params(**kw)
        q = orc.params(**kw)
        for f, _ in orc.OrParams._fields_:
            if f in ("n_g", "n_t", "n_sub"):
                assert getattr(p, f) == getattr(q, f)
            else:
                assert getattr(p, f) == getattr(q, f), (kw, f)
        assert p.wave_cos[0] == p.v0 and all(p.wave_cos[h] == 0 for h in range(1, 16))


def test_params_derive_rejects_bad_input():
    import cudapic_amd as ca
    from cudapic_amd import _abi
    p = _abi.PicParams()
    _abi.lib().pic_params_default(C.byref(p))
    p.n_g = 2
    assert _abi.lib().pic_params_derive(C.byref(p)) == _abi.PIC_ERR_ARG
    p.n_g = 5000
    assert _abi.lib().pic_params_derive(C.byref(p)) == _abi.PIC_ERR_ARG
    with pytest.raises(ca.PicError):
        ca.params(pressure=-1.0)


def test_host_cross_sections_equal_oracle(orc):
    import cudapic_amd as ca
    for pressure in (10.0, 3.0):
        p = ca.params(pressure=pressure)
        cs, se, si = ca.kernels.cross_sections(p.gas_density)
        c2, s2, i2 = orc.cross_sections(p.gas_density)
        np.testing.assert_array_equal(cs, c2)
        np.testing.assert_array_equal(se, s2)
        np.testing.assert_array_equal(si, i2)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from cudapic_amd import _abi
    monkeypatch.setattr(_abi, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_abi, "_LIB", None)
    with pytest.raises(RuntimeError, match="must be built"):
        _abi.lib()


def test_step_entry_points_reject_bad_arguments_before_the_gpu():
    """pic_init validates the discharge (n_t a multiple of 2*n_sub, n_g in
    range) and pic_step / pic_step_graph a null context before any HIP call."""
    import cudapic_amd as ca
    from cudapic_amd import _abi
    lib = _abi.lib()
    h = C.c_void_p()
    p = ca.params()
    p.n_t = 4010                      # not a multiple of 2 * n_sub = 40
    assert lib.pic_init(C.byref(p), None, C.byref(h)) == _abi.PIC_ERR_ARG
    p = ca.params()
    p.n_g = 5000
    assert lib.pic_init(C.byref(p), None, C.byref(h)) == _abi.PIC_ERR_ARG
    cfg = _abi.PicStepConfig(compact_every=100)   # not a multiple of 4 * n_sub
    assert lib.pic_init(C.byref(ca.params()), C.byref(cfg), C.byref(h)) == _abi.PIC_ERR_ARG
    st = _abi.PicState()
    assert lib.pic_step(None, C.byref(st), 0, 0, None) == _abi.PIC_ERR_ARG
    assert lib.pic_step_graph(None, C.byref(st), 0, 80, 0, None) == _abi.PIC_ERR_ARG
    assert b"null" in lib.pic_last_error()


def test_unsupported_workgroup_sizes_are_rejected_before_the_gpu():
    """block_size is honoured by the fused particle kernels (128, 256, 512)
    and rejected otherwise -- never silently ignored."""
    import cudapic_amd as ca
    from cudapic_amd import _abi
    lib = _abi.lib()
    h = C.c_void_p()
    for path, block in ((1, 300), (1, 1024), (0, 512)):
        cfg = _abi.PicSimConfig()
        lib.pic_sim_config_default(C.byref(cfg))
        cfg.path, cfg.block_size = path, block
        assert lib.pic_sim_create(C.byref(cfg), C.byref(h)) == _abi.PIC_ERR_ARG, (path, block)
        assert b"block_size" in lib.pic_last_error()
    p = (_abi.PicParams * 1)(ca.params())
    caps = (C.c_uint32 * 1)(1000)
    ec = _abi.PicEnsembleConfig(n_instances=1, params=p, capacity=caps, block_size=192)
    assert lib.pic_ens_create(C.byref(ec), C.byref(h)) == _abi.PIC_ERR_ARG
    assert b"block_size" in lib.pic_last_error()


def test_capacities_beyond_the_fused_kernels_offsets_are_rejected_before_the_gpu():
    """The fused kernels address particle arrays by 32-bit byte offsets, so a
    species holds at most 2^30 slots: pic_sim_create (2^30 - 4096 before
    segment rounding), the caller-owned step and the ensemble's total all
    refuse more before any HIP call."""
    import cudapic_amd as ca
    from cudapic_amd import _abi
    lib = _abi.lib()
    h = C.c_void_p()
    cfg = _abi.PicSimConfig()
    lib.pic_sim_config_default(C.byref(cfg))
    cfg.capacity = (1 << 30) - 4095
    assert lib.pic_sim_create(C.byref(cfg), C.byref(h)) == _abi.PIC_ERR_ARG
    assert b"2^30" in lib.pic_last_error()
    st = _abi.PicState()
    for name, _ in st._fields_:
        if name not in ("e_density", "i_density", "cumul_e_density", "cumul_i_density", "eepf_counts",
                        "capacity", "pad"):
            setattr(st, name, 0x1000)  # never dereferenced: the check comes first
    st.capacity = (1 << 30) + 1
    p = ca.params()
    assert lib.pic_step_noalloc(C.byref(p), C.byref(st), 0, 0, None) == _abi.PIC_ERR_ARG
    assert b"2^30" in lib.pic_last_error()
    pp = (_abi.PicParams * 2)(ca.params(), ca.params())
    caps = (C.c_uint32 * 2)(1 << 29, (1 << 29) + 1)
    ec = _abi.PicEnsembleConfig(n_instances=2, params=pp, capacity=caps)
    assert lib.pic_ens_create(C.byref(ec), C.byref(h)) == _abi.PIC_ERR_ARG
    assert b"2^30" in lib.pic_last_error()
