"""pic_step_noalloc (the caller-owned RF step on the library's fixed device
workspace, for callers that cannot allocate) against pic_step (the same step
on a pic_init workspace), through the C ABI: two calls in a row, the second
on a smaller discharge, each a collisionless FP64 step whose particles must
come out as the same sorted multiset (bit-exact) with the same counts,
potential and field within 1e-5 of the maximum, and equal measurement
accumulators; the second call must not see the first one's state."""
import numpy as np
import pytest

from helpers import assert_same_multiset, synthetic_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


def _zero_tables(orc, po):
    cs, se, si = orc.cross_sections(po.gas_density)
    return cs, np.zeros_like(se), np.zeros_like(si)


def _same_step(a, b):
    for sp in (0, 1):
        assert a.count(sp) == b.count(sp), sp
        assert_same_multiset(a.download(sp)[0], b.download(sp)[0])
    ga, gb = a.grid(), b.grid()
    for x, y in zip(ga, gb):
        assert np.abs(x - y).max() <= 1e-5 * np.abs(x).max()
    ma, mb = a.measurements(), b.measurements()
    np.testing.assert_array_equal(ma["eepf"], mb["eepf"])
    for key in ("cumul_e", "cumul_i"):
        np.testing.assert_allclose(ma[key], mb[key], rtol=1e-12, atol=0)


def test_noalloc_step_equals_pic_step(ca, orc):
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    tables = _zero_tables(orc, po)
    cases = [(0, 120_000, 125_000, 61), (1, 40_000, 43_000, 62)]   # second call: smaller, t = 1
    ref = ca.StepState(p, 200_000, tables=tables)
    nal = ca.StepState(p, 200_000, tables=tables, context=False)
    try:
        for t, n_e, n_i, seed in cases:
            print(f"noalloc vs pic_step: t={t} n_e={n_e} n_i={n_i}", flush=True)
            E, I = synthetic_state(n_e, n_i, seed=seed)
            for st in (ref, nal):
                st.upload(0, *E)
                st.upload(1, *I)
            ref.step(t, measure=True, reset=True)
            nal.step_noalloc(t, measure=True, reset=True)
            _same_step(ref, nal)
            # and the next step, which reads the workspace the first one left
            ref.step(t + 1, measure=True)
            nal.step_noalloc(t + 1, measure=True)
            _same_step(ref, nal)
    finally:
        ref.close()
        nal.close()


def test_noalloc_rejects_bad_arguments(ca):
    from cudapic_amd._abi import PIC_ERR_ARG, lib
    import ctypes as C
    p = ca.params(weight=2.0e3)
    st = ca.StepState(p, 1000, context=False)
    try:
        assert lib().pic_step_noalloc(C.byref(p), C.byref(st.st), 4000, 2, st.stream) == PIC_ERR_ARG
        q = ca.params(weight=2.0e3)
        q.n_t = 4010
        assert lib().pic_step_noalloc(C.byref(q), C.byref(st.st), 0, 2, st.stream) == PIC_ERR_ARG
    finally:
        st.close()
