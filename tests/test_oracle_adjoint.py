"""Pins the FP64 CPU adjoint (oracle/adjoint.py) that checks the engine's
adjoint kernels: its gradient equals central finite differences of its own
FP64 forward on a tape with collisions, absorptions and ionisations, and its
dJ/dV(t) equals sum_k dJ/dE_k(t) / L (the boundary voltage adds the uniform
field V/L at every node)."""
import numpy as np
import pytest

from oracle import adjoint as oa


@pytest.fixture(scope="module")
def case():
    import cudapic_amd as ca
    from cudapic_amd import synthetic
    n = 2000
    p = ca.params(weight=7.0e4 * 110_000 / n)
    m = oa.Model.from_params(p, t0=0)
    E, I = synthetic.state(n, int(n * 1.05), seed=4)
    state = {"xe": E[0], "vxe": E[1], "xi": I[0], "vxi": I[1]}
    wc = np.zeros(16)
    ws = np.zeros(16)
    wc[0] = 250.0
    rng = np.random.default_rng(1)
    ws[1:] = rng.normal(0.0, 10.0, 15)
    steps = 200
    tape, full = oa.random_tape(m, state, wc, ws, steps, seed=2, p_coll=0.005, p_ion=0.001)
    wt = wc.copy()
    wt[0] *= 1.05
    _, target = oa.forward(m, full, tape, wt, ws, np.ones(p.n_g), steps)
    return m, full, tape, wc, ws, target, steps


def test_tape_has_all_event_kinds(case):
    m, full, tape, *_ = case
    assert len(tape["slot_e"]) > 100 and len(tape["slot_i"]) > 10 and len(tape["parent"]) > 5
    assert (tape["death_e"][:2000] != 0xffffffff).sum() > 0   # absorptions


def test_adjoint_equals_central_differences(case):
    m, full, tape, wc, ws, target, steps = case
    J, gc, gs, dv, dv_sum = oa.gradient(m, full, tape, wc, ws, target, steps)
    assert J > 1e-4
    np.testing.assert_allclose(dv, dv_sum, rtol=1e-7, atol=1e-12 * np.abs(dv).max())
    # FP64 differences of the FP64 forward: O(h^2) and rounding are far
    # below the tolerance at h = 1e-4 V
    h = 1e-4
    for kind, k in (("cos", 0), ("sin", 1), ("cos", 5), ("sin", 15)):
        wp, wm = (wc.copy(), wc.copy()) if kind == "cos" else (ws.copy(), ws.copy())
        wp[k] += h
        wm[k] -= h
        if kind == "cos":
            jp = oa.forward(m, full, tape, wp, ws, target, steps)[0]
            jm = oa.forward(m, full, tape, wm, ws, target, steps)[0]
            g = gc[k]
        else:
            jp = oa.forward(m, full, tape, wc, wp, target, steps)[0]
            jm = oa.forward(m, full, tape, wc, wm, target, steps)[0]
            g = gs[k]
        fd = (jp - jm) / (2 * h)
        assert abs(fd - g) <= 1e-5 * max(abs(g), 1e-2 * np.abs(gc).max()), (kind, k, fd, g)


def test_planted_sign_error_is_caught(case):
    """A mistake in the field adjoint (the voltage's path through the
    right-hand side f_1 -= V dropped) must show up against the finite
    differences."""
    m, full, tape, wc, ws, target, steps = case
    orig = m.field_adjoint

    def broken(lE):
        lrho, dv = orig(lE)
        s = 0.5 / m.dx
        lphi0 = lE[1] * s + lE[0] / m.dx
        return lrho, lphi0   # dJ/dV without the -u_1 term
    m.field_adjoint = broken
    try:
        _, gc, _, _, _ = oa.gradient(m, full, tape, wc, ws, target, steps)
    finally:
        m.field_adjoint = orig
    h = 1e-4
    wp, wm = wc.copy(), wc.copy()
    wp[0] += h
    wm[0] -= h
    fd = (oa.forward(m, full, tape, wp, ws, target, steps)[0] - oa.forward(m, full, tape, wm, ws, target, steps)[0]) / (2 * h)
    assert abs(fd - gc[0]) > 1e-3 * abs(fd)
