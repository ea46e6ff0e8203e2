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


def test_forward_without_events_equals_the_oracle_step(orc):
    """The adjoint oracle's forward model (deposit, Thomas solve and field,
    leapfrog push, absorption at the walls, the drive's phase) against the
    C restatement of the reference step (or_step, main.rs:1725-1808 without
    collisions): a collisionless tape whose only events are the particles
    that leave the gap, 60 steps with three ion steps.  or_step pushes f32
    particles in an f32 field from an FP64 solve, the model stays in FP64,
    so the two agree to the f32 resolution: fields to 1e-5 of their maximum,
    densities to 1e-4 of theirs (the f32 CIC weights of or_step round the
    position in cells to ~2e-5 at the far end of the grid), positions to
    1e-7 m (electrons) and 1e-8 m (ions) after 60 steps, the same survivors.
    A drive of the wrong phase (sine for cosine) fails the field check."""
    from cudapic_amd import synthetic
    po = orc.params(weight=7.0e4 * 110_000 / 3000)
    m = oa.Model.from_params(po, t0=0)
    E, I = synthetic.state(3000, 3150, seed=9, te_ev=20.0)
    # electrons spread over the whole gap, so that some leave it
    E = (np.random.default_rng(9).uniform(0.0, po.gap, 3000).astype(np.float32),) + tuple(E[1:])
    cap = 8000
    er = orc.seed_streams([0x12345678, 0x11112222, 0x22223333, 0x33334444], cap)
    ir = orc.seed_streams([0x44445555, 0x55556666, 0x66667777, 0x77778888], cap)
    sim = orc.Sim(po, cap, *E, *I, er, ir)
    wc, ws = np.zeros(16), np.zeros(16)
    wc[0] = po.v0
    steps = 60
    state = {"xe": E[0].astype(np.float64), "vxe": E[1].astype(np.float64),
             "xi": I[0].astype(np.float64), "vxi": I[1].astype(np.float64)}
    tape, full = oa.random_tape(m, state, wc, ws, steps, seed=0, p_coll=0.0, p_ion=0.0)
    assert len(tape["slot_e"]) == len(tape["slot_i"]) == len(tape["parent"]) == 0
    _, nbar, traj = oa.forward(m, full, tape, wc, ws, np.ones(m.n_g), steps, keep=True)
    ion_dens = []
    lost = 0
    dmax = dmax_i = 0.0
    for t in range(steps):
        sim.step(t, collisions=False)
        tr = traj[t]
        # the field and densities of step t
        assert np.abs(sim.efield - tr["E"]).max() <= 1e-5 * np.abs(tr["E"]).max(), t
        want_e = m.deposit(tr["xe"][tr["ae"]])
        assert np.abs(sim.e_dens - want_e).max() <= 1e-4 * want_e.max(), t
        if t % m.n_sub == 0:
            ion_dens.append(sim.i_dens.copy())
        # the survivors of step t, in slot order (or_step's compaction keeps the order)
        if t + 1 < steps:
            nx = traj[t + 1]
            got_e, got_i = sim.electrons()[0], sim.ions()[0]
            want_e_x = nx["xe"][nx["ae"]]
            assert len(got_e) == len(want_e_x), t
            dmax = max(dmax, np.abs(got_e - want_e_x).max())
            if nx["xi"] is not None:
                want_i_x = nx["xi"][nx["ai"]]
                assert len(got_i) == len(want_i_x)
                dmax_i = max(dmax_i, np.abs(got_i - want_i_x).max())
        lost = 3000 - sim.n_e
    print("max position difference: electrons", dmax, "ions", dmax_i)
    assert dmax <= 1e-7 and dmax_i <= 1e-8
    assert lost > 0   # absorptions happened and were matched
    # the window-averaged ion density of the objective
    np.testing.assert_allclose(nbar, np.mean(ion_dens, axis=0), rtol=1e-4, atol=1e-4 * nbar.max())
    # planted mistake: the drive as a sine
    _, _, bad = oa.forward(m, full, tape, ws, wc, np.ones(m.n_g), 1, keep=True)
    sim2 = orc.Sim(po, cap, *E, *I, er, ir)
    sim2.step(0, collisions=False)
    assert np.abs(sim2.efield - bad[0]["E"]).max() > 1e-2 * np.abs(bad[0]["E"]).max()
