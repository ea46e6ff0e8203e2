"""The adjoint kernels (pic_adj_*) against the FP64 CPU adjoint of the same
frozen-event model on the GPU's own event tape, against central finite
differences (h = 1e-3 V) of the FP64 forward of that model on the same tape
(frozen events: common random numbers by construction), and the determinism
the checkpointed recomputation relies on.

The window is 400 steps (20 ion steps, two checkpoint segments), at 20 k +
21 k and at 200 k + 210 k particles.  Measured
on the MI355X: the GPU gradient agrees with the FP64 CPU adjoint to 5.4e-4,
and the CPU adjoint with central differences at h = 1e-3 V to 2e-5; at
h = 1e-2 V the differences themselves are 0.3-1.3 % off the model's exact
gradient (the frozen-event model is only piecewise smooth: a larger step
moves particles across cells and clamps), so that step is not used.
Central differences of the GPU's own
FP32 forward are not used: a 1e-2 V change of the drive moves an ion by
~1e-10 m in a window, below the 2e-9 m resolution of an f32 position in a
25 mm gap, so they measure rounding (3-13 % off at h = 1e-2 V).  Beyond ~500
steps the frozen-event model is ill-conditioned: at 1000 steps even its FP64
central differences at h = 1e-2 V and at 1 V differ from each other and from
the adjoint by 10-100 % (electron trajectories diverge exponentially and
cross cell boundaries), so longer windows are measured for cost only."""
import numpy as np
import pytest

from oracle import adjoint as oa

pytestmark = pytest.mark.gpu
STEPS, CHECKPOINT = 400, 200


def _waveform(seed=3):
    wc = np.zeros(16)
    ws = np.zeros(16)
    ws[0] = 250.0
    ws[1:] = np.random.default_rng(seed).normal(0.0, 10.0, 15)
    return wc, ws


def _make_run(n, seed):
    import cudapic_amd as ca
    from cudapic_amd import synthetic
    p = ca.params(weight=7.0e4 * 110_000 / n)
    wc, ws = _waveform()
    for h in range(16):
        p.wave_cos[h], p.wave_sin[h] = wc[h], ws[h]
    E, I = synthetic.state(n, int(n * 1.05), seed=seed)
    adj = ca.Adjoint(p, 2 * n, 2 * n, STEPS, checkpoint=CHECKPOINT)
    adj.upload(0, *E)
    adj.upload(1, *I)
    adj.set_target(np.full(p.n_g, 1.0e15))
    j_rec = adj.record()
    # the target: the window-averaged ion density of a 5 % stronger drive
    wt = ws.copy()
    wt[0] *= 1.05
    _, target = adj.forward(wc, wt, nbar=True)
    adj.set_target(target)
    j0, nbar0 = adj.forward(wc, ws, nbar=True)
    grads = adj.backward()
    return {"ca": ca, "p": p, "adj": adj, "E": E, "I": I, "wc": wc, "ws": ws, "target": target,
            "j_rec": j_rec, "j0": j0, "nbar0": nbar0, "grads": grads}


@pytest.fixture(scope="module")
def run():
    r = _make_run(20_000, 21)
    yield r
    r["adj"].close()


@pytest.fixture(scope="module")
def run_large():
    """Ten times the particles: 200 k electrons + 210 k ions."""
    r = _make_run(200_000, 22)
    yield r
    r["adj"].close()


def test_replay_is_bit_exact(run):
    """Fixed-point densities: a replay of the recorded events reproduces the
    record pass exactly, and so does every later replay."""
    adj, wc, ws = run["adj"], run["wc"], run["ws"]
    adj.set_target(np.full(run["p"].n_g, 1.0e15))
    j_a, nb_a = adj.forward(wc, ws, nbar=True)
    j_b, nb_b = adj.forward(wc, ws, nbar=True)
    adj.set_target(run["target"])
    assert j_a == run["j_rec"] and j_b == j_a
    np.testing.assert_array_equal(nb_a, nb_b)
    tape = adj.tape()
    print("tape: collisions e", len(tape["slot_e"]), "i", len(tape["slot_i"]), "ionisations", len(tape["parent"]),
          "absorbed e", int((tape["death_e"] != 0xffffffff).sum()), "i", int((tape["death_i"] != 0xffffffff).sum()))
    assert len(tape["slot_e"]) > 1000 and len(tape["slot_i"]) > 100 and len(tape["parent"]) > 0


def _cpu_case(run):
    adj, p = run["adj"], run["p"]
    tape = adj.tape()
    n_e, n_i = len(run["E"][0]), len(run["I"][0])
    cap = 2 * n_e
    state = {"xe": np.zeros(cap), "vxe": np.zeros(cap), "xi": np.zeros(cap), "vxi": np.zeros(cap)}
    state["xe"][:n_e], state["vxe"][:n_e] = run["E"][0], run["E"][1]
    state["xi"][:n_i], state["vxi"][:n_i] = run["I"][0], run["I"][1]
    tape["death_e"] = np.concatenate([tape["death_e"], np.zeros(cap - len(tape["death_e"]), np.uint32)])
    tape["death_i"] = np.concatenate([tape["death_i"], np.zeros(cap - len(tape["death_i"]), np.uint32)])
    return oa.Model.from_params(p), state, tape


def test_gradient_matches_fp64_cpu_adjoint(run):
    m, state, tape = _cpu_case(run)
    J, gc, gs, dv, _ = oa.gradient(m, state, tape, run["wc"], run["ws"], run["target"], STEPS)
    ggc, ggs, gdv = run["grads"]
    print("J gpu", run["j0"], "cpu", J)
    print("grad_sin gpu", ggs[:4], "cpu", gs[:4])
    assert J > 0 and abs(run["j0"] / J - 1.0) < 1e-3
    scale = max(np.abs(gs).max(), np.abs(gc).max())
    assert np.abs(ggs - gs).max() < 5e-3 * scale and np.abs(ggc - gc).max() < 5e-3 * scale
    assert abs(ggs[0] / gs[0] - 1.0) < 5e-3
    assert np.linalg.norm(gdv - dv) < 1e-2 * np.linalg.norm(dv)


def test_gradient_matches_central_differences(run):
    """Central differences of the frozen-event model (FP64 forward on the
    GPU's tape) at h = 1e-3 V, within 1 % of the GPU adjoint."""
    m, state, tape = _cpu_case(run)
    wc, ws, target = run["wc"], run["ws"], run["target"]
    ggc, ggs, _ = run["grads"]
    h = 1e-3
    for kind, k in (("sin", 0), ("sin", 1), ("cos", 0), ("sin", 6)):
        wp, wm = (ws.copy(), ws.copy()) if kind == "sin" else (wc.copy(), wc.copy())
        wp[k] += h
        wm[k] -= h
        if kind == "sin":
            fd = (oa.forward(m, state, tape, wc, wp, target, STEPS)[0]
                  - oa.forward(m, state, tape, wc, wm, target, STEPS)[0]) / (2 * h)
            g = ggs[k]
        else:
            fd = (oa.forward(m, state, tape, wp, ws, target, STEPS)[0]
                  - oa.forward(m, state, tape, wm, ws, target, STEPS)[0]) / (2 * h)
            g = ggc[k]
        print(kind, k, "fd", fd, "adjoint", g)
        assert abs(fd - g) <= 1e-2 * abs(g), (kind, k, fd, g)


def _fd(m, state, tape, wc, ws, target, kind, k, h):
    wp, wm = (ws.copy(), ws.copy()) if kind == "sin" else (wc.copy(), wc.copy())
    wp[k] += h
    wm[k] -= h
    if kind == "sin":
        return (oa.forward(m, state, tape, wc, wp, target, STEPS)[0]
                - oa.forward(m, state, tape, wc, wm, target, STEPS)[0]) / (2 * h)
    return (oa.forward(m, state, tape, wp, ws, target, STEPS)[0]
            - oa.forward(m, state, tape, wm, ws, target, STEPS)[0]) / (2 * h)


def test_large_window_gradient_matches_cpu_adjoint_and_differences(run_large):
    """200 k + 210 k particles, the same 400-step window: the GPU gradient
    against the FP64 CPU adjoint on the GPU's tape, and against central
    differences at h = 1e-3 V, each coefficient within 1 % (purely
    relative), and at the differentiation configuration's h = 1e-2 V, the
    four coefficients as one vector within 1 % and each within 3 %.  The
    h = 1e-2 V differences average the derivative over a step that moves
    particles across cell boundaries, and the frozen-event model is only
    piecewise smooth there: how far they land from the exact gradient
    depends on the realisation (measured on the MI355X: 0.01-0.09 % on one
    tape, 0.4-1.5 % per coefficient and 0.5 % as a vector on another), so
    the configuration's 1 % per coefficient at h = 1e-2 V is not met on
    every tape at this size."""
    run = run_large
    m, state, tape = _cpu_case(run)
    wc, ws, target = run["wc"], run["ws"], run["target"]
    J, gc, gs, dv, _ = oa.gradient(m, state, tape, wc, ws, target, STEPS)
    ggc, ggs, gdv = run["grads"]
    print("J gpu", run["j0"], "cpu", J)
    assert J > 0 and abs(run["j0"] / J - 1.0) < 1e-3
    scale = max(np.abs(gs).max(), np.abs(gc).max())
    assert np.abs(ggs - gs).max() < 5e-3 * scale and np.abs(ggc - gc).max() < 5e-3 * scale
    assert np.linalg.norm(gdv - dv) < 1e-2 * np.linalg.norm(dv)
    gv, fd2v = [], []
    for kind, k in (("sin", 0), ("cos", 0), ("sin", 1), ("sin", 6)):
        g = ggs[k] if kind == "sin" else ggc[k]
        fd3 = _fd(m, state, tape, wc, ws, target, kind, k, 1e-3)
        fd2 = _fd(m, state, tape, wc, ws, target, kind, k, 1e-2)
        print(f"{kind}_{k}: adjoint {g:.6e} fd(1e-3) {fd3:.6e} ({fd3 / g - 1:+.2e}) "
              f"fd(1e-2) {fd2:.6e} ({fd2 / g - 1:+.2e})")
        assert abs(fd3 - g) <= 1e-2 * abs(g), (kind, k, fd3, g)
        assert abs(fd2 - g) <= 3e-2 * abs(g), (kind, k, fd2, g)
        gv.append(g)
        fd2v.append(fd2)
    gv, fd2v = np.array(gv), np.array(fd2v)
    rel = np.linalg.norm(fd2v - gv) / np.linalg.norm(gv)
    print(f"h = 1e-2 V, four coefficients as a vector: {rel:.2e}")
    assert rel <= 1e-2
