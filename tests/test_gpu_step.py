"""GPU parity of the RF step on caller-owned buffers (pic_init / pic_step /
pic_step_graph): the reference's GpuSimState stays the caller's and one call
runs the loop body of Rust-CUDA/src/main.rs:1725-1808 on the fused kernels.

  * one step against the CPU oracle, stage by stage (densities, FP64 field,
    per-slot collision replay, ionisation products, EEPF), at a compacting
    ion step, an in-place electron-only step and an in-place ion step;
  * many steps against the simulation object on the same kernels
    (pic_sim_run with one segment): eager steps and graph periods, across a
    compaction, with the struct's pointer swaps, the measurement
    accumulators and the wall-loss counters.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_same_multiset, no_ionisation_tables, sort_rows, synthetic_state
from replay import _check_new_ions, _push_ref, _replay_species, _split_ions

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


@pytest.mark.parametrize("t0,lag,lag_out", [(0, None, 0), (1, None, 0), (20, None, 0),
                                             (1, 3, 4), (21, 5, 0), (20, 2, 3), (41, 0, 1)])
def test_one_step_vs_oracle(ca, orc, t0, lag, lag_out):
    """lag None: pic_step.  Otherwise pic_step_lagged, the step inside a run
    of deferred electron RNG write-backs: the stored words of the live
    electron slots start `lag` draws behind their streams and end `lag_out`
    behind.  The test brings them to their streams' positions with the
    oracle's generator (live slots only; holes and unused slots must hold
    exact words) and replays the step against them."""
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    n_e, n_i, cap = 200_000, 210_000, 300_000
    E, I = synthetic_state(n_e, n_i, seed=61 + t0 + 7 * (lag or 0), te_ev=6.0)
    if lag is not None:
        # 2000 of them next to the walls, moving into them: lost in the step
        r = np.random.default_rng(t0)
        near = r.uniform(0.0, 5e-6, 2000).astype(np.float32)
        E[0][:1000] = near[:1000]
        E[1][:1000] = -2e6
        E[0][1000:2000] = np.float32(0.025) - near[1000:]
        E[1][1000:2000] = 2e6
    cs, se, si = orc.cross_sections(po.gas_density)
    with ca.StepState(p, cap, tables=(cs, se, si)) as st:
        st.upload(0, *E)
        st.upload(1, *I)
        tmp_e, tmp_i = st.st.tmp_x, st.st.tmp_i_x
        rng_e0, rng_i0 = st.rng_download(0), st.rng_download(1)
        print(f"pic_step one step t0={t0} lag={lag} lag_out={lag_out}", flush=True)
        if lag is None:
            st.step(t0, measure=True, reset=True)
        else:
            st.step_lagged(t0, lag, lag_out, measure=True, reset=True)
            # every uploaded electron is live: its stream is `lag` draws on
            k0 = np.zeros(cap, np.uint32)
            k0[:n_e] = lag
            rng_e0 = orc.xo_skip_soa(rng_e0, k0)
        pot, ef, ed, idn = st.grid()
        for got, parts in ((ed, E), (idn, I)):
            want = orc.deposit(parts[0], po)
            assert np.abs(got - want).max() <= 2e-6 * want.max()
        pw, ew = orc.poisson_scan_f64(ed, idn, 250.0 * np.cos(2.0 * np.pi * t0 / 4000.0), po)
        assert np.abs(pot - pw).max() <= 1e-5 * np.abs(pw).max()
        assert np.abs(ef - ew).max() <= 1e-5 * np.abs(ew).max()
        compacting = t0 % 800 == 0
        # main.rs:1764-1767: after a compacting step the particles live in
        # the arrays that were the spares
        assert (st.st.e_x == tmp_e) == compacting
        assert (st.st.i_x == tmp_i) == compacting
        ge, slots_e = st.download(0)
        gi, slots_i = st.download(1)
        rng_e1, rng_i1 = st.rng_download(0), st.rng_download(1)
        if lag is not None:
            # live slots (newborns included) are `lag_out` behind; holes and
            # the slots past the watermark hold exact words
            k1 = np.zeros(cap, np.uint32)
            k1[slots_e] = lag_out
            rng_e1 = orc.xo_skip_soa(rng_e1, k1)
            x_slots = st.slots_in_use(0)[0]
            holes = np.nonzero(x_slots < 0.0)[0]
            unused = np.arange(len(x_slots), cap)
            assert len(holes) >= 1900
            for w in range(4):
                # a particle lost in the step drew nothing: its stream's words before the step
                np.testing.assert_array_equal(rng_e1[w][holes], rng_e0[w][holes])
                np.testing.assert_array_equal(rng_e1[w][unused], rng_e0[w][unused])
        born_x, ions = _replay_species(orc, po, se, cs, E, ef, ge, slots_e, rng_e0, rng_e1, "e")
        assert len(born_x) > 0
        ion_step = t0 % 20 == 0
        if ion_step:
            cnt_i, pushed_i, _ = _push_ref(orc, po, ef, I, "i")
        else:
            cnt_i, pushed_i = len(I[0]), I
        assert len(gi[0]) == cnt_i + len(born_x)
        _, old_ion, n_amb, amb = _split_ions(gi, pushed_i[0])
        if ion_step:
            _replay_species(orc, po, si, cs, I, ef, [a[old_ion] for a in gi], slots_i[old_ion],
                            rng_i0, rng_i1, "i", missing=n_amb)
        else:
            keep = ~np.isin(I[0], amb)
            assert_same_multiset([a[old_ion] for a in gi], [a[keep] for a in I])
            for w in range(4):
                np.testing.assert_array_equal(rng_i1[w], rng_i0[w])
        _check_new_ions(gi, pushed_i[0], born_x, ions, check_v=True)
        # watermarks: in-place steps keep the holes, a compaction squeezes them out
        n_lost_e = n_e + len(born_x) - len(ge[0])
        assert st.count(0) == (len(ge[0]) if compacting else n_e + len(born_x))
        assert n_lost_e >= 0
        counters, ifed, overflow = st.diagnostics()
        assert overflow == 0
        assert counters[0] + counters[1] == n_lost_e
        assert counters[4] > 1000 and (counters[5] > 100) == ion_step
        assert counters[6] == 1 and counters[7] == int(ion_step)
        # absorbed ions (ion steps only) are binned below 200 eV
        assert ifed.sum() <= counters[2] + counters[3]
        assert ion_step or counters[2] + counters[3] == 0
        m = st.measurements()
        np.testing.assert_array_equal(m["eepf"], orc.accumulate_eepf(*ge, ef, po))
        np.testing.assert_array_equal(m["cumul_e"], ed)
        if ion_step:
            np.testing.assert_array_equal(m["cumul_i"], idn)


def _zero_tables(orc, po):
    cs, se, si = orc.cross_sections(po.gas_density)
    return cs, np.zeros_like(se), np.zeros_like(si)


def _compare(sim, st, what):
    """A pic_sim run and a caller-owned run of the same collisionless FP64
    steps see the same fields up to the summation order of the density
    atomics (the test of graph replays against eager steps uses the same
    bounds)."""
    n_sim = sim.counts()
    got = [st.download(s)[0] for s in (0, 1)]
    assert abs(n_sim[0] - len(got[0][0])) <= 2 and abs(n_sim[1] - len(got[1][0])) <= 2, (what, n_sim)
    for s in (0, 1):
        a = sort_rows(*sim.download(s))
        b = sort_rows(*got[s])
        m = min(len(a), len(b))
        assert np.abs(a[:m, 0] - b[:m, 0]).max() < 1e-6, what
    pa, ea, _, _ = sim.grid()
    pb, eb, _, _ = st.grid()
    assert np.abs(pa - pb).max() < 1e-3 * np.abs(pa).max(), what
    assert np.abs(ea - eb).max() < 1e-3 * np.abs(ea).max(), what


def test_eager_and_graph_steps_match_sim_run(ca, orc):
    """200 collisionless FP64 steps from t = 760 (a compaction at t = 800)
    with measurement: pic_sim_run (one segment, no graph), pic_step one step
    at a time, and pic_step_graph in calls that start on and off the
    80-step period, with eager steps between graph replays."""
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    E, I = synthetic_state(100_000, 110_000, seed=71)
    tables = _zero_tables(orc, po)
    t0 = 760
    sim = ca.Sim(p, capacity=200_000, path=ca.PIC_PATH_FUSED, use_graph=False, segments=1)
    sim.set_tables(*tables)
    sim.set_time(t0)
    sim.upload(0, *E)
    sim.upload(1, *I)
    eager = ca.StepState(p, 200_000, tables=tables)
    graph = ca.StepState(p, 200_000, tables=tables)
    try:
        for st in (eager, graph):
            st.upload(0, *E)
            st.upload(1, *I)
        ptr0 = (graph.st.e_x, graph.st.tmp_x, graph.st.i_x, graph.st.tmp_i_x)
        print("pic_step: 200 eager steps", flush=True)
        sim.run(200, measure=True)
        for i in range(200):
            eager.step((t0 + i) % 4000, measure=True, reset=(i == 0))
        print("pic_step_graph: 40 + 80 + 3 + 77 steps", flush=True)
        graph.run(760, 40, measure=True, reset=True)   # eager up to the period start
        graph.run(800, 80, measure=True)                # compacting period, captured
        graph.run(880, 3, measure=True)                 # eager
        graph.run(883, 77, measure=True)                # eager (off the period)
        sim.synchronize()
        for st in (eager, graph):
            _compare(sim, st, "eager" if st is eager else "graph")
        # one compaction: the live particles are in the former spares
        assert (graph.st.e_x, graph.st.tmp_x, graph.st.i_x, graph.st.tmp_i_x) == (
            ptr0[1], ptr0[0], ptr0[3], ptr0[2])
        ms = sim.measurements()
        for st in (eager, graph):
            m = st.measurements()
            for key in ("cumul_e", "cumul_i"):
                np.testing.assert_allclose(m[key], ms[key], rtol=1e-6, atol=0)
            assert np.abs(m["eepf"].astype(np.int64) - ms["eepf"].astype(np.int64)).sum() <= 8
            c, ifed, ov = st.diagnostics()
            cs_ = ms["counters"]
            assert ov == 0
            assert abs(c[0] + c[1] - cs_["e_lost_powered"] - cs_["e_lost_grounded"]) <= 2
            assert abs(c[2] + c[3] - cs_["i_lost_powered"] - cs_["i_lost_grounded"]) <= 2
            assert c[6] == 200 and c[7] == 10
            assert abs(int(ifed.sum()) - int(ms["ifed"].sum())) <= 2
        # graph replays of a period already captured, then more steps, again equal
        sim.run(160)
        for i in range(160):
            eager.step((960 + i) % 4000)
        graph.run(960, 160)
        sim.synchronize()
        for st in (eager, graph):
            _compare(sim, st, "second window")
    finally:
        sim.close()
        eager.close()
        graph.close()


def test_deferred_rng_write_back_equals_every_step(ca, orc):
    """213 steps from t = 1 of pic_step_graph (eager steps, a captured period,
    eager steps ending off a write-back boundary) with the default deferred
    RNG write-back (runs of 8 steps) and with a write-back every step
    (rng_defer = 1): the same collisions from the same streams, so every
    slot below the watermark -- x, v and the four RNG words -- must be equal
    when the calls return.  Without ionisation the runs are deterministic
    (the density flush's floating-point atomics can only change the last bit
    of a sum, which may move an f32 field value in rare cases: up to 1e-4 of
    the slots may differ)."""
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    E, I = synthetic_state(200_000, 210_000, seed=91, te_ev=6.0)
    tables = no_ionisation_tables(orc, po)
    runs = []
    for defer in (1, 8):
        with ca.StepState(p, 300_000, tables=tables, rng_defer=defer) as st:
            st.upload(0, *E)
            st.upload(1, *I)
            print(f"deferred write-back: rng_defer={defer}, 213 steps", flush=True)
            st.run(1, 213, measure=True, reset=True)
            st.synchronize()
            counters, _, overflow = st.diagnostics()
            n = (st.count(0), st.count(1))
            runs.append((n, st.slots_in_use(0), st.rng_download(0, n[0]), st.slots_in_use(1),
                         st.rng_download(1, n[1]), counters, overflow))
    (na, ea, rea, ia, ria, ca_, ova), (nb, eb, reb, ib, rib, cb, ovb) = runs
    assert na == nb and ova == ovb == 0
    assert ca_[4] > 10_000   # electron collisions happened (and ion ones: index 5)
    assert ca_[5] > 0
    for a, b, ra, rb, n in ((ea, eb, rea, reb, na[0]), (ia, ib, ria, rib, na[1])):
        differ = np.zeros(n, bool)
        for u, v in zip(a + ra, b + rb):
            differ |= np.asarray(u).view(np.uint32) != np.asarray(v).view(np.uint32)
        assert differ.sum() <= max(2, n // 10_000), (differ.sum(), n)
    assert (np.asarray(ea[0]) < 0).sum() > 0   # holes were made and kept exact words


def test_reset_restarts_from_new_particles(ca, orc):
    """PIC_STEP_RESET after the caller replaced the particles: the step sees
    only the new ones (densities of the new state)."""
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    E, I = synthetic_state(50_000, 52_000, seed=81)
    E2, I2 = synthetic_state(30_000, 33_000, seed=82)
    with ca.StepState(p, 100_000, tables=_zero_tables(orc, po)) as st:
        st.upload(0, *E)
        st.upload(1, *I)
        st.run(0, 100, reset=True)
        st.upload(0, *E2)
        st.upload(1, *I2)
        st.step(100, reset=True)
        _, _, ed, idn = st.grid()
        for got, parts in ((ed, E2), (idn, I2)):
            want = orc.deposit(parts[0], po)
            assert np.abs(got - want).max() <= 2e-6 * want.max()


def test_step_rejects_bad_arguments(ca):
    from cudapic_amd._abi import PIC_ERR_ARG, PicState, lib
    p = ca.params(weight=2.0e3)
    with ca.StepState(p, 1000) as st:
        # t beyond the cycle
        assert lib().pic_step(st.ctx, C.byref(st.st), 4000, 0, st.stream) == PIC_ERR_ARG
        # a state with a missing buffer
        bad = PicState.from_buffer_copy(st.st)
        bad.tmp_i_vz = None
        assert lib().pic_step(st.ctx, C.byref(bad), 0, 2, st.stream) == PIC_ERR_ARG
        # measurement without its accumulators
        bad = PicState.from_buffer_copy(st.st)
        bad.eepf_counts = None
        assert lib().pic_step(st.ctx, C.byref(bad), 0, 3, st.stream) == PIC_ERR_ARG
        # one density output without the other
        bad = PicState.from_buffer_copy(st.st)
        bad.i_density = None
        assert lib().pic_step_graph(st.ctx, C.byref(bad), 0, 80, 2, st.stream) == PIC_ERR_ARG
        # a capacity overflow is reported, not hidden
        E, I = synthetic_state(1000, 1000, seed=3, te_ev=20.0)
        st.upload(0, *E)
        st.upload(1, *I)
        st.run(0, 400, reset=True)
        st.synchronize()
        _, _, ov = st.diagnostics()
        assert ov == 1 or max(st.count(0), st.count(1)) <= 1000
