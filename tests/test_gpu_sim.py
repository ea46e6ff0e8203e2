"""GPU parity of the whole RF step (both paths) against the CPU oracle.

One step is checked stage by stage, each stage fed with the GPU's own
output of the stage before it:
  densities      vs the oracle deposit of the initial state (tolerance),
  pot / efield   vs the oracle FP64 scan of the GPU densities (1e-5),
  particles      vs the oracle push of the initial state in the GPU field:
                 bit-exact multiset for a collisionless step; with
                 collisions every surviving particle is matched to its
                 source by its pushed position and its collision is replayed
                 with the RNG words of the slot it landed in.
Whole runs are compared statistically with the oracle and with the
published eduPIC init cycle.
"""
import os

import numpy as np
import pytest

from helpers import SEED_E, SEED_I, assert_same_multiset, no_ionisation_tables, sort_rows, synthetic_state
from replay import _check_new_ions, _check_slot_draws, _push_ref, _replay_species, _split_ions

pytestmark = pytest.mark.gpu
PATHS = [0, 1]   # PIC_PATH_REFERENCE, PIC_PATH_FUSED


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


def make_sim(ca, p, cap, path, graph=False, **kw):
    return ca.Sim(p, capacity=cap, path=path, use_graph=graph, **kw)


def zero_tables(orc, po):
    cs, se, si = orc.cross_sections(po.gas_density)
    return cs, np.zeros_like(se), np.zeros_like(si)


# start steps: 0 is an ion step on both paths and a compacting step on the
# fused path; on the fused path 1 updates electrons in place (no ion step)
# and 20 (= N_SUB) updates both species in place
STARTS = [(0, 0), (1, 0), (1, 1), (1, 20)]
# fused-path segment counts: the default (one) and 32, the most the
# autotuner tries
SEGMENTS = [0, 32]
# (path, t0, segments); the reference path keeps one segment
CASES = [(path, t0, seg) for path, t0 in STARTS for seg in SEGMENTS if path or not seg]


def _pot0(t0):
    return float(np.float32(250.0 * np.cos(2.0 * np.pi * t0 / 4000.0)))


@pytest.mark.parametrize("path,t0,segments", CASES)
@pytest.mark.parametrize("prec", [0, 2])
def test_collisionless_step_stage_by_stage(ca, orc, path, t0, prec, segments):
    p = ca.params(weight=2.0e3, precision=prec)
    po = orc.params(weight=2.0e3)
    n_e, n_i, cap = 300_000, 320_000, 400_000
    E, I = synthetic_state(n_e, n_i, seed=31 + t0)
    sim = make_sim(ca, p, cap, path, segments=segments)
    sim.set_time(t0)
    cs, se0, si0 = zero_tables(orc, po)
    sim.set_tables(cs, se0, si0)
    sim.upload(0, *E)
    sim.upload(1, *I)
    rng_e0 = sim.rng_download(0)
    rng_i0 = sim.rng_download(1)
    sim.run(1)
    assert sim.time()[0] == t0 + 1
    pot, ef, ed, idn = sim.grid()
    # densities of the initial state
    for got, parts in ((ed, E), (idn, I)):
        want = orc.deposit(parts[0], po)
        assert np.abs(got - want).max() <= 2e-6 * want.max()
    # field from the GPU densities (pot0 = V0 cos(2 pi t0 / N_T))
    pw, ew = orc.poisson_scan_f64(ed, idn, _pot0(t0), po)
    # FP64: 1e-5 of the max norm; FP32: the n_g*eps*|pot| error of the f32
    # scan, which the central difference divides by 2 dx
    eps = np.finfo(np.float32).eps
    if prec == 2:
        assert np.abs(pot - pw).max() <= 1e-5 * np.abs(pw).max()
        assert np.abs(ef - ew).max() <= 1e-5 * np.abs(ew).max()
    else:
        assert np.abs(pot - pw).max() <= 2 * p.n_g * eps * np.abs(pw).max()
        assert np.abs(ef - ew).max() <= 2 * p.n_g * eps * np.abs(pw).max() / p.dx
    # particles: bit-exact multisets of the push in the GPU field
    ne, ni = sim.counts()
    ion_step = t0 % 20 == 0
    for sp, parts, rng0, name in ((0, E, rng_e0, "e"), (1, I, rng_i0, "i")):
        got = sim.download(sp)
        rng1 = sim.rng_download(sp)
        if name == "i" and not ion_step:
            assert_same_multiset(got, parts)
            for w in range(4):
                np.testing.assert_array_equal(rng1[w], rng0[w])
            continue
        cnt, want, _ = _push_ref(orc, po, ef, parts, name)
        assert len(got[0]) == cnt
        assert_same_multiset(got, want)
        # every occupied slot drew exactly one (ions: five) numbers
        _check_slot_draws(orc, rng0, rng1, sim.slots(sp), cnt, 1 if name == "e" else 5)
    assert sim.counts() == (ne, ni)
    sim.check()
    sim.close()


@pytest.mark.parametrize("path,t0,segments", CASES)
def test_step_with_collisions_slot_replay(ca, orc, path, t0, segments):
    _collision_step_replay(ca, orc, path, t0, segments=segments)


@pytest.mark.parametrize("block", [128, 512])
@pytest.mark.parametrize("t0", [0, 1, 20])
def test_fused_workgroup_sizes_slot_replay(ca, orc, block, t0):
    """The fused kernels at 128 and 512 lanes per workgroup (the autotuner's
    other sizes): the same per-slot replay as at 256, at a compacting step,
    an electron-only in-place step and an in-place ion step."""
    _collision_step_replay(ca, orc, 1, t0, segments=32, block_size=block)


def _collision_step_replay(ca, orc, path, t0, **sim_kw):
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    n_e, n_i, cap = 200_000, 210_000, 300_000
    E, I = synthetic_state(n_e, n_i, seed=41 + t0, te_ev=6.0)
    sim = make_sim(ca, p, cap, path, **sim_kw)
    sim.set_time(t0)
    cs, se, si = orc.cross_sections(po.gas_density)
    sim.upload(0, *E)
    sim.upload(1, *I)
    rng_e0 = sim.rng_download(0)
    rng_i0 = sim.rng_download(1)
    sim.run(1, measure=True)
    _, ef, _, _ = sim.grid()
    ne, ni = sim.counts()
    ge, gi = sim.download(0), sim.download(1)
    rng_e1, rng_i1 = sim.rng_download(0), sim.rng_download(1)
    born_x, ions = _replay_species(orc, po, se, cs, E, ef, ge, sim.slots(0), rng_e0, rng_e1, "e")
    assert len(born_x) > 0
    ion_step = t0 % 20 == 0
    if ion_step:
        cnt_i, pushed_i, _ = _push_ref(orc, po, ef, I, "i")
    else:
        cnt_i, pushed_i = len(I[0]), I
    assert ni == cnt_i + len(born_x)
    slots_i = sim.slots(1)
    _, old_ion, n_amb, amb = _split_ions(gi, pushed_i[0])
    if ion_step:
        # the ion replay covers the pushed ions (new ions have no pushed source)
        _replay_species(orc, po, si, cs, I, ef, [a[old_ion] for a in gi], slots_i[old_ion],
                        rng_i0, rng_i1, "i", missing=n_amb)
    else:
        keep = ~np.isin(I[0], amb)
        assert_same_multiset([a[old_ion] for a in gi], [a[keep] for a in I])
        for w in range(4):
            np.testing.assert_array_equal(rng_i1[w], rng_i0[w])
    # new ions: at their parents' positions and thermal (on the reference path
    # they may collide within the same ion step, so only positions are compared)
    _check_new_ions(gi, pushed_i[0], born_x, ions, check_v=(path == 1 or not ion_step))
    m = sim.measurements()
    c = m["counters"]
    assert c["e_meas_steps"] == 1 and c["i_meas_steps"] == int(ion_step)
    assert c["e_collisions"] > 1000
    assert (c["i_collisions"] > 100) == ion_step
    # EEPF of the post-collision electrons in the step's field (main.rs:1802-1807)
    want = orc.accumulate_eepf(*ge, ef, po)
    np.testing.assert_array_equal(m["eepf"], want)
    np.testing.assert_allclose(m["cumul_e"], sim.grid()[2], rtol=0, atol=0)
    sim.check()
    sim.close()


def _init_cycle_counts(ca, orc, path, seed, graph):
    """One RF cycle from 1000 electrons + 1000 ions at rest (the eduPIC init,
    C/eduPIC.cc:283-294)."""
    p = ca.params()
    rng = np.random.default_rng(500 + seed)
    z = np.zeros(1000, np.float32)
    ex = (rng.random(1000) * p.gap).astype(np.float32)
    ix = (rng.random(1000) * p.gap).astype(np.float32)
    se = list(SEED_E)
    se[0] += seed
    si = list(SEED_I)
    si[0] += seed
    sim = make_sim(ca, p, 20_000, path, graph=graph, seed_e=se, seed_i=si)
    sim.upload(0, ex, z, z, z)
    sim.upload(1, ix, z, z, z)
    sim.run(p.n_t)
    out = sim.counts()
    sim.check()
    sim.close()
    return out


@pytest.mark.parametrize("path,graph", [(0, False), (1, False), (1, True)])
def test_init_cycle_statistics(ca, orc, path, graph):
    """The published eduPIC init cycle ends at 888 electrons / 3345 ions
    (C/ReferenceCase_RawData/Figure9_conv.dat row 0); the oracle over 4 seeds
    gives 769-949 / 3154-3416.  Mean over 6 GPU seeds within 15%."""
    res = np.array([_init_cycle_counts(ca, orc, path, s, graph) for s in range(6)], float)
    assert abs(res[:, 0].mean() - 888) < 0.15 * 888, res
    assert abs(res[:, 1].mean() - 3345) < 0.15 * 3345, res


def test_graph_replays_match_eager(ca, orc):
    """Graph replays (4*N_SUB steps each) against the same steps run eagerly,
    collisionless and in FP64 so that both runs see the same fields up to the
    summation order of the density atomics; eager work and synchronisations
    between the replays."""
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    E, I = synthetic_state(100_000, 110_000, seed=51)
    cs, se0, si0 = zero_tables(orc, po)
    sims = []
    for graph in (False, True):
        s = make_sim(ca, p, 200_000, 1, graph=graph)
        s.set_tables(cs, se0, si0)
        s.upload(0, *E)
        s.upload(1, *I)
        sims.append(s)
    for rep in range(3):
        for s in sims:
            s.run(160)
            s.synchronize()
        ce = [s.counts() for s in sims]
        assert abs(ce[0][0] - ce[1][0]) <= 2 and abs(ce[0][1] - ce[1][1]) <= 2, ce
        a = sort_rows(*sims[0].download(0))
        b = sort_rows(*sims[1].download(0))
        m = min(len(a), len(b))
        assert np.abs(a[:m, 0] - b[:m, 0]).max() < 1e-6
        pa = sims[0].grid()[0]
        pb = sims[1].grid()[0]
        assert np.abs(pa - pb).max() < 1e-3 * np.abs(pa).max()
        # eager work between replays
        for s in sims:
            s.run(3)
            s.run(77)
    for s in sims:
        assert s.time()[0] == 720
        s.close()


def test_timed_window_equals_plain_run(ca, orc):
    """A timed window first captures the graphs it will replay (a dry pass
    over the same periods) and must leave the host state where a plain run
    of the same steps leaves it: collisionless FP64 runs, one timed and one
    not, agree after the window and after more steps."""
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    E, I = synthetic_state(100_000, 110_000, seed=52)
    cs, se0, si0 = zero_tables(orc, po)
    sims = []
    for _ in range(4):
        s = make_sim(ca, p, 200_000, 1, graph=True)
        s.set_tables(cs, se0, si0)
        s.upload(0, *E)
        s.upload(1, *I)
        sims.append(s)
    r = sims[0].time_steps(880)
    sims[1].run(880)
    # the electron kernel timed with events around each eager launch, from a
    # step that is not on a period boundary
    sims[2].run(3)
    k = sims[2].time_steps(877, kernel_events=True)
    # timed windows that start and end inside periods: captured partial
    # periods (the head of one, whole ones, the tail of another)
    sims[3].run(3)
    sims[3].time_steps(150)
    sims[3].time_steps(727)
    assert r["particle_steps"] > 0 and r["ms_total"] > 0
    assert k["kernel_launches"] == 877 and 0 < k["ms_kernel"] < k["ms_total"]
    for extra in (0, 80):
        for s in sims:
            s.run(extra)
            s.synchronize()
        ce = [s.counts() for s in sims]
        b = sort_rows(*sims[1].download(0))
        pb = sims[1].grid()[0]
        for other in (0, 2, 3):
            assert abs(ce[other][0] - ce[1][0]) <= 2 and abs(ce[other][1] - ce[1][1]) <= 2, ce
            a = sort_rows(*sims[other].download(0))
            m = min(len(a), len(b))
            assert np.abs(a[:m, 0] - b[:m, 0]).max() < 1e-6
            pa = sims[other].grid()[0]
            assert np.abs(pa - pb).max() < 1e-3 * np.abs(pa).max()
    for s in sims:
        assert s.time()[0] == 960
        s.check()
        s.close()


def test_deferred_rng_write_back_equals_every_step(ca, orc):
    """The simulation object (segments, captured periods, timed eager steps)
    with the default deferred electron RNG write-back and with a write-back
    every step (rng_defer = 1).  Without ionisation the slot contents are
    deterministic, so after pic_sim_run (eager steps, a captured period,
    eager steps), a timed window inside one period (a captured partial
    period that ends off a write-back boundary) and pic_sim_time_steps with
    kernel events (eager steps) every slot -- particles and RNG words -- must
    be equal (up to 1e-4 of the slots, for the last bit that the density
    flush's floating-point atomics may move in an f32 field value)."""
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    E, I = synthetic_state(200_000, 210_000, seed=93, te_ev=6.0)
    tables = no_ionisation_tables(orc, po)
    out = []
    for defer in (1, 8):
        s = ca.Sim(p, capacity=300_000, rng_defer=defer)
        try:
            s.set_tables(*tables)
            s.set_time(1)
            s.upload(0, *E)
            s.upload(1, *I)
            print(f"pic_sim rng_defer={defer}: 170 steps run, 27 + 37 timed", flush=True)
            s.run(170, measure=True)
            s.time_steps(27)
            s.time_steps(37, kernel_events=True)
            s.synchronize()
            s.check()
            out.append((s.counts(), s.download(0) + s.rng_download(0), s.download(1) + s.rng_download(1),
                        s.measurements()["counters"]))
        finally:
            s.close()
    (na, ea, ia, ca_), (nb, eb, ib, cb) = out
    assert na == nb
    assert ca_["e_collisions"] > 10_000 and ca_["i_collisions"] > 0, ca_
    for a, b in ((ea, eb), (ia, ib)):
        n = len(a[0])
        assert [len(u) for u in a] == [len(v) for v in b]
        differ = np.zeros(len(a[4]), bool)
        for w in range(4):
            differ[:n] |= np.asarray(a[w]).view(np.uint32) != np.asarray(b[w]).view(np.uint32)
            differ |= a[4 + w] != b[4 + w]
        assert differ.sum() <= max(2, n // 10_000), (differ.sum(), n)


def test_candidate_masks_equal_plain_runs(ca, orc):
    """Collision-candidate masks (a live slot reads no RNG words on a step
    of a deferred run whose draw cannot fall below the masks' threshold)
    change no draw and no outcome.  Without ionisation the slot contents are
    deterministic, so the simulation object with the masks at the default
    threshold, at a threshold below most electrons' collision probability
    (nearly every in-gap electron loads its words late, after the push), at
    a threshold that makes a third of the steps candidates, and with the
    masks off must leave every slot -- particles and RNG words -- as a run
    that writes the words back every step, after eager steps, captured
    periods, measured steps (which run without masks) and timed windows (up
    to 1e-4 of the slots, for the last bit the density flush's
    floating-point atomics may move in an f32 field value)."""
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    E, I = synthetic_state(200_000, 210_000, seed=97, te_ev=6.0)
    tables = no_ionisation_tables(orc, po)
    out = []
    variants = ((1, -1, 0.0), (8, 0, 0.0), (8, 1, 0.002), (8, 1, 0.05), (8, -1, 0.0))
    for defer, masks, thr in variants:
        s = ca.Sim(p, capacity=300_000, rng_defer=defer, cand_masks=masks, cand_threshold=thr)
        try:
            s.set_tables(*tables)
            s.set_time(3)
            s.upload(0, *E)
            s.upload(1, *I)
            print(f"pic_sim rng_defer={defer} cand_masks={masks} threshold={thr}: 330 steps", flush=True)
            s.run(170)
            s.run(45, measure=True)
            s.run(51)
            s.time_steps(27)
            s.time_steps(37, kernel_events=True)
            s.synchronize()
            s.check()
            out.append((s.counts(), s.download(0) + s.rng_download(0), s.download(1) + s.rng_download(1),
                        s.measurements()["counters"]))
        finally:
            s.close()
    nb, eb, ib, cb = out[0]
    assert cb["e_collisions"] > 50_000 and cb["i_collisions"] > 0, cb
    for (na, ea, ia, ca_), v in zip(out[1:], variants[1:]):
        assert na == nb, v
        assert abs(int(ca_["e_collisions"]) - int(cb["e_collisions"])) <= max(2, cb["e_collisions"] // 10_000), \
            (v, ca_, cb)
        for a, b in ((ea, eb), (ia, ib)):
            n = len(a[0])
            assert [len(u) for u in a] == [len(w) for w in b]
            differ = np.zeros(len(a[4]), bool)
            for w in range(4):
                differ[:n] |= np.asarray(a[w]).view(np.uint32) != np.asarray(b[w]).view(np.uint32)
                differ |= a[4 + w] != b[4 + w]
            assert differ.sum() <= max(2, n // 10_000), (v, differ.sum(), n)


def test_checkpoint_round_trip(ca, orc, tmp_path):
    p = ca.params(weight=2.0e3)
    E, I = synthetic_state(50_000, 52_000, seed=61)
    sim = make_sim(ca, p, 100_000, 1)
    sim.upload(0, *E)
    sim.upload(1, *I)
    sim.run(57, measure=True)
    path = os.path.join(tmp_path, "state.ckpt")
    sim.save(path)
    e0, i0 = sim.download(0), sim.download(1)
    r0 = [sim.rng_download(s) for s in (0, 1)]
    slots0 = [sim.slots(s) for s in (0, 1)]
    m0 = sim.measurements()
    sim2 = make_sim(ca, p, 100_000, 1)
    sim2.load(path)
    for a, b in zip(sim2.download(0), e0):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(sim2.download(1), i0):
        np.testing.assert_array_equal(a, b)
    # every particle keeps its RNG stream, although the load packs the
    # particles into other slots (the holes of the saved run are gone)
    for s in (0, 1):
        r1, slots1 = sim2.rng_download(s), sim2.slots(s)
        for w in range(4):
            np.testing.assert_array_equal(r1[w][slots1], r0[s][w][slots0[s]])
    m1 = sim2.measurements()
    np.testing.assert_array_equal(m1["cumul_e"], m0["cumul_e"])
    np.testing.assert_array_equal(m1["eepf"], m0["eepf"])
    assert m1["counters"] == m0["counters"]
    assert sim2.time() == sim.time()
    sim2.run(43, measure=True)
    sim2.check()
    assert sim2.counts()[0] > 0
    sim.close()
    sim2.close()


def test_checkpoint_continuation_and_corrupt_files(ca, orc, tmp_path):
    """A collisionless FP64 run saved and resumed continues as the run that
    was not interrupted (up to the summation order of the density atomics);
    truncated or padded files are refused before anything is allocated."""
    p = ca.params(weight=2.0e3, precision=2)
    po = orc.params(weight=2.0e3)
    cs, se0, si0 = zero_tables(orc, po)
    E, I = synthetic_state(60_000, 63_000, seed=62)
    sims = [make_sim(ca, p, 100_000, 1) for _ in range(2)]
    path = os.path.join(tmp_path, "c.ckpt")
    try:
        for s in sims:
            s.set_tables(cs, se0, si0)
        sims[0].upload(0, *E)
        sims[0].upload(1, *I)
        sims[0].run(57)
        sims[0].save(path)
        sims[1].load(path)
        for s in sims:
            s.run(103)
        ca_, cb = sims[0].counts(), sims[1].counts()
        assert abs(ca_[0] - cb[0]) <= 2 and abs(ca_[1] - cb[1]) <= 2, (ca_, cb)
        for sp in (0, 1):
            a = sort_rows(*sims[0].download(sp))
            b = sort_rows(*sims[1].download(sp))
            m = min(len(a), len(b))
            assert np.abs(a[:m, 0] - b[:m, 0]).max() < 1e-6
        raw = open(path, "rb").read()
        for bad in (raw[:-8], raw + b"\0" * 8, raw[:100]):
            with open(path, "wb") as f:
                f.write(bad)
            with pytest.raises(ca.PicError, match="status 5"):
                sims[1].load(path)
    finally:
        for s in sims:
            s.close()


def test_capacity_overflow_is_reported(ca, orc):
    p = ca.params(weight=2.0e3)
    E, I = synthetic_state(20_000, 20_000, seed=71, te_ev=30.0)
    sim = make_sim(ca, p, 20_010, 1)
    sim.upload(0, *E)
    sim.upload(1, *I)
    sim.run(20)
    with pytest.raises(ca.PicError):
        sim.check()
    sim.close()


def test_upload_above_capacity_is_an_error(ca):
    p = ca.params()
    sim = make_sim(ca, p, 10, 1)
    z = np.zeros(11, np.float32)
    with pytest.raises(ca.PicError):
        sim.upload(0, z, z, z, z)
    sim.close()
