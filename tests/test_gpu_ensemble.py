"""Ensembles of independent discharges in one launch per stage: each
instance's step equals the single-discharge step of its own parameters
(checked stage by stage against the oracle as in test_gpu_sim.py), and
whole runs reproduce the reference's init-cycle statistics per instance."""
import numpy as np
import pytest

from helpers import assert_same_multiset, sort_rows, synthetic_state
from replay import _check_new_ions, _push_ref, _replay_species, _split_ions

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


def _source_slots(orc, po, ef, parts, name):
    """Slot (upload index) of every pushed survivor: the push leaves vy
    untouched and keeps the order, so an index carried in vy comes back."""
    x, vx, _, vz = parts
    idx = np.arange(len(x), dtype=np.float32)
    return _push_ref(orc, po, ef, [x, vx, idx, vz], name)[1][2].astype(np.int64)


# start steps: 1 updates electrons in place, 20 (= N_SUB) both species in
# place, 0 both species in place followed by the compaction pass; the
# instances stepped as one group, or each as a group of its own (every
# per-instance array then addressed from the group's first instance)
# grid 8: eight particle workgroups, so that each takes ~45 tiles and
# flushes its histograms and scatters its collision list several times
# within an instance
@pytest.mark.parametrize("grid", [0, 8], ids=["grid-default", "grid-8"])
@pytest.mark.parametrize("grouped", [False, True], ids=["one-group", "group-per-instance"])
@pytest.mark.parametrize("t0", [0, 1, 20])
def test_ensemble_step_per_instance_replay(ca, orc, t0, grouped, grid):
    specs = [(250.0, 10.0, 2.0e3, 60_000), (120.0, 4.0, 5.0e3, 30_000), (380.0, 25.0, 1.0e3, 90_000)]
    ps = [ca.params(v0=v, pressure=pr, weight=w, precision=2) for v, pr, w, _ in specs]
    caps = [int(n * 1.3) for *_, n in specs]
    ens = ca.Ensemble(ps, caps, use_graph=False, group_slots=1 if grouped else 0xFFFFFFFF, grid_blocks=grid)
    ens.set_time(t0)
    states = []
    for i, (v, pr, w, n) in enumerate(specs):
        E, I = synthetic_state(n, int(n * 1.05), seed=80 + i + t0, te_ev=6.0)
        ens.upload(i, 0, *E)
        ens.upload(i, 1, *I)
        states.append((E, I))
    rng0 = [(ens.rng_download(i, 0, caps[i]), ens.rng_download(i, 1, caps[i])) for i in range(3)]
    ens.run(1, measure=True)
    ne, ni = ens.counts()
    ion_step = t0 % 20 == 0
    compacted = t0 % 40 == 0
    born_total = 0
    for i, (v, pr, w, n) in enumerate(specs):
        po = orc.params(v0=v, pressure=pr, weight=w)
        E, I = states[i]
        pot, ef = ens.grid(i)
        ce, ci, cnt = ens.measurements(i)
        # densities of the initial state (cumul after one measured step)
        want = orc.deposit(E[0], po)
        assert np.abs(ce - want).max() <= 2e-6 * want.max()
        if ion_step:
            want = orc.deposit(I[0], po)
            assert np.abs(ci - want).max() <= 2e-6 * want.max()
        pw, ew = orc.poisson_scan_f64(ce, orc.deposit(I[0], po), v * np.cos(2 * np.pi * t0 / 4000), po)
        assert np.abs(pot - pw).max() <= 1e-5 * np.abs(pw).max()
        assert np.abs(ef - ew).max() <= 1e-5 * np.abs(ew).max()
        cs, se, si = orc.cross_sections(po.gas_density)
        ge, gi = ens.download(i, 0), ens.download(i, 1)
        cnt_e = _push_ref(orc, po, ef, E, "e")[0]
        re1 = ens.rng_download(i, 0, caps[i])
        ri1 = ens.rng_download(i, 1, caps[i])
        if compacted:
            born_x, ions = _replay_species(orc, po, se, cs, E, ef, ge, None, rng0[i][0], re1, "e",
                                           src_slots=_source_slots(orc, po, ef, E, "e"))
        else:
            born_x, ions = _replay_species(orc, po, se, cs, E, ef, ge, ens.slots(i, 0), rng0[i][0], re1, "e")
        assert ne[i] == len(ge[0]) == cnt_e + len(born_x)
        if ion_step:
            cnt_i, pushed_i, _ = _push_ref(orc, po, ef, I, "i")
        else:
            cnt_i, pushed_i = len(I[0]), I
        assert ni[i] == len(gi[0]) == cnt_i + len(born_x)
        _, old_ion, n_amb, amb = _split_ions(gi, pushed_i[0])
        if ion_step:
            sel = [a[old_ion] for a in gi]
            if compacted:
                _replay_species(orc, po, si, cs, I, ef, sel, None, rng0[i][1], ri1, "i", missing=n_amb,
                                src_slots=_source_slots(orc, po, ef, I, "i"))
            else:
                _replay_species(orc, po, si, cs, I, ef, sel, ens.slots(i, 1)[old_ion], rng0[i][1], ri1, "i",
                                missing=n_amb)
        else:
            keep = ~np.isin(I[0], amb)
            assert_same_multiset([a[old_ion] for a in gi], [a[keep] for a in I])
            for q in range(4):
                np.testing.assert_array_equal(ri1[q], rng0[i][1][q])
        _check_new_ions(gi, pushed_i[0], born_x, ions, check_v=True)
        born_total += len(born_x)
        assert cnt[6] == 1 and cnt[7] == int(ion_step) and cnt[4] > 0
        assert (cnt[5] > 0) == ion_step
    assert born_total > 0
    ens.check()
    ens.close()


def test_ensemble_graph_replays_match_eager(ca):
    """Graph replays (4*N_SUB steps each) against the same steps run eagerly,
    with eager work between the replays: one group of all instances, and
    each instance a group of its own that runs two periods before the
    next one (temporal blocking), on one stream and dealt to two streams.
    At 1e-9 Pa the discharges are
    collisionless (collision probability ~1e-13 per particle-step), so all
    runs see the same fields up to the summation order of the density
    atomics.  The weights put both at the reference case's plasma density,
    so that the plasma shields the RF voltage and keeps its electrons."""
    specs = [(250.0, 1.5e5, 40_000), (150.0, 3.0e5, 20_000)]
    ps = [ca.params(v0=v, pressure=1e-9, weight=w, precision=2) for v, w, _ in specs]
    caps = [int(n * 1.5) for *_, n in specs]
    ens = []
    for graph, slots, periods, streams in ((False, 0, 0, 0), (True, 0xFFFFFFFF, 0, 0), (True, 1, 2, 1),
                                           (True, 1, 2, 2)):
        e = ca.Ensemble(ps, caps, use_graph=graph, group_slots=slots, group_periods=periods,
                        group_streams=streams)
        for i, (*_, n) in enumerate(specs):
            E, I = synthetic_state(n, int(n * 1.05), seed=7 + i)
            e.upload(i, 0, *E)
            e.upload(i, 1, *I)
        ens.append(e)
    for rep in range(3):
        for e in ens:
            e.run(160)
            e.synchronize()
        ca_ = ens[0].counts()
        for other in ens[1:]:
            cb = other.counts()
            print("rep", rep, "counts eager", ca_, "graph", cb, flush=True)
            for i in range(len(specs)):
                assert ca_[0][i] > 0.5 * specs[i][2]
                assert abs(int(ca_[0][i]) - int(cb[0][i])) <= 2 and abs(int(ca_[1][i]) - int(cb[1][i])) <= 2
                a = sort_rows(*ens[0].download(i, 0))
                b = sort_rows(*other.download(i, 0))
                m = min(len(a), len(b))
                assert np.abs(a[:m, 0] - b[:m, 0]).max() < 1e-6
                pa, pb = ens[0].grid(i)[0], other.grid(i)[0]
                assert np.abs(pa - pb).max() < 1e-3 * np.abs(pa).max()
        for e in ens:
            e.run(3)
            e.run(77)
    for e in ens:
        e.check()
        e.close()


def test_ensemble_timed_window_equals_plain_run(ca):
    """A timed window of a temporally blocked ensemble first captures every
    group's graphs (a dry pass) and must leave the host state where a plain
    run of the same steps leaves it (collisionless, as the graph test)."""
    specs = [(250.0, 1.5e5, 40_000), (150.0, 3.0e5, 20_000), (300.0, 2.0e5, 30_000)]
    ps = [ca.params(v0=v, pressure=1e-9, weight=w, precision=2) for v, w, _ in specs]
    caps = [int(n * 1.5) for *_, n in specs]
    ens = []
    for _ in range(2):
        e = ca.Ensemble(ps, caps, group_slots=80_000, group_periods=2)
        for i, (*_, n) in enumerate(specs):
            E, I = synthetic_state(n, int(n * 1.05), seed=17 + i)
            e.upload(i, 0, *E)
            e.upload(i, 1, *I)
        ens.append(e)
    r = ens[0].time_steps(240)
    ens[1].run(240)
    assert r["particle_steps"] > 0 and r["ms_total"] > 0
    for extra in (0, 160):
        for e in ens:
            e.run(extra)
            e.synchronize()
        ca_, cb = ens[0].counts(), ens[1].counts()
        for i in range(len(specs)):
            assert abs(int(ca_[0][i]) - int(cb[0][i])) <= 2 and abs(int(ca_[1][i]) - int(cb[1][i])) <= 2
            a = sort_rows(*ens[0].download(i, 0))
            b = sort_rows(*ens[1].download(i, 0))
            m = min(len(a), len(b))
            assert np.abs(a[:m, 0] - b[:m, 0]).max() < 1e-6
            pa, pb = ens[0].grid(i)[0], ens[1].grid(i)[0]
            assert np.abs(pa - pb).max() < 1e-3 * np.abs(pa).max()
    for e in ens:
        e.check()
        e.close()


def test_ensemble_deferred_rng_write_back_equals_every_step(ca):
    """The ensemble with the default deferred electron RNG write-back and
    with a write-back every step (rng_defer = 1): cold electrons (1 eV) and
    a 0.01 V drive, so electrons collide elastically but cannot ionise and
    every slot's contents are deterministic.  From t = 1 (no compaction in
    the window): eager steps, a captured period per instance group, eager
    steps, then a timed window ending off a write-back boundary.  Every
    slot -- particles and RNG words -- of every instance must be equal (up to
    1e-4 of the slots, for the last bit that the density flush's
    floating-point atomics may move in an f32 field value)."""
    specs = [(0.01, 2.0e3, 60_000), (0.01, 4.0e3, 30_000)]
    ps = [ca.params(v0=v, weight=w, precision=2) for v, w, _ in specs]
    caps = [int(n * 1.5) for *_, n in specs]
    runs = []
    for defer in (1, 8):
        e = ca.Ensemble(ps, caps, group_slots=1, group_periods=1, rng_defer=defer)
        try:
            e.set_time(1)
            for i, (*_, n) in enumerate(specs):
                E, I = synthetic_state(n, int(n * 1.05), seed=27 + i, te_ev=1.0)
                e.upload(i, 0, *E)
                e.upload(i, 1, *I)
            print(f"ensemble rng_defer={defer}: 170 steps run, 27 timed", flush=True)
            e.run(170)
            e.time_steps(27)
            e.synchronize()
            e.check()
            runs.append([(e.download(i, 0), e.rng_download(i, 0, caps[i]), e.download(i, 1),
                          e.rng_download(i, 1, caps[i]), e.measurements(i)[2]) for i in range(len(specs))])
        finally:
            e.close()
    for i in range(len(specs)):
        ea, rea, ia, ria, ca_ = runs[0][i]
        eb, reb, ib, rib, cb = runs[1][i]
        assert ca_[4] > 1000, ca_   # electron collisions happened
        for a, b, ra, rb in ((ea, eb, rea, reb), (ia, ib, ria, rib)):
            assert [len(u) for u in a] == [len(v) for v in b]
            n = len(a[0])
            differ = np.zeros(len(ra[0]), bool)
            for w in range(4):
                differ[:n] |= np.asarray(a[w]).view(np.uint32) != np.asarray(b[w]).view(np.uint32)
                differ |= ra[w] != rb[w]
            assert differ.sum() <= max(2, n // 10_000), (i, differ.sum(), n)


def test_ensemble_candidate_masks_equal_plain_runs(ca):
    """The ensemble with the collision-candidate masks (the default) and
    without them (cand_masks = -1), both with the deferred RNG write-back:
    two instances at 3 Pa and 30 Pa, so that their mask thresholds (each
    instance's largest collision probability) differ by ten, cold electrons
    (1 eV) and a 0.01 V drive as in the test above, so that every slot's
    contents are deterministic.  From t = 1: eager steps, captured periods,
    eager steps, then a timed window; every slot of every instance --
    particles and RNG words -- must be equal (the same slack as above)."""
    specs = [(3.0, 2.0e3, 60_000), (30.0, 4.0e3, 30_000)]
    ps = [ca.params(v0=0.01, pressure=p, weight=w, precision=2) for p, w, _ in specs]
    caps = [int(n * 1.5) for *_, n in specs]
    runs = []
    for masks in (-1, 0):
        e = ca.Ensemble(ps, caps, group_slots=1, group_periods=1, cand_masks=masks)
        try:
            e.set_time(1)
            for i, (*_, n) in enumerate(specs):
                E, I = synthetic_state(n, int(n * 1.05), seed=31 + i, te_ev=1.0)
                e.upload(i, 0, *E)
                e.upload(i, 1, *I)
            print(f"ensemble cand_masks={masks}: 250 steps run, 27 timed", flush=True)
            e.run(250)
            e.time_steps(27)
            e.synchronize()
            e.check()
            runs.append([(e.download(i, 0), e.rng_download(i, 0, caps[i]), e.download(i, 1),
                          e.rng_download(i, 1, caps[i]), e.measurements(i)[2]) for i in range(len(specs))])
        finally:
            e.close()
    for i in range(len(specs)):
        ea, rea, ia, ria, ca_ = runs[0][i]
        eb, reb, ib, rib, cb = runs[1][i]
        assert ca_[4] > 1000 and abs(int(ca_[4]) - int(cb[4])) <= 2, (ca_, cb)  # the same collisions
        for a, b, ra, rb in ((ea, eb, rea, reb), (ia, ib, ria, rib)):
            assert [len(u) for u in a] == [len(v) for v in b]
            n = len(a[0])
            differ = np.zeros(len(ra[0]), bool)
            for w in range(4):
                differ[:n] |= np.asarray(a[w]).view(np.uint32) != np.asarray(b[w]).view(np.uint32)
                differ |= ra[w] != rb[w]
            assert differ.sum() <= max(2, n // 10_000), (i, differ.sum(), n)


def test_ensemble_with_an_empty_instance(ca):
    """An instance without particles between two populated ones, each a
    group of its own on the graph path and all in one group eagerly: the
    empty one stays empty with a zero field apart from the drive, and the
    others are unaffected by it."""
    specs = [(250.0, 1.5e5, 20_000), (250.0, 1.5e5, 0), (150.0, 3.0e5, 20_000)]
    ps = [ca.params(v0=v, pressure=1e-9, weight=w, precision=2) for v, w, _ in specs]
    caps = [max(4096, int(n * 1.5)) for *_, n in specs]
    runs = []
    for graph, slots in ((True, 1), (False, 0xFFFFFFFF)):
        e = ca.Ensemble(ps, caps, use_graph=graph, group_slots=slots)
        for i, (*_, n) in enumerate(specs):
            if n:
                E, I = synthetic_state(n, int(n * 1.05), seed=41 + i)
                e.upload(i, 0, *E)
                e.upload(i, 1, *I)
        e.run(160)
        e.synchronize()
        e.check()
        ne, ni = e.counts()
        assert ne[1] == 0 and ni[1] == 0
        assert ne[0] > 10_000 and ne[2] > 10_000
        pot, ef = e.grid(1)
        assert np.isfinite(pot).all() and np.isfinite(ef).all()
        # no space charge: the potential is the linear drop of the drive
        lin = np.linspace(pot[0], 0.0, len(pot))
        assert np.abs(pot - lin).max() < 1e-4 * max(1.0, abs(pot[0]))
        runs.append((ne.copy(), ni.copy()))
        e.close()
    for i in (0, 2):
        assert abs(int(runs[0][0][i]) - int(runs[1][0][i])) <= 2


def test_ensemble_init_cycle_statistics(ca):
    """Four copies of the reference case from 1000 e + 1000 i at rest end the
    init cycle near eduPIC's published 888 e / 3345 i; a fifth instance at
    100 V / 3 Pa runs alongside without disturbing them."""
    p_ref = ca.params()
    ps = [p_ref] * 4 + [ca.params(v0=100.0, pressure=3.0)]
    ens = ca.Ensemble(ps, [20_000] * 5, use_graph=True)
    rng = np.random.default_rng(9)
    z = np.zeros(1000, np.float32)
    for i in range(5):
        ens.upload(i, 0, (rng.random(1000) * 0.025).astype(np.float32), z, z, z)
        ens.upload(i, 1, (rng.random(1000) * 0.025).astype(np.float32), z, z, z)
    ens.run(4000)
    ens.check()
    ne, ni = ens.counts()
    assert abs(ne[:4].mean() - 888) < 0.15 * 888, ne
    assert abs(ni[:4].mean() - 3345) < 0.15 * 3345, ni
    assert ne[4] != ne[0] and ne[4] > 0
    ens.close()


def test_ensemble_rejects_mismatched_grids(ca):
    with pytest.raises(ca.PicError):
        ca.Ensemble([ca.params(), ca.params(n_g=200)], [1000, 1000])
