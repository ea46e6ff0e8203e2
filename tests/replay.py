"""Per-slot replay of one RF step against the CPU oracle, shared by the GPU
parity tests of the simulation object and of the caller-owned step: every
GPU particle is matched to its source by its pushed position and its
collision replayed with the RNG words of its slot."""
import numpy as np

from helpers import assert_same_multiset  # noqa: F401  (re-exported for the tests)


def _push_ref(orc, po, efield, parts, species):
    x, vx, vy, vz = parts
    if species == "e":
        return orc.move_and_compact(efield, x, vx, vy, vz, po.factor_e, po.dt_e_f, po)
    return orc.move_and_compact(efield, x, vx, vy, vz, po.factor_i, po.dt_i_f, po,
                                mass_over_2e=np.float32(0.5 * 6.63352090e-26 / 1.60217662e-19))


def _check_slot_draws(orc, rng0, rng1, slots, cnt, draws):
    """Every occupied slot advanced its stream `draws` times, every other slot
    is untouched."""
    s4 = np.zeros(4, np.uint32)
    for j in slots[:: max(1, cnt // 2000)]:
        s4[:] = [rng0[w][j] for w in range(4)]
        for _ in range(draws):
            orc.xo_next(s4)
        assert [rng1[w][j] for w in range(4)] == s4.tolist()
    free = np.ones(len(rng0[0]), bool)
    free[slots] = False
    for w in range(4):
        np.testing.assert_array_equal(rng1[w][free], rng0[w][free])


def _replay_species(orc, po, stot, cs, parts, efield, got, slots, rng0, rng1, name, missing=0,
                    src_slots=None):
    """Match every GPU particle to its source by its pushed position and
    replay its collision with the RNG words of its slot.

    f32 positions repeat (a few hundred coincidences in 2e5 particles), so the
    check runs over positions held by exactly one pushed particle.  There a
    single GPU particle is a survivor (replayed); two are an ionisation: the
    parent's slot drew its numbers, the newborn's slot (or staging entry,
    slot -1) did not.  Returns the positions at which electrons were born
    (one entry per newborn) and the ion velocities of the checked pairs.
    `missing`: pushed particles whose GPU rows were left out of `got`.
    `src_slots`: for a step that compacted after colliding, the slot whose
    stream pushed source p used (its slot before the compaction); the parent
    of a pair is then the one with the replayed velocity."""
    cnt, pushed, _ = _push_ref(orc, po, efield, parts, name)
    px, pvx, pvy, pvz = pushed
    ux, uidx, ucnt = np.unique(px, return_index=True, return_counts=True)
    gx, gvx, gvy, gvz = got
    n = len(gx)
    gux, ginv, gcnt = np.unique(gx, return_inverse=True, return_counts=True)
    k = np.searchsorted(ux, gux)
    k = np.minimum(k, len(ux) - 1)
    in_pushed = ux[k] == gux
    assert in_pushed.all(), "a GPU particle has no pushed source"
    mo = ucnt[k]                      # pushed sources per GPU position
    born = gcnt - mo                  # newborns per GPU position
    assert (born >= 0).all()
    assert born.sum() == n - cnt + missing

    def moved(j):
        sl = slots[j]
        return sl >= 0 and any(rng1[w][sl] != rng0[w][sl] for w in range(4))

    def replay(src, slot):
        v = np.array([pvx[src], pvy[src], pvz[src]], np.float32)
        st = np.array([rng0[w][slot] for w in range(4)], np.uint32)
        if name == "e":
            kind, vo, so, sev, siv = orc.collide_e_one(stot, cs, v, st, po)
            return kind, v, vo, so, sev, siv
        kind, vo, so = orc.collide_i_one(stot, cs, v, st, po)
        return kind, v, vo, so, None, None

    def rng_slot(j, g):
        return src_slots[uidx[k[g]]] if src_slots is not None else slots[j]

    single = (gcnt == 1) & (mo == 1)
    checked = mism = 0
    step = max(1, n // 20000)
    for j in range(0, n, step):
        g = ginv[j]
        if not single[g]:
            continue
        rs = rng_slot(j, g)
        kind, v, vo, so, _, _ = replay(uidx[k[g]], rs)
        checked += 1
        if so.tolist() != [rng1[w][rs] for w in range(4)]:
            mism += 1
            continue
        gv = np.array([gvx[j], gvy[j], gvz[j]], np.float64)
        assert np.linalg.norm(gv - vo) <= 1e-4 * max(np.linalg.norm(v), 2000.0), (name, j, gv, vo)
    assert checked > 0.9 * (n // step) - 0.02 * n
    assert mism <= max(2, checked // 10000)
    ions = {}
    if name == "e":
        pair_groups = np.nonzero((gcnt == 2) & (mo == 1))[0]
        members = {}
        for j in np.nonzero(np.isin(ginv, pair_groups))[0]:
            members.setdefault(ginv[j], []).append(j)
        for g, pair in members.items():
            if src_slots is None:
                par = [j for j in pair if moved(j)]
                kid = [j for j in pair if not moved(j)]
                assert len(par) == 1 and len(kid) == 1, (gux[g], pair)
                kind, v, vo, so, sev, siv = replay(uidx[k[g]], slots[par[0]])
            else:
                rs = src_slots[uidx[k[g]]]
                kind, v, vo, so, sev, siv = replay(uidx[k[g]], rs)
                assert so.tolist() == [rng1[w][rs] for w in range(4)]
                gv = [np.array([gvx[j], gvy[j], gvz[j]], np.float64) for j in pair]
                swap = (np.linalg.norm(gv[1] - vo) + np.linalg.norm(gv[0] - sev)
                        < np.linalg.norm(gv[0] - vo) + np.linalg.norm(gv[1] - sev))
                par, kid = ([pair[1]], [pair[0]]) if swap else ([pair[0]], [pair[1]])
            assert kind == 3
            tol = 1e-4 * max(np.linalg.norm(v), 2000.0)
            assert np.linalg.norm(np.array([gvx[par[0]], gvy[par[0]], gvz[par[0]]]) - vo) <= tol
            assert np.linalg.norm(np.array([gvx[kid[0]], gvy[kid[0]], gvz[kid[0]]]) - sev) <= tol
            ions[float(gux[g])] = siv
        assert len(members) >= 0.8 * born.sum()
    born_x = np.repeat(gux, born)
    return born_x, ions


def _split_ions(gi, old_x):
    """Separate the GPU ions into the old ones (positions old_x) and the
    ionisation products.  f32 positions can coincide, so this works on
    multisets: returns the positions of the new ions (the multiset
    difference), a mask of the GPU ions that are certainly old (their
    position is held by old ions only), the number of old ions left out of
    that mask and the ambiguous positions."""
    ux, uc = np.unique(old_x, return_counts=True)
    gx, gc = np.unique(gi[0], return_counts=True)
    k = np.clip(np.searchsorted(ux, gx), 0, max(0, len(ux) - 1))
    have = np.where(ux[k] == gx, uc[k], 0) if len(ux) else np.zeros_like(gc)
    surplus = gc - have
    assert (surplus >= 0).all(), "an old ion is missing"
    amb_sel = (surplus > 0) & (have > 0)
    amb = gx[amb_sel]
    old_mask = np.isin(gi[0], old_x) & ~np.isin(gi[0], amb)
    return np.repeat(gx, surplus), old_mask, int(have[amb_sel].sum()), amb


def _check_new_ions(gi, pushed_ion_x, born_x, ions, check_v):
    """Ions that are not pushed survivors are the ionisation products: one
    per newborn electron, at the same position, with a thermal velocity."""
    new_x = _split_ions(gi, pushed_ion_x)[0]
    np.testing.assert_array_equal(np.sort(new_x), np.sort(born_x))
    new = ~np.isin(gi[0], pushed_ion_x)
    if check_v:
        for j in np.nonzero(new)[0]:
            x = float(gi[0][j])
            if x in ions:
                v = np.array([gi[1][j], gi[2][j], gi[3][j]])
                assert np.linalg.norm(v - ions[x]) <= 1e-3 * 270 * 6
