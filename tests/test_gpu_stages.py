"""GPU parity of every stage entry point against the CPU oracle.

Integer work (xoshiro streams, jump seeding, compaction, histograms) is
compared bit-exactly; the push is bit-exact too because the kernel and the
oracle spell the same FP32 operations with explicit fmaf; the Poisson solve
is compared with the FP64 Thomas solve at the tolerance each precision
mode guarantees; collisions are replayed slot by slot.
"""
import numpy as np
import pytest

from helpers import SEED_E, SEED_I, assert_same_multiset, sort_rows, synthetic_state

pytestmark = pytest.mark.gpu

E_CHARGE = 1.60217662e-19


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


def dev(ca, a, dtype=None):
    return ca.DeviceBuffer.from_host(np.asarray(a) if dtype is None else np.asarray(a, dtype))


def counter(ca, v=0):
    return dev(ca, np.array([v], np.uint32))


# ---------------------------------------------------------------- RNG
def test_xoshiro_matches_reference_golden(ca):
    from test_oracle_rng import golden
    out = ca.DeviceBuffer.zeroed(10000, np.uint32)
    st = ca.DeviceBuffer.zeroed(4, np.uint32)
    ca.kernels.xoshiro128p_next_n([1, 1, 1, 1], 10000, out, st)
    np.testing.assert_array_equal(out.to_host_vec(), golden())
    s = np.ones(4, np.uint32)
    from oracle import oracle as orc
    for _ in range(10000):
        orc.xo_next(s)
    np.testing.assert_array_equal(st.to_host_vec(), s)


@pytest.mark.parametrize("n", [1, 257, 70_000])
def test_device_seeding_equals_sequential_jumps(ca, orc, n):
    rng = [ca.DeviceBuffer.zeroed(n, np.uint32) for _ in range(4)]
    for master in (SEED_E, SEED_I):
        ca.kernels.seed_streams(master, n, rng)
        want = orc.seed_streams(master, n)
        for w in range(4):
            np.testing.assert_array_equal(rng[w].to_host_vec(), want[w])


def test_seeding_rejects_zero_master(ca):
    rng = [ca.DeviceBuffer.zeroed(4, np.uint32) for _ in range(4)]
    with pytest.raises(ca.PicError):
        ca.kernels.seed_streams([0, 0, 0, 0], 4, rng)


# ---------------------------------------------------------------- deposition
@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("n", [0, 1, 1000, 300_000])
def test_get_density_vs_oracle(ca, orc, f64, n):
    p = ca.params()
    po = orc.params()
    rng = np.random.default_rng(n)
    x = (rng.random(max(n, 1)) * p.gap).astype(np.float32)[:n]
    if n >= 2:
        x[0] = p.gap_f      # x == L lands on the last node
        x[1] = 0.0
    d_x = dev(ca, x if n else np.zeros(1, np.float32))
    dens = ca.DeviceBuffer.zeroed(p.n_g, np.float64 if f64 else np.float32)
    ca.kernels.get_density(d_x, dens, counter(ca, n), p, density_f64=f64)
    got = dens.to_host_vec().astype(np.float64)
    want = orc.deposit(x, po)
    # each workgroup sums its particles in f32 LDS before the (f32 or f64)
    # global flush: a few f32 roundings per node
    scale = max(want.max(), 1.0)
    assert np.abs(got - want).max() <= 2e-6 * scale
    # charge invariant
    wf = float(np.float32(p.weight_factor))
    total = got[1:-1].sum() + 0.5 * (got[0] + got[-1])
    assert abs(total - n * wf) <= 1e-5 * max(n, 1) * wf


# ---------------------------------------------------------------- Poisson
def _densities(p, seed=0):
    rng = np.random.default_rng(seed)
    from helpers import refcase_density
    _, ne, ni = refcase_density()
    xs = np.linspace(0, 1, p.n_g)
    x0 = np.linspace(0, 1, len(ne))
    n_e = np.interp(xs, x0, ne) * (1 + 0.02 * rng.standard_normal(p.n_g))
    n_i = np.interp(xs, x0, ni) * (1 + 0.02 * rng.standard_normal(p.n_g))
    return np.maximum(n_e, 0), np.maximum(n_i, 0)


@pytest.mark.parametrize("n_g", [400, 4000])
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_poisson_vs_fp64_thomas(ca, orc, n_g, prec):
    """FP64 and compensated scans: phi and E within 1e-5 relative (of their
    max norm) of the FP64 Thomas solve; FP32 within its i*eps*|pot0| bound."""
    p = ca.params(n_g=n_g, precision=prec)
    po = orc.params(n_g=n_g)
    n_e, n_i = _densities(p, n_g + prec)
    dt = np.float64 if prec == 2 else np.float32
    ne32, ni32 = n_e.astype(dt), n_i.astype(dt)
    pot = ca.DeviceBuffer.zeroed(n_g, np.float32)
    ef = ca.DeviceBuffer.zeroed(n_g, np.float32)
    for pot0 in (250.0, -173.25, 0.0):
        ca.kernels.solve_poisson_scan(dev(ca, ne32), dev(ca, ni32), pot, ef, pot0, p)
        rho = E_CHARGE * (ni32.astype(np.float64) - ne32.astype(np.float64))
        p0 = pot0 if prec == 2 else float(np.float32(pot0))
        pw, ew = orc.poisson_thomas(rho, p0, po)
        pg, eg = pot.to_host_vec().astype(np.float64), ef.to_host_vec().astype(np.float64)
        perr = np.abs(pg - pw).max() / np.abs(pw).max()
        eerr = np.abs(eg - ew).max() / np.abs(ew).max()
        if prec in (1, 2):
            assert perr < 1e-5 and eerr < 1e-5, (perr, eerr)
        else:
            bound = 4 * n_g * np.finfo(np.float32).eps
            assert perr < bound, perr
        assert pg[0] == np.float32(p0) and pg[-1] == 0.0


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_poisson_flexible_hakim_order_two(ca, prec):
    """The reference's Hakim case (main.rs:1887-1981) on the GPU solver."""
    errs, dxs = [], []
    for n_elem in (8, 16, 32, 64):
        n = n_elem + 1
        dx = 1.0 / n_elem
        x = np.arange(n) * dx
        src = dev(ca, (1.0 - 2.0 * x * x).astype(np.float32))
        pot = ca.DeviceBuffer.zeroed(n, np.float32)
        ca.kernels.solve_poisson_dps_flexible(src, pot, n, 0.0, 0.0, dx, precision=prec)
        exact = x * x / 2 - x ** 4 / 6 - x / 3
        errs.append(np.abs(pot.to_host_vec()[1:-1] - exact[1:-1]).mean())
        dxs.append(dx)
    orders = [np.log(errs[i] / errs[i + 1]) / np.log(dxs[i] / dxs[i + 1]) for i in range(3)]
    assert all(abs(o - 2.0) < 0.15 for o in orders), orders


# ---------------------------------------------------------------- push + compaction
def _push_case(ca, orc, n, species, seed, field_scale=1.0):
    p = ca.params()
    po = orc.params()
    (ex, evx, evy, evz), (ix, ivx, ivy, ivz) = synthetic_state(max(n, 1), max(n, 1), seed=seed)
    if n >= 100:   # a tenth of the particles within a few cells of either wall
        k = n // 10
        wall = np.random.default_rng(seed + 2).random(k) * 4 * p.dx
        ex = ex.copy()
        ix = ix.copy()
        ex[:k // 2] = wall[:k // 2]
        ex[k // 2:k] = p.gap - wall[k // 2:]
        ix[:k // 2] = wall[:k // 2]
        ix[k // 2:k] = p.gap - wall[k // 2:]
    if species == "e":
        x, vx, vy, vz = ex[:n], evx[:n] * 8, evy[:n], evz[:n]   # hot electrons cross the walls
        factor, dt, hm = p.factor_e, p.dt_e_f, 0.0
    else:
        x, vx, vy, vz = ix[:n], ivx[:n] * 300, ivy[:n], ivz[:n]
        factor, dt, hm = p.factor_i, p.dt_i_f, p.half_mi_over_e
    rng = np.random.default_rng(seed + 1)
    efield = (rng.standard_normal(p.n_g) * 2e4 * field_scale).astype(np.float32)
    src = [dev(ca, a if n else np.zeros(1, np.float32)) for a in (x, vx, vy, vz)]
    dst = [ca.DeviceBuffer.zeroed(max(n, 1), np.float32) for _ in range(4)]
    alive = counter(ca, 0)
    absc = ca.DeviceBuffer.zeroed(2, np.uint64)
    ifed = ca.DeviceBuffer.zeroed(400, np.uint32)
    ca.kernels.move_and_compact(dev(ca, efield), src, dst, alive, counter(ca, n), factor, dt, p,
                                half_m_over_e=hm, abs_counts=absc, ifed=ifed if species == "i" else None)
    cnt = int(alive.to_host_vec()[0])
    ifed_w = np.zeros(400, np.uint32)
    wc, want, wabs = orc.move_and_compact(efield, x, vx, vy, vz, factor, dt, po, mass_over_2e=hm,
                                          ifed=ifed_w if species == "i" else None)
    return cnt, [b.to_host_vec(cnt) for b in dst], absc.to_host_vec(), ifed.to_host_vec(), wc, want, wabs, ifed_w


@pytest.mark.parametrize("species", ["e", "i"])
@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 4097, 250_000])
def test_move_and_compact_bit_exact_multiset(ca, orc, species, n):
    cnt, got, gabs, gifed, wc, want, wabs, wifed = _push_case(ca, orc, n, species, seed=n)
    assert cnt == wc
    assert_same_multiset(got, want)
    np.testing.assert_array_equal(gabs, wabs)
    if species == "i":
        np.testing.assert_array_equal(gifed, wifed)
    assert n == 0 or wabs.sum() > 0 or n < 100


def test_move_and_compact_all_lost(ca, orc):
    cnt, got, gabs, _, wc, want, wabs, _ = _push_case(ca, orc, 5000, "e", seed=3, field_scale=1e6)
    assert cnt == wc
    assert_same_multiset(got, want)
    np.testing.assert_array_equal(gabs, wabs)


# ---------------------------------------------------------------- collisions
def _energetic_electrons(n, seed):
    rng = np.random.default_rng(seed)
    en = rng.uniform(0.0, 80.0, n)                 # eV, covers all three processes
    v = np.sqrt(2 * en * E_CHARGE / 9.10938356e-31)
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    vx, vy, vz = (v * d).astype(np.float32)
    vx[: n // 50] = 0.0                             # exercise the gx == 0 / gy == 0 branches
    vy[n // 50: n // 25] = 0.0
    x = (rng.random(n) * 0.025).astype(np.float32)
    return x, vx, vy, vz


def test_check_collisions_e_slot_replay(ca, orc):
    p = ca.params()
    po = orc.params()
    n, n_i0 = 200_000, 1000
    cap = n + 50_000
    x, vx, vy, vz = _energetic_electrons(n, 11)
    rng_words = orc.seed_streams(SEED_E, cap)
    cs, se, si = orc.cross_sections(po.gas_density)
    E = ca.kernels.Species(cap)
    I = ca.kernels.Species(cap)
    E.upload(x, vx, vy, vz, rng_words)
    alive_e, alive_i = counter(ca, n), counter(ca, n_i0)
    ovf = counter(ca, 0)
    ncoll = ca.DeviceBuffer.zeroed(1, np.uint64)
    d_se, d_cs = dev(ca, se), dev(ca, cs)
    ca.kernels.check_collisions_e(d_se, d_cs, counter(ca, n), E, I, alive_e, alive_i, ovf, ncoll, p)
    ne_gpu, ni_gpu = int(alive_e.to_host_vec()[0]), int(alive_i.to_host_vec()[0])
    gx, gvx, gvy, gvz = E.download(ne_gpu)
    grng = E.rng_host(n)
    # oracle on copies
    ox, ovx, ovy, ovz = (a.copy() for a in (x, vx, vy, vz))
    ox = np.concatenate([ox, np.zeros(cap - n, np.float32)])
    ovx, ovy, ovz = (np.concatenate([a, np.zeros(cap - n, np.float32)]) for a in (ovx, ovy, ovz))
    orng = [w.copy() for w in rng_words]
    oi = [np.zeros(cap, np.float32) for _ in range(4)]
    ne_o, ni_o, ncoll_o = orc.collide_e(se, cs, ox, ovx, ovy, ovz, orng, *oi, n, n_i0, po)
    # per-slot: RNG state bit-exact where the same number of draws happened
    same = np.ones(n, bool)
    for w in range(4):
        same &= grng[w] == orng[w][:n]
    assert (~same).sum() <= max(2, n // 100000), (~same).sum()
    v_in = np.sqrt(vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2 + vz.astype(np.float64) ** 2)
    dv = np.sqrt((gvx[:n] - ovx[:n]) ** 2 + (gvy[:n] - ovy[:n]) ** 2 + (gvz[:n] - ovz[:n]) ** 2)
    ok = dv[same] <= 1e-4 * v_in[same] + 1.0
    assert ok.all(), (dv[same][~ok][:5], v_in[same][~ok][:5])
    assert abs(int(ncoll.to_host_vec()[0]) - ncoll_o) <= (~same).sum()
    assert ncoll_o > 1000
    # spawned electrons and ions: same multisets (within the transcendental tolerance)
    assert ne_gpu - n == ne_o - n and ni_gpu == ni_o
    assert int(ovf.to_host_vec()[0]) == 0
    sg = sort_rows(gx[n:ne_gpu], gvx[n:ne_gpu], gvy[n:ne_gpu], gvz[n:ne_gpu])
    so = sort_rows(ox[n:ne_o], ovx[n:ne_o], ovy[n:ne_o], ovz[n:ne_o])
    np.testing.assert_array_equal(sg[:, 0], so[:, 0])          # x is copied, bit-exact
    np.testing.assert_allclose(sg[:, 1:], so[:, 1:], rtol=0, atol=1e-4 * 6e6)
    gi = I.download(ni_gpu)
    ig = sort_rows(*(a[n_i0:ni_gpu] for a in gi))
    io = sort_rows(*(a[n_i0:ni_o] for a in oi))
    np.testing.assert_array_equal(ig[:, 0], io[:, 0])
    np.testing.assert_allclose(ig[:, 1:], io[:, 1:], rtol=0, atol=1e-3 * 270 * 6)


def test_check_collisions_e_capacity_guard(ca, orc):
    """Appends beyond MAX_PARTICLES are dropped and raise the overflow flag
    (the reference does not guard them, main.rs:1133,1142)."""
    p = ca.params()
    po = orc.params()
    n = 50_000
    x, vx, vy, vz = _energetic_electrons(n, 12)
    cs, se, si = orc.cross_sections(po.gas_density)
    E = ca.kernels.Species(n + 3)
    I = ca.kernels.Species(5)
    E.upload(x, vx, vy, vz, orc.seed_streams(SEED_E, n + 3))
    alive_e, alive_i, ovf = counter(ca, n), counter(ca, 0), counter(ca, 0)
    ca.kernels.check_collisions_e(dev(ca, se), dev(ca, cs), counter(ca, n), E, I, alive_e, alive_i,
                                  ovf, None, p)
    assert int(ovf.to_host_vec()[0]) == 1
    assert int(alive_e.to_host_vec()[0]) > n + 3


def test_check_collisions_i_slot_replay(ca, orc):
    p = ca.params()
    po = orc.params()
    n = 200_000
    rng = np.random.default_rng(13)
    en = rng.uniform(0.0, 100.0, n)
    v = np.sqrt(2 * en * E_CHARGE / 6.63352090e-26)
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    vx, vy, vz = (v * d).astype(np.float32)
    words = orc.seed_streams(SEED_I, n)
    cs, se, si = orc.cross_sections(po.gas_density)
    I = ca.kernels.Species(n)
    I.upload(np.zeros(n, np.float32), vx, vy, vz, words)
    ncoll = ca.DeviceBuffer.zeroed(1, np.uint64)
    ca.kernels.check_collisions_i(dev(ca, si), dev(ca, cs), counter(ca, n), I, ncoll, p)
    _, gvx, gvy, gvz = I.download(n)
    grng = I.rng_host(n)
    ov = [a.copy() for a in (vx, vy, vz)]
    orng = [w.copy() for w in words]
    nco = orc.collide_i(si, cs, *ov, orng, n, po)
    same = np.ones(n, bool)
    for w in range(4):
        same &= grng[w] == orng[w]
    assert (~same).sum() <= max(2, n // 100000)
    dv = np.sqrt((gvx - ov[0]) ** 2 + (gvy - ov[1]) ** 2 + (gvz - ov[2]) ** 2)
    assert (dv[same] <= 1e-4 * (v[same] + 270.0 * 4)).all()
    assert abs(int(ncoll.to_host_vec()[0]) - nco) <= (~same).sum()
    assert nco > 500


# ---------------------------------------------------------------- measurement
def test_accumulate_eepf_bit_exact(ca, orc):
    p = ca.params()
    po = orc.params()
    (x, vx, vy, vz), _ = synthetic_state(300_000, 1, seed=21)
    rng = np.random.default_rng(22)
    efield = (rng.standard_normal(p.n_g) * 3e3).astype(np.float32)
    eepf = ca.DeviceBuffer.zeroed(2000, np.uint32)
    bufs = [dev(ca, a) for a in (x, vx, vy, vz)]
    ca.kernels.accumulate_eepf(*bufs, dev(ca, efield), counter(ca, len(x)), eepf, p)
    want = orc.accumulate_eepf(x, vx, vy, vz, efield, po)
    np.testing.assert_array_equal(eepf.to_host_vec(), want)
    assert want.sum() > 10000


def test_accumulate_density(ca):
    d = np.random.default_rng(1).random(400).astype(np.float32)
    cum = dev(ca, np.arange(400, dtype=np.float64))
    ca.kernels.accumulate_density(dev(ca, d), cum, 400)
    ca.kernels.accumulate_density(dev(ca, d), cum, 400)
    np.testing.assert_array_equal(cum.to_host_vec(), np.arange(400) + 2 * d.astype(np.float64))


def test_null_arguments_are_errors(ca):
    p = ca.params()
    with pytest.raises(ca.PicError):
        ca.kernels.get_density(None, None, None, p)
