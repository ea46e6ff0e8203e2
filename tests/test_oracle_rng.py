"""Pins of the oracle's RNG functions (CPU only).

xoshiro128+ next(): the reference's own golden file (Rust-CUDA/xorshiro128c/
10000-c.txt, copied to tests/golden/xoshiro128p_10000.txt) and the
reference's xoshiro128p.c compiled into oracle/_ref.
jump(): closed form -- the state update is linear over GF(2), so jump() must
equal the 128x128 transition matrix raised to 2^64.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "xoshiro128p_10000.txt")


def golden():
    return np.loadtxt(GOLDEN, dtype=np.uint64).astype(np.uint32)


def test_next_matches_reference_golden(orc):
    s = np.ones(4, np.uint32)
    out = np.array([orc.xo_next(s) for _ in range(10000)], np.uint32)
    g = golden()
    assert len(g) == 10000
    assert out[:5].tolist() == [2, 1, 2048, 5245440, 2153779203]
    np.testing.assert_array_equal(out, g)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                    "libxoshiro128p.so")), reason="oracle/_ref not built")
def test_next_matches_reference_c_build(orc):
    ref = orc.ref_xoshiro()
    ref_out = np.array([ref.next() for _ in range(10000)], np.uint32)
    np.testing.assert_array_equal(ref_out, golden())


# ---- GF(2) closed form for jump() ------------------------------------------
def _state_to_bits(s):
    s = np.asarray(s, np.uint32)
    return np.array([(int(s[w]) >> b) & 1 for w in range(4) for b in range(32)], np.uint8)


def _bits_to_state(v):
    out = np.zeros(4, np.uint32)
    for w in range(4):
        acc = 0
        for b in range(32):
            acc |= int(v[w * 32 + b]) << b
        out[w] = acc
    return out


def _transition_matrix(orc):
    """Column c = next() applied to the unit state e_c (next() is pinned by
    the golden vectors above, so T is too)."""
    T = np.zeros((128, 128), np.uint8)
    for c in range(128):
        e = np.zeros(128, np.uint8)
        e[c] = 1
        s = _bits_to_state(e)
        orc.xo_next(s)
        T[:, c] = _state_to_bits(s)
    return T


def _matmul2(a, b):
    return (a.astype(np.int64) @ b.astype(np.int64) % 2).astype(np.uint8)


def _pow2k(T, k):
    M = T.copy()
    for _ in range(k):
        M = _matmul2(M, M)
    return M


def test_matrix_power_code_against_brute_force(orc):
    """T^(2^k) from repeated squaring equals 2^k explicit next() calls."""
    T = _transition_matrix(orc)
    rng = np.random.default_rng(1)
    for k in (0, 1, 5, 12, 16):
        M = _pow2k(T, k)
        s = rng.integers(1, 2**32, size=4, dtype=np.uint64).astype(np.uint32)
        want = s.copy()
        for _ in range(2**k):
            orc.xo_next(want)
        got = _bits_to_state(_matmul2(M, _state_to_bits(s)[:, None])[:, 0])
        np.testing.assert_array_equal(got, want)


def test_jump_equals_T_to_the_2_64(orc):
    T = _transition_matrix(orc)
    J = _pow2k(T, 64)
    rng = np.random.default_rng(7)
    for _ in range(8):
        s = rng.integers(0, 2**32, size=4, dtype=np.uint64).astype(np.uint32)
        want = _bits_to_state(_matmul2(J, _state_to_bits(s)[:, None])[:, 0])
        got = s.copy()
        orc.xo_jump(got)
        np.testing.assert_array_equal(got, want)


def test_seed_streams_are_successive_jumps(orc):
    master = [0x12345678, 0x11112222, 0x22223333, 0x33334444]  # main.rs:1667
    s0, s1, s2, s3 = orc.seed_streams(master, 6)
    st = np.array(master, np.uint32)
    for k in range(6):
        assert [s0[k], s1[k], s2[k], s3[k]] == st.tolist()
        orc.xo_jump(st)


def test_u32_to_real_closed_form(orc):
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.integers(0, 2**32, 2000, dtype=np.uint64),
                         [0, 1, 511, 512, 2**31, 2**32 - 1]]).astype(np.uint64)
    for x in xs:
        x = int(x)
        assert orc.u32_to_real(x) == (x >> 9) * 2.0**-23
    assert orc.u32_to_real(2**32 - 1) == 1.0 - 2.0**-23


def test_three_normal_is_standard_normal(orc):
    """Moments and a Kolmogorov-Smirnov test against N(0, sigma^2)."""
    from scipy import stats
    sigma = 269.90040554976775
    s = np.array([0x44445555, 0x55556666, 0x66667777, 0x77778888], np.uint32)
    z = np.array([orc.three_normal(s, sigma) for _ in range(20000)], np.float64) / sigma
    for col in range(3):
        assert abs(z[:, col].mean()) < 0.03
        assert abs(z[:, col].std() - 1.0) < 0.02
        assert stats.kstest(z[:, col], "norm").pvalue > 1e-3
