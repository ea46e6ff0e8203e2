// pic_device.h -- device-side building blocks shared by the gfx950 kernels:
// xoshiro128+ in registers, wave64 ballot compaction, CIC weights, and the
// electron / ion Monte Carlo collision bodies.
//
// All floating point here is compiled with -ffp-contract=off: the only fused
// multiply-adds are the __builtin_fmaf calls written out below, so the push
// arithmetic is reproducible operation for operation on the CPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pic {

constexpr int kWave = 64;
constexpr int kCsRanges = 1000000;
constexpr int kEepfBins = 2000;
constexpr int kIfedBins = 200;
constexpr float kPiF = 3.14159265358979f;
constexpr float kTwoPiF = 6.28318530717959f;
constexpr float kHalfPiF = 1.57079632679490f;
constexpr float kLn2F = 0.693147180559945f;
constexpr float kLog2eF = 1.4426950408889634f;  // log2(e) (main.rs:97)

// Physics factors the particle kernels need, all in f32 as the reference
// casts them (Rust-CUDA/src/main.rs:80-109).
struct KPhys {
    int n_g;
    float inv_dx, gap, wf;
    float factor_e, factor_i, dt_e, dt_i;
    float normal_range, f1, f2, half_me_over_e, half_mu_over_e, half_mi_over_e;
    float e_charge, me, exc_th, ion_th, de_cs;
    float inv_de_cs;  // 1 / de_cs rounded to nearest (energy_index's reciprocal)
    float min_x, max_x, de_eepf, eepf_kick;
    // multiplies the total cross-section table: 1 for the per-discharge
    // tables (sigma_tot * n_gas precomputed), the gas density for a shared
    // table of summed cross sections (ensembles of discharges at different
    // pressures)
    float stot_scale;
};

// ---------------------------------------------------------------- RNG
// xoshiro128+ 1.0 (main.rs:860-881); the state lives in VGPRs for the whole
// particle body and is loaded / stored once per particle.
struct Xo {
    uint32_t s0, s1, s2, s3;
};

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

__device__ __forceinline__ uint32_t xo_next(Xo &s) {
    const uint32_t result = s.s0 + s.s3;
    const uint32_t t = s.s1 << 9;
    s.s2 ^= s.s0;
    s.s3 ^= s.s1;
    s.s1 ^= s.s2;
    s.s0 ^= s.s3;
    s.s2 ^= t;
    s.s3 = rotl32(s.s3, 11);
    return result;
}

// n state transitions of xo_next, outputs discarded.
__device__ __forceinline__ void xo_skip(Xo &s, uint32_t n) {
    for (uint32_t r = 0; r < n; ++r) (void)xo_next(s);
}

// The inverse transition: the state one draw earlier.  Forward, with t =
// s1 << 9: s0' = s0 ^ s3 ^ s1, s1' = s1 ^ s2 ^ s0, s2' = s2 ^ s0 ^ t,
// s3' = rotl(s3 ^ s1, 11).  So s3 ^ s1 = rotr(s3', 11), s0 follows from s0',
// s1' ^ s2' = s1 ^ (s1 << 9), which y ^ (y << 9) ^ (y << 18) ^ (y << 27)
// inverts, and s2, s3 follow.
__device__ __forceinline__ void xo_prev(Xo &s) {
    const uint32_t d1 = rotl32(s.s3, 21);
    const uint32_t a = s.s0 ^ d1;
    const uint32_t y = s.s1 ^ s.s2;
    const uint32_t b = y ^ (y << 9) ^ (y << 18) ^ (y << 27);
    s.s2 = s.s1 ^ b ^ a;
    s.s3 = d1 ^ b;
    s.s0 = a;
    s.s1 = b;
}

__device__ __forceinline__ void xo_back(Xo &s, uint32_t n) {
    for (uint32_t r = 0; r < n; ++r) xo_prev(s);
}

// main.rs:883-885: the top 23 bits as the mantissa of [1,2), minus 1.
__device__ __forceinline__ float u32_to_real(uint32_t x) {
    return __uint_as_float(0x3F800000u | (x >> 9)) - 1.0f;
}

__device__ __forceinline__ float draw(Xo &s) { return u32_to_real(xo_next(s)); }

// Element i of a particle array as its base plus a 32-bit byte offset.
// The offset i * sizeof(T) is formed in 32 bits (for arrays of at most
// 2^30 elements of up to 4 bytes), so an access takes the global
// instructions' scalar-base form: one vector offset, shared by the arrays
// indexed alike, instead of 64-bit vector address arithmetic per array.
template <typename T>
__device__ __forceinline__ T &el(T *base, uint32_t i) {
    return *(T *)((char *)base + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T &el(const T *base, uint32_t i) {
    return *(const T *)((const char *)base + i * (uint32_t)sizeof(T));
}

__device__ __forceinline__ Xo rng_load(const uint32_t *__restrict__ r0, const uint32_t *__restrict__ r1,
                                       const uint32_t *__restrict__ r2, const uint32_t *__restrict__ r3,
                                       uint32_t i) {
    return Xo{r0[i], r1[i], r2[i], r3[i]};
}

__device__ __forceinline__ void rng_store(uint32_t *__restrict__ r0, uint32_t *__restrict__ r1,
                                          uint32_t *__restrict__ r2, uint32_t *__restrict__ r3,
                                          uint32_t i, const Xo &s) {
    r0[i] = s.s0;
    r1[i] = s.s1;
    r2[i] = s.s2;
    r3[i] = s.s3;
}

// The same through el(): for arrays of at most 2^30 slots (the fused
// kernels, whose callers are checked against that bound).
__device__ __forceinline__ Xo rng_load32(const uint32_t *__restrict__ r0, const uint32_t *__restrict__ r1,
                                         const uint32_t *__restrict__ r2, const uint32_t *__restrict__ r3,
                                         uint32_t i) {
    return Xo{el(r0, i), el(r1, i), el(r2, i), el(r3, i)};
}

__device__ __forceinline__ void rng_store32(uint32_t *__restrict__ r0, uint32_t *__restrict__ r1,
                                            uint32_t *__restrict__ r2, uint32_t *__restrict__ r3,
                                            uint32_t i, const Xo &s) {
    el(r0, i) = s.s0;
    el(r1, i) = s.s1;
    el(r2, i) = s.s2;
    el(r3, i) = s.s3;
}

// Three N(0, sigma^2) samples from four uniforms (main.rs:908-969): two
// Box-Muller pairs, the fourth normal discarded.  The angle goes straight to
// the hardware sine/cosine, which take their argument in revolutions, so
// cos(2 pi u) needs no range reduction; log uses the hardware log2.
__device__ __forceinline__ void three_normal(Xo &s, float sigma, float &a, float &b, float &c) {
    const float u1 = fmaxf(draw(s), 1e-9f);
    const float u2 = fmaxf(draw(s), 1e-9f);
    const float u3 = fmaxf(draw(s), 1e-9f);
    const float u4 = fmaxf(draw(s), 1e-9f);
    const float r1 = __builtin_sqrtf(-2.0f * kLn2F * __builtin_amdgcn_logf(u1));
    const float r2 = __builtin_sqrtf(-2.0f * kLn2F * __builtin_amdgcn_logf(u3));
    a = r1 * __builtin_amdgcn_cosf(u2) * sigma;
    b = r1 * __builtin_amdgcn_sinf(u2) * sigma;
    c = r2 * __builtin_amdgcn_cosf(u4) * sigma;
}

// ---------------------------------------------------------------- wave64
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (kWave - 1); }

// Number of set bits of `mask` below this lane (v_mbcnt_lo/hi).
__device__ __forceinline__ uint32_t mbcnt64(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// ---------------------------------------------------------------- grid
// CIC cell of a position, clamped to [0, n_g-2] so that x == L lands on the
// last two nodes (main.rs:385-386 and :600 would index N_G there).
__device__ __forceinline__ int cell_of(float pos, int n_g) {
    int q = (int)pos;
    q = q < 0 ? 0 : q;
    return q > n_g - 2 ? n_g - 2 : q;
}

// Field at x by linear interpolation and the leapfrog update (main.rs:597-603),
// spelled with explicit fmaf.
__device__ __forceinline__ float field_at(const float *ef, float x, float inv_dx, int n_g) {
    const float pos = x * inv_dx;
    const int q = cell_of(pos, n_g);
    const float c = pos - (float)q;
    return __builtin_fmaf(c, ef[q + 1], (1.0f - c) * ef[q]);
}

// ---------------------------------------------------------------- collisions
// The cross-section bin of an energy, min(CS_RANGES - 1, (int)(energy /
// DE_CS + 0.5)) as main.rs:1002-1005 computes it with an IEEE division.  The
// quotient is formed from the rounded reciprocal y = RN(1/DE_CS) with one
// fused correction (Markstein's method): q = RN(e*y), r = fma(-q, DE_CS, e)
// (exact), q' = RN(q + r*y), which is the correctly rounded quotient
// whenever it neither underflows nor overflows -- three operations instead of
// the division's ten.  A tiny energy gives bin 0 either way, and the clamp is
// written so that an overflowed (NaN) quotient lands in the last bin, as the
// division's infinity does; the bins are equal to the division's for every
// non-negative float (checked exhaustively on the CPU).
__device__ __forceinline__ int energy_index(float energy, const KPhys &k) {
    float q = energy * k.inv_de_cs;
    const float r = __builtin_fmaf(-q, k.de_cs, energy);
    q = __builtin_fmaf(r, k.inv_de_cs, q);
    const float c = q + 0.5f;
    return c < (float)(kCsRanges - 1) ? (int)(uint32_t)c : kCsRanges - 1;
}

struct Frame {
    float st, ct, sp, cp;
};

// Euler angles of the relative velocity (main.rs:1068-1093).
__device__ __forceinline__ Frame frame_of(float gx, float gy, float gz) {
    float theta, phi;
    if (gx == 0.0f) theta = kHalfPiF;
    else theta = atan2f(__builtin_sqrtf(gy * gy + gz * gz), gx);
    if (gy == 0.0f) phi = gz > 0.0f ? kHalfPiF : -kHalfPiF;
    else phi = atan2f(gz, gy);
    Frame f;
    sincosf(theta, &f.st, &f.ct);
    sincosf(phi, &f.sp, &f.cp);
    return f;
}

// Rotate a relative speed g by scattering angle chi and azimuth eta in the
// frame f (main.rs:1152-1158).
__device__ __forceinline__ void rotate(float g, float chi, float eta, const Frame &f, float &ox,
                                       float &oy, float &oz) {
    float sc, cc, se, ce;
    sincosf(chi, &sc, &cc);
    sincosf(eta, &se, &ce);
    ox = g * (f.ct * cc - f.st * sc * ce);
    oy = g * (f.st * f.cp * cc + f.ct * f.cp * sc * ce - f.sp * sc * se);
    oz = g * (f.st * f.sp * cc + f.ct * f.sp * sc * ce + f.cp * sc * se);
}

enum CollKind { kNone = 0, kElastic = 1, kExcitation = 2, kIonisation = 3 };

// The collision of one electron (check_collisions_e + collision_e,
// main.rs:996-1165) in three parts, so that a kernel can gather the table
// entries of several particles at once: e_index (energy bin and speed),
// e_test (collision probability; one draw) and e_scatter (process choice and
// scattering with the three cross sections of the bin).  v is updated in
// place; on ionisation the ejected electron and the new ion velocities are
// returned.  RNG draw order: test, process, then (elastic/excitation) chi,
// eta or (ionisation) energy share, eta, and four draws for the ion's
// thermal velocity.
__device__ __forceinline__ int e_index(const KPhys &k, float vx, float vy, float vz, float &velocity) {
    const float v2 = vx * vx + vy * vy + vz * vz;
    const float energy = k.half_me_over_e * v2;
    // the hardware square root (1 ulp), as the reference's sqrt.approx.f32
    // (main.rs:1007, 695-706): the speed only scales the collision
    // frequency, where an ulp moves the collision probability by ~1e-10
    velocity = __builtin_amdgcn_sqrtf(v2);
    return energy_index(energy, k);
}

// 1 - exp(-a) as the reference evaluates it (main.rs:1012, 779-781): the
// exponential as the hardware base-2 exponential of a * log2(e) (v_exp_f32,
// the counterpart of the reference's ex2.approx.f32).
__device__ __forceinline__ float one_minus_exp_neg(float a) {
    return 1.0f - __builtin_amdgcn_exp2f(-a * kLog2eF);
}

// Collision probability of an electron in one step, 1 - exp(-nu dt_e) with
// nu = sigma_tot n_gas |v| (main.rs:1007-1012); e_test compares the step's
// draw with it.
__device__ __forceinline__ float e_prob(const KPhys &k, float stot, float velocity) {
    const float nu = stot * k.stot_scale * velocity;
    return one_minus_exp_neg(nu * k.dt_e);
}

__device__ __forceinline__ bool e_test(const KPhys &k, float stot, float velocity, Xo &s) {
    const float rv = draw(s);
    return rv < e_prob(k, stot, velocity);
}

// Collision-candidate mask of a slot for the lags [from, 8) of a deferred
// write-back run: bit j is set when the draw the slot makes at entry lag j
// (the (j - from)-th output from state s) is below thr.  A step whose bit is
// clear and whose collision probability is at most thr cannot collide, so
// it needs no RNG words.
__device__ __forceinline__ uint32_t cand_mask(Xo s, uint32_t from, float thr) {
    uint32_t m = 0;
    for (uint32_t j = from; j < 8u; ++j) m |= (draw(s) < thr ? 1u : 0u) << j;
    return m;
}

__device__ __forceinline__ int e_scatter(const KPhys &k, float c0, float c1, float c2, float &vx,
                                         float &vy, float &vz, Xo &s, float *se, float *si) {
    float gx = vx, gy = vy, gz = vz;
    float g = __builtin_sqrtf(gx * gx + gy * gy + gz * gz);
    const float wx = k.f1 * vx, wy = k.f1 * vy, wz = k.f1 * vz;
    const float t0 = c0;
    const float t1 = t0 + c1;
    const float t2 = t1 + c2;
    const Frame f = frame_of(gx, gy, gz);
    const float rnd = draw(s);
    const int kind = rnd < t0 / t2 ? kElastic : (rnd < t1 / t2 ? kExcitation : kIonisation);
    // every process draws two numbers next: the scattering angle (elastic,
    // excitation) or the energy share (ionisation), then the azimuth; the
    // processes share one acosf and one rotation of the scattered electron
    const float d1 = draw(s);
    const float d2 = draw(s);
    float chi_arg = 1.0f - 2.0f * d1;
    const float en0 = 0.5f * k.me * g * g;
    float g_new = 0.0f, chi_new_arg = 0.0f;
    if (kind == kExcitation) {
        const float en = fabsf(en0 - k.exc_th * k.e_charge);
        g = __builtin_sqrtf(2.0f * en / k.me);
    } else if (kind == kIonisation) {
        const float en = fabsf(en0 - k.ion_th * k.e_charge);
        const float e_new = 10.0f * tanf(d1 * atanf(en / k.e_charge / 20.0f)) * k.e_charge;
        const float e_orig = fabsf(en - e_new);
        g = __builtin_sqrtf(2.0f * e_orig / k.me);
        g_new = __builtin_sqrtf(2.0f * e_new / k.me);
        chi_arg = __builtin_sqrtf(e_orig / en);
        chi_new_arg = __builtin_sqrtf(e_new / en);
    }
    const float chi = acosf(chi_arg);
    const float eta = kTwoPiF * d2;
    if (kind == kIonisation) {
        const float chi_new = acosf(chi_new_arg);
        const float eta_new = eta + kPiF;
        float nx, ny, nz;
        rotate(g_new, chi_new, eta_new, f, nx, ny, nz);
        se[0] = wx + k.f2 * nx;
        se[1] = wy + k.f2 * ny;
        se[2] = wz + k.f2 * nz;
        three_normal(s, k.normal_range, si[0], si[1], si[2]);
    }
    float ox, oy, oz;
    rotate(g, chi, eta, f, ox, oy, oz);
    vx = wx + k.f2 * ox;
    vy = wy + k.f2 * oy;
    vz = wz + k.f2 * oz;
    return kind;
}

__device__ __forceinline__ int collide_electron(const float *__restrict__ stot_e,
                                                const float *__restrict__ cs, const KPhys &k,
                                                float &vx, float &vy, float &vz, Xo &s,
                                                float *se, float *si) {
    float velocity;
    const int idx = e_index(k, vx, vy, vz, velocity);
    if (!e_test(k, stot_e[idx], velocity, s)) return kNone;
    return e_scatter(k, cs[idx], cs[kCsRanges + idx], cs[2 * kCsRanges + idx], vx, vy, vz, s, se, si);
}

// The collision of one ion against a Maxwellian atom (check_collisions_i +
// collision_i, main.rs:1185-1292) in the same three parts: i_index draws the
// atom's velocity a (three normals) and bins the relative energy, i_test
// draws the collision test, i_scatter picks isotropic or backward
// scattering with the two cross sections of the bin.
__device__ __forceinline__ int i_index(const KPhys &k, float vx, float vy, float vz, Xo &s, float &ax,
                                       float &ay, float &az, float &g) {
    three_normal(s, k.normal_range, ax, ay, az);
    const float gx = vx - ax, gy = vy - ay, gz = vz - az;
    const float g2 = gx * gx + gy * gy + gz * gz;
    const float energy = k.half_mu_over_e * g2;
    g = __builtin_sqrtf(g2);
    return energy_index(energy, k);
}

__device__ __forceinline__ bool i_test(const KPhys &k, float stot, float g, Xo &s) {
    const float nu = stot * k.stot_scale * g;
    const float rv = draw(s);
    const float pc = one_minus_exp_neg(nu * k.dt_i);
    return rv < pc;
}

__device__ __forceinline__ int i_scatter([[maybe_unused]] const KPhys &k, float c0, float c1, float &vx, float &vy,
                                         float &vz, float ax, float ay, float az, float g, Xo &s) {
    const float gx = vx - ax, gy = vy - ay, gz = vz - az;
    const float t0 = c0;
    const float t1 = t0 + c1;
    const float wx = 0.5f * (vx + ax), wy = 0.5f * (vy + ay), wz = 0.5f * (vz + az);
    const Frame f = frame_of(gx, gy, gz);
    float chi;
    int kind;
    const float rnd = draw(s);
    if (rnd < t0 / t1) {
        chi = acosf(1.0f - 2.0f * draw(s));
        kind = kElastic;
    } else {
        chi = kPiF;
        kind = kExcitation;  // backward (charge exchange)
    }
    const float eta = kTwoPiF * draw(s);
    float ox, oy, oz;
    rotate(g, chi, eta, f, ox, oy, oz);
    vx = wx + 0.5f * ox;
    vy = wy + 0.5f * oy;
    vz = wz + 0.5f * oz;
    return kind;
}

__device__ __forceinline__ int collide_ion(const float *__restrict__ stot_i,
                                           const float *__restrict__ cs, const KPhys &k,
                                           float &vx, float &vy, float &vz, Xo &s) {
    float ax, ay, az, g;
    const int idx = i_index(k, vx, vy, vz, s, ax, ay, az, g);
    if (!i_test(k, stot_i[idx], g, s)) return kNone;
    return i_scatter(k, cs[3 * kCsRanges + idx], cs[4 * kCsRanges + idx], vx, vy, vz, ax, ay, az, g, s);
}

}  // namespace pic
