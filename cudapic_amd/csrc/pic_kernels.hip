// pic_kernels.hip -- gfx950 kernels of the RF step and the stage entry
// points of the C ABI (one per kernel of the reference's `mod kernels`,
// Rust-CUDA/src/main.rs:352-1358).
//
// Launch model: every particle kernel is a persistent grid (a fixed number of
// workgroups, by default 8 per compute unit) that reads the live count from
// device memory and strides over tiles, so no host round trip is needed and
// the work scales with the live count, not with MAX_PARTICLES.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pic_internal.h"

namespace pic {

// ============================================================ errors / device
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
int hip_fail(hipError_t e, const char *what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return PIC_ERR_HIP;
}

namespace {
// One integer attribute of the current device, cached per device (a fixed
// table: launch paths that must not allocate query it too); `dflt` if the
// query fails.
int device_attr(hipDeviceAttribute_t attr, int slot, int dflt) {
    constexpr int kDevices = 64;
    static std::mutex mu;
    static int cache[kDevices][4] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return dflt;
    int v = 0;
    if (dev < 0 || dev >= kDevices) {
        if (hipDeviceGetAttribute(&v, attr, dev) != hipSuccess || v <= 0) v = dflt;
        return v;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (cache[dev][slot] == 0) {
        if (hipDeviceGetAttribute(&v, attr, dev) != hipSuccess || v <= 0) v = dflt;
        cache[dev][slot] = v;
    }
    return cache[dev][slot];
}
}  // namespace

int device_cus() { return device_attr(hipDeviceAttributeMultiprocessorCount, 0, 256); }

int device_lds_per_cu() {
    return device_attr(hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, 1, 64 * 1024);
}

int device_lds_per_block() { return device_attr(hipDeviceAttributeMaxSharedMemoryPerBlock, 2, 64 * 1024); }

Shape resolve(Shape s) {
    if (s.block <= 0) s.block = 256;
    s.block = std::min(1024, std::max(64, (s.block / 64) * 64));
    if (s.grid <= 0) s.grid = device_cus() * std::max(1, 2048 / s.block);
    if (s.ppt <= 0) s.ppt = 4;
    if (s.ppt_inplace <= 0) s.ppt_inplace = 2;
    if (s.replicas <= 0) s.replicas = 1;
    return s;
}

KPhys make_kphys(const pic_params_t &p) {
    KPhys k;
    k.n_g = p.n_g;
    k.inv_dx = p.inv_dx_f;
    k.gap = p.gap_f;
    k.wf = p.weight_factor;
    k.factor_e = p.factor_e;
    k.factor_i = p.factor_i;
    k.dt_e = p.dt_e_f;
    k.dt_i = p.dt_i_f;
    k.normal_range = p.normal_range;
    k.f1 = p.f1;
    k.f2 = p.f2;
    k.half_me_over_e = p.half_me_over_e;
    k.half_mu_over_e = p.half_mu_over_e;
    k.half_mi_over_e = p.half_mi_over_e;
    k.e_charge = p.e_charge_f;
    k.me = p.me_f;
    k.exc_th = p.exc_th_f;
    k.ion_th = p.ion_th_f;
    k.de_cs = p.de_cs_f;
    k.inv_de_cs = 1.0f / p.de_cs_f;
    k.min_x = p.min_x_f;
    k.max_x = p.max_x_f;
    k.de_eepf = p.de_eepf_f;
    k.eepf_kick = p.eepf_kick_f;
    k.stot_scale = 1.0f;
    return k;
}

float coll_threshold(const float *stot_e, double stot_scale, const pic_params_t &p) {
    const double e = 1.60217662e-19, me = 9.10938356e-31;
    double pmax = 0.0;
    for (int b = 0; b < PIC_CS_RANGES; ++b) {
        const double v = std::sqrt(2.0 * e * ((double)b + 0.5) * (double)p.de_cs_f / me);
        pmax = std::max(pmax, -std::expm1(-(double)stot_e[b] * stot_scale * v * (double)p.dt_e_f));
    }
    // a margin over the f32 rounding of the kernel's probability; the
    // kernel compares its own probability with the threshold, so any value
    // is correct and only the share of candidates depends on it
    return (float)std::min(1.0, pmax * (1.0 + 1e-4) + 1e-7);
}

// LDS budget (bytes) one particle workgroup may take for its histograms
// so that several workgroups stay resident per compute unit.
static constexpr int kLdsBudget = 48 * 1024;

static inline int lds_replicas(int block, int n_g, int fixed_bytes) {
    int waves = block / kWave;
    int fit = (kLdsBudget - fixed_bytes) / (n_g * 4);
    return std::max(1, std::min(waves, fit));
}

// ============================================================ small kernels
__global__ void k_fill(uint32_t *__restrict__ p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int launch_fill(void *ptr, uint32_t value, uint64_t words, hipStream_t s) {
    if (words == 0) return PIC_OK;
    int grid = (int)std::min<uint64_t>((words + 255) / 256, (uint64_t)device_cus() * 4);
    k_fill<<<grid, 256, 0, s>>>((uint32_t *)ptr, value, words);
    PIC_LAUNCHED("k_fill");
    return PIC_OK;
}

__global__ void k_copy_u32(uint32_t *dst, const uint32_t *src, uint32_t clamp) {
    if (threadIdx.x == 0) *dst = src ? min(*src, clamp) : 0u;
}

int launch_copy_u32(uint32_t *dst, const uint32_t *src, hipStream_t s, uint32_t clamp) {
    k_copy_u32<<<1, 64, 0, s>>>(dst, src, clamp);
    PIC_LAUNCHED("k_copy_u32");
    return PIC_OK;
}

// Two counter copies in one launch (both sources read before either store).
__global__ void k_copy2_u32(uint32_t *d0, const uint32_t *s0, uint32_t *d1, const uint32_t *s1, uint32_t clamp) {
    if (threadIdx.x == 0) {
        const uint32_t a = min(*s0, clamp), b = min(*s1, clamp);
        *d0 = a;
        *d1 = b;
    }
}

int launch_copy2_u32(uint32_t *d0, const uint32_t *s0, uint32_t *d1, const uint32_t *s1, hipStream_t s,
                     uint32_t clamp) {
    k_copy2_u32<<<1, 64, 0, s>>>(d0, s0, d1, s1, clamp);
    PIC_LAUNCHED("k_copy2_u32");
    return PIC_OK;
}

// ============================================================ deposition
// get_density (main.rs:356-406).  Each wave adds into an LDS histogram
// replica (fixed point, integer LDS atomics); the workgroup sums its
// replicas once and flushes each non-zero node with one global atomic (f32
// or f64), spread over `replicas` global copies.
//
// One workgroup of 1024 lanes per compute unit (16 waves, four per SIMD)
// and 16-byte loads of four positions, two in flight per lane per
// iteration: the kernel reads 4 B per particle, so it is bound by how many
// loads are in flight (one wave per SIMD with scalar loads reached 0.9 TB/s
// at 3.8 M particles), not by the LDS atomics.  Fewer workgroups keep the
// flush's same-address global atomics few.
constexpr int kDepositBlock = 1024;
constexpr int kDepositPerLane = 8;  // positions per lane per iteration (two float4)

__device__ __forceinline__ void deposit_one(uint32_t *h, float x, float inv_dx, int n_g, float scale) {
    if (x < 0.0f) return;  // a hole
    const float pos = x * inv_dx;
    const int q = cell_of(pos, n_g);
    const float rem = pos - (float)q;
    const uint32_t u2 = (uint32_t)(rem * scale + 0.5f);
    atomicAdd(&h[q], (uint32_t)scale - u2);
    atomicAdd(&h[q + 1], u2);
}

template <typename Acc>
__global__ __launch_bounds__(kDepositBlock) void k_deposit(const float *__restrict__ x,
                                                           const uint32_t *__restrict__ n_active,
                                                           uint32_t n_cap, Acc *__restrict__ dens, int reps,
                                                           int rep_stride, float inv_dx, float wf, int n_g,
                                                           int lds_reps) {
    // Fixed-point LDS histograms: a particle adds exactly `scale` units split
    // by its CIC weights.  The scale is the largest power of two (at most
    // 2^20) for which a replica cannot overflow: its waves deposit at most
    // iters * kDepositPerLane * 64 particles each.
    extern __shared__ uint32_t hs[];
    const int wave = threadIdx.x / kWave;
    for (int k = threadIdx.x; k < lds_reps * n_g; k += blockDim.x) hs[k] = 0u;
    __syncthreads();
    uint32_t *h = hs + (wave % lds_reps) * n_g;
    const uint32_t n = min(*n_active, n_cap);
    // 16-byte loads need a 16-byte aligned start; a misaligned array takes
    // its first (4 - offset) positions one at a time
    const uint32_t lead = min(n, (uint32_t)((16u - ((uintptr_t)x & 15u)) & 15u) / 4u);
    const uint32_t n4 = (n - lead) / 4u;          // float4 groups after the lead
    const float4 *x4 = reinterpret_cast<const float4 *>(x + lead);
    const uint32_t stride = gridDim.x * blockDim.x;  // float4 groups per half-iteration
    const uint64_t iters = ((uint64_t)n4 + 2ull * stride - 1) / (2ull * stride) + 1;
    const uint64_t waves_per_rep = (blockDim.x / kWave + lds_reps - 1) / lds_reps;
    // per lane: kDepositPerLane positions an iteration, plus one of the
    // lead and one of the tail
    const uint64_t bound = (iters * (uint64_t)kDepositPerLane + 2u) * kWave * waves_per_rep + 1u;
    int sh_bits = 20;
    while (sh_bits > 0 && (bound << sh_bits) >= (1ull << 32)) --sh_bits;
    const float scale = (float)(1u << sh_bits);
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t g = gid; g < n4; g += 2u * stride) {
        const float4 a = x4[g];
        const bool second = g + stride < n4;
        const float4 b = x4[second ? g + stride : g];
        deposit_one(h, a.x, inv_dx, n_g, scale);
        deposit_one(h, a.y, inv_dx, n_g, scale);
        deposit_one(h, a.z, inv_dx, n_g, scale);
        deposit_one(h, a.w, inv_dx, n_g, scale);
        if (second) {
            deposit_one(h, b.x, inv_dx, n_g, scale);
            deposit_one(h, b.y, inv_dx, n_g, scale);
            deposit_one(h, b.z, inv_dx, n_g, scale);
            deposit_one(h, b.w, inv_dx, n_g, scale);
        }
    }
    // the misaligned lead and the tail of fewer than four positions
    const uint32_t tail0 = lead + 4u * n4;
    if (gid < lead) deposit_one(h, x[gid], inv_dx, n_g, scale);
    if (gid < n - tail0) deposit_one(h, x[tail0 + gid], inv_dx, n_g, scale);
    __syncthreads();
    if ((uint64_t)blockIdx.x * blockDim.x >= (uint64_t)n4 && blockIdx.x != 0) return;  // saw nothing
    Acc *out = dens + (blockIdx.x % reps) * rep_stride;
    const Acc unit = (Acc)wf / (Acc)scale;
    for (int k = threadIdx.x; k < n_g; k += blockDim.x) {
        uint64_t s = 0;
        for (int r = 0; r < lds_reps; ++r) s += hs[r * n_g + k];
        if (s) {
            Acc v = (Acc)s * unit;
            if (k == 0 || k == n_g - 1) v *= (Acc)2;
            unsafeAtomicAdd(&out[k], v);
        }
    }
}

int launch_deposit(const float *x, const uint32_t *n_active, void *dens, bool f64, const KPhys &k,
                   const Shape &sh0, hipStream_t s, uint32_t n_cap) {
    Shape sh = resolve(sh0);
    const int block = kDepositBlock;
    const int lr = lds_replicas(block, k.n_g, 0);
    const size_t lds = (size_t)lr * k.n_g * sizeof(uint32_t);
    const int stride = sh.rep_stride > 0 ? sh.rep_stride : k.n_g;
    const int grid = device_cus();
    if (f64)
        k_deposit<double><<<grid, block, lds, s>>>(x, n_active, n_cap, (double *)dens, sh.replicas, stride,
                                                   k.inv_dx, k.wf, k.n_g, lr);
    else
        k_deposit<float><<<grid, block, lds, s>>>(x, n_active, n_cap, (float *)dens, sh.replicas, stride,
                                                  k.inv_dx, k.wf, k.n_g, lr);
    PIC_LAUNCHED("k_deposit");
    return PIC_OK;
}

// ============================================================ Poisson
// Scan arithmetic for the three precisions: float (the reference), double,
// and float-float ("compensated FP32": an unevaluated sum hi + lo carried
// through error-free transformations, Dekker/Knuth TwoSum and fma TwoProd).
struct ff {
    float hi, lo;
};
__device__ __forceinline__ ff ff_fast2(float a, float b) {
    float s = a + b;
    return ff{s, b - (s - a)};
}
__device__ __forceinline__ ff ff_two_sum(float a, float b) {
    float s = a + b;
    float bb = s - a;
    return ff{s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ ff operator+(ff a, ff b) {
    ff s = ff_two_sum(a.hi, b.hi);
    ff t = ff_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = ff_fast2(s.hi, s.lo);
    s.lo += t.lo;
    return ff_fast2(s.hi, s.lo);
}
__device__ __forceinline__ ff operator-(ff a) { return ff{-a.hi, -a.lo}; }
__device__ __forceinline__ ff operator-(ff a, ff b) { return a + (-b); }
__device__ __forceinline__ ff operator*(ff a, ff b) {
    float p = a.hi * b.hi;
    float e = __builtin_fmaf(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return ff_fast2(p, e);
}
__device__ __forceinline__ ff operator/(ff a, ff b) {
    float q1 = a.hi / b.hi;
    ff r = a - b * ff{q1, 0.0f};
    float q2 = r.hi / b.hi;
    r = r - b * ff{q2, 0.0f};
    float q3 = r.hi / b.hi;
    ff q = ff_fast2(q1, q2);
    return q + ff{q3, 0.0f};
}

// A lane's value from another lane of the wave by DPP (row_shr:n within
// rows of 16 lanes, row_bcast:15 / :31 across rows); lanes with no source
// read 0.  CTRL is the DPP control word, ROWMASK the rows that receive.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWMASK, 0xF, true);
}

template <typename T> struct Num;
template <> struct Num<float> {
    __device__ static float from(double d) { return (float)d; }
    __device__ static float fromf(float f) { return f; }
    __device__ static float zero() { return 0.0f; }
    __device__ static float to_f(float v) { return v; }
    __device__ static float shfl_up(float v, int d) { return __shfl_up(v, d, kWave); }
    template <int C, int M>
    __device__ static float dpp(float v) { return __int_as_float(dpp_i<C, M>(__float_as_int(v))); }
};
template <> struct Num<double> {
    __device__ static double from(double d) { return d; }
    __device__ static double fromf(float f) { return (double)f; }
    __device__ static double zero() { return 0.0; }
    __device__ static float to_f(double v) { return (float)v; }
    __device__ static double shfl_up(double v, int d) { return __shfl_up(v, d, kWave); }
    template <int C, int M>
    __device__ static double dpp(double v) {
        const long long b = __double_as_longlong(v);
        const int lo = dpp_i<C, M>((int)(uint32_t)b), hi = dpp_i<C, M>((int)(b >> 32));
        return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
    }
};
template <> struct Num<ff> {
    __device__ static ff from(double d) {
        float h = (float)d;
        return ff{h, (float)(d - (double)h)};
    }
    __device__ static ff fromf(float f) { return ff{f, 0.0f}; }
    __device__ static ff zero() { return ff{0.0f, 0.0f}; }
    __device__ static float to_f(ff v) { return v.hi + v.lo; }
    __device__ static ff shfl_up(ff v, int d) {
        return ff{__shfl_up(v.hi, d, kWave), __shfl_up(v.lo, d, kWave)};
    }
    template <int C, int M>
    __device__ static ff dpp(ff v) { return ff{Num<float>::dpp<C, M>(v.hi), Num<float>::dpp<C, M>(v.lo)}; }
};

// Inclusive scan over the 64 lanes of a wave on DPP moves: prefixes within
// each row of 16 lanes (row_shr 1, 2, 4, 8), then row 0's total into row 1
// and row 2's into row 3 (row_bcast:15), then lanes 0-31's total into rows 2
// and 3 (row_bcast:31).
template <typename T>
__device__ __forceinline__ T wave_scan_incl(T v) {
    v = v + Num<T>::template dpp<0x111, 0xF>(v);
    v = v + Num<T>::template dpp<0x112, 0xF>(v);
    v = v + Num<T>::template dpp<0x114, 0xF>(v);
    v = v + Num<T>::template dpp<0x118, 0xF>(v);
    v = v + Num<T>::template dpp<0x142, 0xA>(v);
    v = v + Num<T>::template dpp<0x143, 0xC>(v);
    return v;
}

// Inclusive scan over the lanes of the workgroup (up to 16 waves): DPP scan
// inside each wave, then every wave adds the totals of the waves before it,
// read from LDS (wtot: 16 entries).  REUSE: a last barrier lets the next
// scan reuse wtot (a caller that gives each scan its own wtot goes without
// it).
template <typename T, bool REUSE = true>
__device__ __forceinline__ T block_scan_incl(T v, T *wtot) {
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    v = wave_scan_incl(v);
    if (lane == kWave - 1) wtot[wave] = v;
    __syncthreads();
    if (wave > 0) {
        T base = wtot[0];
        for (int w = 1; w < wave; ++w) base = base + wtot[w];
        v = base + v;
    }
    if (REUSE) __syncthreads();
    return v;
}

constexpr int kPoissonThreads = 1024;
constexpr int kPoissonMaxPer = 4;  // n_g <= 4096
constexpr int kPoissonWtot = 33;   // LDS entries of poisson_solve beyond the n_g potentials

// Nodes per lane and lanes of a Poisson workgroup: as few waves as cover the
// grid (their barriers and the inter-wave scans are the solve's latency).
inline int poisson_per(int n_g) { return (n_g + kPoissonThreads - 1) / kPoissonThreads; }
inline int poisson_threads(int n_g) {
    const int lanes = (n_g + poisson_per(n_g) - 1) / poisson_per(n_g);
    return std::min(kPoissonThreads, (lanes + kWave - 1) / kWave * kWave);
}

struct PoissonArgs {
    const void *e_dens, *i_dens;
    PoissonGeom g;
    double pot0;
    const uint32_t *time;
    const double *pot0_in;
    double *pot0_next;
    const double *pot0_tab;
    const double *i_sum;
    double *dens_e_out, *dens_i_out;
    int n_t, n_sub, n_harm;
    double wave_cos[PIC_MAX_HARMONICS], wave_sin[PIC_MAX_HARMONICS];
    float *pot, *efield;
    double *cumul_e, *cumul_i;
    unsigned long long *meas_steps;
    uint32_t *zero_a;
    int zero_a_words;
    uint32_t *zero_b;
    int zero_b_words;
    uint32_t *zero_cnt;
    int zero_cnt_words;
    uint32_t *zero_cnt_ion;
    int zero_cnt_ion_words;
    const uint32_t *copy_cnt_src;
    uint32_t *copy_cnt_dst;
    int copy_cnt_words;
    uint32_t *advance_time;
    const uint32_t *copy_cnt_ion_src;
    uint32_t *copy_cnt_ion_dst;
    int64_t t_value;
    double *xt_grid;
    unsigned long long *xt_count;
    uint32_t xt_bin;
    int xt_n;
};

// Node k of one density summed over the global replicas, in replica order.
// The replica count is 1 or a multiple of 8; the loads of a group of 8
// replicas are independent, so a group costs one memory latency.
template <typename D>
__device__ __forceinline__ double sum_replicas(const D *d, int reps, int stride, int k) {
    if (reps == 1) return (double)d[k];
    double sum = 0.0;
    for (int r0 = 0; r0 < reps; r0 += 8) {
        D v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = d[(r0 + j) * stride + k];
#pragma unroll
        for (int j = 0; j < 8; ++j) sum += (double)v[j];
    }
    return sum;
}

// Densities of the (up to 4) nodes of this lane, summed over the replicas.
// With i_sum the ion densities are taken from it (n_g doubles that an
// earlier solve summed from the same replicas: the same values, one load).
template <typename D>
__device__ __forceinline__ void load_node_densities(const PoissonGeom &pg, const D *ed, const D *id,
                                                    const double *i_sum, double *dne, double *dni) {
    const int k0 = threadIdx.x * pg.per;
#pragma unroll
    for (int u = 0; u < kPoissonMaxPer; ++u) {
        dne[u] = dni[u] = 0.0;
        if (u < pg.per && k0 + u < pg.n_g) {
            dne[u] = sum_replicas(ed, pg.reps, pg.rep_stride, k0 + u);
            dni[u] = i_sum ? i_sum[k0 + u] : sum_replicas(id, pg.reps, pg.rep_stride, k0 + u);
        }
    }
}

// pot0 of step t: sum_h a_h cos(2 pi (h+1) t / N_T) + b_h sin(...), the
// phase reduced exactly in integers before sincospi (main.rs:1741).
__device__ double waveform(uint32_t t, int n_t, int n_harm, const double *wc, const double *ws) {
    double p0 = 0.0;
    for (int h = 0; h < n_harm; ++h) {
        const uint64_t m = (2ull * (uint64_t)(h + 1) * t) % (2ull * (uint64_t)n_t);
        double sv, cv;
        sincospi((double)m / (double)n_t, &sv, &cv);
        p0 += wc[h] * cv + ws[h] * sv;
    }
    return p0;
}

// pot0 = v0 cos(2 pi t / n_t) from the tabulated cos / sin of step t, with
// the arithmetic of waveform() for one harmonic with amplitudes (v0, 0)
__device__ __forceinline__ double pot0_from_trig(double v0, double c, double s) {
    double p0 = 0.0;
    p0 += v0 * c + 0.0 * s;
    return p0;
}

// solve_poisson_scan_f32 (main.rs:408-494) generalised to three
// precisions and n_g <= 4096 (up to four nodes per lane), for one discharge
// by one workgroup of 1024 lanes.  The tridiagonal Dirichlet system with
// A = C = 1, B = -2 is solved as two inclusive scans:
//   G_i = sum_{j<=i} j f_j,  r_i = -G_i / ((i+1) i),  R = scan(r),
//   pot_i = i (R_{N-2} - R_i + r_i),
// then E by central differences with the electrode formulas of
// C/eduPIC.cc:491-493.  pot_s holds n_g values of T, wtot kPoissonWtot
// (the two scans' wave totals, then R_{N-2}).
// The densities of the lane's nodes come in dne / dni (load_node_densities).
template <typename T>
__device__ void poisson_solve(const PoissonGeom &pg, const double *dne, const double *dni, double pot0_d, bool ion,
                              float *pot, float *efield, double *cumul_e, double *cumul_i, T *pot_s,
                              T *wtot) {
    const int tid = threadIdx.x;
    const int n_g = pg.n_g;
    const int per = pg.per;
    const T pot0 = Num<T>::from(pot0_d);
    const T alpha = Num<T>::from(pg.alpha);
    const T echarge = Num<T>::from(pg.e_charge);
    T rho[kPoissonMaxPer], g[kPoissonMaxPer];
    const int k0 = tid * per;
    // charge density, g_i = i f_i
    T run = Num<T>::zero();
#pragma unroll
    for (int u = 0; u < kPoissonMaxPer; ++u) {
        const int k = k0 + u;
        rho[u] = Num<T>::zero();
        g[u] = Num<T>::zero();
        if (u < per && k < n_g) {
            const T ne = Num<T>::from(dne[u]);
            const T ni = Num<T>::from(dni[u]);
            rho[u] = echarge * (ni - ne);
            if (k >= 1 && k <= n_g - 2) {
                T f = alpha * rho[u];
                if (k == 1) f = f - pot0;
                g[u] = Num<T>::from((double)k) * f;
            }
            run = run + g[u];
            g[u] = run;  // chunk-local inclusive prefix
        }
    }
    T excl = block_scan_incl<T, false>(run, wtot) - run;
    T r[kPoissonMaxPer], R[kPoissonMaxPer];
    run = Num<T>::zero();
#pragma unroll
    for (int u = 0; u < kPoissonMaxPer; ++u) {
        const int k = k0 + u;
        r[u] = Num<T>::zero();
        if (u < per && k >= 1 && k <= n_g - 2) {
            const T G = excl + g[u];
            const T h = -G / Num<T>::from((double)k + 1.0);
            r[u] = h / Num<T>::from((double)k);
        }
        if (u < per && k < n_g) {
            run = run + r[u];
            R[u] = run;
        }
    }
    excl = block_scan_incl<T, false>(run, wtot + 16) - run;
    // R_{N-2}
#pragma unroll
    for (int u = 0; u < kPoissonMaxPer; ++u)
        if (u < per && k0 + u == n_g - 2) wtot[32] = excl + R[u];
    __syncthreads();
    const T r_total = wtot[32];
#pragma unroll
    for (int u = 0; u < kPoissonMaxPer; ++u) {
        const int k = k0 + u;
        if (u < per && k < n_g) {
            T pk;
            if (k == 0) pk = pot0;
            else if (k < n_g - 1) pk = Num<T>::from((double)k) * (r_total - (excl + R[u]) + r[u]);
            else pk = Num<T>::zero();
            pot_s[k] = pk;
        }
    }
    __syncthreads();
    const T inv_dx = Num<T>::from(pg.inv_dx);
    const T half_inv_dx = Num<T>::from(pg.half_inv_dx);
    const T hdx_eps = Num<T>::from(pg.half_dx_over_eps);
#pragma unroll
    for (int u = 0; u < kPoissonMaxPer; ++u) {
        const int k = k0 + u;
        if (u < per && k < n_g) {
            const T pk = pot_s[k];
            T e;
            if (k == 0) e = (pk - pot_s[1]) * inv_dx - rho[u] * hdx_eps;
            else if (k == n_g - 1) e = (pot_s[n_g - 2] - pk) * inv_dx + rho[u] * hdx_eps;
            else e = (pot_s[k - 1] - pot_s[k + 1]) * half_inv_dx;
            pot[k] = Num<T>::to_f(pk);
            efield[k] = Num<T>::to_f(e);
            if (cumul_e) {
                cumul_e[k] += dne[u];
                if (ion) cumul_i[k] += dni[u];
            }
        }
    }
}

template <typename T, typename D>
__global__ __launch_bounds__(kPoissonThreads) void k_poisson(PoissonArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    T *pot_s = reinterpret_cast<T *>(smem);  // n_g
    T *wtot = pot_s + a.g.n_g;               // kPoissonWtot
    __shared__ double s_pot0;
    __shared__ int s_ion;
    __shared__ uint32_t s_t;
    const int tid = threadIdx.x;
    // the densities, the clock, the step's pot0 and the counters to copy are
    // all requested before any of them is waited for.  One node per lane and
    // one group of eight replicas (the default) is loaded here; other shapes
    // go through load_node_densities.  The clock and pot0 are vector loads:
    // as scalar-cache loads the compiler would wait for each in turn.
    const bool one_group = a.g.per == 1 && a.g.reps == 8;
    const bool my_node = one_group && tid < a.g.n_g;
    D ve[8], vi[8];
    double is = 0.0;
    if (my_node) {
        const D *ed = (const D *)a.e_dens, *id = (const D *)a.i_dens;
        const int st = a.g.rep_stride;
#pragma unroll
        for (int j = 0; j < 8; ++j) ve[j] = ed[j * st + tid];
        if (a.i_sum) {
            is = a.i_sum[tid];
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) vi[j] = id[j * st + tid];
        }
    }
    uint32_t t = 0;
    double p0_in = 0.0;
    if (tid == 0 && a.t_value >= 0) {
        t = (uint32_t)a.t_value;
        if (a.pot0_tab) p0_in = __hip_atomic_load(a.pot0_tab + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (tid == 0 && a.time) {
        t = __hip_atomic_load(a.time, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (a.pot0_in) p0_in = __hip_atomic_load(a.pot0_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    uint32_t cnt_copy = 0, cnt_copy_ion = 0;
    if (a.copy_cnt_dst && tid < a.copy_cnt_words) cnt_copy = a.copy_cnt_src[tid];
    if (a.copy_cnt_ion_dst && tid == 0)
        cnt_copy_ion = __hip_atomic_load(a.copy_cnt_ion_src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double dne[kPoissonMaxPer], dni[kPoissonMaxPer];
    if (one_group) {
        // sum_replicas' order
#pragma unroll
        for (int u = 0; u < kPoissonMaxPer; ++u) dne[u] = dni[u] = 0.0;
        if (my_node) {
            double se = 0.0;
#pragma unroll
            for (int j = 0; j < 8; ++j) se += (double)ve[j];
            if (!a.i_sum) {
#pragma unroll
                for (int j = 0; j < 8; ++j) is += (double)vi[j];
            }
            dne[0] = se;
            dni[0] = is;
        }
    } else {
        load_node_densities<D>(a.g, (const D *)a.e_dens, (const D *)a.i_dens, a.i_sum, dne, dni);
    }
    double p0_next = 0.0;
    if (tid == 0) {
        double p0 = a.pot0;
        int ion = 1;
        if (a.t_value >= 0 || a.time) {
            ion = (t % (uint32_t)a.n_sub) == 0;
            const bool have = a.t_value >= 0 ? a.pot0_tab != nullptr : a.pot0_in != nullptr;
            p0 = have ? p0_in : waveform(t, a.n_t, a.n_harm, a.wave_cos, a.wave_sin);
            // the next step's pot0 from the table: the load is in flight
            // during the solve
            if (a.pot0_tab && a.pot0_next) p0_next = a.pot0_tab[(t + 1u) % (uint32_t)a.n_t];
        }
        s_pot0 = p0;
        s_ion = ion;
        s_t = t;
    }
    __syncthreads();
    poisson_solve<T>(a.g, dne, dni, s_pot0, s_ion != 0, a.pot, a.efield, a.cumul_e, a.cumul_i, pot_s, wtot);
    if (a.xt_grid && a.cumul_e) {
        // XT maps of the grid quantities (C/eduPIC.cc:705-715); each lane
        // reads back the pot and E it wrote
        const uint32_t tb = s_t / a.xt_bin;
        const size_t plane = (size_t)a.g.n_g * a.xt_n;
        const int k0 = tid * a.g.per;
#pragma unroll
        for (int u = 0; u < kPoissonMaxPer; ++u)
            if (u < a.g.per && k0 + u < a.g.n_g) {
                const size_t o = (size_t)(k0 + u) * a.xt_n + tb;
                a.xt_grid[o] += (double)a.pot[k0 + u];
                a.xt_grid[plane + o] += (double)a.efield[k0 + u];
                a.xt_grid[2 * plane + o] += dne[u];
                a.xt_grid[3 * plane + o] += dni[u];
            }
        if (tid == 0) a.xt_count[tb] += 1ull;
    }
    if (a.dens_e_out) {
        const int k0 = tid * a.g.per;
#pragma unroll
        for (int u = 0; u < kPoissonMaxPer; ++u)
            if (u < a.g.per && k0 + u < a.g.n_g) {
                a.dens_e_out[k0 + u] = dne[u];
                a.dens_i_out[k0 + u] = dni[u];
            }
    }
    // fused-path bookkeeping for the particle kernels that follow
    for (int w = tid; w < a.zero_a_words; w += blockDim.x) a.zero_a[w] = 0u;
    if (s_ion)
        for (int w = tid; w < a.zero_b_words; w += blockDim.x) a.zero_b[w] = 0u;
    if (a.zero_cnt && tid < a.zero_cnt_words) a.zero_cnt[tid] = 0u;
    if (s_ion && a.zero_cnt_ion && tid < a.zero_cnt_ion_words) a.zero_cnt_ion[tid] = 0u;
    if (a.copy_cnt_dst && tid < a.copy_cnt_words) a.copy_cnt_dst[tid] = cnt_copy;
    if (s_ion && a.copy_cnt_ion_dst && tid == 0) *a.copy_cnt_ion_dst = cnt_copy_ion;
    if (tid == (int)blockDim.x - 1 && a.meas_steps && a.cumul_e) {
        a.meas_steps[0] += 1ull;
        if (s_ion) a.meas_steps[1] += 1ull;
    }
    // the clock was read by this lane (the same word when advance_time is set)
    if (tid == 0 && a.advance_time) {
        const uint32_t tn = (t + 1u) % (uint32_t)a.n_t;
        *a.advance_time = tn;
        if (a.pot0_next)
            *a.pot0_next = a.pot0_tab ? p0_next : waveform(tn, a.n_t, a.n_harm, a.wave_cos, a.wave_sin);
    }
}

__global__ void k_pot0(const uint32_t *time, int n_t, int n_harm, PoissonArgs a, double *out) {
    *out = waveform(*time, n_t, n_harm, a.wave_cos, a.wave_sin);
}

// tab[t] = pot0 of step t for t < n_t (the values k_pot0 computes)
__global__ void k_pot0_table(int n_t, int n_harm, PoissonArgs a, double *tab) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_t) tab[t] = waveform((uint32_t)t, n_t, n_harm, a.wave_cos, a.wave_sin);
}

// out[k] = node k of dens summed over the replicas, as the Poisson kernel sums it
template <typename D>
__global__ void k_sum_replicas(PoissonGeom g, const D *dens, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < g.n_g) out[k] = sum_replicas(dens, g.reps, g.rep_stride, k);
}

int launch_poisson(const PoissonJob &j, hipStream_t s) {
    if (j.n_g < 3 || j.n_g > kPoissonThreads * kPoissonMaxPer) {
        set_error("poisson: n_g must be in [3, 4096]");
        return PIC_ERR_ARG;
    }
    if (j.replicas != 1 && (j.replicas % 8) != 0) {
        set_error("poisson: density replicas must be 1 or a multiple of 8");
        return PIC_ERR_ARG;
    }
    if (j.advance_time && j.advance_time != j.time) {
        set_error("poisson: the clock advanced must be the clock read");
        return PIC_ERR_ARG;
    }
    if ((j.pot0_in || j.pot0_next) && !j.time) {
        set_error("poisson: a precomputed pot0 needs the device clock");
        return PIC_ERR_ARG;
    }
    if ((j.dens_e_out == nullptr) != (j.dens_i_out == nullptr)) {
        set_error("poisson: both density outputs or neither");
        return PIC_ERR_ARG;
    }
    PoissonArgs a;
    memset(&a, 0, sizeof(a));
    a.e_dens = j.e_dens;
    a.i_dens = j.i_dens;
    a.g.reps = std::max(1, j.replicas);
    a.g.rep_stride = j.rep_stride > 0 ? j.rep_stride : j.n_g;
    a.g.n_g = j.n_g;
    a.g.per = poisson_per(j.n_g);
    a.pot0 = j.pot0;
    a.time = j.time;
    a.pot0_in = j.pot0_in;
    a.pot0_next = j.pot0_next;
    a.pot0_tab = j.pot0_tab;
    a.i_sum = j.i_sum;
    a.xt_grid = j.xt_grid;
    a.xt_count = j.xt_count;
    a.xt_bin = j.xt_bin;
    a.xt_n = j.xt_n;
    if (a.xt_grid && (!a.time || !a.xt_count || a.xt_bin == 0 || a.xt_n <= 0 ||
                      (uint32_t)a.xt_n * a.xt_bin != (uint32_t)j.n_t)) {
        set_error("poisson: XT maps need the device clock and xt_bin * xt_n == n_t");
        return PIC_ERR_ARG;
    }
    a.dens_e_out = j.dens_e_out;
    a.dens_i_out = j.dens_i_out;
    a.n_t = j.n_t;
    a.n_sub = j.n_sub;
    a.n_harm = 0;
    for (int h = 0; h < PIC_MAX_HARMONICS; ++h) {
        a.wave_cos[h] = j.wave_cos[h];
        a.wave_sin[h] = j.wave_sin[h];
        if (j.wave_cos[h] != 0.0 || j.wave_sin[h] != 0.0) a.n_harm = h + 1;
    }
    a.g.e_charge = j.e_charge;
    a.g.alpha = j.alpha;
    a.g.inv_dx = j.inv_dx;
    a.g.half_inv_dx = j.half_inv_dx;
    a.g.half_dx_over_eps = j.half_dx_over_eps;
    a.pot = j.pot;
    a.efield = j.efield;
    a.cumul_e = j.cumul_e;
    a.cumul_i = j.cumul_i;
    a.meas_steps = j.meas_steps;
    a.zero_a = (uint32_t *)j.zero_a;
    a.zero_a_words = j.zero_a_words;
    a.zero_b = (uint32_t *)j.zero_b;
    a.zero_b_words = j.zero_b_words;
    a.zero_cnt = j.zero_cnt;
    a.zero_cnt_words = j.zero_cnt_words;
    a.zero_cnt_ion = j.zero_cnt_ion;
    a.zero_cnt_ion_words = j.zero_cnt_ion_words;
    a.copy_cnt_src = j.copy_cnt_src;
    a.copy_cnt_dst = j.copy_cnt_dst;
    a.copy_cnt_words = j.copy_cnt_words;
    a.advance_time = j.advance_time;
    a.copy_cnt_ion_src = j.copy_cnt_ion_src;
    a.copy_cnt_ion_dst = j.copy_cnt_ion_dst;
    a.t_value = j.t_value;
    if (j.t_value >= j.n_t || (j.copy_cnt_ion_dst && !j.copy_cnt_ion_src)) {
        set_error("poisson: step index beyond the cycle or an ion counter copy without a source");
        return PIC_ERR_ARG;
    }
    switch (j.precision) {
    case PIC_PREC_FP32: {
        size_t lds = (j.n_g + kPoissonWtot) * sizeof(float);
        k_poisson<float, float><<<1, poisson_threads(j.n_g), lds, s>>>(a);
        break;
    }
    case PIC_PREC_COMP: {
        size_t lds = (j.n_g + kPoissonWtot) * sizeof(ff);
        k_poisson<ff, float><<<1, poisson_threads(j.n_g), lds, s>>>(a);
        break;
    }
    case PIC_PREC_FP64: {
        size_t lds = (j.n_g + kPoissonWtot) * sizeof(double);
        k_poisson<double, double><<<1, poisson_threads(j.n_g), lds, s>>>(a);
        break;
    }
    default:
        set_error("poisson: unknown precision");
        return PIC_ERR_ARG;
    }
    PIC_LAUNCHED("k_poisson");
    return PIC_OK;
}

int launch_pot0(const PoissonJob &j, double *out, hipStream_t s) {
    if (!j.time || !out || j.n_t <= 0) {
        set_error("pot0: needs the device clock and an output");
        return PIC_ERR_ARG;
    }
    PoissonArgs a;
    memset(&a, 0, sizeof(a));
    int n_harm = 0;
    for (int h = 0; h < PIC_MAX_HARMONICS; ++h) {
        a.wave_cos[h] = j.wave_cos[h];
        a.wave_sin[h] = j.wave_sin[h];
        if (j.wave_cos[h] != 0.0 || j.wave_sin[h] != 0.0) n_harm = h + 1;
    }
    k_pot0<<<1, 1, 0, s>>>(j.time, j.n_t, n_harm, a, out);
    PIC_LAUNCHED("k_pot0");
    return PIC_OK;
}

int launch_pot0_table(const PoissonJob &j, double *tab, hipStream_t s) {
    if (!tab || j.n_t <= 0) {
        set_error("pot0 table: needs n_t > 0 and an output");
        return PIC_ERR_ARG;
    }
    PoissonArgs a;
    memset(&a, 0, sizeof(a));
    int n_harm = 0;
    for (int h = 0; h < PIC_MAX_HARMONICS; ++h) {
        a.wave_cos[h] = j.wave_cos[h];
        a.wave_sin[h] = j.wave_sin[h];
        if (j.wave_cos[h] != 0.0 || j.wave_sin[h] != 0.0) n_harm = h + 1;
    }
    k_pot0_table<<<(j.n_t + 255) / 256, 256, 0, s>>>(j.n_t, n_harm, a, tab);
    PIC_LAUNCHED("k_pot0_table");
    return PIC_OK;
}

int launch_sum_replicas(const PoissonJob &j, const void *dens, double *out, hipStream_t s) {
    if (!dens || !out || j.n_g <= 0 || (j.replicas != 1 && (j.replicas % 8) != 0)) {
        set_error("sum_replicas: bad arguments");
        return PIC_ERR_ARG;
    }
    PoissonGeom g;
    memset(&g, 0, sizeof(g));
    g.n_g = j.n_g;
    g.reps = std::max(1, j.replicas);
    g.rep_stride = j.rep_stride > 0 ? j.rep_stride : j.n_g;
    if (j.precision == PIC_PREC_FP64) k_sum_replicas<double><<<(j.n_g + 255) / 256, 256, 0, s>>>(g, (const double *)dens, out);
    else k_sum_replicas<float><<<(j.n_g + 255) / 256, 256, 0, s>>>(g, (const float *)dens, out);
    PIC_LAUNCHED("k_sum_replicas");
    return PIC_OK;
}

// trig[2t], trig[2t+1] = cos, sin of the phase of step t, as waveform()
// evaluates them for the first harmonic
__global__ void k_trig_table(int n_t, double *trig) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_t) return;
    const uint64_t m = (2ull * (uint64_t)t) % (2ull * (uint64_t)n_t);
    double sv, cv;
    sincospi((double)m / (double)n_t, &sv, &cv);
    trig[2 * t] = cv;
    trig[2 * t + 1] = sv;
}

__global__ void k_ens_pot0(const EnsInst *inst, int n_inst, const double *trig, uint32_t t, double *pot0) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n_inst) pot0[b] = pot0_from_trig(inst[b].v0, trig[2 * t], trig[2 * t + 1]);
}

template <typename D>
__global__ void k_sum_replicas_ens(PoissonGeom g, const D *dens, int inst_stride, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (k < g.n_g) out[(size_t)b * g.n_g + k] = sum_replicas(dens + (size_t)b * inst_stride, g.reps, g.rep_stride, k);
}

int launch_ens_pot0_init(const EnsInst *inst, int n_inst, int n_t, uint32_t t, double *trig, double *pot0,
                         hipStream_t s) {
    if (!inst || n_inst <= 0 || n_t <= 0 || t >= (uint32_t)n_t || !trig || !pot0) {
        set_error("ens_pot0_init: bad arguments");
        return PIC_ERR_ARG;
    }
    k_trig_table<<<(n_t + 255) / 256, 256, 0, s>>>(n_t, trig);
    k_ens_pot0<<<(n_inst + 255) / 256, 256, 0, s>>>(inst, n_inst, trig, t, pot0);
    PIC_LAUNCHED("k_ens_pot0");
    return PIC_OK;
}

int launch_sum_replicas_ens(const PoissonGeom &g, int precision, const void *dens, int inst_stride, int n_inst,
                            double *out, hipStream_t s) {
    if (!dens || !out || n_inst <= 0 || g.n_g <= 0 || (g.reps != 1 && (g.reps % 8) != 0)) {
        set_error("sum_replicas_ens: bad arguments");
        return PIC_ERR_ARG;
    }
    const dim3 grid((g.n_g + 255) / 256, n_inst);
    if (precision == PIC_PREC_FP64) k_sum_replicas_ens<double><<<grid, 256, 0, s>>>(g, (const double *)dens, inst_stride, out);
    else k_sum_replicas_ens<float><<<grid, 256, 0, s>>>(g, (const float *)dens, inst_stride, out);
    PIC_LAUNCHED("k_sum_replicas_ens");
    return PIC_OK;
}

// Ensemble: one workgroup per discharge, all discharges in one launch.
struct EnsPoissonArgs {
    PoissonGeom g;
    const EnsInst *inst;
    const void *e_dens, *i_dens;
    int inst_stride;
    const uint32_t *time;
    int n_t, n_sub;
    float *pot, *efield;
    double *cumul_e, *cumul_i;
    unsigned long long *counters;
    uint32_t *zero_a, *zero_b;
    int zero_words;  // 32-bit words per instance
    uint32_t *cnt;
    int copy_src, copy_dst, zc0, zc1;
    double *pot0;            // [n_inst] pot0 of the step (read, then the next step's written)
    const double *trig;      // [n_t][2] cos, sin of the drive's phase
    const double *i_sum_in;  // [n_inst][n_g] summed ion densities (steps without ions)
    double *i_sum_out;       // [n_inst][n_g] written (ion steps)
};


template <typename T, typename D>
__global__ __launch_bounds__(kPoissonThreads) void k_poisson_ens(EnsPoissonArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    T *pot_s = reinterpret_cast<T *>(smem);
    T *wtot = pot_s + a.g.n_g;
    __shared__ double s_pot0;
    __shared__ int s_ion;
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int n_g = a.g.n_g;
    const size_t off = (size_t)b * a.inst_stride;
    uint32_t t = 0;
    double p0 = 0.0, v0 = 0.0;
    if (tid == 0) {
        t = *a.time;
        if (a.pot0) {
            p0 = a.pot0[b];
            v0 = a.inst[b].v0;
        }
    }
    double dne[kPoissonMaxPer], dni[kPoissonMaxPer];
    load_node_densities<D>(a.g, (const D *)a.e_dens + off, (const D *)a.i_dens + off,
                           a.i_sum_in ? a.i_sum_in + (size_t)b * n_g : nullptr, dne, dni);
    double nc = 0.0, ns = 0.0;
    if (tid == 0) {
        s_ion = (t % (uint32_t)a.n_sub) == 0;
        if (a.pot0) {
            // the next step's cos / sin: the loads are in flight during the solve
            const uint32_t tn = (t + 1u) % (uint32_t)a.n_t;
            nc = a.trig[2 * tn];
            ns = a.trig[2 * tn + 1];
        } else {
            const double wc = a.inst[b].v0, ws = 0.0;
            p0 = waveform(t, a.n_t, 1, &wc, &ws);
        }
        s_pot0 = p0;
    }
    __syncthreads();
    const bool ion = s_ion != 0;
    poisson_solve<T>(a.g, dne, dni, s_pot0, ion,
                        a.pot + (size_t)b * n_g, a.efield + (size_t)b * n_g,
                        a.cumul_e ? a.cumul_e + (size_t)b * n_g : nullptr,
                        a.cumul_i ? a.cumul_i + (size_t)b * n_g : nullptr, pot_s, wtot);
    if (a.i_sum_out) {
        const int k0 = tid * a.g.per;
#pragma unroll
        for (int u = 0; u < kPoissonMaxPer; ++u)
            if (u < a.g.per && k0 + u < n_g) a.i_sum_out[(size_t)b * n_g + k0 + u] = dni[u];
    }
    const size_t zoff = (size_t)b * a.zero_words;
    if (a.zero_a)
        for (int w = tid; w < a.zero_words; w += blockDim.x) a.zero_a[zoff + w] = 0u;
    if (ion && a.zero_b)
        for (int w = tid; w < a.zero_words; w += blockDim.x) a.zero_b[zoff + w] = 0u;
    if (tid == 0 && a.pot0) a.pot0[b] = pot0_from_trig(v0, nc, ns);
    uint32_t *c = a.cnt + (size_t)b * EC_WORDS;
    if (tid == 0) {
        if (a.copy_dst >= 0) c[a.copy_dst] = c[a.copy_src];
        if (a.zc0 >= 0) c[a.zc0] = 0u;
        if (a.zc1 >= 0) c[a.zc1] = 0u;
    }
    if (tid == 3 && a.cumul_e) {
        a.counters[(size_t)b * 8 + 6] += 1ull;
        if (ion) a.counters[(size_t)b * 8 + 7] += 1ull;
    }
}

int launch_poisson_ens(const EnsPoissonJob &j, hipStream_t s) {
    if (j.g.n_g < 3 || j.g.n_g > kPoissonThreads * kPoissonMaxPer || j.n_inst <= 0 ||
        (j.g.reps != 1 && (j.g.reps % 8) != 0)) {
        set_error("poisson_ens: bad arguments");
        return PIC_ERR_ARG;
    }
    EnsPoissonArgs a;
    memset(&a, 0, sizeof(a));
    a.g = j.g;
    a.g.per = poisson_per(j.g.n_g);
    a.inst = j.inst;
    a.e_dens = j.e_dens;
    a.i_dens = j.i_dens;
    a.inst_stride = j.inst_stride;
    a.time = j.time;
    a.n_t = j.n_t;
    a.n_sub = j.n_sub;
    a.pot = j.pot;
    a.efield = j.efield;
    a.cumul_e = j.cumul_e;
    a.cumul_i = j.cumul_i;
    a.counters = j.counters;
    a.zero_a = (uint32_t *)j.zero_a;
    a.zero_b = (uint32_t *)j.zero_b;
    a.cnt = j.cnt;
    a.copy_src = j.copy_src;
    a.copy_dst = j.copy_dst;
    a.zc0 = j.zc0;
    a.zc1 = j.zc1;
    a.pot0 = j.pot0;
    a.trig = j.trig;
    a.i_sum_in = j.i_sum_in;
    a.i_sum_out = j.i_sum_out;
    if (a.pot0 && !a.trig) {
        set_error("poisson_ens: pot0 needs the phase table");
        return PIC_ERR_ARG;
    }
    switch (j.precision) {
    case PIC_PREC_FP32:
        a.zero_words = j.inst_stride;
        k_poisson_ens<float, float><<<j.n_inst, poisson_threads(j.g.n_g), (j.g.n_g + kPoissonWtot) * sizeof(float), s>>>(a);
        break;
    case PIC_PREC_COMP:
        a.zero_words = j.inst_stride;
        k_poisson_ens<ff, float><<<j.n_inst, poisson_threads(j.g.n_g), (j.g.n_g + kPoissonWtot) * sizeof(ff), s>>>(a);
        break;
    default:
        a.zero_words = 2 * j.inst_stride;
        k_poisson_ens<double, double><<<j.n_inst, poisson_threads(j.g.n_g), (j.g.n_g + kPoissonWtot) * sizeof(double), s>>>(a);
        break;
    }
    PIC_LAUNCHED("k_poisson_ens");
    return PIC_OK;
}

// solve_poisson_dps_flexible (main.rs:496-564): psi'' = src on n nodes.
template <typename T>
__global__ __launch_bounds__(kPoissonThreads) void k_poisson_flexible(const float *src, int n, int per,
                                                                        float psi_l, float psi_r,
                                                                        float dx, float *pot) {
    __shared__ T wtot[17];
    const int tid = threadIdx.x;
    const int k0 = tid * per;
    const T dx2 = Num<T>::fromf(dx) * Num<T>::fromf(dx);
    T g[kPoissonMaxPer];
    T run = Num<T>::zero();
#pragma unroll
    for (int u = 0; u < kPoissonMaxPer; ++u) {
        const int k = k0 + u;
        g[u] = Num<T>::zero();
        if (u < per && k >= 1 && k <= n - 2) {
            T f = dx2 * Num<T>::fromf(src[k]);
            if (k == 1) f = f - Num<T>::fromf(psi_l);
            if (k == n - 2) f = f - Num<T>::fromf(psi_r);
            g[u] = Num<T>::from((double)k) * f;
        }
        run = run + g[u];
        g[u] = run;
    }
    T excl = block_scan_incl<T>(run, wtot) - run;
    T r[kPoissonMaxPer], R[kPoissonMaxPer];
    run = Num<T>::zero();
#pragma unroll
    for (int u = 0; u < kPoissonMaxPer; ++u) {
        const int k = k0 + u;
        r[u] = Num<T>::zero();
        if (u < per && k >= 1 && k <= n - 2) {
            const T h = -(excl + g[u]) / Num<T>::from((double)k + 1.0);
            r[u] = h / Num<T>::from((double)k);
        }
        run = run + r[u];
        R[u] = run;
    }
    excl = block_scan_incl<T>(run, wtot) - run;
#pragma unroll
    for (int u = 0; u < kPoissonMaxPer; ++u)
        if (u < per && k0 + u == n - 2) wtot[16] = excl + R[u];
    __syncthreads();
    const T r_total = wtot[16];
#pragma unroll
    for (int u = 0; u < kPoissonMaxPer; ++u) {
        const int k = k0 + u;
        if (u < per && k < n) {
            T pk;
            if (k == 0) pk = Num<T>::fromf(psi_l);
            else if (k < n - 1) pk = Num<T>::from((double)k) * (r_total - (excl + R[u]) + r[u]);
            else pk = Num<T>::fromf(psi_r);
            pot[k] = Num<T>::to_f(pk);
        }
    }
}

// ============================================================ push + compaction
struct MoveArgs {
    const float *efield;
    const float *sx, *svx, *svy, *svz;
    float *dx, *dvx, *dvy, *dvz;
    uint32_t *alive;
    const uint32_t *n_active;
    float factor, dt, half_m_over_e;
    unsigned long long *abs_counts;
    uint32_t *ifed;
    uint32_t n_cap;
    int n_g;
    float inv_dx, gap;
};

// Workgroup-wide slot reservation for one tile: every wave reports its
// survivor count, thread 0 prefix-sums the (<= 16) wave counts and takes the
// tile's range with a single global atomic, and every wave reads back its
// base.  `buf` alternates between two LDS slots so that one tile costs two
// barriers.
__device__ __forceinline__ uint32_t tile_reserve(uint32_t wave_count, uint32_t *s_wave,
                                                 uint32_t *s_base, int buf, uint32_t *counter) {
    const int wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (lane_id() == 0) s_wave[buf * 16 + wave] = wave_count;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < nw; ++w) {
            uint32_t c = s_wave[buf * 16 + w];
            s_wave[buf * 16 + w] = acc;
            acc += c;
        }
        s_base[buf] = acc ? atomicAdd(counter, acc) : 0u;
    }
    __syncthreads();
    return s_base[buf] + s_wave[buf * 16 + wave];
}

// move_and_compact (main.rs:566-678) on wave64.  A persistent grid of one
// 1024-lane workgroup per compute unit walks tiles of 1024 * PPT particles;
// each thread takes PPT particles of a tile (strided by the workgroup size,
// so every load is coalesced), survivors are ranked with 64-bit ballots and
// v_mbcnt, and the tile reserves its output range with one global atomic.
// The next tile's loads are issued before the current tile's reservation
// and stores, so the reservations -- same-address atomics on the single
// alive counter, which serialise in the memory-side atomic unit -- overlap
// with the streaming instead of forming a phase of their own.  The efield
// array is staged in LDS.  Lost particles are counted per electrode (one
// atomic per workgroup) and, for ions, binned by energy (IFED,
// C/eduPIC.cc:625-667).
constexpr int kMoveBlock = 1024;

// Sum v over the workgroup and add the total to *dst with one atomic: the
// counters of these kernels are single words, and thousands of waves adding
// to one address serialise in the memory-side atomic unit.  Every thread of
// the workgroup calls it; only thread 0 reads the LDS sum, so consecutive
// calls need no barrier between them.
__device__ void wg_atomic_add(unsigned long long *dst, uint32_t v) {
    __shared__ uint32_t s_sum;
    if (threadIdx.x == 0) s_sum = 0u;
    __syncthreads();
    const uint32_t w = wave_sum(v);
    if (lane_id() == 0 && w) atomicAdd(&s_sum, w);
    __syncthreads();
    if (threadIdx.x == 0 && s_sum) atomicAdd(dst, (unsigned long long)s_sum);
}

template <int PPT>
__global__ __launch_bounds__(kMoveBlock) void k_move_compact(MoveArgs a) {
    extern __shared__ float ef[];  // n_g
    __shared__ uint32_t s_wave[32], s_base[2];
    __shared__ uint32_t s_ifed[2 * kIfedBins];
    const bool do_ifed = a.ifed != nullptr;
    for (int k = threadIdx.x; k < a.n_g; k += blockDim.x) ef[k] = a.efield[k];
    if (do_ifed)
        for (int k = threadIdx.x; k < 2 * kIfedBins; k += blockDim.x) s_ifed[k] = 0u;
    __syncthreads();
    const uint32_t n = min(*a.n_active, a.n_cap);
    const uint32_t tile_sz = blockDim.x * PPT;
    const uint32_t n_tiles = (n + tile_sz - 1) / tile_sz;
    uint32_t lost_pow = 0, lost_gnd = 0;
    // loads of one tile; lanes past the count read slot 0 (the count test
    // is applied where the values are used, so no wait sits at the issue)
    auto load = [&](uint32_t tile, float (&x)[PPT], float (&vx)[PPT], float (&vy)[PPT], float (&vz)[PPT]) {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const uint32_t i0 = tile * tile_sz + threadIdx.x + j * blockDim.x;
            const uint32_t i = tile < n_tiles && i0 < n ? i0 : 0u;
            x[j] = a.sx[i];
            vx[j] = a.svx[i];
            vy[j] = a.svy[i];
            vz[j] = a.svz[i];
        }
    };
    float cx[PPT], cvx[PPT], cvy[PPT], cvz[PPT];
    if (n_tiles > blockIdx.x) load(blockIdx.x, cx, cvx, cvy, cvz);
    int buf = 0;
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x, buf ^= 1) {
        float px[PPT], pvx[PPT], pvy[PPT], pvz[PPT];
        load(tile + gridDim.x, px, pvx, pvy, pvz);
        float nx[PPT], nvx[PPT];
        bool keep[PPT];
        uint32_t off[PPT];
        const uint32_t base = tile * tile_sz + threadIdx.x;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            keep[j] = false;
            nx[j] = nvx[j] = 0.0f;
            if (base + j * blockDim.x < n) {
                const float e = field_at(ef, cx[j], a.inv_dx, a.n_g);
                nvx[j] = __builtin_fmaf(a.factor, e, cvx[j]);
                nx[j] = __builtin_fmaf(nvx[j], a.dt, cx[j]);
                keep[j] = nx[j] >= 0.0f && nx[j] <= a.gap;
                if (!keep[j]) {
                    const bool pw = nx[j] < 0.0f;
                    lost_pow += pw;
                    lost_gnd += !pw;
                    if (do_ifed) {
                        const float en = a.half_m_over_e * (nvx[j] * nvx[j] + cvy[j] * cvy[j] + cvz[j] * cvz[j]);
                        const int b = (int)en;
                        if (en >= 0.0f && b < kIfedBins) atomicAdd(&s_ifed[(pw ? 0 : 1) * kIfedBins + b], 1u);
                    }
                }
            }
        }
        uint32_t wcount = 0;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const uint64_t m = ballot64(keep[j]);
            off[j] = wcount + mbcnt64(m);
            wcount += (uint32_t)__popcll(m);
        }
        const uint32_t wbase = tile_reserve(wcount, s_wave, s_base, buf, a.alive);
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            if (keep[j]) {
                const uint32_t slot = wbase + off[j];
                a.dx[slot] = nx[j];
                a.dvx[slot] = nvx[j];
                a.dvy[slot] = cvy[j];
                a.dvz[slot] = cvz[j];
            }
        }
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            cx[j] = px[j];
            cvx[j] = pvx[j];
            cvy[j] = pvy[j];
            cvz[j] = pvz[j];
        }
    }
    if (a.abs_counts) {
        wg_atomic_add(&a.abs_counts[0], lost_pow);
        wg_atomic_add(&a.abs_counts[1], lost_gnd);
    }
    if (do_ifed) {
        __syncthreads();
        for (int k = threadIdx.x; k < 2 * kIfedBins; k += blockDim.x)
            if (s_ifed[k]) atomicAdd(&a.ifed[k], s_ifed[k]);
    }
}

int launch_move_compact(const MoveJob &j, const KPhys &k, const Shape &sh0, hipStream_t s) {
    Shape sh = resolve(sh0);
    MoveArgs a;
    a.efield = j.efield;
    a.sx = j.sx;
    a.svx = j.svx;
    a.svy = j.svy;
    a.svz = j.svz;
    a.dx = j.dx;
    a.dvx = j.dvx;
    a.dvy = j.dvy;
    a.dvz = j.dvz;
    a.alive = j.alive;
    a.n_active = j.n_active;
    a.factor = j.factor;
    a.dt = j.dt;
    a.half_m_over_e = j.half_m_over_e;
    a.abs_counts = j.abs_counts;
    a.ifed = j.ifed;
    a.n_cap = j.n_cap;
    a.n_g = k.n_g;
    a.inv_dx = k.inv_dx;
    a.gap = k.gap;
    const size_t lds = (size_t)k.n_g * sizeof(float);
    const int block = kMoveBlock;
    const int grid = device_cus();
    switch (sh.ppt) {
    case 1: k_move_compact<1><<<grid, block, lds, s>>>(a); break;
    case 2: k_move_compact<2><<<grid, block, lds, s>>>(a); break;
    case 8: k_move_compact<8><<<grid, block, lds, s>>>(a); break;
    default: k_move_compact<4><<<grid, block, lds, s>>>(a); break;
    }
    PIC_LAUNCHED("k_move_compact");
    return PIC_OK;
}

// ============================================================ collisions
// Both stage collision kernels (check_collisions_e / _i with collision_e /
// _i, main.rs:971-1292) run as persistent workgroups over tiles of
// 256 * kCollidePPT slots, the next tile's loads in flight while the
// current one is tested.  The collision test only lists a colliding particle
// (its slot, velocity, RNG state after the test draw and energy bin; ions
// also the partner atom's velocity and the relative speed) in LDS; after the
// last tile the workgroup scatters its list 64 entries per wave pass, one
// entry per lane.  With ~0.5 % of electrons colliding per step, a wave that
// ran the long scattering code for its one or two colliding lanes in every
// tile spent most of the kernel there.  Each particle's stream continues
// from the listed state, so its draws are those of the sequential order.  A
// full list (far above the tens of entries a workgroup sees) scatters the
// excess in place.  Ionisations append the new electron and ion with one
// atomic per species per wave (lanes ranked by v_mbcnt).
constexpr int kCollideBlock = 256;
constexpr int kCollidePPT = 2;
constexpr uint32_t kStageList = 256;

struct CollEArgs {
    const float *stot_e, *cs;
    const uint32_t *n_active;
    float *x, *vx, *vy, *vz;
    uint32_t *r0, *r1, *r2, *r3;
    float *ix, *ivx, *ivy, *ivz;
    uint32_t *alive_e, *alive_i;
    uint32_t cap_e, cap_i;
    uint32_t *overflow;
    unsigned long long *n_coll;
    KPhys k;
};

// The ionisation products of the lanes with kind == kIonisation: the new
// electron at (*alive_e)++ (position of the parent, velocity se), the new
// ion at (*alive_i)++ (velocity si); appends past the capacities are
// dropped and flagged.
__device__ __forceinline__ void stage_ionisation(const CollEArgs &a, int kind, uint32_t parent,
                                                 const float *se, const float *si) {
    const uint64_t sp = ballot64(kind == kIonisation);
    if (!sp) return;
    const int leader = __builtin_ctzll(sp);
    uint32_t be = 0, bi = 0;
    if ((int)lane_id() == leader) {
        const uint32_t cnt = (uint32_t)__popcll(sp);
        be = atomicAdd(a.alive_e, cnt);
        bi = atomicAdd(a.alive_i, cnt);
    }
    be = __shfl(be, leader, kWave);
    bi = __shfl(bi, leader, kWave);
    if (kind != kIonisation) return;
    const uint32_t o = mbcnt64(sp);
    const float xi = a.x[parent];
    const uint32_t je = be + o, ji = bi + o;
    if (je < a.cap_e) {
        a.x[je] = xi;
        a.vx[je] = se[0];
        a.vy[je] = se[1];
        a.vz[je] = se[2];
    } else {
        *a.overflow = 1u;
    }
    if (ji < a.cap_i) {
        a.ix[ji] = xi;
        a.ivx[ji] = si[0];
        a.ivy[ji] = si[1];
        a.ivz[ji] = si[2];
    } else {
        *a.overflow = 1u;
    }
}

__global__ __launch_bounds__(kCollideBlock) void k_collide_e(CollEArgs a) {
    // list fields: slot, vx, vy, vz, 4 RNG words, bin
    __shared__ uint32_t list[9 * kStageList];
    __shared__ uint32_t s_n;
    const KPhys &k = a.k;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const uint32_t n = min(*a.n_active, a.cap_e);
    const uint32_t tile_sz = blockDim.x * kCollidePPT;
    const uint32_t n_tiles = (n + tile_sz - 1) / tile_sz;
    uint32_t ncoll = 0;
    auto load = [&](uint32_t tile, float (&vx)[kCollidePPT], float (&vy)[kCollidePPT], float (&vz)[kCollidePPT],
                    Xo (&st)[kCollidePPT]) {
#pragma unroll
        for (int u = 0; u < kCollidePPT; ++u) {
            const uint32_t i0 = tile * tile_sz + threadIdx.x + u * blockDim.x;
            const uint32_t i = tile < n_tiles && i0 < n ? i0 : 0u;
            vx[u] = a.vx[i];
            vy[u] = a.vy[i];
            vz[u] = a.vz[i];
            st[u] = rng_load(a.r0, a.r1, a.r2, a.r3, i);
        }
    };
    float vx[kCollidePPT], vy[kCollidePPT], vz[kCollidePPT];
    Xo st[kCollidePPT];
    if (n_tiles > blockIdx.x) load(blockIdx.x, vx, vy, vz, st);
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        float pvx[kCollidePPT], pvy[kCollidePPT], pvz[kCollidePPT];
        Xo pst[kCollidePPT];
        load(tile + gridDim.x, pvx, pvy, pvz, pst);
#pragma unroll
        for (int u = 0; u < kCollidePPT; ++u) {
            const uint32_t i = tile * tile_sz + threadIdx.x + u * blockDim.x;
            const bool valid = i < n;
            bool hit = false;
            int idx = 0;
            if (valid) {
                float speed;
                idx = e_index(k, vx[u], vy[u], vz[u], speed);
                hit = e_test(k, a.stot_e[idx], speed, st[u]);
                rng_store(a.r0, a.r1, a.r2, a.r3, i, st[u]);
            }
            const uint64_t m = ballot64(hit);
            if (!m) continue;
            const int leader = __builtin_ctzll(m);
            uint32_t base = 0;
            if ((int)lane_id() == leader) base = atomicAdd(&s_n, (uint32_t)__popcll(m));
            base = __shfl(base, leader, kWave);
            const uint32_t pos = base + mbcnt64(m);
            const bool listed = hit && pos < kStageList;
            if (listed) {
                uint32_t *e = list + pos;
                e[0 * kStageList] = i;
                e[1 * kStageList] = __float_as_uint(vx[u]);
                e[2 * kStageList] = __float_as_uint(vy[u]);
                e[3 * kStageList] = __float_as_uint(vz[u]);
                e[4 * kStageList] = st[u].s0;
                e[5 * kStageList] = st[u].s1;
                e[6 * kStageList] = st[u].s2;
                e[7 * kStageList] = st[u].s3;
                e[8 * kStageList] = (uint32_t)idx;
            }
            const bool local = hit && !listed;
            if (ballot64(local)) {  // a full list: scatter here
                int kind = kNone;
                float se[3], si[3];
                if (local) {
                    float w = vx[u], y = vy[u], z = vz[u];
                    Xo s2 = st[u];
                    kind = e_scatter(k, a.cs[idx], a.cs[kCsRanges + idx], a.cs[2 * kCsRanges + idx], w, y, z, s2,
                                     se, si);
                    rng_store(a.r0, a.r1, a.r2, a.r3, i, s2);
                    a.vx[i] = w;
                    a.vy[i] = y;
                    a.vz[i] = z;
                    ++ncoll;
                }
                stage_ionisation(a, kind, i, se, si);
            }
        }
#pragma unroll
        for (int u = 0; u < kCollidePPT; ++u) {
            vx[u] = pvx[u];
            vy[u] = pvy[u];
            vz[u] = pvz[u];
            st[u] = pst[u];
        }
    }
    __syncthreads();
    const uint32_t n_list = min(s_n, kStageList);
    const uint32_t wave = threadIdx.x / kWave;
    for (uint32_t b0 = wave * kWave; b0 < n_list; b0 += blockDim.x) {
        const uint32_t e = b0 + lane_id();
        int kind = kNone;
        float se[3], si[3];
        uint32_t slot = 0;
        if (e < n_list) {
            const uint32_t *q = list + e;
            slot = q[0];
            float w = __uint_as_float(q[1 * kStageList]), y = __uint_as_float(q[2 * kStageList]),
                  z = __uint_as_float(q[3 * kStageList]);
            Xo s2{q[4 * kStageList], q[5 * kStageList], q[6 * kStageList], q[7 * kStageList]};
            const int b = (int)q[8 * kStageList];
            kind = e_scatter(k, a.cs[b], a.cs[kCsRanges + b], a.cs[2 * kCsRanges + b], w, y, z, s2, se, si);
            rng_store(a.r0, a.r1, a.r2, a.r3, slot, s2);
            a.vx[slot] = w;
            a.vy[slot] = y;
            a.vz[slot] = z;
            ++ncoll;
        }
        stage_ionisation(a, kind, slot, se, si);
    }
    if (a.n_coll) wg_atomic_add(a.n_coll, ncoll);
}

int launch_collide_e(const CollideEJob &j, const KPhys &k, const Shape &sh0, hipStream_t s) {
    (void)sh0;
    CollEArgs a;
    a.stot_e = j.stot_e;
    a.cs = j.cs;
    a.n_active = j.n_active;
    a.x = j.e.x;
    a.vx = j.e.vx;
    a.vy = j.e.vy;
    a.vz = j.e.vz;
    a.r0 = j.e.rng0;
    a.r1 = j.e.rng1;
    a.r2 = j.e.rng2;
    a.r3 = j.e.rng3;
    a.ix = j.i.x;
    a.ivx = j.i.vx;
    a.ivy = j.i.vy;
    a.ivz = j.i.vz;
    a.alive_e = j.alive_e;
    a.alive_i = j.alive_i;
    a.cap_e = j.e.capacity;
    a.cap_i = j.i.capacity;
    a.overflow = j.overflow;
    a.n_coll = j.n_coll;
    a.k = k;
    // a persistent grid of four 256-lane workgroups per compute unit
    k_collide_e<<<device_cus() * 4, kCollideBlock, 0, s>>>(a);
    PIC_LAUNCHED("k_collide_e");
    return PIC_OK;
}

struct CollIArgs {
    const float *stot_i, *cs;
    const uint32_t *n_active;
    uint32_t n_cap;
    float *vx, *vy, *vz;
    uint32_t *r0, *r1, *r2, *r3;
    unsigned long long *n_coll;
    KPhys k;
};

__global__ __launch_bounds__(kCollideBlock) void k_collide_i(CollIArgs a) {
    // list fields: slot, vx, vy, vz, 4 RNG words, bin, atom velocity (3), relative speed
    __shared__ uint32_t list[13 * kStageList];
    __shared__ uint32_t s_n;
    const KPhys &k = a.k;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const uint32_t n = min(*a.n_active, a.n_cap);
    const uint32_t tile_sz = blockDim.x * kCollidePPT;
    const uint32_t n_tiles = (n + tile_sz - 1) / tile_sz;
    uint32_t ncoll = 0;
    auto load = [&](uint32_t tile, float (&vx)[kCollidePPT], float (&vy)[kCollidePPT], float (&vz)[kCollidePPT],
                    Xo (&st)[kCollidePPT]) {
#pragma unroll
        for (int u = 0; u < kCollidePPT; ++u) {
            const uint32_t i0 = tile * tile_sz + threadIdx.x + u * blockDim.x;
            const uint32_t i = tile < n_tiles && i0 < n ? i0 : 0u;
            vx[u] = a.vx[i];
            vy[u] = a.vy[i];
            vz[u] = a.vz[i];
            st[u] = rng_load(a.r0, a.r1, a.r2, a.r3, i);
        }
    };
    float vx[kCollidePPT], vy[kCollidePPT], vz[kCollidePPT];
    Xo st[kCollidePPT];
    if (n_tiles > blockIdx.x) load(blockIdx.x, vx, vy, vz, st);
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        float pvx[kCollidePPT], pvy[kCollidePPT], pvz[kCollidePPT];
        Xo pst[kCollidePPT];
        load(tile + gridDim.x, pvx, pvy, pvz, pst);
#pragma unroll
        for (int u = 0; u < kCollidePPT; ++u) {
            const uint32_t i = tile * tile_sz + threadIdx.x + u * blockDim.x;
            const bool valid = i < n;
            bool hit = false;
            int idx = 0;
            float ax = 0.0f, ay = 0.0f, az = 0.0f, g = 0.0f;
            if (valid) {
                idx = i_index(k, vx[u], vy[u], vz[u], st[u], ax, ay, az, g);
                hit = i_test(k, a.stot_i[idx], g, st[u]);
                rng_store(a.r0, a.r1, a.r2, a.r3, i, st[u]);
            }
            const uint64_t m = ballot64(hit);
            if (!m) continue;
            const int leader = __builtin_ctzll(m);
            uint32_t base = 0;
            if ((int)lane_id() == leader) base = atomicAdd(&s_n, (uint32_t)__popcll(m));
            base = __shfl(base, leader, kWave);
            const uint32_t pos = base + mbcnt64(m);
            if (hit && pos < kStageList) {
                uint32_t *e = list + pos;
                e[0 * kStageList] = i;
                e[1 * kStageList] = __float_as_uint(vx[u]);
                e[2 * kStageList] = __float_as_uint(vy[u]);
                e[3 * kStageList] = __float_as_uint(vz[u]);
                e[4 * kStageList] = st[u].s0;
                e[5 * kStageList] = st[u].s1;
                e[6 * kStageList] = st[u].s2;
                e[7 * kStageList] = st[u].s3;
                e[8 * kStageList] = (uint32_t)idx;
                e[9 * kStageList] = __float_as_uint(ax);
                e[10 * kStageList] = __float_as_uint(ay);
                e[11 * kStageList] = __float_as_uint(az);
                e[12 * kStageList] = __float_as_uint(g);
            } else if (hit) {  // a full list: scatter here
                float w = vx[u], y = vy[u], z = vz[u];
                Xo s2 = st[u];
                i_scatter(k, a.cs[3 * kCsRanges + idx], a.cs[4 * kCsRanges + idx], w, y, z, ax, ay, az, g, s2);
                rng_store(a.r0, a.r1, a.r2, a.r3, i, s2);
                a.vx[i] = w;
                a.vy[i] = y;
                a.vz[i] = z;
                ++ncoll;
            }
        }
#pragma unroll
        for (int u = 0; u < kCollidePPT; ++u) {
            vx[u] = pvx[u];
            vy[u] = pvy[u];
            vz[u] = pvz[u];
            st[u] = pst[u];
        }
    }
    __syncthreads();
    const uint32_t n_list = min(s_n, kStageList);
    for (uint32_t e = threadIdx.x; e < n_list; e += blockDim.x) {
        const uint32_t *q = list + e;
        const uint32_t slot = q[0];
        float w = __uint_as_float(q[1 * kStageList]), y = __uint_as_float(q[2 * kStageList]),
              z = __uint_as_float(q[3 * kStageList]);
        Xo s2{q[4 * kStageList], q[5 * kStageList], q[6 * kStageList], q[7 * kStageList]};
        const int b = (int)q[8 * kStageList];
        i_scatter(k, a.cs[3 * kCsRanges + b], a.cs[4 * kCsRanges + b], w, y, z, __uint_as_float(q[9 * kStageList]),
                  __uint_as_float(q[10 * kStageList]), __uint_as_float(q[11 * kStageList]),
                  __uint_as_float(q[12 * kStageList]), s2);
        rng_store(a.r0, a.r1, a.r2, a.r3, slot, s2);
        a.vx[slot] = w;
        a.vy[slot] = y;
        a.vz[slot] = z;
        ++ncoll;
    }
    if (a.n_coll) wg_atomic_add(a.n_coll, ncoll);
}

int launch_collide_i(const CollideIJob &j, const KPhys &k, const Shape &sh0, hipStream_t s) {
    (void)sh0;
    CollIArgs a;
    a.stot_i = j.stot_i;
    a.cs = j.cs;
    a.n_active = j.n_active;
    a.n_cap = j.i.capacity;
    a.vx = j.i.vx;
    a.vy = j.i.vy;
    a.vz = j.i.vz;
    a.r0 = j.i.rng0;
    a.r1 = j.i.rng1;
    a.r2 = j.i.rng2;
    a.r3 = j.i.rng3;
    a.n_coll = j.n_coll;
    a.k = k;
    k_collide_i<<<device_cus() * 4, kCollideBlock, 0, s>>>(a);
    PIC_LAUNCHED("k_collide_i");
    return PIC_OK;
}

// ============================================================ measurement
// accumulate_density (main.rs:1294-1299)
template <typename D>
__global__ void k_accum_density(const D *__restrict__ d, double *__restrict__ c, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) c[k] += (double)d[k];
}

int launch_accum_density(const void *dens, bool f64, double *cumul, int n_g, hipStream_t s) {
    int grid = (n_g + 255) / 256;
    if (f64) k_accum_density<double><<<grid, 256, 0, s>>>((const double *)dens, cumul, n_g);
    else k_accum_density<float><<<grid, 256, 0, s>>>((const float *)dens, cumul, n_g);
    PIC_LAUNCHED("k_accum_density");
    return PIC_OK;
}

// EEPF bin of one electron: time-centred v_x at its (post-push) position
// (main.rs:1316-1334).  Returns -1 outside the centre window or the range.
__device__ __forceinline__ int eepf_bin(const float *ef, float x, float vx, float vy, float vz,
                                        const KPhys &k) {
    if (x <= k.min_x || x >= k.max_x) return -1;
    const float e = field_at(ef, x, k.inv_dx, k.n_g);
    const float mvx = vx - 0.5f * e * k.eepf_kick;
    const float v2 = mvx * mvx + vy * vy + vz * vz;
    const float en = k.half_me_over_e * v2;
    const uint32_t b = (uint32_t)(en / k.de_eepf);
    return b < (uint32_t)kEepfBins ? (int)b : -1;
}

// accumulate_eepf (main.rs:1301-1339) with an LDS histogram per workgroup.
__global__ __launch_bounds__(256) void k_eepf(const float *__restrict__ x, const float *__restrict__ vx,
                                              const float *__restrict__ vy, const float *__restrict__ vz,
                                              const float *__restrict__ efield,
                                              const uint32_t *__restrict__ n_active, uint32_t n_cap,
                                              uint32_t *__restrict__ eepf, KPhys k) {
    __shared__ uint32_t h[kEepfBins];
    for (int b = threadIdx.x; b < kEepfBins; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    const uint32_t n = min(*n_active, n_cap);
    bool any = false;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int b = eepf_bin(efield, x[i], vx[i], vy[i], vz[i], k);
        if (b >= 0) atomicAdd(&h[b], 1u);
        any = true;
    }
    if (!__syncthreads_or(any)) return;
    for (int b = threadIdx.x; b < kEepfBins; b += blockDim.x)
        if (h[b]) atomicAdd(&eepf[b], h[b]);
}

// Fixed-point scales of the XT partial sums (exact 64-bit integer sums in
// LDS: float LDS atomics run about ten times slower than integer ones on
// gfx950): CIC weight 2^32, weighted velocity 2^16 per m/s, weighted energy
// 2^28 per eV, weighted ionisation probability 2^44.  A workgroup sums at
// most the particles of its share of one segment.
__device__ constexpr float kXtScale[4] = {4294967296.0f, 65536.0f, 268435456.0f, 17592186044416.0f};

template <bool ELECTRON>
__global__ __launch_bounds__(256) void k_xt(XtJob j, KPhys k) {
    extern __shared__ unsigned long long acc[];  // [NQ][n_g], two's complement
    constexpr int NQ = ELECTRON ? 4 : 3;
    const int n_g = k.n_g;
    for (int q = threadIdx.x; q < NQ * n_g; q += blockDim.x) acc[q] = 0ull;
    __syncthreads();
    const float half_kick = 0.5f * (ELECTRON ? k.factor_e : k.factor_i);
    const float to_ev = ELECTRON ? k.half_me_over_e : k.half_mi_over_e;
    const float rate_f = k.dt_e * j.gas_density;
    auto add = [&](int q, int node, float v) {
        atomicAdd(&acc[q * n_g + node], (unsigned long long)__float2ll_rn(v * kXtScale[q]));
    };
    {
        // blockIdx.y: the segment; the x workgroups of a segment stride over it
        const int sg = blockIdx.y;
        const uint32_t n = min(j.cnt[sg], j.seg_size);
        const size_t base = (size_t)sg * j.seg_size;
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
            const float x = j.x[base + i];
            if (!(x >= 0.0f)) continue;  // a hole
            const float pos = x * k.inv_dx;
            const int p = cell_of(pos, n_g);
            const float c2 = pos - (float)p, c1 = 1.0f - c2;
            const float ex = c1 * j.efield[p] + c2 * j.efield[p + 1];
            const float u = j.vx[base + i] + half_kick * ex;
            const float vy = j.vy[base + i], vz = j.vz[base + i];
            const float v2 = u * u + vy * vy + vz * vz;
            const float en = to_ev * v2;
            add(0, p, c1);
            add(0, p + 1, c2);
            add(1, p, c1 * u);
            add(1, p + 1, c2 * u);
            add(2, p, c1 * en);
            add(2, p + 1, c2 * en);
            if (ELECTRON) {
                const float rate = j.sigma_ion[energy_index(en, k)] * sqrtf(v2) * rate_f;
                add(3, p, c1 * rate);
                add(3, p + 1, c2 * rate);
            }
        }
    }
    __syncthreads();
    // the workgroup's partial maps, reduced over the workgroups by k_xt_reduce
    // (atomics from every workgroup onto the same few thousand map entries
    // would serialise)
    float *st = j.stage + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * NQ * n_g;
    for (int q = threadIdx.x; q < NQ * n_g; q += blockDim.x)
        st[q] = (float)((double)(long long)acc[q] / (double)kXtScale[q / n_g]);
}

// map entry e of the bin of this step += the sum of the workgroups'
// partials: workgroup (x, y) sums slice y of the partials for 256 entries
// (coalesced over the entries) and adds it with one atomic per entry
constexpr int kXtBlocksPerCu = 8, kXtSlices = 32;
__global__ __launch_bounds__(256) void k_xt_reduce(XtJob j, int nq, int n_g, int parts) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = nq * n_g;
    if (e >= n) return;
    const int w0 = (int)((int64_t)parts * blockIdx.y / gridDim.y);
    const int w1 = (int)((int64_t)parts * (blockIdx.y + 1) / gridDim.y);
    double sum = 0.0;
#pragma unroll 8
    for (int w = w0; w < w1; ++w) sum += (double)j.stage[(size_t)w * n + e];
    if (sum == 0.0) return;
    const uint32_t t = (*j.clock + (uint32_t)j.n_t - 1u) % (uint32_t)j.n_t;
    const uint32_t tb = t / j.bin_steps;
    const int q = e / n_g, node = e - q * n_g;
    atomicAdd(&j.xt[((size_t)q * n_g + node) * j.n_xt + tb], sum);
}

size_t xt_stage_floats(int n_g) { return (size_t)device_cus() * kXtBlocksPerCu * 4 * n_g; }

int launch_xt(const XtJob &j, const KPhys &k, hipStream_t s) {
    const int nq = j.electrons ? 4 : 3;
    const size_t lds = (size_t)nq * k.n_g * sizeof(unsigned long long);
    if (!j.x || !j.vx || !j.vy || !j.vz || !j.cnt || !j.efield || !j.clock || !j.xt || !j.stage || j.n_seg < 1 ||
        j.bin_steps == 0 || j.n_xt <= 0 || (j.electrons && !j.sigma_ion) || k.n_g < 2) {
        set_error("xt: bad arguments");
        return PIC_ERR_ARG;
    }
    if (lds > (size_t)device_lds_per_block()) {
        set_error("xt: n_g too large for the per-workgroup accumulators");
        return PIC_ERR_UNSUPPORTED;
    }
    // several workgroups per compute unit to hide the gathers' latency,
    // shared out over the segments (each a y row of the grid); the partial
    // maps fit xt_stage_floats
    const int gx = std::max(1, device_cus() * kXtBlocksPerCu / j.n_seg);
    const dim3 grid(gx, j.n_seg);
    if (j.electrons) k_xt<true><<<grid, 256, lds, s>>>(j, k);
    else k_xt<false><<<grid, 256, lds, s>>>(j, k);
    PIC_LAUNCHED("k_xt");
    k_xt_reduce<<<dim3((nq * k.n_g + 255) / 256, kXtSlices), 256, 0, s>>>(j, nq, k.n_g, gx * j.n_seg);
    PIC_LAUNCHED("k_xt_reduce");
    return PIC_OK;
}

int launch_eepf(const float *x, const float *vx, const float *vy, const float *vz,
                const float *efield, const uint32_t *n_active, uint32_t *eepf, const KPhys &k,
                const Shape &sh0, hipStream_t s, uint32_t n_cap) {
    Shape sh = resolve(sh0);
    const int block = 256;
    const int grid = std::max(1, sh.grid * sh.block / block / 4);
    k_eepf<<<grid, block, 0, s>>>(x, vx, vy, vz, efield, n_active, n_cap, eepf, k);
    PIC_LAUNCHED("k_eepf");
    return PIC_OK;
}

// ============================================================ RNG utilities
// xoshiro128p_test_wrapper (main.rs:1341-1357), n outputs in one launch.
__global__ void k_xo_next_n(uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3, uint32_t n,
                            uint32_t *out, uint32_t *state_out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    Xo s{s0, s1, s2, s3};
    for (uint32_t i = 0; i < n; ++i) out[i] = xo_next(s);
    if (state_out) {
        state_out[0] = s.s0;
        state_out[1] = s.s1;
        state_out[2] = s.s2;
        state_out[3] = s.s3;
    }
}

// Stream k = J^k * master over GF(2), J = jump() as a 128x128 bit matrix.
// mats holds J^(2^m) for m = 0..31 column-major: mats[(m*128 + c)*4 + w] is
// word w of column c.  A lane applies the matrices of the set bits of k.
__global__ __launch_bounds__(256) void k_seed_streams(const uint32_t *__restrict__ mats, uint32_t m0,
                                                      uint32_t m1, uint32_t m2, uint32_t m3,
                                                      uint32_t n, uint32_t *r0, uint32_t *r1,
                                                      uint32_t *r2, uint32_t *r3) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        uint32_t v[4] = {m0, m1, m2, m3};
        for (int m = 0; m < 32; ++m) {
            if (!((k >> m) & 1u)) continue;
            const uint32_t *M = mats + (size_t)m * 128 * 4;
            uint32_t o[4] = {0u, 0u, 0u, 0u};
            for (int c = 0; c < 128; ++c) {
                const uint32_t bit = (v[c >> 5] >> (c & 31)) & 1u;
                const uint32_t msk = 0u - bit;
                o[0] ^= M[c * 4 + 0] & msk;
                o[1] ^= M[c * 4 + 1] & msk;
                o[2] ^= M[c * 4 + 2] & msk;
                o[3] ^= M[c * 4 + 3] & msk;
            }
            v[0] = o[0];
            v[1] = o[1];
            v[2] = o[2];
            v[3] = o[3];
        }
        r0[k] = v[0];
        r1[k] = v[1];
        r2[k] = v[2];
        r3[k] = v[3];
    }
}

// Host side: the jump of main.rs:1473-1487 as a linear map, and its powers.
static void host_xo_next(uint32_t s[4]) {
    const uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = (s[3] << 11) | (s[3] >> 21);
}
static void host_xo_jump(uint32_t s[4]) {
    static const uint32_t JUMP[4] = {0x8764000b, 0xf542d2d3, 0x6fa035c3, 0x77f2db5b};
    uint32_t acc[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; ++w)
        for (int b = 0; b < 32; ++b) {
            if (JUMP[w] & (1u << b))
                for (int q = 0; q < 4; ++q) acc[q] ^= s[q];
            host_xo_next(s);
        }
    for (int q = 0; q < 4; ++q) s[q] = acc[q];
}
// C = A * B over GF(2), column-major 128x128 with 4 words per column.
static void gf2_mul(const uint32_t *A, const uint32_t *B, uint32_t *C) {
    for (int c = 0; c < 128; ++c) {
        uint32_t o[4] = {0, 0, 0, 0};
        for (int k = 0; k < 128; ++k)
            if ((B[c * 4 + (k >> 5)] >> (k & 31)) & 1u)
                for (int w = 0; w < 4; ++w) o[w] ^= A[k * 4 + w];
        for (int w = 0; w < 4; ++w) C[c * 4 + w] = o[w];
    }
}

static const std::vector<uint32_t> &jump_powers() {
    static std::vector<uint32_t> mats;
    static std::once_flag once;
    std::call_once(once, [] {
        mats.assign(32 * 128 * 4, 0u);
        for (int c = 0; c < 128; ++c) {
            uint32_t s[4] = {0, 0, 0, 0};
            s[c >> 5] = 1u << (c & 31);
            host_xo_jump(s);
            for (int w = 0; w < 4; ++w) mats[c * 4 + w] = s[w];
        }
        for (int m = 1; m < 32; ++m)
            gf2_mul(&mats[(m - 1) * 512], &mats[(m - 1) * 512], &mats[m * 512]);
    });
    return mats;
}

}  // namespace pic

// ============================================================ C ABI
using namespace pic;

static inline hipStream_t S(pic_stream_t s) { return (hipStream_t)s; }

extern "C" {

const char *pic_last_error(void) { return pic::g_err.c_str(); }

void pic_params_default(pic_params_t *p) {
    memset(p, 0, sizeof(*p));
    p->n_g = 400;
    p->n_t = 4000;
    p->n_sub = 20;
    p->precision = PIC_PREC_FP32;
    p->freq = 13.56e6;
    p->v0 = 250.0;
    p->gap = 0.025;
    p->pressure = 10.0;
    p->t_gas = 350.0;
    p->weight = 7.0e4;
    p->area = 1.0e-4;
}

int pic_params_derive(pic_params_t *p) {
    if (!p || p->n_g < 3 || p->n_g > 4096 || p->n_t <= 0 || p->n_sub <= 0 || p->freq <= 0 ||
        p->gap <= 0 || p->pressure <= 0 || p->t_gas <= 0 || p->weight <= 0 || p->area <= 0) {
        set_error("pic_params_derive: invalid parameters");
        return PIC_ERR_ARG;
    }
    const double E = 1.60217662e-19, ME = 9.10938356e-31, MAR = 6.63352090e-26;
    const double KB = 1.38064852e-23, EPS0 = 8.85418781e-12;
    const double period = 1.0 / p->freq;
    p->dt_e = period / (double)p->n_t;
    p->dt_i = (double)p->n_sub * p->dt_e;
    p->dx = p->gap / (double)(p->n_g - 1);
    p->inv_dx = 1.0 / p->dx;
    p->gas_density = p->pressure / (KB * p->t_gas);
    p->inv_dx_f = (float)p->inv_dx;
    p->factor_e = (float)(p->dt_e / ME * (-E));
    p->factor_i = (float)(p->dt_i / MAR * E);
    p->weight_factor = (float)(p->weight / (p->area * p->dx));
    p->dt_e_f = (float)p->dt_e;
    p->dt_i_f = (float)p->dt_i;
    p->gap_f = (float)p->gap;
    p->normal_range = (float)std::sqrt(KB * p->t_gas / MAR);
    p->f1 = (float)(ME / (ME + MAR));
    p->f2 = (float)(MAR / (ME + MAR));
    p->half_me_over_e = (float)(0.5 * ME / E);
    p->half_mu_over_e = (float)(0.5 * (MAR / 2.0) / E);
    p->half_mi_over_e = (float)(0.5 * MAR / E);
    p->e_charge_f = (float)E;
    p->alpha_f = (float)(-p->dx * p->dx / EPS0);
    p->half_dx_over_eps_f = (float)(p->dx / (2.0 * EPS0));
    p->half_inv_dx_f = (float)(0.5 * p->inv_dx);
    p->min_x_f = (float)(0.45 * p->gap);
    p->max_x_f = (float)(0.55 * p->gap);
    p->de_eepf_f = (float)0.05;
    p->de_cs_f = (float)0.001;
    p->me_f = (float)ME;
    p->exc_th_f = (float)11.5;
    p->ion_th_f = (float)15.8;
    p->eepf_kick_f = (float)(p->dt_e * E / ME);
    bool any = false;
    for (int h = 0; h < PIC_MAX_HARMONICS; ++h) any |= p->wave_cos[h] != 0.0 || p->wave_sin[h] != 0.0;
    if (!any) p->wave_cos[0] = p->v0;
    return PIC_OK;
}

int pic_get_density(const float *x, const uint32_t *n_active, void *density, int density_f64,
                    const pic_params_t *p, pic_stream_t stream) {
    if (!x || !n_active || !density || !p) {
        set_error("pic_get_density: null argument");
        return PIC_ERR_ARG;
    }
    return launch_deposit(x, n_active, density, density_f64 != 0, make_kphys(*p), Shape{}, S(stream));
}

static PoissonJob poisson_job_from(const pic_params_t &p, int precision) {
    PoissonJob j;
    j.precision = precision;
    j.n_g = p.n_g;
    j.n_t = p.n_t;
    j.n_sub = p.n_sub;
    const double E = 1.60217662e-19, EPS0 = 8.85418781e-12;
    j.e_charge = E;
    j.alpha = -p.dx * p.dx / EPS0;
    j.inv_dx = p.inv_dx;
    j.half_inv_dx = 0.5 * p.inv_dx;
    j.half_dx_over_eps = p.dx / (2.0 * EPS0);
    for (int h = 0; h < PIC_MAX_HARMONICS; ++h) {
        j.wave_cos[h] = p.wave_cos[h];
        j.wave_sin[h] = p.wave_sin[h];
    }
    return j;
}

int pic_solve_poisson(const void *e_density, const void *i_density, double pot0, int precision,
                      float *pot, float *efield, const pic_params_t *p, pic_stream_t stream) {
    if (!e_density || !i_density || !pot || !efield || !p) {
        set_error("pic_solve_poisson: null argument");
        return PIC_ERR_ARG;
    }
    PoissonJob j = poisson_job_from(*p, precision);
    j.e_dens = e_density;
    j.i_dens = i_density;
    j.pot0 = precision == PIC_PREC_FP64 ? pot0 : (double)(float)pot0;
    j.pot = pot;
    j.efield = efield;
    return launch_poisson(j, S(stream));
}

int pic_solve_poisson_flexible(const float *source, int n, float psi_left, float psi_right, float dx,
                               int precision, float *pot, pic_stream_t stream) {
    if (!source || !pot || n < 3 || n > kPoissonThreads * kPoissonMaxPer) {
        set_error("pic_solve_poisson_flexible: bad argument");
        return PIC_ERR_ARG;
    }
    const int per = (n + kPoissonThreads - 1) / kPoissonThreads;
    hipStream_t s = S(stream);
    switch (precision) {
    case PIC_PREC_FP32:
        k_poisson_flexible<float><<<1, kPoissonThreads, 0, s>>>(source, n, per, psi_left, psi_right, dx, pot);
        break;
    case PIC_PREC_COMP:
        k_poisson_flexible<ff><<<1, kPoissonThreads, 0, s>>>(source, n, per, psi_left, psi_right, dx, pot);
        break;
    case PIC_PREC_FP64:
        k_poisson_flexible<double><<<1, kPoissonThreads, 0, s>>>(source, n, per, psi_left, psi_right, dx, pot);
        break;
    default:
        set_error("pic_solve_poisson_flexible: unknown precision");
        return PIC_ERR_ARG;
    }
    PIC_LAUNCHED("k_poisson_flexible");
    return PIC_OK;
}

int pic_move_and_compact(const float *efield, const float *src_x, const float *src_vx,
                         const float *src_vy, const float *src_vz, float *dst_x, float *dst_vx,
                         float *dst_vy, float *dst_vz, uint32_t *alive_counter,
                         const uint32_t *n_active, float factor, float dt, float half_m_over_e,
                         unsigned long long *abs_counts, uint32_t *ifed, const pic_params_t *p,
                         pic_stream_t stream) {
    if (!efield || !src_x || !src_vx || !src_vy || !src_vz || !dst_x || !dst_vx || !dst_vy ||
        !dst_vz || !alive_counter || !n_active || !p) {
        set_error("pic_move_and_compact: null argument");
        return PIC_ERR_ARG;
    }
    MoveJob j{efield, src_x, src_vx, src_vy, src_vz, dst_x, dst_vx, dst_vy, dst_vz,
              alive_counter, n_active, factor, dt, half_m_over_e, abs_counts, ifed, 0xffffffffu};
    return launch_move_compact(j, make_kphys(*p), Shape{}, S(stream));
}

static bool species_ok(const pic_species_t *s, bool need_rng) {
    return s && s->x && s->vx && s->vy && s->vz &&
           (!need_rng || (s->rng0 && s->rng1 && s->rng2 && s->rng3));
}

int pic_check_collisions_e(const float *stot_e, const float *cs, const uint32_t *n_active,
                           const pic_species_t *electrons, const pic_species_t *ions,
                           uint32_t *alive_e, uint32_t *alive_i, uint32_t *overflow,
                           unsigned long long *n_coll, const pic_params_t *p, pic_stream_t stream) {
    if (!stot_e || !cs || !n_active || !species_ok(electrons, true) || !species_ok(ions, false) ||
        !alive_e || !alive_i || !overflow || !p) {
        set_error("pic_check_collisions_e: null argument");
        return PIC_ERR_ARG;
    }
    CollideEJob j{stot_e, cs, n_active, *electrons, *ions, alive_e, alive_i, overflow, n_coll};
    return launch_collide_e(j, make_kphys(*p), Shape{}, S(stream));
}

int pic_check_collisions_i(const float *stot_i, const float *cs, const uint32_t *n_active,
                           const pic_species_t *ions, unsigned long long *n_coll,
                           const pic_params_t *p, pic_stream_t stream) {
    if (!stot_i || !cs || !n_active || !species_ok(ions, true) || !p) {
        set_error("pic_check_collisions_i: null argument");
        return PIC_ERR_ARG;
    }
    CollideIJob j{stot_i, cs, n_active, *ions, n_coll};
    return launch_collide_i(j, make_kphys(*p), Shape{}, S(stream));
}

int pic_accumulate_density(const void *density, int density_f64, double *cumul, int n_g,
                           pic_stream_t stream) {
    if (!density || !cumul || n_g <= 0) {
        set_error("pic_accumulate_density: bad argument");
        return PIC_ERR_ARG;
    }
    return launch_accum_density(density, density_f64 != 0, cumul, n_g, S(stream));
}

int pic_accumulate_eepf(const float *x, const float *vx, const float *vy, const float *vz,
                        const float *efield, const uint32_t *n_active, uint32_t *eepf,
                        const pic_params_t *p, pic_stream_t stream) {
    if (!x || !vx || !vy || !vz || !efield || !n_active || !eepf || !p) {
        set_error("pic_accumulate_eepf: null argument");
        return PIC_ERR_ARG;
    }
    return launch_eepf(x, vx, vy, vz, efield, n_active, eepf, make_kphys(*p), Shape{}, S(stream));
}

int pic_xoshiro_next_n(const uint32_t state[4], uint32_t n, uint32_t *out, uint32_t *state_out,
                       pic_stream_t stream) {
    if (!state || (n && !out)) {
        set_error("pic_xoshiro_next_n: null argument");
        return PIC_ERR_ARG;
    }
    k_xo_next_n<<<1, 64, 0, S(stream)>>>(state[0], state[1], state[2], state[3], n, out, state_out);
    PIC_LAUNCHED("k_xo_next_n");
    return PIC_OK;
}

int pic_seed_streams(const uint32_t master[4], uint32_t n, uint32_t *rng0, uint32_t *rng1,
                     uint32_t *rng2, uint32_t *rng3, pic_stream_t stream) {
    if (!master || !rng0 || !rng1 || !rng2 || !rng3) {
        set_error("pic_seed_streams: null argument");
        return PIC_ERR_ARG;
    }
    if (!(master[0] | master[1] | master[2] | master[3])) {
        set_error("pic_seed_streams: the master state must not be all zero");
        return PIC_ERR_ARG;
    }
    if (n == 0) return PIC_OK;
    // The 64 KiB of jump matrices live in one device buffer per device,
    // uploaded on first use and kept for the life of the process.
    static std::mutex mu;
    static std::vector<uint32_t *> per_dev;
    int dev = 0;
    PIC_HIP(hipGetDevice(&dev));
    uint32_t *mats = nullptr;
    {
        std::lock_guard<std::mutex> lk(mu);
        if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1, nullptr);
        if (!per_dev[dev]) {
            const std::vector<uint32_t> &h = jump_powers();
            uint32_t *d = nullptr;
            PIC_HIP(hipMalloc(&d, h.size() * sizeof(uint32_t)));
            PIC_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
            per_dev[dev] = d;
        }
        mats = per_dev[dev];
    }
    const int grid = (int)std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cus() * 8);
    k_seed_streams<<<grid, 256, 0, S(stream)>>>(mats, master[0], master[1], master[2], master[3], n,
                                                rng0, rng1, rng2, rng3);
    PIC_LAUNCHED("k_seed_streams");
    return PIC_OK;
}

// Streaming device copy (16-byte vector loads and stores): the calibration
// kernel for the HBM traffic counters and a measure of achievable bandwidth.
__global__ __launch_bounds__(256) void k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                     uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

// Read-only stream (16-byte loads, four in flight per lane), summed so the
// loads cannot be dropped: the read half of the bandwidth on its own.
__global__ __launch_bounds__(256) void k_stream_read(const uint4 *__restrict__ src, uint64_t n16,
                                                     unsigned long long *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t acc = 0;
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        acc += (a.x ^ a.y ^ a.z ^ a.w) + (b.x ^ b.y ^ b.z ^ b.w) + (c.x ^ c.y ^ c.z ^ c.w) + (d.x ^ d.y ^ d.z ^ d.w);
    }
    for (; i < n16; i += stride) {
        const uint4 a = src[i];
        acc += a.x ^ a.y ^ a.z ^ a.w;
    }
    acc = wave_sum(acc);
    if (lane_id() == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

int pic_stream_copy(void *dst, const void *src, uint64_t bytes, pic_stream_t stream) {
    if (!dst || !src || (bytes % 16) != 0 || ((uintptr_t)dst % 16) != 0 || ((uintptr_t)src % 16) != 0) {
        set_error("pic_stream_copy: pointers and size must be 16-byte aligned");
        return PIC_ERR_ARG;
    }
    const uint64_t n16 = bytes / 16;
    if (n16 == 0) return PIC_OK;
    const int grid = (int)std::min<uint64_t>((n16 + 255) / 256, (uint64_t)device_cus() * 16);
    k_stream_copy<<<grid, 256, 0, S(stream)>>>((const uint4 *)src, (uint4 *)dst, n16);
    PIC_LAUNCHED("k_stream_copy");
    return PIC_OK;
}

int pic_stream_read(const void *src, uint64_t bytes, unsigned long long *sum, pic_stream_t stream) {
    if (!src || !sum || (bytes % 16) != 0 || ((uintptr_t)src % 16) != 0) {
        set_error("pic_stream_read: pointer and size must be 16-byte aligned");
        return PIC_ERR_ARG;
    }
    const uint64_t n16 = bytes / 16;
    if (n16 == 0) return PIC_OK;
    const int grid = (int)std::min<uint64_t>((n16 + 255) / 256, (uint64_t)device_cus() * 16);
    k_stream_read<<<grid, 256, 0, S(stream)>>>((const uint4 *)src, n16, sum);
    PIC_LAUNCHED("k_stream_read");
    return PIC_OK;
}

int pic_memcpy_dtod_async(void *dst, const void *src, uint64_t bytes, pic_stream_t stream) {
    if (bytes) PIC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)));
    return PIC_OK;
}

int pic_malloc(void **ptr, uint64_t bytes) {
    if (!ptr) return PIC_ERR_ARG;
    PIC_HIP(hipMalloc(ptr, bytes ? bytes : 4));
    return PIC_OK;
}
int pic_free(void *ptr) {
    if (ptr) PIC_HIP(hipFree(ptr));
    return PIC_OK;
}
int pic_memcpy_htod(void *dst, const void *src, uint64_t bytes) {
    if (bytes) PIC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return PIC_OK;
}
int pic_memcpy_dtoh(void *dst, const void *src, uint64_t bytes) {
    if (bytes) PIC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PIC_OK;
}
int pic_memset(void *dst, int value, uint64_t bytes) {
    if (bytes) PIC_HIP(hipMemset(dst, value, bytes));
    return PIC_OK;
}
int pic_device_synchronize(void) {
    PIC_HIP(hipDeviceSynchronize());
    return PIC_OK;
}
int pic_stream_create(pic_stream_t *out) {
    if (!out) return PIC_ERR_ARG;
    hipStream_t st = nullptr;
    PIC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *out = st;
    return PIC_OK;
}
int pic_stream_destroy(pic_stream_t stream) {
    if (stream) PIC_HIP(hipStreamDestroy((hipStream_t)stream));
    return PIC_OK;
}
int pic_stream_synchronize(pic_stream_t stream) {
    PIC_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PIC_OK;
}
int pic_device_info(int *compute_units, int *wave_size, char *arch, int arch_len) {
    int dev = 0;
    PIC_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    PIC_HIP(hipGetDeviceProperties(&prop, dev));
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (wave_size) *wave_size = prop.warpSize;
    if (arch && arch_len > 0) {
        strncpy(arch, prop.gcnArchName, arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return PIC_OK;
}

}  // extern "C"
