// pic_fused.hip -- the fast path: one persistent kernel per species per RF
// step that does everything the reference does in move_and_compact,
// check_collisions_e/_i, get_density (of the next step) and accumulate_eepf
// (Rust-CUDA/src/main.rs:566-678, 971-1292, 356-406, 1301-1339) in a single
// pass over the particles.
//
// A species is stored in segments (each dense from its start, with its own
// live count).  Most steps update the particles in place (k_inplace): a
// particle stays in its slot and collides with that slot's RNG stream, a
// particle that leaves the gap becomes a hole (x = -1).  Every compact_every steps
// the compacting kernel (k_fused) squeezes the holes out into the other
// array of the pair; there, as in the reference (main.rs:1785-1791), a
// survivor collides with the RNG stream of the slot it lands in.  Electrons
// born in a step are appended after their segment's watermark and are not
// collision-tested in that step (the loop bound is the watermark read at
// launch, main.rs:992); the next step's density includes every survivor and
// every newborn particle.
//
// HBM traffic per in-place particle-step: read x, v and the four RNG words
// (32 B), write x, vx and the RNG words (24 B), vy and vz only after a
// collision.  Electrons defer the RNG write-back: within a run of in-place
// steps a live slot that does not collide keeps its stored words, which
// fall one draw behind its stream per step, and catches up on its next
// load (the deferred draws cost a few ALU operations, the spared store
// 16 B); the run's last step writes every slot's exact words again, so
// outside a run the words are the reference's.  The field, the histograms
// and the cross-section gathers stay in LDS / L2.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pic_internal.h"

namespace pic {

struct FusedArgs {
    FusedJob j;
    KPhys k;
    int lds_reps;
    int reps, rep_stride;
};

// Workgroup density histograms of the single-discharge kernels are fixed
// point: LDS integer atomics run at full rate, float ones at a small
// fraction of it.  A particle adds exactly `scale` units split by its CIC
// weights (weight wf = scale units), so the sum is exact; each weight is
// rounded to 1/scale.  The scale is the largest power of two (at most 2^20)
// for which no replica can overflow: a bound on the deposits of one
// replica (two per particle -- the particle and an ionisation product --
// for each tile of its waves, plus the collision list) times the scale
// stays below 2^32.
__device__ __forceinline__ void deposit_fix(uint32_t *h, float x, const KPhys &k, float scale) {
    const float pos = x * k.inv_dx;
    const int q = cell_of(pos, k.n_g);
    const float rem = pos - (float)q;
    const uint32_t u2 = (uint32_t)(rem * scale + 0.5f);
    atomicAdd(&h[q], (uint32_t)scale - u2);
    atomicAdd(&h[q + 1], u2);
}

__device__ __forceinline__ float fix_scale(uint32_t n_tiles, int ppt, int lds_reps) {
    const uint32_t my_tiles = n_tiles > blockIdx.x ? (n_tiles - blockIdx.x + gridDim.x - 1) / gridDim.x : 0u;
    const uint32_t waves_per_rep = (blockDim.x / kWave + lds_reps - 1) / lds_reps;
    const uint64_t bound = (uint64_t)my_tiles * waves_per_rep * 2u * ppt * kWave + 256u + 1u;
    int s = 20;
    while (s > 0 && (bound << s) >= (1ull << 32)) --s;
    return (float)(1u << s);
}

// Flush of fixed-point workgroup histograms into a global density replica.
template <typename Acc>
__device__ __forceinline__ void flush_fix(const uint32_t *hist, int lds_reps, int n_g, float scale, float wf,
                                          Acc *out) {
    const Acc unit = (Acc)wf / (Acc)scale;
    for (int q = threadIdx.x; q < n_g; q += blockDim.x) {
        uint64_t sum = 0;
        for (int r = 0; r < lds_reps; ++r) sum += hist[r * n_g + q];
        if (sum) {
            Acc v = (Acc)sum * unit;
            if (q == 0 || q == n_g - 1) v *= (Acc)2;
            unsafeAtomicAdd(&out[q], v);
        }
    }
}

// flush_fix by the threads t0, t0 + stride, ... (t = threadIdx.x - first):
// part of the workgroup flushes while the rest does other work.
template <typename Acc>
__device__ __forceinline__ void flush_fix_part(const uint32_t *hist, int lds_reps, int n_g, float scale, float wf,
                                               Acc *out, int t, int stride) {
    const Acc unit = (Acc)wf / (Acc)scale;
    for (int q = t; q < n_g; q += stride) {
        uint64_t sum = 0;
        for (int r = 0; r < lds_reps; ++r) sum += hist[r * n_g + q];
        if (sum) {
            Acc v = (Acc)sum * unit;
            if (q == 0 || q == n_g - 1) v *= (Acc)2;
            unsafeAtomicAdd(&out[q], v);
        }
    }
}

// This workgroup's share of a buffer the particle kernels clear while they
// run.
__device__ __forceinline__ void clear_slice(uint32_t *buf, uint32_t words) {
    if (!buf) return;
    const uint32_t per = (words + gridDim.x - 1) / gridDim.x;
    const uint32_t w0 = blockIdx.x * per, w1 = min(words, w0 + per);
    for (uint32_t w = w0 + threadIdx.x; w < w1; w += blockDim.x) buf[w] = 0u;
}

// The field into LDS in two halves: field_issue loads this lane's first
// kFieldPre nodes into registers (issued with the other set-up loads, not
// waited for), field_to_lds stores them and copies any nodes beyond.
constexpr int kFieldPre = 2;
__device__ __forceinline__ void field_issue(const float *efield, int n_g, float (&v)[kFieldPre]) {
#pragma unroll
    for (int r = 0; r < kFieldPre; ++r) {
        const int q = (int)threadIdx.x + r * (int)blockDim.x;
        v[r] = efield[min(q, n_g - 1)];  // unconditional: no branch holds a wait
    }
}
__device__ __forceinline__ void field_to_lds(const float *efield, int n_g, const float (&v)[kFieldPre], float *ef) {
#pragma unroll
    for (int r = 0; r < kFieldPre; ++r) {
        const int q = (int)threadIdx.x + r * (int)blockDim.x;
        if (q < n_g) ef[q] = v[r];
    }
    for (int q = (int)threadIdx.x + kFieldPre * (int)blockDim.x; q < n_g; q += blockDim.x) ef[q] = efield[q];
}

// Live count of input segment q = threadIdx.x < K (K <= 64: lanes of wave
// 0), 0 for the other lanes; the caller issues this load with its other
// set-up loads, before any of them is waited for.
__device__ __forceinline__ uint32_t segment_count(const uint32_t *cnt_in, int K) {
    // unconditional (lanes past K reread the last count), so that no branch
    // holds the wait for it
    return cnt_in[min(threadIdx.x, (unsigned)K - 1u)];
}

// The segments' live counts (segment_count; clamped: a counter that
// overflowed its segment counts dropped particles) into s_cnt[K] and the
// inclusive prefix of their tile counts into s_tiles[0..K] (s_tiles[0] = 0),
// by wave 0 in registers: a few lane shuffles instead of a serial loop over
// LDS.  The caller's barrier publishes them.
__device__ __forceinline__ void segment_tiles(uint32_t cnt, int K, uint32_t seg_size, uint32_t tile_sz,
                                              uint32_t *s_cnt, uint32_t *s_tiles) {
    if (threadIdx.x >= (unsigned)kWave) return;
    const int q = (int)threadIdx.x;
    const uint32_t c = q < K ? min(cnt, seg_size) : 0u;
    uint32_t incl = (c + tile_sz - 1) / tile_sz;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, kWave);
        if (q >= off) incl += v;
    }
    if (q < K) {
        s_cnt[q] = c;
        s_tiles[q + 1] = incl;
    }
    if (q == 0) s_tiles[0] = 0u;
}

// Wall losses and collisions of the workgroup: summed in LDS, then one
// atomic per counter into the workgroup's copy of the counters.  Contains
// barriers: every thread of the workgroup calls it, and a second call needs
// a barrier between (thread 0 reads the LDS sums after the last one).
// `c`: the counter words to add into (null: none), `overflow_flag` set when
// any thread saw an overflow.
template <bool ELECTRON>
__device__ __forceinline__ void add_diagnostics(unsigned long long *c, uint32_t *overflow_flag, uint32_t lost_pow,
                                                uint32_t lost_gnd, uint32_t ncoll, bool overflow) {
    __shared__ uint32_t s_diag[3];
    if (threadIdx.x < 3) s_diag[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t wp = wave_sum(lost_pow), wg = wave_sum(lost_gnd), wc = wave_sum(ncoll);
    if (lane_id() == 0) {
        if (wp) atomicAdd(&s_diag[0], wp);
        if (wg) atomicAdd(&s_diag[1], wg);
        if (wc) atomicAdd(&s_diag[2], wc);
    }
    if (overflow) *overflow_flag = 1u;
    __syncthreads();
    if (threadIdx.x == 0 && c) {
        const int b = ELECTRON ? 0 : 2;
        if (s_diag[0]) atomicAdd(&c[b], (unsigned long long)s_diag[0]);
        if (s_diag[1]) atomicAdd(&c[b + 1], (unsigned long long)s_diag[1]);
        if (s_diag[2]) atomicAdd(&c[ELECTRON ? 4 : 5], (unsigned long long)s_diag[2]);
    }
}

template <typename Acc>
__device__ __forceinline__ void deposit_global(Acc *d, float x, const KPhys &k) {
    const float pos = x * k.inv_dx;
    const int q = cell_of(pos, k.n_g);
    const float rem = pos - (float)q;
    Acc c1 = (Acc)((1.0f - rem) * k.wf), c2 = (Acc)(rem * k.wf);
    if (q == 0) c1 *= (Acc)2;
    if (q + 1 == k.n_g - 1) c2 *= (Acc)2;
    unsafeAtomicAdd(&d[q], c1);
    unsafeAtomicAdd(&d[q + 1], c2);
}

__device__ __forceinline__ void eepf_lds(uint32_t *h, const float *ef, float x, float vx, float vy,
                                         float vz, const KPhys &k) {
    if (x <= k.min_x || x >= k.max_x) return;
    const float e = field_at(ef, x, k.inv_dx, k.n_g);
    const float mvx = vx - 0.5f * e * k.eepf_kick;
    const float en = k.half_me_over_e * (mvx * mvx + vy * vy + vz * vz);
    const uint32_t b = (uint32_t)(en / k.de_eepf);
    if (b < (uint32_t)kEepfBins) atomicAdd(&h[b], 1u);
}

__device__ uint32_t tile_reserve_f(uint32_t wave_count, uint32_t *s_wave, uint32_t *s_base, int buf,
                                   uint32_t *counter) {
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

// Where an electron step puts its ionisation products: the new electron
// after the watermark *e_cnt of the electron region that starts at slot
// e_base (e_cap slots), the new ion after *i_cnt of the ion region at
// i_base (i_cap slots).  e_back > 0 (deferred RNG write-back): the new
// electron's slot becomes live with its RNG words e_back draws behind its
// stream, so they are rewound (er0..er3: the electron RNG words).
struct Appends {
    float *ex, *evx, *evy, *evz;
    uint32_t *e_cnt;
    uint32_t e_base, e_cap;
    float *ix, *ivx, *ivy, *ivz;
    uint32_t *i_cnt;
    uint32_t i_base, i_cap;
    uint32_t *er0, *er1, *er2, *er3;
    uint32_t e_back;
    uint8_t *e_mask;  // collision-candidate masks (null: none)
};

// Apply an electron collision's products: the new electron (into the LDS
// density, or the global one when e_dens_g is given, and the EEPF) and the
// new ion (into the global ion density).  Wave-aggregated slot reservation;
// `kind` per lane.
template <bool MEASURE, typename Acc>
__device__ __forceinline__ void ionisation_products(const Appends &a, const KPhys &k, int kind, float px,
                                                    const float *se, const float *si, uint32_t *h, float scale,
                                                    Acc *i_dens_g, uint32_t *eepf_h, const float *ef,
                                                    bool &overflow, Acc *e_dens_g = nullptr) {
    const uint64_t sp = ballot64(kind == kIonisation);
    if (!sp) return;
    const int leader = __builtin_ctzll(sp);
    uint32_t be = 0, bi = 0;
    if ((int)lane_id() == leader) {
        const uint32_t cnt = (uint32_t)__popcll(sp);
        be = atomicAdd(a.e_cnt, cnt);
        bi = atomicAdd(a.i_cnt, cnt);
    }
    be = __shfl(be, leader, kWave);
    bi = __shfl(bi, leader, kWave);
    if (kind != kIonisation) return;
    const uint32_t o = mbcnt64(sp);
    const uint32_t eo = be + o, il = bi + o;
    if (eo < a.e_cap) {
        const uint32_t es = a.e_base + eo;
        el(a.ex, es) = px;
        el(a.evx, es) = se[0];
        el(a.evy, es) = se[1];
        el(a.evz, es) = se[2];
        if (a.e_back) {
            Xo r = rng_load32(a.er0, a.er1, a.er2, a.er3, es);
            xo_back(r, a.e_back);
            rng_store32(a.er0, a.er1, a.er2, a.er3, es, r);
        }
        if (a.e_mask) el(a.e_mask, es) = 0xFFu;  // loads its words on every step of the run
        if (e_dens_g) deposit_global<Acc>(e_dens_g, px, k);
        else deposit_fix(h, px, k, scale);
        if (MEASURE) eepf_lds(eepf_h, ef, px, se[0], se[1], se[2], k);
    } else {
        overflow = true;
    }
    if (il < a.i_cap) {
        const uint32_t is = a.i_base + il;
        el(a.ix, is) = px;
        el(a.ivx, is) = si[0];
        el(a.ivy, is) = si[1];
        el(a.ivz, is) = si[2];
        deposit_global<Acc>(i_dens_g, px, k);
    } else {
        overflow = true;
    }
}

// Wave priority from the tiles a workgroup has left, its current one
// included: 3 down to 0 over its last four.  The workgroups that share a
// compute unit otherwise progress oldest first, so the unit's youngest one
// would run its last tiles alone, at a quarter of the unit's occupancy;
// lowering the priority of the workgroups ahead lets the others catch up
// and the unit's workgroups end together.
__device__ __forceinline__ void tile_priority(uint32_t left) {
    left = __builtin_amdgcn_readfirstlane(left);
    if (left >= 4u) __builtin_amdgcn_s_setprio(3);
    else if (left == 3u) __builtin_amdgcn_s_setprio(2);
    else if (left == 2u) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// Segmented particle arrays: a species lives in K segments of seg_size
// slots, each dense from its start with its own live counter.  The kernel
// cuts every input segment into tiles (an LDS prefix of per-segment tile
// counts maps a global tile to its segment), and workgroup w writes its
// survivors -- and, for electrons, the electrons it creates -- into output
// segment w % K.  Each counter therefore sees 1/K of the slot reservations:
// same-address atomics serialise in the memory-side atomic units, and one
// counter for 4 M particles was the bottleneck of the whole step.
template <bool ELECTRON, bool MEASURE, int PPT, typename Acc>
__global__ __launch_bounds__(kMaxParticleBlock) void k_fused(FusedArgs A) {
    const FusedJob &j = A.j;
    const KPhys &k = A.k;
    extern __shared__ __align__(16) float sh[];
    __shared__ uint32_t s_wave[32], s_base[2];
    __shared__ uint32_t s_tiles[kMaxSegments + 1], s_cnt[kMaxSegments];
    const int n_g = k.n_g;
    const int K = j.n_seg;
    float *ef = sh;
    uint32_t *hist = (uint32_t *)(sh + n_g);
    uint32_t *aux = hist + A.lds_reps * n_g;  // EEPF (electrons) or IFED (ions)
    const int aux_n = ELECTRON ? (MEASURE ? kEepfBins : 0) : 2 * kIfedBins;
    const uint32_t tile_sz = blockDim.x * PPT;
    const uint32_t seg_cnt = segment_count(j.cnt_in, K);
    float efv[kFieldPre];
    field_issue(j.efield, n_g, efv);
    segment_tiles(seg_cnt, K, j.seg_size, tile_sz, s_cnt, s_tiles);
    field_to_lds(j.efield, n_g, efv, ef);
    for (int q = threadIdx.x; q < A.lds_reps * n_g; q += blockDim.x) hist[q] = 0u;
    for (int q = threadIdx.x; q < aux_n; q += blockDim.x) aux[q] = 0u;
    clear_slice(j.zero, j.zero_words);
    __syncthreads();
    const int wave = threadIdx.x / kWave;
    uint32_t *h = hist + (wave % A.lds_reps) * n_g;
    const uint32_t n_tiles = s_tiles[K];
    const float scale = fix_scale(n_tiles, PPT, A.lds_reps);
    const uint32_t ko = blockIdx.x % (uint32_t)K;               // output segment
    uint32_t *const cnt_out = j.cnt_out + ko;
    const uint32_t out_base = ko * j.seg_size;
    const float factor = ELECTRON ? k.factor_e : k.factor_i;
    const float dt = ELECTRON ? k.dt_e : k.dt_i;
    Acc *const i_dens_g = (Acc *)j.ion_dens + (blockIdx.x % A.reps) * A.rep_stride;
    const Appends app{j.dx,  j.dvx, j.dvy, j.dvz, cnt_out,        out_base,            j.seg_size,
                      j.ix,  j.ivx, j.ivy, j.ivz, j.ion_cnt + ko, ko * j.ion_seg_size, j.ion_seg_size};
    uint32_t lost_pow = 0, lost_gnd = 0, ncoll = 0;
    bool overflow = false;
    int buf = 0;
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x, buf ^= 1) {
        tile_priority((n_tiles - tile + gridDim.x - 1) / gridDim.x);
        // segment of this tile: largest q with s_tiles[q] <= tile
        int lo = 0, hi = K - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_tiles[mid] <= tile) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t seg_n = s_cnt[lo];
        const uint32_t in_base = (uint32_t)lo * j.seg_size;
        const uint32_t base = (tile - s_tiles[lo]) * tile_sz + threadIdx.x;
        float nx[PPT], nvx[PPT], vy[PPT], vz[PPT];
        bool keep[PPT];
        uint32_t off[PPT];
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const uint32_t v = base + u * blockDim.x;
            keep[u] = false;
            const uint32_t i = in_base + v;
            const float x = v < seg_n ? el(j.sx, i) : -1.0f;
            if (x >= 0.0f) {  // x < 0 marks a hole left by an in-place step
                const float e = field_at(ef, x, k.inv_dx, n_g);
                nvx[u] = __builtin_fmaf(factor, e, el(j.svx, i));
                nx[u] = __builtin_fmaf(nvx[u], dt, x);
                keep[u] = nx[u] >= 0.0f && nx[u] <= k.gap;
                if (keep[u]) {
                    vy[u] = el(j.svy, i);
                    vz[u] = el(j.svz, i);
                } else {
                    const bool pw = nx[u] < 0.0f;
                    lost_pow += pw;
                    lost_gnd += !pw;
                    if (!ELECTRON) {
                        const float y = el(j.svy, i), z = el(j.svz, i);
                        const float en = k.half_mi_over_e * (nvx[u] * nvx[u] + y * y + z * z);
                        const int b = (int)en;
                        if (en >= 0.0f && b < kIfedBins) atomicAdd(&aux[(pw ? 0 : 1) * kIfedBins + b], 1u);
                    }
                }
            }
        }
        uint32_t wcount = 0;
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const uint64_t m = ballot64(keep[u]);
            off[u] = wcount + mbcnt64(m);
            wcount += (uint32_t)__popcll(m);
        }
        const uint32_t wbase = tile_reserve_f(wcount, s_wave, s_base, buf, cnt_out);
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            int kind = kNone;
            float se[3], si[3];
            if (keep[u]) {
                const uint32_t local = wbase + off[u];
                if (local < j.seg_size) {
                    const uint32_t slot = out_base + local;
                    Xo s = rng_load32(j.rng0, j.rng1, j.rng2, j.rng3, slot);
                    float vx = nvx[u], y = vy[u], z = vz[u];
                    if (ELECTRON) kind = collide_electron(j.stot, j.cs, k, vx, y, z, s, se, si);
                    else kind = collide_ion(j.stot, j.cs, k, vx, y, z, s);
                    rng_store32(j.rng0, j.rng1, j.rng2, j.rng3, slot, s);
                    el(j.dx, slot) = nx[u];
                    el(j.dvx, slot) = vx;
                    el(j.dvy, slot) = y;
                    el(j.dvz, slot) = z;
                    deposit_fix(h, nx[u], k, scale);
                    if (MEASURE && ELECTRON) eepf_lds(aux, ef, nx[u], vx, y, z, k);
                    ncoll += kind != kNone;
                } else {
                    overflow = true;
                }
            }
            // ionisation products: the new electron joins this output
            // segment (it is not collision-tested until the next step,
            // main.rs:992), the new ion the same-numbered ion segment
            if (ELECTRON)
                ionisation_products<MEASURE, Acc>(app, k, kind, nx[u], se, si, h, scale, i_dens_g, aux, ef,
                                                  overflow);
        }
    }
    add_diagnostics<ELECTRON>(
        j.counters ? j.counters + (size_t)(blockIdx.x % kCounterLanes) * kCounterWords : nullptr, j.overflow,
        lost_pow, lost_gnd, ncoll, overflow);
    if (!__syncthreads_or(n_tiles > blockIdx.x)) return;
    // one flush of the workgroup histogram into the next step's density
    flush_fix<Acc>(hist, A.lds_reps, n_g, scale, k.wf, (Acc *)j.dens + (blockIdx.x % A.reps) * A.rep_stride);
    uint32_t *aux_out = ELECTRON ? j.eepf : j.ifed;
    if (aux_out)
        for (int q = threadIdx.x; q < aux_n; q += blockDim.x)
            if (aux[q]) atomicAdd(&aux_out[q], aux[q]);
}

// The in-place step (every step that does not compact): each particle is
// updated in the slot it occupies, so all of its loads -- x, v and the four
// RNG words of the slot -- are independent and issued together, and there
// is no slot reservation.  A particle that leaves the gap becomes a hole
// (x = -1); electrons born in the step are appended after the segment's
// watermark (input count cnt_in, copied into cnt_out before the launch),
// so they are neither pushed nor collision-tested until the next step
// (main.rs:992).  Per particle the traffic is read x, v, RNG (32 B), write
// x, vx (8 B), the RNG words (16 B) only on a step that leaves them exact
// (FusedJob::rng_next_lag == 0) and vy, vz only after a collision.
//
// Collisions are rare (well under 1% of particle-steps) and their
// scattering code is long: a wave that ran it for its one or two colliding
// lanes in every tile would spend most of the step there, and four waves per
// SIMD doing that at once saturate the vector ALUs.  The collision test
// therefore only lists the colliding particle (slot, position, velocity,
// RNG state after the test draw, energy bin; ions also the partner velocity
// and relative speed) in a workgroup list in LDS, and after the last tile
// the workgroup's waves scatter the list 64 entries at a time, one entry per
// lane.  The RNG stream of a particle continues from the listed state, so
// its draws are those of the sequential order.  A list that fills up (far
// beyond the expected tens of entries per workgroup) scatters the excess in
// place.
constexpr uint32_t kCollPerWg = 256;  // list entries per workgroup
enum CollField {
    CF_SLOT = 0, CF_X, CF_VX, CF_VY, CF_VZ, CF_R0, CF_R1, CF_R2, CF_R3, CF_BIN, CF_AX, CF_AY, CF_AZ,
    CF_G = CF_X  // ions: relative speed in the (unused) position field; CF_AX.. unused by electrons
};
constexpr int kCollFieldsE = CF_BIN + 1;
constexpr int kCollFieldsI = CF_AZ + 1;

// RM: collision-candidate masks of deferred electron runs.  0: none (every
// live slot loads its RNG words); 1: masks read -- a live slot loads its
// words at the start of its tile only when its bit for this step's lag is
// set, and late (after the push) when its collision probability exceeds
// the masks' threshold or when it leaves the gap; 2: a write-back step
// that writes the next run's masks.
template <bool ELECTRON, bool MEASURE, int PPT, typename Acc, int RM>
__global__ __launch_bounds__(kMaxParticleBlock) void k_inplace(FusedArgs A) {
    static_assert(RM == 0 || (ELECTRON && !MEASURE), "candidate masks: plain electron steps only");
    const FusedJob &j = A.j;
    const KPhys &k = A.k;
    extern __shared__ __align__(16) float sh[];
    __shared__ uint32_t s_tiles[kMaxSegments + 1], s_cnt[kMaxSegments];
    __shared__ uint32_t s_ncoll;
    const int n_g = k.n_g;
    const int K = j.n_seg;
    float *ef = sh;
    uint32_t *hist = (uint32_t *)(sh + n_g);
    uint32_t *aux = hist + A.lds_reps * n_g;
    const int aux_n = ELECTRON ? (MEASURE ? kEepfBins : 0) : 2 * kIfedBins;
    uint32_t *list = aux + aux_n;  // [NF][kCollPerWg]
    const uint32_t tile_sz = blockDim.x * PPT;
    const uint32_t seg_cnt = segment_count(j.cnt_in, K);
    float efv[kFieldPre];
    field_issue(j.efield, n_g, efv);
    segment_tiles(seg_cnt, K, j.seg_size, tile_sz, s_cnt, s_tiles);
    field_to_lds(j.efield, n_g, efv, ef);
    for (int q = threadIdx.x; q < A.lds_reps * n_g; q += blockDim.x) hist[q] = 0u;
    for (int q = threadIdx.x; q < aux_n; q += blockDim.x) aux[q] = 0u;
    clear_slice(j.zero, j.zero_words);
    if (threadIdx.x == 0) s_ncoll = 0;
    __syncthreads();
    const int wave = threadIdx.x / kWave;
    uint32_t *h = hist + (wave % A.lds_reps) * n_g;
    const uint32_t n_tiles = s_tiles[K];
    // fixed-point scale 2^20; a replica takes at most cap tiles of deposits
    // (two per particle) plus the collision list's before it could overflow,
    // so the histograms are flushed and cleared every cap tiles
    const float scale = 1048576.0f;
    const uint32_t waves_per_rep = (blockDim.x / kWave + A.lds_reps - 1) / A.lds_reps;
    const uint32_t cap = (4096u - kCollPerWg - 1u) / (waves_per_rep * 2u * PPT * kWave);
    uint32_t since_flush = 0;
    const uint32_t ko = blockIdx.x % (uint32_t)K;  // segment that receives this workgroup's newborns
    const float factor = ELECTRON ? k.factor_e : k.factor_i;
    const float dt = ELECTRON ? k.dt_e : k.dt_i;
    float *const X = (float *)j.sx, *const VX = (float *)j.svx, *const VY = (float *)j.svy,
                 *const VZ = (float *)j.svz;
    Acc *const i_dens_g = (Acc *)j.ion_dens + (blockIdx.x % A.reps) * A.rep_stride;
    // deferred RNG write-back (electrons): the words of live slots are `lag`
    // draws behind their streams on entry and are left `lag_out` behind
    const uint32_t lag = ELECTRON ? j.rng_lag : 0u;
    const uint32_t lag_out = ELECTRON ? j.rng_next_lag : 0u;
    const Appends app{X,        VX,       VY,       VZ,       j.cnt_out + ko, ko * j.seg_size,     j.seg_size,
                      j.ix,     j.ivx,    j.ivy,    j.ivz,    j.ion_cnt + ko, ko * j.ion_seg_size, j.ion_seg_size,
                      j.rng0,   j.rng1,   j.rng2,   j.rng3,   lag_out,        RM ? j.mask : nullptr};
    uint8_t *const MK = j.mask;
    const float thr = j.coll_thr;
    uint32_t lost_pow = 0, lost_gnd = 0, ncoll = 0;
    bool overflow = false;

    // (1) particle and RNG loads of all PPT particles of a tile; tile t + G
    // (G = grid size) is loaded while tile t is processed, so each wave keeps
    // a tile of loads in flight behind its compute and gathers.  The loads
    // are unconditional (lanes past the segment's count read the segment's
    // first slot) and the count test is applied where the values are used,
    // so that no wait for them is needed at the point of issue.  With the
    // masks read (RM 1) the tile brings the mask bytes instead of the RNG
    // words, which only candidates load.
    auto load_tile = [&](uint32_t tile, uint32_t &s0, uint32_t &avail, float (&lx)[PPT], float (&lvx)[PPT],
                         float (&lvy)[PPT], float (&lvz)[PPT], Xo (&lst)[PPT], uint32_t (&lmk)[PPT]) {
        uint32_t safe = 0;
        avail = 0;
        s0 = 0;
        if (tile < n_tiles) {
            int lo = 0, hi = K - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_tiles[mid] <= tile) lo = mid;
                else hi = mid - 1;
            }
            const uint32_t local0 = (tile - s_tiles[lo]) * tile_sz + threadIdx.x;
            safe = (uint32_t)lo * j.seg_size;
            s0 = safe + local0;
            avail = s_cnt[lo] > local0 ? s_cnt[lo] - local0 : 0u;
        }
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const uint32_t i = u * blockDim.x < avail ? s0 + u * blockDim.x : safe;
            lx[u] = el(X, i);
            lvx[u] = el(VX, i);
            lvy[u] = el(VY, i);
            lvz[u] = el(VZ, i);
            if (RM == 1) lmk[u] = el(MK, i);
            else lst[u] = rng_load32(j.rng0, j.rng1, j.rng2, j.rng3, i);
        }
    };
    float x[PPT], vx[PPT], vy[PPT], vz[PPT];
    Xo st[PPT];
    uint32_t mk[PPT];
    uint32_t slot0 = 0, avail = 0;
    // workgroup b takes tiles b, b + G, b + 2G, ... (G = grid size)
    uint32_t tile = blockIdx.x;
    load_tile(tile, slot0, avail, x, vx, vy, vz, st, mk);
    while (tile < n_tiles) {
        const uint32_t tile_next = tile + gridDim.x;
        tile_priority((n_tiles - tile + gridDim.x - 1) / gridDim.x);
        bool valid[PPT];
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            valid[u] = u * blockDim.x < avail;
            if (!valid[u]) x[u] = -1.0f;  // past the segment's count
        }
        // a live slot's stream catches up with the draws it was spared; with
        // the masks read only candidates load their words here (the catch-up
        // follows the late loads after the push)
        bool has[PPT];
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            has[u] = x[u] >= 0.0f;
            if (RM == 1) {
                has[u] = has[u] && (lag >= 8u || ((mk[u] >> lag) & 1u));
                if (has[u]) st[u] = rng_load32(j.rng0, j.rng1, j.rng2, j.rng3, slot0 + u * blockDim.x);
            }
        }
        if (RM != 1 && lag) {
#pragma unroll
            for (int u = 0; u < PPT; ++u)
                if (x[u] >= 0.0f) xo_skip(st[u], lag);
        }
        // (2) push; in-gap particles bin their energy and gather stot
        float nx[PPT], spd[PPT], sig[PPT];
        float ax[PPT], ay[PPT], az[PPT];  // ions: velocity of the collision partner
        int idx[PPT];
        bool in[PPT];
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            in[u] = false;
            nx[u] = 0.0f;
            idx[u] = 0;
            spd[u] = 0.0f;
            ax[u] = ay[u] = az[u] = 0.0f;
            if (x[u] >= 0.0f) {
                const float e = field_at(ef, x[u], k.inv_dx, n_g);
                vx[u] = __builtin_fmaf(factor, e, vx[u]);
                nx[u] = __builtin_fmaf(vx[u], dt, x[u]);
                in[u] = nx[u] >= 0.0f && nx[u] <= k.gap;
            }
            if (in[u]) {
                if (ELECTRON) idx[u] = e_index(k, vx[u], vy[u], vz[u], spd[u]);
                else idx[u] = i_index(k, vx[u], vy[u], vz[u], st[u], ax[u], ay[u], az[u], spd[u]);
            }
        }
        // the gather is issued for every lane (the others read bin 0, one
        // shared line) so that it stays in flight while the next tile's
        // loads are issued
#pragma unroll
        for (int u = 0; u < PPT; ++u) sig[u] = el(j.stot, in[u] ? (uint32_t)idx[u] : 0u);
        float px[PPT], pvx[PPT], pvy[PPT], pvz[PPT];
        Xo pst[PPT];
        uint32_t pmk[PPT];
        uint32_t pslot0, pavail;
        load_tile(tile_next, pslot0, pavail, px, pvx, pvy, pvz, pst, pmk);
        // the deposits and the x / vx write-back do not depend on the
        // collision test: they go first, while the gather is in flight
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            if (in[u]) deposit_fix(h, nx[u], k, scale);
            if (valid[u]) {
                el(X, slot0 + u * blockDim.x) = in[u] ? nx[u] : -1.0f;
                el(VX, slot0 + u * blockDim.x) = vx[u];
            }
        }
        // masks read: a live slot that is no candidate loads its words now
        // if it may collide above the threshold, or if it leaves the gap
        // behind on its words (they are stored exact); then the loaded
        // streams catch up
        float pc[PPT];
        if (RM == 1) {
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                pc[u] = in[u] ? e_prob(k, sig[u], spd[u]) : 0.0f;
                if (!has[u] && x[u] >= 0.0f && (in[u] ? pc[u] > thr : lag != 0u)) {
                    st[u] = rng_load32(j.rng0, j.rng1, j.rng2, j.rng3, slot0 + u * blockDim.x);
                    has[u] = true;
                }
            }
            if (lag) {
#pragma unroll
                for (int u = 0; u < PPT; ++u)
                    if (has[u]) xo_skip(st[u], lag);
            }
        }
        // (3) collision tests and the RNG write-back; colliding particles
        // are listed; lost ones leave a hole.  Every slot of the tile has
        // written x and vx back (above), and on a step that leaves the RNG
        // words exact (lag_out == 0) also writes its RNG words (a colliding
        // particle's values
        // are overwritten by its scattering after the last tile, holes and
        // lost particles write their unchanged RNG words): stores that skip
        // lanes leave partly written lines, which cost a read-modify-write
        // in memory.  Otherwise only a particle lost in this step writes its
        // RNG words (when they were behind): a hole's words are exact.
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const uint32_t i = slot0 + u * blockDim.x;
            bool hit = false;
            if (in[u]) {
                if (RM == 1) hit = has[u] && draw(st[u]) < pc[u];  // no candidate: no collision
                else hit = ELECTRON ? e_test(k, sig[u], spd[u], st[u]) : i_test(k, sig[u], spd[u], st[u]);
                if (MEASURE && ELECTRON && !hit) eepf_lds(aux, ef, nx[u], vx[u], vy[u], vz[u], k);
            }
            if (valid[u]) {
                if (lag_out == 0u || (lag && !in[u] && x[u] >= 0.0f))
                    rng_store32(j.rng0, j.rng1, j.rng2, j.rng3, i, st[u]);
                // the next run's masks from the exact words (a colliding
                // particle's are rewritten after its scattering)
                if (RM == 2) el(MK, i) = (uint8_t)(in[u] && !hit ? cand_mask(st[u], 0u, thr) : 0xFFu);
            }
            if (!in[u] && x[u] >= 0.0f) {  // lost in this step
                const bool pw = nx[u] < 0.0f;
                lost_pow += pw;
                lost_gnd += !pw;
                if (!ELECTRON) {
                    const float en = k.half_mi_over_e * (vx[u] * vx[u] + vy[u] * vy[u] + vz[u] * vz[u]);
                    const int b = (int)en;
                    if (en >= 0.0f && b < kIfedBins) atomicAdd(&aux[(pw ? 0 : 1) * kIfedBins + b], 1u);
                }
            }
            const uint64_t m = ballot64(hit);
            if (m) {
                const int leader = __builtin_ctzll(m);
                uint32_t base = 0;
                if ((int)lane_id() == leader) base = atomicAdd(&s_ncoll, (uint32_t)__popcll(m));
                base = __shfl(base, leader, kWave);
                const uint32_t pos = base + mbcnt64(m);
                const bool listed = hit && pos < kCollPerWg;
                if (listed) {
                    uint32_t *e = list + pos;
                    e[CF_SLOT * kCollPerWg] = i;
                    e[CF_VX * kCollPerWg] = __float_as_uint(vx[u]);
                    e[CF_VY * kCollPerWg] = __float_as_uint(vy[u]);
                    e[CF_VZ * kCollPerWg] = __float_as_uint(vz[u]);
                    e[CF_R0 * kCollPerWg] = st[u].s0;
                    e[CF_R1 * kCollPerWg] = st[u].s1;
                    e[CF_R2 * kCollPerWg] = st[u].s2;
                    e[CF_R3 * kCollPerWg] = st[u].s3;
                    e[CF_BIN * kCollPerWg] = (uint32_t)idx[u];
                    if (ELECTRON) {
                        e[CF_X * kCollPerWg] = __float_as_uint(nx[u]);
                    } else {
                        e[CF_AX * kCollPerWg] = __float_as_uint(ax[u]);
                        e[CF_AY * kCollPerWg] = __float_as_uint(ay[u]);
                        e[CF_AZ * kCollPerWg] = __float_as_uint(az[u]);
                        e[CF_G * kCollPerWg] = __float_as_uint(spd[u]);
                    }
                }
                // a full list: scatter here
                const bool local = hit && !listed;
                if (ballot64(local)) {
                    int kind = kNone;
                    float se[3], si[3];
                    if (local) {
                        float w = vx[u], y = vy[u], z = vz[u];
                        Xo s = st[u];
                        const int b = idx[u];
                        if (ELECTRON)
                            kind = e_scatter(k, j.cs[b], j.cs[kCsRanges + b], j.cs[2 * kCsRanges + b], w, y, z, s,
                                             se, si);
                        else
                            kind = i_scatter(k, j.cs[3 * kCsRanges + b], j.cs[4 * kCsRanges + b], w, y, z, ax[u],
                                             ay[u], az[u], spd[u], s);
                        if (RM) el(MK, i) = (uint8_t)cand_mask(s, lag_out, thr);
                        xo_back(s, lag_out);
                        rng_store32(j.rng0, j.rng1, j.rng2, j.rng3, i, s);
                        el(VX, i) = w;
                        el(VY, i) = y;
                        el(VZ, i) = z;
                        ++ncoll;
                        if (MEASURE && ELECTRON) eepf_lds(aux, ef, nx[u], w, y, z, k);
                    }
                    if (ELECTRON)
                        ionisation_products<MEASURE, Acc>(app, k, kind, nx[u], se, si, h, scale, i_dens_g, aux,
                                                          ef, overflow);
                }
            }
        }
        slot0 = pslot0;
        avail = pavail;
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            x[u] = px[u];
            vx[u] = pvx[u];
            vy[u] = pvy[u];
            vz[u] = pvz[u];
            if (RM == 1) mk[u] = pmk[u];
            else st[u] = pst[u];
        }
        tile = tile_next;
        if (++since_flush == cap && tile < n_tiles) {  // uniform over the workgroup
            since_flush = 0;
            __syncthreads();
            flush_fix<Acc>(hist, A.lds_reps, n_g, scale, k.wf, (Acc *)j.dens + (blockIdx.x % A.reps) * A.rep_stride);
            __syncthreads();
            for (int q = threadIdx.x; q < A.lds_reps * n_g; q += blockDim.x) hist[q] = 0u;
            __syncthreads();
        }
    }
    // (4) scatter the listed collisions, 64 per wave pass.  When they fit
    // the first waves, the other waves flush the density histograms
    // meanwhile and the scattering waves deposit newborn electrons straight
    // into the global density replica.
    __syncthreads();
    const uint32_t n_list = min(s_ncoll, kCollPerWg);
    const uint32_t n_waves = blockDim.x / kWave;
    const uint32_t scat_waves = min(n_waves, (n_list + kWave - 1) / kWave);
    const bool early_flush = scat_waves < n_waves && n_tiles > blockIdx.x;  // uniform
    Acc *const dens_g = (Acc *)j.dens + (blockIdx.x % A.reps) * A.rep_stride;
    if (early_flush && (uint32_t)wave >= scat_waves)
        flush_fix_part<Acc>(hist, A.lds_reps, n_g, scale, k.wf, dens_g, (int)(threadIdx.x - scat_waves * kWave),
                            (int)(blockDim.x - scat_waves * kWave));
    for (uint32_t b0 = (uint32_t)wave * kWave; b0 < n_list; b0 += blockDim.x) {
        const uint32_t e = b0 + lane_id();
        int kind = kNone;
        float se[3], si[3], px = 0.0f;
        if (e < n_list) {
            const uint32_t *q = list + e;
            const uint32_t slot = q[CF_SLOT * kCollPerWg];
            float w = __uint_as_float(q[CF_VX * kCollPerWg]), y = __uint_as_float(q[CF_VY * kCollPerWg]),
                  z = __uint_as_float(q[CF_VZ * kCollPerWg]);
            Xo s{q[CF_R0 * kCollPerWg], q[CF_R1 * kCollPerWg], q[CF_R2 * kCollPerWg], q[CF_R3 * kCollPerWg]};
            const int b = (int)q[CF_BIN * kCollPerWg];
            if (ELECTRON) {
                px = __uint_as_float(q[CF_X * kCollPerWg]);
                kind = e_scatter(k, j.cs[b], j.cs[kCsRanges + b], j.cs[2 * kCsRanges + b], w, y, z, s, se, si);
                if (MEASURE) eepf_lds(aux, ef, px, w, y, z, k);
            } else {
                kind = i_scatter(k, j.cs[3 * kCsRanges + b], j.cs[4 * kCsRanges + b], w, y, z,
                                 __uint_as_float(q[CF_AX * kCollPerWg]), __uint_as_float(q[CF_AY * kCollPerWg]),
                                 __uint_as_float(q[CF_AZ * kCollPerWg]), __uint_as_float(q[CF_G * kCollPerWg]), s);
            }
            // the rest of the run's masks from the state after the scattering
            if (RM) el(MK, slot) = (uint8_t)cand_mask(s, lag_out, thr);
            xo_back(s, lag_out);
            rng_store32(j.rng0, j.rng1, j.rng2, j.rng3, slot, s);
            el(VX, slot) = w;
            el(VY, slot) = y;
            el(VZ, slot) = z;
            ++ncoll;
        }
        if (ELECTRON)
            ionisation_products<MEASURE, Acc>(app, k, kind, px, se, si, h, scale, i_dens_g, aux, ef, overflow,
                                              early_flush ? dens_g : nullptr);
    }
    add_diagnostics<ELECTRON>(
        j.counters ? j.counters + (size_t)(blockIdx.x % kCounterLanes) * kCounterWords : nullptr, j.overflow,
        lost_pow, lost_gnd, ncoll, overflow);
    if (!__syncthreads_or(n_tiles > blockIdx.x)) return;
    if (!early_flush) flush_fix<Acc>(hist, A.lds_reps, n_g, scale, k.wf, dens_g);
    uint32_t *aux_out = ELECTRON ? j.eepf : j.ifed;
    if (aux_out)
        for (int q2 = threadIdx.x; q2 < aux_n; q2 += blockDim.x)
            if (aux[q2]) atomicAdd(&aux_out[q2], aux[q2]);
}

// Live particles (x >= 0) of a segmented species, for the host's counts.
__global__ __launch_bounds__(256) void k_count_live(const float *__restrict__ x, const uint32_t *__restrict__ cnt,
                                                   int n_seg, uint32_t seg_size, uint32_t *__restrict__ out) {
    uint32_t c = 0;
    for (int q = 0; q < n_seg; ++q) {
        const uint32_t n = min(cnt[q], seg_size);
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
            c += x[(size_t)q * seg_size + i] >= 0.0f;
    }
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(out, c);
}

int launch_count_live(const float *x, const uint32_t *cnt, int n_seg, uint32_t seg_size, uint32_t *out,
                      hipStream_t s) {
    k_count_live<<<device_cus() * 4, 256, 0, s>>>(x, cnt, n_seg, seg_size, out);
    PIC_LAUNCHED("k_count_live");
    return PIC_OK;
}

template <bool E, bool M, int PPT, int RM>
static void launch_ip(const FusedArgs &a, bool f64, int grid, int block, size_t lds, hipStream_t s) {
    if (f64) k_inplace<E, M, PPT, double, RM><<<grid, block, lds, s>>>(a);
    else k_inplace<E, M, PPT, float, RM><<<grid, block, lds, s>>>(a);
}

template <bool E, bool M, int PPT>
static void launch_t(const FusedArgs &a, bool f64, bool inplace, int grid, int block, size_t lds, hipStream_t s) {
    if (inplace) {
        const int rm = a.j.mask_read ? 1 : a.j.mask_write ? 2 : 0;
        if constexpr (E && !M) {
            if (rm == 1) launch_ip<E, M, PPT, 1>(a, f64, grid, block, lds, s);
            else if (rm == 2) launch_ip<E, M, PPT, 2>(a, f64, grid, block, lds, s);
            else launch_ip<E, M, PPT, 0>(a, f64, grid, block, lds, s);
        } else {
            launch_ip<E, M, PPT, 0>(a, f64, grid, block, lds, s);
        }
    } else {
        if (f64) k_fused<E, M, PPT, double><<<grid, block, lds, s>>>(a);
        else k_fused<E, M, PPT, float><<<grid, block, lds, s>>>(a);
    }
}

template <bool E, bool M>
static void launch_ppt(const FusedArgs &a, bool f64, bool inplace, int ppt, int grid, int block, size_t lds,
                       hipStream_t s) {
    switch (ppt) {
    case 1: launch_t<E, M, 1>(a, f64, inplace, grid, block, lds, s); break;
    case 2: launch_t<E, M, 2>(a, f64, inplace, grid, block, lds, s); break;
    default: launch_t<E, M, 4>(a, f64, inplace, grid, block, lds, s); break;
    }
}

int launch_fused(const FusedJob &j, const KPhys &k, const Shape &sh0, hipStream_t s) {
    if (j.n_seg < 1 || j.n_seg > kMaxSegments || j.seg_size == 0) {
        set_error("fused: 1 to 64 segments of non-zero size");
        return PIC_ERR_ARG;
    }
    // a live slot that does not collide keeps its words: one draw further
    // behind after the step, or exact if the step writes them all back
    const bool deferred = j.rng_lag || j.rng_next_lag;
    if (deferred && (!j.electrons || !j.inplace || j.rng_lag >= kMaxRngLag ||
                     (j.rng_next_lag != 0 && j.rng_next_lag != j.rng_lag + 1))) {
        set_error("fused: deferred RNG write-back is for in-place electron steps, next lag 0 or lag + 1");
        return PIC_ERR_ARG;
    }
    if ((j.mask_read || j.mask_write) &&
        (!j.mask || !j.electrons || !j.inplace || j.eepf || (j.mask_read && j.mask_write) ||
         (j.mask_read && j.rng_next_lag == 0) || (j.mask_write && j.rng_next_lag != 0) || !(j.coll_thr >= 0.0f))) {
        set_error("fused: candidate masks are read on deferred in-place electron steps and written on their "
                  "write-back steps (no measurement)");
        return PIC_ERR_ARG;
    }
    Shape sh = resolve(sh0);
    if (!particle_block_ok(sh.block)) {
        set_error("fused: workgroup size must be 128, 256 or 512");
        return PIC_ERR_ARG;
    }
    FusedArgs a;
    a.j = j;
    a.k = k;
    const int block = sh.block;
    const bool measure = j.electrons && j.eepf != nullptr;
    const int aux_words = j.electrons ? (measure ? kEepfBins : 0) : 2 * kIfedBins;
    const int grid = std::max(1, sh.grid);
    // LDS per workgroup: what the resident workgroups of a compute unit can
    // share (2 KB kept for the static arrays), at most what one may allocate.
    // It holds the field, the EEPF / IFED bins, the collision list of the
    // in-place kernel and one histogram replica per wave (fewer when space is
    // short).
    const int per_cu = std::max(1, (grid + device_cus() - 1) / device_cus());
    const int budget = std::min(device_lds_per_block(), device_lds_per_cu() / per_cu - 2048);
    const int list_words = j.inplace ? (j.electrons ? kCollFieldsE : kCollFieldsI) * (int)kCollPerWg : 0;
    const int fixed = (k.n_g + aux_words + list_words) * 4;
    a.lds_reps = std::max(1, std::min(block / kWave, (budget - fixed) / (k.n_g * 4)));
    if (sh.lds_reps > 0) a.lds_reps = std::min(a.lds_reps, sh.lds_reps);
    a.reps = sh.replicas;
    a.rep_stride = sh.rep_stride > 0 ? sh.rep_stride : k.n_g;
    const size_t lds = (size_t)fixed + (size_t)a.lds_reps * k.n_g * 4;
    const int ppt = j.inplace ? sh.ppt_inplace : sh.ppt;
    if (j.electrons) {
        if (measure) launch_ppt<true, true>(a, j.dens_f64, j.inplace, ppt, grid, block, lds, s);
        else launch_ppt<true, false>(a, j.dens_f64, j.inplace, ppt, grid, block, lds, s);
    } else {
        launch_ppt<false, false>(a, j.dens_f64, j.inplace, ppt, grid, block, lds, s);
    }
    PIC_LAUNCHED(j.inplace ? "k_inplace" : "k_fused");
    return PIC_OK;
}

}  // namespace pic

// ============================================================ ensembles
// The in-place particle step for an ensemble of independent discharges in
// one persistent launch, and the compaction pass that squeezes out holes
// every compact_every steps.  Every instance's slots are cut into tiles; a
// per-instance tile prefix (in LDS) maps a global tile to its instance.
// Each workgroup takes a contiguous range of global tiles, so it meets few
// instance boundaries.  At a boundary it finishes the instance it leaves --
// scatters its listed collisions, adds its diagnostics, flushes its density
// histogram -- and stages the next instance's field and physics.  The tile
// body is that of the single-discharge in-place kernel: prefetch of the
// next tile, collision test, full-line write-back, colliding particles
// listed in LDS and scattered when the workgroup leaves the instance.
namespace pic {

constexpr int kMaxEnsemble = 256;

struct EnsArgs {
    EnsStepJob j;
    int lds_reps, reps, rep_stride;
    // tile order: 0 = each workgroup a contiguous range of tiles; c > 0 =
    // chunks of c tiles dealt round-robin to the workgroups
    int chunk;
};

// RM: collision-candidate masks of the deferred electron runs, as in the
// single-discharge kernel (0 none, 1 read, 2 written by a write-back step);
// the threshold is the instance's own (its gas density scales the
// probabilities).
template <bool ELECTRON, int PPT, typename Acc, int RM>
__global__ __launch_bounds__(kMaxParticleBlock) void k_inplace_ens(EnsArgs A) {
    static_assert(RM == 0 || ELECTRON, "candidate masks: electron steps only");
    const EnsStepJob &j = A.j;
    extern __shared__ __align__(16) float sh[];
    __shared__ uint32_t s_tiles[kMaxEnsemble + 1], s_cnt[kMaxEnsemble], s_off[kMaxEnsemble];
    __shared__ uint32_t s_ncoll;
    const int n_g = j.k.n_g;
    const int E = j.n_inst;
    float *ef = sh;
    uint32_t *hist = (uint32_t *)(sh + n_g);
    uint32_t *list = hist + A.lds_reps * n_g;  // [NF][kCollPerWg]
    const uint32_t tile_sz = blockDim.x * PPT;
    for (int i = threadIdx.x; i < E; i += blockDim.x) {
        const EnsInst &in = j.inst[i];
        const uint32_t c = min(j.cnt[(size_t)i * EC_WORDS + j.c_in], ELECTRON ? in.e_cap : in.i_cap);
        s_cnt[i] = c;
        s_off[i] = ELECTRON ? in.e_off : in.i_off;
        s_tiles[i + 1] = (c + tile_sz - 1) / tile_sz;
    }
    for (int q = threadIdx.x; q < A.lds_reps * n_g; q += blockDim.x) hist[q] = 0u;
    if (threadIdx.x == 0) s_ncoll = 0;
    clear_slice(j.zero, j.zero_words);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        s_tiles[0] = 0;
        for (int i = 0; i < E; ++i) {
            acc += s_tiles[i + 1];
            s_tiles[i + 1] = acc;
        }
        if (ELECTRON && j.time && blockIdx.x == 0) *j.time = (*j.time + 1u) % (uint32_t)j.n_t;
    }
    __syncthreads();
    const uint32_t T = s_tiles[E];
    const uint32_t C = (uint32_t)A.chunk;
    const uint32_t t_begin = C ? blockIdx.x * C : (uint32_t)((uint64_t)T * blockIdx.x / gridDim.x);
    const uint32_t t_end = C ? T : (uint32_t)((uint64_t)T * (blockIdx.x + 1) / gridDim.x);
    if (t_begin >= t_end) return;  // uniform over the workgroup
    // the workgroup's next tile (tiles only grow)
    auto next_tile = [&](uint32_t t) {
        return (C == 0 || (t + 1) % C != 0) ? t + 1 : t + 1 + (gridDim.x - 1) * C;
    };

    const int wave = threadIdx.x / kWave;
    uint32_t *h = hist + (wave % A.lds_reps) * n_g;
    const float scale = 1048576.0f;  // fixed point 2^20, flushed every cap tiles
    const uint32_t waves_per_rep = (blockDim.x / kWave + A.lds_reps - 1) / A.lds_reps;
    const uint32_t cap = (4096u - kCollPerWg - 1u) / (waves_per_rep * 2u * PPT * kWave);
    const float factor = ELECTRON ? j.k.factor_e : j.k.factor_i;
    const float dt = ELECTRON ? j.k.dt_e : j.k.dt_i;
    float *const X = j.x, *const VX = j.vx, *const VY = j.vy, *const VZ = j.vz;
    const int rep = blockIdx.x % A.reps;
    // deferred RNG write-back (electrons), as in the single-discharge kernel
    const uint32_t lag = ELECTRON ? j.rng_lag : 0u;
    const uint32_t lag_out = ELECTRON ? j.rng_next_lag : 0u;
    uint8_t *const MK = j.mask;
    float thr = 1.0f;  // the current instance's mask threshold

    // instance of a tile: the largest i with s_tiles[i] <= tile (instances
    // without tiles are skipped); tiles only grow, so the search walks on
    auto inst_of = [&](uint32_t tile, int from) {
        int i = from;
        while (i + 1 < E && s_tiles[i + 1] <= tile) ++i;
        return i;
    };
    // loads of one tile (unconditional: lanes past the count read the
    // instance's first slot; the count test is applied at use)
    // li_*: first tile, end tile, first slot and live count of the instance
    // of the last tile loaded, kept in scalar registers so that a tile of
    // the same instance (the common case: contiguous ranges) needs no LDS
    // lookup before its loads are issued
    uint32_t li_t0 = 0, li_t1 = 0, li_off = 0, li_cnt = 0;
    auto inst_regs = [&](int i) {
        li_t0 = __builtin_amdgcn_readfirstlane(s_tiles[i]);
        li_t1 = __builtin_amdgcn_readfirstlane(s_tiles[i + 1]);
        li_off = __builtin_amdgcn_readfirstlane(s_off[i]);
        li_cnt = __builtin_amdgcn_readfirstlane(s_cnt[i]);
    };
    auto load_tile = [&](uint32_t tile, uint32_t &s0, uint32_t &avail, float (&lx)[PPT], float (&lvx)[PPT],
                         float (&lvy)[PPT], float (&lvz)[PPT], Xo (&lst)[PPT], uint32_t (&lmk)[PPT]) {
        uint32_t safe = 0;
        avail = 0;
        s0 = 0;
        if (tile < t_end) {
            safe = li_off;
            const uint32_t local0 = (tile - li_t0) * tile_sz + threadIdx.x;
            s0 = safe + local0;
            avail = li_cnt > local0 ? li_cnt - local0 : 0u;
        }
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const uint32_t g = u * blockDim.x < avail ? s0 + u * blockDim.x : safe;
            lx[u] = el(X, g);
            lvx[u] = el(VX, g);
            lvy[u] = el(VY, g);
            lvz[u] = el(VZ, g);
            if (RM == 1) lmk[u] = el(MK, g);
            else lst[u] = rng_load32(j.rng0, j.rng1, j.rng2, j.rng3, g);
        }
    };

    // the instance being processed and its physics
    int cur = -1;
    KPhys k = j.k;
    Appends app{};
    Acc *i_dens_g = nullptr;
    uint32_t lost_pow = 0, lost_gnd = 0, ncoll = 0, since_flush = 0;
    bool overflow = false;

    // leave instance `cur`: scatter its listed collisions, add its
    // diagnostics, flush and clear the histograms
    auto finish = [&]() {
        __syncthreads();
        const uint32_t n_list = min(s_ncoll, kCollPerWg);
        for (uint32_t b0 = (uint32_t)wave * kWave; b0 < n_list; b0 += blockDim.x) {
            const uint32_t e = b0 + lane_id();
            int kind = kNone;
            float se[3], si[3], px = 0.0f;
            if (e < n_list) {
                const uint32_t *q = list + e;
                const uint32_t slot = q[CF_SLOT * kCollPerWg];
                float w = __uint_as_float(q[CF_VX * kCollPerWg]), y = __uint_as_float(q[CF_VY * kCollPerWg]),
                      z = __uint_as_float(q[CF_VZ * kCollPerWg]);
                Xo s{q[CF_R0 * kCollPerWg], q[CF_R1 * kCollPerWg], q[CF_R2 * kCollPerWg], q[CF_R3 * kCollPerWg]};
                const int b = (int)q[CF_BIN * kCollPerWg];
                if (ELECTRON) {
                    px = __uint_as_float(q[CF_X * kCollPerWg]);
                    kind = e_scatter(k, j.cs[b], j.cs[kCsRanges + b], j.cs[2 * kCsRanges + b], w, y, z, s, se, si);
                } else {
                    kind = i_scatter(k, j.cs[3 * kCsRanges + b], j.cs[4 * kCsRanges + b], w, y, z,
                                     __uint_as_float(q[CF_AX * kCollPerWg]), __uint_as_float(q[CF_AY * kCollPerWg]),
                                     __uint_as_float(q[CF_AZ * kCollPerWg]), __uint_as_float(q[CF_G * kCollPerWg]),
                                     s);
                }
                if (RM) el(MK, slot) = (uint8_t)cand_mask(s, lag_out, thr);
                xo_back(s, lag_out);
                rng_store32(j.rng0, j.rng1, j.rng2, j.rng3, slot, s);
                el(VX, slot) = w;
                el(VY, slot) = y;
                el(VZ, slot) = z;
                ++ncoll;
            }
            if (ELECTRON)
                ionisation_products<false, Acc>(app, k, kind, px, se, si, h, scale, i_dens_g, nullptr, ef,
                                                overflow);
        }
        add_diagnostics<ELECTRON>(j.counters ? j.counters + (size_t)cur * 8 : nullptr,
                                  j.cnt + (size_t)cur * EC_WORDS + EC_OVERFLOW, lost_pow, lost_gnd, ncoll, overflow);
        lost_pow = lost_gnd = ncoll = 0;
        overflow = false;
        flush_fix<Acc>(hist, A.lds_reps, n_g, scale, k.wf,
                       (Acc *)j.dens + (size_t)cur * j.inst_stride + rep * A.rep_stride);
        __syncthreads();
        for (int q = threadIdx.x; q < A.lds_reps * n_g; q += blockDim.x) hist[q] = 0u;
        if (threadIdx.x == 0) s_ncoll = 0;
        since_flush = 0;
    };
    // enter instance i: its field and physics, its append regions
    auto enter = [&](int i) {
        for (int q = threadIdx.x; q < n_g; q += blockDim.x) ef[q] = j.efield[(size_t)i * n_g + q];
        const EnsInst &in = j.inst[i];
        k.wf = in.k.wf;
        k.stot_scale = in.k.stot_scale;
        k.normal_range = in.k.normal_range;
        thr = in.coll_thr;
        if (ELECTRON) {
            uint32_t *c = j.cnt + (size_t)i * EC_WORDS;
            app = Appends{X,      VX,     VY,     VZ,     c + j.c_out, in.e_off, in.e_cap,
                          j.ix,   j.ivx,  j.ivy,  j.ivz,  c + j.c_ion, in.i_off, in.i_cap,
                          j.rng0, j.rng1, j.rng2, j.rng3, lag_out, RM ? MK : nullptr};
            i_dens_g = (Acc *)j.ion_dens + (size_t)i * j.inst_stride + rep * A.rep_stride;
        }
        cur = i;
        __syncthreads();
    };

    float x[PPT], vx[PPT], vy[PPT], vz[PPT];
    Xo st[PPT];
    uint32_t mk[PPT];
    uint32_t slot0 = 0, avail = 0;
    uint32_t tile = t_begin;
    int ti = 0;
    {
        int lo = 0, hi = E - 1;  // binary search for the first tile's instance
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_tiles[mid] <= tile) lo = mid;
            else hi = mid - 1;
        }
        ti = inst_of(tile, lo);
    }
    ti = __builtin_amdgcn_readfirstlane(ti);
    inst_regs(ti);
    load_tile(tile, slot0, avail, x, vx, vy, vz, st, mk);
    while (tile < t_end) {
        tile_priority(C ? 4u : t_end - tile);
        if (ti != cur) {
            if (cur >= 0) finish();
            enter(ti);
        }
        const uint32_t tile_next = next_tile(tile);
        // the next tile's instance: the current one unless the tile is past
        // its end (li_* describe instance ti here)
        int ti_next = ti;
        if (tile_next < t_end && tile_next >= li_t1) {  // uniform
            ti_next = __builtin_amdgcn_readfirstlane(inst_of(tile_next, ti));
            inst_regs(ti_next);
        }
        bool valid[PPT];
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            valid[u] = u * blockDim.x < avail;
            if (!valid[u]) x[u] = -1.0f;
        }
        // a live slot's stream catches up with the draws it was spared;
        // with the masks read only candidates load their words here
        bool has[PPT];
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            has[u] = x[u] >= 0.0f;
            if (RM == 1) {
                has[u] = has[u] && (lag >= 8u || ((mk[u] >> lag) & 1u));
                if (has[u]) st[u] = rng_load32(j.rng0, j.rng1, j.rng2, j.rng3, slot0 + u * blockDim.x);
            }
        }
        if (RM != 1 && lag) {
#pragma unroll
            for (int u = 0; u < PPT; ++u)
                if (x[u] >= 0.0f) xo_skip(st[u], lag);
        }
        float nx[PPT], spd[PPT], sig[PPT];
        float ax[PPT], ay[PPT], az[PPT];
        int idx[PPT];
        bool in[PPT];
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            in[u] = false;
            nx[u] = 0.0f;
            idx[u] = 0;
            spd[u] = 0.0f;
            ax[u] = ay[u] = az[u] = 0.0f;
            if (x[u] >= 0.0f) {
                const float e = field_at(ef, x[u], k.inv_dx, n_g);
                vx[u] = __builtin_fmaf(factor, e, vx[u]);
                nx[u] = __builtin_fmaf(vx[u], dt, x[u]);
                in[u] = nx[u] >= 0.0f && nx[u] <= k.gap;
            }
            if (in[u]) {
                if (ELECTRON) idx[u] = e_index(k, vx[u], vy[u], vz[u], spd[u]);
                else idx[u] = i_index(k, vx[u], vy[u], vz[u], st[u], ax[u], ay[u], az[u], spd[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < PPT; ++u) sig[u] = el(j.stot, in[u] ? (uint32_t)idx[u] : 0u);
        float px[PPT], pvx[PPT], pvy[PPT], pvz[PPT];
        Xo pst[PPT];
        uint32_t pmk[PPT];
        uint32_t pslot0, pavail;
        load_tile(tile_next, pslot0, pavail, px, pvx, pvy, pvz, pst, pmk);
        // deposits and the x / vx write-back first, while the gather is in
        // flight (as in the single-discharge kernel)
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            if (in[u]) deposit_fix(h, nx[u], k, scale);
            if (valid[u]) {
                el(X, slot0 + u * blockDim.x) = in[u] ? nx[u] : -1.0f;
                el(VX, slot0 + u * blockDim.x) = vx[u];
            }
        }
        // masks read: a live slot that is no candidate loads its words now
        // if it may collide above the threshold or leaves the gap behind on
        // its words; then the loaded streams catch up
        float pc[PPT];
        if (RM == 1) {
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                pc[u] = in[u] ? e_prob(k, sig[u], spd[u]) : 0.0f;
                if (!has[u] && x[u] >= 0.0f && (in[u] ? pc[u] > thr : lag != 0u)) {
                    st[u] = rng_load32(j.rng0, j.rng1, j.rng2, j.rng3, slot0 + u * blockDim.x);
                    has[u] = true;
                }
            }
            if (lag) {
#pragma unroll
                for (int u = 0; u < PPT; ++u)
                    if (has[u]) xo_skip(st[u], lag);
            }
        }
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const uint32_t i = slot0 + u * blockDim.x;
            bool hit = false;
            if (in[u]) {
                if (RM == 1) hit = has[u] && draw(st[u]) < pc[u];  // no candidate: no collision
                else hit = ELECTRON ? e_test(k, sig[u], spd[u], st[u]) : i_test(k, sig[u], spd[u], st[u]);
            }
            if (valid[u]) {
                if (lag_out == 0u || (lag && !in[u] && x[u] >= 0.0f))
                    rng_store32(j.rng0, j.rng1, j.rng2, j.rng3, i, st[u]);
                // the next run's masks from the exact words (a colliding
                // particle's are rewritten after its scattering)
                if (RM == 2) el(MK, i) = (uint8_t)(in[u] && !hit ? cand_mask(st[u], 0u, thr) : 0xFFu);
            }
            if (!in[u] && x[u] >= 0.0f) {
                const bool pw = nx[u] < 0.0f;
                lost_pow += pw;
                lost_gnd += !pw;
            }
            const uint64_t m = ballot64(hit);
            if (m) {
                const int leader = __builtin_ctzll(m);
                uint32_t base = 0;
                if ((int)lane_id() == leader) base = atomicAdd(&s_ncoll, (uint32_t)__popcll(m));
                base = __shfl(base, leader, kWave);
                const uint32_t pos = base + mbcnt64(m);
                const bool listed = hit && pos < kCollPerWg;
                if (listed) {
                    uint32_t *e = list + pos;
                    e[CF_SLOT * kCollPerWg] = i;
                    e[CF_VX * kCollPerWg] = __float_as_uint(vx[u]);
                    e[CF_VY * kCollPerWg] = __float_as_uint(vy[u]);
                    e[CF_VZ * kCollPerWg] = __float_as_uint(vz[u]);
                    e[CF_R0 * kCollPerWg] = st[u].s0;
                    e[CF_R1 * kCollPerWg] = st[u].s1;
                    e[CF_R2 * kCollPerWg] = st[u].s2;
                    e[CF_R3 * kCollPerWg] = st[u].s3;
                    e[CF_BIN * kCollPerWg] = (uint32_t)idx[u];
                    if (ELECTRON) {
                        e[CF_X * kCollPerWg] = __float_as_uint(nx[u]);
                    } else {
                        e[CF_AX * kCollPerWg] = __float_as_uint(ax[u]);
                        e[CF_AY * kCollPerWg] = __float_as_uint(ay[u]);
                        e[CF_AZ * kCollPerWg] = __float_as_uint(az[u]);
                        e[CF_G * kCollPerWg] = __float_as_uint(spd[u]);
                    }
                }
                const bool local = hit && !listed;  // a full list: scatter here
                if (ballot64(local)) {
                    int kind = kNone;
                    float se[3], si[3];
                    if (local) {
                        float w = vx[u], y = vy[u], z = vz[u];
                        Xo s = st[u];
                        const int b = idx[u];
                        if (ELECTRON)
                            kind = e_scatter(k, j.cs[b], j.cs[kCsRanges + b], j.cs[2 * kCsRanges + b], w, y, z, s,
                                             se, si);
                        else
                            kind = i_scatter(k, j.cs[3 * kCsRanges + b], j.cs[4 * kCsRanges + b], w, y, z, ax[u],
                                             ay[u], az[u], spd[u], s);
                        if (RM) el(MK, i) = (uint8_t)cand_mask(s, lag_out, thr);
                        xo_back(s, lag_out);
                        rng_store32(j.rng0, j.rng1, j.rng2, j.rng3, i, s);
                        el(VX, i) = w;
                        el(VY, i) = y;
                        el(VZ, i) = z;
                        ++ncoll;
                    }
                    if (ELECTRON)
                        ionisation_products<false, Acc>(app, k, kind, nx[u], se, si, h, scale, i_dens_g, nullptr,
                                                        ef, overflow);
                }
            }
        }
        slot0 = pslot0;
        avail = pavail;
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            x[u] = px[u];
            vx[u] = pvx[u];
            vy[u] = pvy[u];
            vz[u] = pvz[u];
            if (RM == 1) mk[u] = pmk[u];
            else st[u] = pst[u];
        }
        tile = tile_next;
        ti = ti_next;
        if (++since_flush == cap && tile < t_end && ti == cur) {  // uniform over the workgroup
            // flush the histograms and scatter the listed collisions, so
            // that a long tile range does not fill the list (its overflow is
            // scattered in place, a divergent wave at a time)
            finish();
            __syncthreads();
        }
    }
    finish();
}

// Compaction pass of an ensemble species (see EnsCompactJob).
template <int PPT>
__global__ __launch_bounds__(256) void k_compact_ens(EnsCompactJob j) {
    __shared__ uint32_t s_tiles[kMaxEnsemble + 1], s_cnt[kMaxEnsemble];
    __shared__ uint32_t s_wave[32], s_base[2];
    const int E = j.n_inst;
    const uint32_t tile_sz = blockDim.x * PPT;
    for (int i = threadIdx.x; i < E; i += blockDim.x) {
        const EnsInst &in = j.inst[i];
        const uint32_t c = min(j.cnt[(size_t)i * EC_WORDS + j.c_in], j.electrons ? in.e_cap : in.i_cap);
        s_cnt[i] = c;
        s_tiles[i + 1] = (c + tile_sz - 1) / tile_sz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        s_tiles[0] = 0;
        for (int i = 0; i < E; ++i) {
            acc += s_tiles[i + 1];
            s_tiles[i + 1] = acc;
        }
    }
    __syncthreads();
    const uint32_t T = s_tiles[E];
    const uint32_t t_begin = (uint32_t)((uint64_t)T * blockIdx.x / gridDim.x);
    const uint32_t t_end = (uint32_t)((uint64_t)T * (blockIdx.x + 1) / gridDim.x);
    int i = 0;
    int buf = 0;
    for (uint32_t tile = t_begin; tile < t_end; ++tile, buf ^= 1) {
        while (i + 1 < E && s_tiles[i + 1] <= tile) ++i;
        const EnsInst &in = j.inst[i];
        const uint32_t off = j.electrons ? in.e_off : in.i_off;
        const uint32_t base = (tile - s_tiles[i]) * tile_sz + threadIdx.x;
        float v[4][PPT];
        bool keep[PPT];
        uint32_t o[PPT];
        uint32_t wcount = 0;
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const uint32_t l = base + u * blockDim.x;
            keep[u] = false;
            if (l < s_cnt[i]) {
#pragma unroll
                for (int c = 0; c < 4; ++c) v[c][u] = j.src[c][off + l];
                keep[u] = v[0][u] >= 0.0f;  // x < 0: a hole
            }
            const uint64_t m = ballot64(keep[u]);
            o[u] = wcount + mbcnt64(m);
            wcount += (uint32_t)__popcll(m);
        }
        const uint32_t wbase = tile_reserve_f(wcount, s_wave, s_base, buf, j.cnt + (size_t)i * EC_WORDS + j.c_out);
#pragma unroll
        for (int u = 0; u < PPT; ++u)
            if (keep[u]) {
                const uint32_t d = off + wbase + o[u];
#pragma unroll
                for (int c = 0; c < 4; ++c) j.dst[c][d] = v[c][u];
            }
    }
}

// Live particles (x >= 0) of every instance: out[i] += count of instance i
// (grid: blocks x instances).
__global__ __launch_bounds__(256) void k_count_live_ens(const float *__restrict__ x, const EnsInst *inst,
                                                       const uint32_t *__restrict__ cnt, int word, bool electrons,
                                                       uint32_t *__restrict__ out) {
    const int i = blockIdx.y;
    const EnsInst &in = inst[i];
    const uint32_t off = electrons ? in.e_off : in.i_off;
    const uint32_t n = min(cnt[(size_t)i * EC_WORDS + word], electrons ? in.e_cap : in.i_cap);
    uint32_t c = 0;
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < n; l += gridDim.x * blockDim.x)
        c += x[off + l] >= 0.0f;
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(out + i, c);
}

int launch_count_live_ens(const float *x, const EnsInst *inst, int n_inst, const uint32_t *cnt, int word,
                          bool electrons, uint32_t *out, hipStream_t s) {
    k_count_live_ens<<<dim3(16, n_inst), 256, 0, s>>>(x, inst, cnt, word, electrons, out);
    PIC_LAUNCHED("k_count_live_ens");
    return PIC_OK;
}

template <bool E, int PPT, int RM>
static void launch_ens_rm(const EnsArgs &a, bool f64, int grid, int block, size_t lds, hipStream_t s) {
    if (f64) k_inplace_ens<E, PPT, double, RM><<<grid, block, lds, s>>>(a);
    else k_inplace_ens<E, PPT, float, RM><<<grid, block, lds, s>>>(a);
}

template <bool E, int PPT>
static void launch_ens_t(const EnsArgs &a, bool f64, int grid, int block, size_t lds, hipStream_t s) {
    const int rm = a.j.mask_read ? 1 : a.j.mask_write ? 2 : 0;
    if constexpr (E) {
        if (rm == 1) launch_ens_rm<E, PPT, 1>(a, f64, grid, block, lds, s);
        else if (rm == 2) launch_ens_rm<E, PPT, 2>(a, f64, grid, block, lds, s);
        else launch_ens_rm<E, PPT, 0>(a, f64, grid, block, lds, s);
    } else {
        launch_ens_rm<E, PPT, 0>(a, f64, grid, block, lds, s);
    }
}

int launch_inplace_ens(const EnsStepJob &j, const Shape &sh0, hipStream_t s) {
    if (j.n_inst <= 0 || j.n_inst > kMaxEnsemble) {
        set_error("ensemble: 1 to 256 instances");
        return PIC_ERR_ARG;
    }
    Shape sh = resolve(sh0);
    if (!particle_block_ok(sh.block)) {
        set_error("ensemble: workgroup size must be 128, 256 or 512");
        return PIC_ERR_ARG;
    }
    if ((j.rng_lag || j.rng_next_lag) &&
        (!j.electrons || j.rng_lag >= kMaxRngLag || (j.rng_next_lag != 0 && j.rng_next_lag != j.rng_lag + 1))) {
        set_error("ensemble: deferred RNG write-back is for electron steps, next lag 0 or lag + 1");
        return PIC_ERR_ARG;
    }
    if ((j.mask_read || j.mask_write) && (!j.mask || !j.electrons || (j.mask_read && j.mask_write) ||
                                          (j.mask_read && j.rng_next_lag == 0) ||
                                          (j.mask_write && j.rng_next_lag != 0))) {
        set_error("ensemble: candidate masks are read on deferred electron steps and written on their "
                  "write-back steps");
        return PIC_ERR_ARG;
    }
    EnsArgs a;
    a.j = j;
    const int n_g = j.k.n_g;
    const int block = sh.block;
    const int grid = std::max(1, sh.grid);
    // LDS as in the single-discharge kernel: the field, the collision list
    // and up to one histogram replica per wave in what the resident
    // workgroups of a compute unit can share
    const int per_cu = std::max(1, (grid + device_cus() - 1) / device_cus());
    const int budget = std::min(device_lds_per_block(), device_lds_per_cu() / per_cu - 2048);
    const int fixed = (n_g + (j.electrons ? kCollFieldsE : kCollFieldsI) * (int)kCollPerWg) * 4;
    a.lds_reps = std::max(1, std::min(block / kWave, (budget - fixed) / (n_g * 4)));
    if (sh.lds_reps > 0) a.lds_reps = std::min(a.lds_reps, sh.lds_reps);
    a.reps = sh.replicas;
    a.rep_stride = sh.rep_stride > 0 ? sh.rep_stride : n_g;
    a.chunk = 0;
    const size_t lds = (size_t)fixed + (size_t)a.lds_reps * n_g * 4;
    const bool f64 = j.dens_f64;
    const int ppt = sh.ppt_inplace;
    if (j.electrons) {
        if (ppt == 1) launch_ens_t<true, 1>(a, f64, grid, block, lds, s);
        else if (ppt == 4) launch_ens_t<true, 4>(a, f64, grid, block, lds, s);
        else launch_ens_t<true, 2>(a, f64, grid, block, lds, s);
    } else {
        if (ppt == 1) launch_ens_t<false, 1>(a, f64, grid, block, lds, s);
        else if (ppt == 4) launch_ens_t<false, 4>(a, f64, grid, block, lds, s);
        else launch_ens_t<false, 2>(a, f64, grid, block, lds, s);
    }
    PIC_LAUNCHED("k_inplace_ens");
    return PIC_OK;
}

int launch_compact_ens(const EnsCompactJob &j, hipStream_t s) {
    if (j.n_inst <= 0 || j.n_inst > kMaxEnsemble) {
        set_error("ensemble: 1 to 256 instances");
        return PIC_ERR_ARG;
    }
    k_compact_ens<4><<<device_cus() * 4, 256, 0, s>>>(j);
    PIC_LAUNCHED("k_compact_ens");
    return PIC_OK;
}

}  // namespace pic
