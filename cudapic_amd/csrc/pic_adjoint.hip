// pic_adjoint.hip -- gradients of the window-averaged ion density misfit
// with respect to the Fourier coefficients of the driving voltage, by
// hand-written adjoint kernels for the push, the deposit and the field solve.
//
// The differentiated model is the RF step with its discrete events frozen:
// a record pass runs the ordinary step (Monte Carlo collisions on the
// per-slot xoshiro128+ streams) and writes an event tape -- for every step
// the particles that collided with their post-collision velocity, the
// particles that left the gap, and the ionisations with the slots and
// velocities of the new electron and ion.  Every later forward pass replays
// the tape instead of drawing: a collided particle's velocity is the taped
// one, an absorbed particle is removed at its taped step, a newborn appears
// at its parent's pushed position with its taped velocity.  With the events
// fixed, x and vx of every particle are differentiable functions of the
// voltage waveform V(t) = sum_h a_h cos(2 pi h t/N_T) + b_h sin(2 pi h t/N_T),
// and (frozen collisions) vy, vz play no part.
//
// Forward step t (the reference's stage order, main.rs:1725-1808):
//   n_e(t) = CIC(x_e(t)); on ion steps n_i(t) = CIC(x_i(t)), else held
//   phi = Poisson(e (n_i - n_e), phi_0 = V(t), phi_{N-1} = 0), E = -D phi
//   v_x += c E(x); x += v_x dt   (electrons every step, ions on ion steps)
//   then the tape of step t.
// Objective: J = sum_k (nbar_k - n*_k)^2 / sum_k n*_k^2 with nbar the ion
// density averaged over the window's ion steps.
//
// The densities are accumulated in 64-bit fixed point (2^31 units per
// particle split by the CIC weights), so that every forward pass -- the
// record pass, a replay, the recomputation of a checkpointed segment -- is
// bit-for-bit deterministic: integer sums do not depend on the order of the
// atomics.  The backward pass restores the state at a checkpoint every K
// steps, replays the K steps storing x(t) and E(t), and runs the adjoint
// steps in reverse:
//   tape (collided: lambda_v = 0; newborns hand lambda_x to their parent)
//   push adjoint: mu = lambda_v + dt lambda_x; lambda_x += mu c E'(x);
//                 lambda_v = mu; lambda_E[nodes of x] += mu c w(x)
//   field adjoint (one workgroup, FP64): lambda_E -> lambda_phi -> one
//                 tridiagonal solve -> lambda_rho, dJ/dV(t)
//   deposit adjoint: lambda_x += wf inv_dx (g[q+1] - g[q]),
//                 g = lambda_n times the node factors
// dJ/da_h = sum_t dJ/dV(t) cos(2 pi h (t0+t)/N_T), likewise b_h with sin.
// No reference counterpart: the reference has no differentiation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pic_internal.h"

using namespace pic;

namespace {

constexpr double kEpsilon0 = 8.85418781e-12;
constexpr double kECharge = 1.60217662e-19;
constexpr float kFixUnit = 2147483648.0f;  // 2^31 fixed-point units per particle
constexpr uint32_t kNever = 0xffffffffu;

// device counters
enum AdjCnt { AC_SLOTS_E = 0, AC_SLOTS_I, AC_COLL_E, AC_COLL_I, AC_ION, AC_OVERFLOW, AC_WORDS = 8 };

struct Tape {
    uint32_t *off_e, *off_i, *off_ion;  // [T+1] event offsets at the start of each step
    uint32_t *slot_e, *slot_i;          // collisions: slot
    float *v_e, *v_i;                   // collisions: vx after the collision
    uint32_t *par, *es, *is;            // ionisations: parent, new electron slot, new ion slot
    float *ve, *vi;                     // ionisations: vx of the new electron and ion
    uint32_t cap;                       // events per kind
};

// Cell, weight of the upper node and whether x is inside the gap, with the
// position clamped to the gap: a particle that leaves it without a taped
// absorption (a replay with another waveform can do that) deposits and
// feels the field at the wall, and its position derivative there is zero.
// Inside the gap this is field_at's and the deposit's cell arithmetic.
__device__ __forceinline__ int adj_cell(float x, float inv_dx, int n_g, float &rem, bool &inside) {
    const float raw = x * inv_dx;
    const float pos = fminf(fmaxf(raw, 0.0f), (float)(n_g - 1));
    const int q = min((int)pos, n_g - 2);
    rem = pos - (float)q;
    inside = raw > 0.0f && raw < (float)(n_g - 1);
    return q;
}

__device__ __forceinline__ float adj_field(const float *ef, float x, float inv_dx, int n_g) {
    float rem;
    bool inside;
    const int q = adj_cell(x, inv_dx, n_g, rem, inside);
    return __builtin_fmaf(rem, ef[q + 1], (1.0f - rem) * ef[q]);
}

// Fixed-point CIC deposit of one particle into an LDS histogram of u64.
__device__ __forceinline__ void dep_fix64(unsigned long long *h, float x, float inv_dx, int n_g) {
    float rem;
    bool inside;
    const int q = adj_cell(x, inv_dx, n_g, rem, inside);
    const unsigned long long u2 = (unsigned long long)(rem * kFixUnit);
    atomicAdd(&h[q], (unsigned long long)kFixUnit - u2);
    atomicAdd(&h[q + 1], u2);
}

__device__ __forceinline__ void flush_fix64(const unsigned long long *h, int n_g, unsigned long long *out) {
    for (int q = threadIdx.x; q < n_g; q += blockDim.x)
        if (h[q]) atomicAdd(&out[q], h[q]);
}

// Inclusive prefix sum over the 256 lanes of the workgroup (tmp: 512
// doubles of LDS).  Every lane calls it.
__device__ double block_scan256(double v, double *tmp) {
    double *a = tmp, *b = tmp + 256;
    const int tid = threadIdx.x;
    a[tid] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        b[tid] = a[tid] + (tid >= off ? a[tid - off] : 0.0);
        __syncthreads();
        double *c = a;
        a = b;
        b = c;
    }
    return a[tid];
}

// Solves tridiag(1, -2, 1) y = f on the interior nodes 1..n_g-2 (f and y in
// LDS, y may alias f) by the prefix-sum form of the Thomas recurrences
// (the engine's Poisson scan): G_i = sum_{j<=i} j f_j, r_i = -G_i / (i (i+1)),
// y_i = i sum_{j>=i} r_j.  Each lane takes a contiguous run of nodes.
// tmp: 512 doubles.  Every lane of a 256-lane workgroup calls it.
__device__ void scan_solve(const double *f, double *y, int n_g, double *tmp) {
    const int m = n_g - 2, per = (m + 255) / 256;
    const int i0 = 1 + threadIdx.x * per, i1 = min(i0 + per, n_g - 1);
    double loc = 0.0;
    for (int i = i0; i < i1; ++i) loc += (double)i * f[i];
    const double excl = block_scan256(loc, tmp) - loc;
    double r[16];
    double G = excl, rs = 0.0;
    for (int i = i0, u = 0; i < i1; ++i, ++u) {
        G += (double)i * f[i];
        r[u] = -G / ((double)i * (double)(i + 1));
        rs += r[u];
    }
    __syncthreads();  // f may be overwritten below (y aliases f)
    const double incl = block_scan256(rs, tmp);
    const double total = tmp[255];  // the scan's last round leaves the total in lane 255's slot
    double suffix = total - (incl - rs);  // sum_{j >= i0} r_j
    for (int i = i0, u = 0; i < i1; ++i, ++u) {
        y[i] = (double)i * suffix;
        suffix -= r[u];
    }
    __syncthreads();
}

struct PushArgs {
    bool record, electrons, store_traj;
    uint32_t t;                  // window step
    float *x, *vx, *vy, *vz;     // vy, vz: record pass only
    uint32_t *rng0, *rng1, *rng2, *rng3;
    uint32_t *death;
    const uint32_t *n_slots;     // [T+1]
    uint32_t *cnt;
    const float *efield;
    unsigned long long *acc;     // next density of this species
    float *traj;                 // x at the start of step t, [cap]
    float c, dt;
    KPhys k;
    const float *stot, *cs;
    Tape tape;
    // record pass, electrons: ionisation products
    float *ix, *ivx, *ivy, *ivz;
    uint32_t *i_death;
    unsigned long long *acc_i;
    uint32_t cap, cap_i;
};

// Push of one species over the particles alive at the start of step t, the
// deposit of the survivors into the next density and, in the record pass,
// absorption and collisions (taped).
__global__ __launch_bounds__(256) void k_adj_push(PushArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n_g = a.k.n_g;
    float *ef = (float *)smem;
    unsigned long long *h = (unsigned long long *)(smem + ((n_g * 4 + 15) / 16) * 16);
    unsigned long long *hi = h + n_g;  // record pass: newborn ions
    for (int q = threadIdx.x; q < n_g; q += blockDim.x) {
        ef[q] = a.efield[q];
        h[q] = 0ull;
        hi[q] = 0ull;
    }
    __syncthreads();
    // the slots in use at the start of the step (newborns of this step are
    // appended after them and wait for the next step)
    const uint32_t n = a.n_slots[a.t];
    bool any_i = false;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < min(n, a.cap); p += gridDim.x * blockDim.x) {
        if (a.death[p] <= a.t) continue;
        const float x0 = a.x[p];
        if (a.store_traj) a.traj[p] = x0;
        float vx = __builtin_fmaf(a.c, adj_field(ef, x0, a.k.inv_dx, n_g), a.vx[p]);
        const float x = __builtin_fmaf(vx, a.dt, x0);
        a.x[p] = x;
        a.vx[p] = vx;
        bool gone;
        if (a.record) {
            gone = !(x >= 0.0f && x <= a.k.gap);
            if (gone) a.death[p] = a.t + 1;
        } else {
            gone = a.death[p] == a.t + 1;
        }
        if (gone) continue;
        dep_fix64(h, x, a.k.inv_dx, n_g);
        if (!a.record) continue;
        float vy = a.vy[p], vz = a.vz[p];
        Xo s = rng_load(a.rng0, a.rng1, a.rng2, a.rng3, p);
        float se[3], si[3];
        int kind;
        if (a.electrons) kind = collide_electron(a.stot, a.cs, a.k, vx, vy, vz, s, se, si);
        else kind = collide_ion(a.stot, a.cs, a.k, vx, vy, vz, s);
        rng_store(a.rng0, a.rng1, a.rng2, a.rng3, p, s);
        if (kind == kNone) continue;
        a.vx[p] = vx;
        a.vy[p] = vy;
        a.vz[p] = vz;
        const uint32_t ev = atomicAdd(&a.cnt[a.electrons ? AC_COLL_E : AC_COLL_I], 1u);
        if (ev < a.tape.cap) {
            (a.electrons ? a.tape.slot_e : a.tape.slot_i)[ev] = p;
            (a.electrons ? a.tape.v_e : a.tape.v_i)[ev] = vx;
        } else {
            a.cnt[AC_OVERFLOW] = 1u;
        }
        if (kind != kIonisation) continue;
        const uint32_t q = atomicAdd(&a.cnt[AC_SLOTS_E], 1u);
        const uint32_t r = atomicAdd(&a.cnt[AC_SLOTS_I], 1u);
        const uint32_t e2 = atomicAdd(&a.cnt[AC_ION], 1u);
        if (q >= a.cap || r >= a.cap_i || e2 >= a.tape.cap) {
            a.cnt[AC_OVERFLOW] = 1u;
            continue;
        }
        a.x[q] = x;
        a.vx[q] = se[0];
        a.vy[q] = se[1];
        a.vz[q] = se[2];
        a.death[q] = kNever;
        a.ix[r] = x;
        a.ivx[r] = si[0];
        a.ivy[r] = si[1];
        a.ivz[r] = si[2];
        a.i_death[r] = kNever;
        a.tape.par[e2] = p;
        a.tape.es[e2] = q;
        a.tape.is[e2] = r;
        a.tape.ve[e2] = se[0];
        a.tape.vi[e2] = si[0];
        dep_fix64(h, x, a.k.inv_dx, n_g);
        dep_fix64(hi, x, a.k.inv_dx, n_g);
        any_i = true;
    }
    __syncthreads();
    flush_fix64(h, n_g, a.acc);
    if (__syncthreads_or(any_i)) flush_fix64(hi, n_g, a.acc_i);
}

struct ReplayArgs {
    uint32_t t;
    Tape tape;
    float *xe, *vxe, *xi, *vxi;
    unsigned long long *acc_e, *acc_i;
    float inv_dx;
    int n_g;
};

// The tape of step t in a replay: taped velocities of collided particles,
// newborn electrons and ions at their parent's pushed position.  Newborn
// densities go straight to the global accumulators (there are few).
__global__ __launch_bounds__(256) void k_adj_replay_tape(ReplayArgs a, bool ion_step) {
    const uint32_t stride = gridDim.x * blockDim.x, g = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t e = a.tape.off_e[a.t] + g; e < a.tape.off_e[a.t + 1]; e += stride)
        a.vxe[a.tape.slot_e[e]] = a.tape.v_e[e];
    if (ion_step)
        for (uint32_t e = a.tape.off_i[a.t] + g; e < a.tape.off_i[a.t + 1]; e += stride)
            a.vxi[a.tape.slot_i[e]] = a.tape.v_i[e];
    for (uint32_t e = a.tape.off_ion[a.t] + g; e < a.tape.off_ion[a.t + 1]; e += stride) {
        const uint32_t p = a.tape.par[e], q = a.tape.es[e], r = a.tape.is[e];
        const float x = a.xe[p];
        a.xe[q] = x;
        a.vxe[q] = a.tape.ve[e];
        a.xi[r] = x;
        a.vxi[r] = a.tape.vi[e];
        float rem;
        bool inside;
        const int c = adj_cell(x, a.inv_dx, a.n_g, rem, inside);
        const unsigned long long u2 = (unsigned long long)(rem * kFixUnit);
        atomicAdd(&a.acc_e[c], (unsigned long long)kFixUnit - u2);
        atomicAdd(&a.acc_e[c + 1], u2);
        atomicAdd(&a.acc_i[c], (unsigned long long)kFixUnit - u2);
        atomicAdd(&a.acc_i[c + 1], u2);
    }
}

// Deposit of the particles alive at step t (the state at a checkpoint).
__global__ __launch_bounds__(256) void k_adj_deposit(const float *x, const uint32_t *death, const uint32_t *n_slots,
                                                     uint32_t t, uint32_t cap, unsigned long long *acc, float inv_dx,
                                                     int n_g) {
    extern __shared__ unsigned long long hd[];
    for (int q = threadIdx.x; q < n_g; q += blockDim.x) hd[q] = 0ull;
    __syncthreads();
    const uint32_t n = min(n_slots[t], cap);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        if (death[p] > t) dep_fix64(hd, x[p], inv_dx, n_g);
    __syncthreads();
    flush_fix64(hd, n_g, acc);
}

struct FieldArgs {
    uint32_t t;
    bool ion_step, record, accumulate;
    int n_g;
    double v, dx, wf_unit;       // wf_unit = weight factor / 2^31
    unsigned long long *acc_e, *acc_i;
    double *n_e, *n_i;           // current electron density, held ion density
    double *nbar;                // objective accumulator (ion steps)
    float *efield, *ehist;       // ehist: E of this step for the backward pass (or null)
    uint32_t *cnt, *n_slots_e, *n_slots_i;
    Tape tape;
};

// Densities of step t from the fixed-point accumulators, the Poisson solve
// (Thomas, FP64, one thread) and the field; the record pass also stamps the
// slot watermarks and tape offsets of the step.
__global__ __launch_bounds__(256) void k_adj_field(FieldArgs a) {
    extern __shared__ double fs[];  // pot, rho, scan scratch
    const int n_g = a.n_g;
    double *pot = fs, *rho = fs + n_g, *tmp = fs + 2 * n_g;
    const double alpha = -a.dx * a.dx / kEpsilon0;
    for (int q = threadIdx.x; q < n_g; q += blockDim.x) {
        const double f = (q == 0 || q == n_g - 1) ? 2.0 : 1.0;
        const double ne = (double)a.acc_e[q] * a.wf_unit * f;
        a.n_e[q] = ne;
        a.acc_e[q] = 0ull;
        double ni;
        if (a.ion_step) {
            ni = (double)a.acc_i[q] * a.wf_unit * f;
            a.n_i[q] = ni;
            a.acc_i[q] = 0ull;
            if (a.accumulate) a.nbar[q] += ni;
        } else {
            ni = a.n_i[q];
        }
        rho[q] = kECharge * (ni - ne);
        pot[q] = alpha * rho[q] - (q == 1 ? a.v : 0.0);  // right-hand side f_q (interior nodes)
    }
    if (a.record && threadIdx.x == 0) {
        a.n_slots_e[a.t] = a.cnt[AC_SLOTS_E];
        a.n_slots_i[a.t] = a.cnt[AC_SLOTS_I];
        a.tape.off_e[a.t] = a.cnt[AC_COLL_E];
        a.tape.off_i[a.t] = a.cnt[AC_COLL_I];
        a.tape.off_ion[a.t] = a.cnt[AC_ION];
    }
    __syncthreads();
    scan_solve(pot, pot, n_g, tmp);
    if (threadIdx.x == 0) {
        pot[0] = a.v;
        pot[n_g - 1] = 0.0;
    }
    __syncthreads();
    const double s = 0.5 / a.dx;
    for (int q = threadIdx.x; q < n_g; q += blockDim.x) {
        double e;
        if (q == 0) e = (pot[0] - pot[1]) / a.dx - rho[0] * a.dx / (2.0 * kEpsilon0);
        else if (q == n_g - 1) e = (pot[n_g - 2] - pot[n_g - 1]) / a.dx + rho[q] * a.dx / (2.0 * kEpsilon0);
        else e = (pot[q - 1] - pot[q + 1]) * s;
        a.efield[q] = (float)e;
        if (a.ehist) a.ehist[q] = (float)e;
    }
}

struct BackPushArgs {
    uint32_t t;
    const float *traj, *ehist;
    float *lx, *lv;
    const uint32_t *death, *n_slots;
    uint32_t cap;
    double *lE;
    float c, dt, inv_dx;
    int n_g;
};

// Adjoint of the push of step t for one species (see the file header).
__global__ __launch_bounds__(256) void k_adj_push_back(BackPushArgs a) {
    // f64 LDS sums: on gfx950 ds_add_f64 runs about five times faster than
    // ds_add_f32 (measured on the same scatter pattern)
    extern __shared__ double lh[];  // n_g
    for (int q = threadIdx.x; q < a.n_g; q += blockDim.x) lh[q] = 0.0;
    __syncthreads();
    const uint32_t n = min(a.n_slots[a.t], a.cap);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        if (a.death[p] <= a.t) continue;
        float rem;
        bool inside;
        const int q = adj_cell(a.traj[p], a.inv_dx, a.n_g, rem, inside);
        const float lxp = a.lx[p];
        const float mu = __builtin_fmaf(a.dt, lxp, a.lv[p]);
        const float de = inside ? (a.ehist[q + 1] - a.ehist[q]) * a.inv_dx : 0.0f;
        a.lx[p] = __builtin_fmaf(mu * a.c, de, lxp);
        a.lv[p] = mu;
        const float m = mu * a.c;
        if (m != 0.0f) {
            atomicAdd(&lh[q], (double)(m * (1.0f - rem)));
            atomicAdd(&lh[q + 1], (double)(m * rem));
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < a.n_g; q += blockDim.x)
        if (lh[q] != 0.0) atomicAdd(&a.lE[q], lh[q]);
}

struct BackTapeArgs {
    uint32_t t;
    bool ion_step;
    Tape tape;
    float *lxe, *lve, *lxi, *lvi;
};

// Adjoint of the tape of step t: a taped velocity does not depend on the
// pushed one; a newborn's position is its parent's.
__global__ __launch_bounds__(256) void k_adj_tape_back(BackTapeArgs a) {
    const uint32_t stride = gridDim.x * blockDim.x, g = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t e = a.tape.off_e[a.t] + g; e < a.tape.off_e[a.t + 1]; e += stride) a.lve[a.tape.slot_e[e]] = 0.0f;
    if (a.ion_step)
        for (uint32_t e = a.tape.off_i[a.t] + g; e < a.tape.off_i[a.t + 1]; e += stride)
            a.lvi[a.tape.slot_i[e]] = 0.0f;
    for (uint32_t e = a.tape.off_ion[a.t] + g; e < a.tape.off_ion[a.t + 1]; e += stride) {
        const uint32_t p = a.tape.par[e], q = a.tape.es[e], r = a.tape.is[e];
        a.lxe[p] += a.lxe[q] + a.lxi[r];
        a.lxe[q] = a.lve[q] = 0.0f;
        a.lxi[r] = a.lvi[r] = 0.0f;
    }
}

struct BackFieldArgs {
    uint32_t t;
    bool ion_step;
    int n_g;
    double dx, wf;
    double *lE;        // consumed and cleared
    double *ln_held;   // adjoint of the held ion density (accumulates over the steps that use it)
    const double *lobj;  // objective's adjoint per ion step
    double *ge, *gi;   // outputs: lambda_n times wf and the node factors
    double *dJdV;
};

// Adjoint of the field solve of step t (FP64, one workgroup, serial solve).
__global__ __launch_bounds__(256) void k_adj_field_back(BackFieldArgs a) {
    extern __shared__ double bs[];
    const int n_g = a.n_g;
    double *lphi = bs, *lrho = bs + n_g, *le = bs + 2 * n_g, *tmp = bs + 3 * n_g;
    for (int q = threadIdx.x; q < n_g; q += blockDim.x) le[q] = a.lE[q];
    __syncthreads();
    // lambda_phi = D^T lambda_E (interior central differences, one-sided
    // boundary formulas) and the boundary charges' part of lambda_rho
    const double s = 0.5 / a.dx;
    for (int q = threadIdx.x; q < n_g; q += blockDim.x) {
        double v = 0.0;
        if (q + 1 <= n_g - 2) v += le[q + 1] * s;
        if (q - 1 >= 1) v -= le[q - 1] * s;
        if (q == 0) v += le[0] / a.dx;
        if (q == 1) v -= le[0] / a.dx;
        if (q == n_g - 2) v += le[n_g - 1] / a.dx;
        if (q == n_g - 1) v -= le[n_g - 1] / a.dx;
        lphi[q] = v;
        lrho[q] = q == 0 ? -le[0] * a.dx / (2.0 * kEpsilon0)
                         : (q == n_g - 1 ? le[n_g - 1] * a.dx / (2.0 * kEpsilon0) : lphi[q]);
    }
    __syncthreads();
    // u = A^-1 lambda_phi on the interior (A symmetric): the forward's solve
    // with f = lambda_phi
    scan_solve(lrho, lrho, n_g, tmp);
    // phi_0 = V enters directly and through f_1 -= V
    if (threadIdx.x == 0) a.dJdV[a.t] = lphi[0] - lrho[1];
    __syncthreads();
    const double alpha = -a.dx * a.dx / kEpsilon0;
    for (int q = threadIdx.x; q < n_g; q += blockDim.x) {
        const double f = ((q == 0 || q == n_g - 1) ? 2.0 : 1.0) * a.wf;
        const double lr = (q == 0 || q == n_g - 1 ? lrho[q] : alpha * lrho[q]) * kECharge;  // rho = e (n_i - n_e)
        a.ge[q] = -lr * f;
        double lh = a.ln_held[q] + lr;
        if (a.ion_step) {
            a.gi[q] = (lh + a.lobj[q]) * f;
            lh = 0.0;
        }
        a.ln_held[q] = lh;
        a.lE[q] = 0.0;
    }
}

struct BackDepArgs {
    uint32_t t;
    const float *traj;
    float *lx;
    const uint32_t *death, *n_slots;
    uint32_t cap;
    const double *g;
    float inv_dx;
    int n_g;
};

// Adjoint of the deposit of step t: lambda_x += d/dx of the CIC weights.
__global__ __launch_bounds__(256) void k_adj_deposit_back(BackDepArgs a) {
    extern __shared__ float gl[];
    for (int q = threadIdx.x; q < a.n_g; q += blockDim.x) gl[q] = (float)a.g[q];
    __syncthreads();
    const uint32_t n = min(a.n_slots[a.t], a.cap);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        if (a.death[p] <= a.t) continue;
        float rem;
        bool inside;
        const int q = adj_cell(a.traj[p], a.inv_dx, a.n_g, rem, inside);
        if (inside) a.lx[p] += a.inv_dx * (gl[q + 1] - gl[q]);
    }
}

}  // namespace

struct pic_adjoint {
    pic_params_t p;
    KPhys k;
    uint32_t T = 0, K = 0, t0 = 0, cap_e = 0, cap_i = 0, n_sub = 20;
    int grid = 0;
    float *xe = nullptr, *vxe = nullptr, *vye = nullptr, *vze = nullptr;
    float *xi = nullptr, *vxi = nullptr, *vyi = nullptr, *vzi = nullptr;
    float *x0e = nullptr, *vx0e = nullptr, *x0i = nullptr, *vx0i = nullptr;  // initial x, vx
    uint32_t n0e = 0, n0i = 0;
    uint32_t *rng_e[4] = {}, *rng_i[4] = {};
    uint32_t *death_e = nullptr, *death_i = nullptr, *death0_e = nullptr, *death0_i = nullptr;
    uint32_t *n_slots_e = nullptr, *n_slots_i = nullptr;  // [T+1]
    uint32_t *cnt = nullptr;
    Tape tape{};
    unsigned long long *acc_e = nullptr, *acc_i = nullptr;
    double *n_e = nullptr, *n_i = nullptr, *nbar = nullptr;
    float *efield = nullptr, *ehist = nullptr;
    float *traj_e = nullptr, *traj_i = nullptr;
    float *ck_xe = nullptr, *ck_vxe = nullptr, *ck_xi = nullptr, *ck_vxi = nullptr;
    float *lxe = nullptr, *lve = nullptr, *lxi = nullptr, *lvi = nullptr;
    double *lE = nullptr, *ln_held = nullptr, *lobj = nullptr, *ge = nullptr, *gi = nullptr, *dJdV = nullptr;
    float *cs = nullptr, *stot_e = nullptr, *stot_i = nullptr;
    std::vector<double> target, nbar_h;
    double wave_cos[PIC_MAX_HARMONICS] = {}, wave_sin[PIC_MAX_HARMONICS] = {};
    bool recorded = false, have_forward = false;
    uint32_t uploaded = 0;  // bit 0 electrons, bit 1 ions
    hipStream_t stream = nullptr;
};

namespace {

#define TRY(x)                         \
    do {                               \
        int rc_ = (x);                 \
        if (rc_ != PIC_OK) return rc_; \
    } while (0)

template <typename T>
int dalloc(T **p, size_t n) {
    PIC_HIP(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    PIC_HIP(hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return PIC_OK;
}

double voltage(const pic_adjoint *a, uint32_t t) {
    const double ph = 2.0 * M_PI * (double)((a->t0 + t) % (uint32_t)a->p.n_t) / (double)a->p.n_t;
    double v = 0.0;
    for (int h = 0; h < PIC_MAX_HARMONICS; ++h)
        if (a->wave_cos[h] != 0.0 || a->wave_sin[h] != 0.0)
            v += a->wave_cos[h] * std::cos((h + 1) * ph) + a->wave_sin[h] * std::sin((h + 1) * ph);
    return v;
}

bool ion_step(const pic_adjoint *a, uint32_t t) { return t % a->n_sub == 0; }

// Window step t, forward: field, ion push (ion steps), electron push, tape.
int fwd_step(pic_adjoint *a, uint32_t t, bool record, bool accumulate, uint32_t seg0, bool store) {
    hipStream_t st = a->stream;
    const int n_g = a->p.n_g;
    const bool ion = ion_step(a, t);
    FieldArgs f{};
    f.t = t;
    f.ion_step = ion;
    f.record = record;
    f.accumulate = accumulate;
    f.n_g = n_g;
    f.v = voltage(a, t);
    f.dx = a->p.dx;
    f.wf_unit = (double)a->k.wf / (double)kFixUnit;
    f.acc_e = a->acc_e;
    f.acc_i = a->acc_i;
    f.n_e = a->n_e;
    f.n_i = a->n_i;
    f.nbar = a->nbar;
    f.efield = a->efield;
    f.ehist = store ? a->ehist + (size_t)(t - seg0) * n_g : nullptr;
    f.cnt = a->cnt;
    f.n_slots_e = a->n_slots_e;
    f.n_slots_i = a->n_slots_i;
    f.tape = a->tape;
    k_adj_field<<<1, 256, (2 * n_g + 512) * sizeof(double), st>>>(f);
    PIC_LAUNCHED("k_adj_field");
    const size_t lds = ((n_g * 4 + 15) / 16) * 16 + 2 * (size_t)n_g * 8;
    PushArgs pa{};
    pa.record = record;
    pa.t = t;
    pa.store_traj = store;
    pa.cnt = a->cnt;
    pa.n_slots = nullptr;
    pa.efield = a->efield;
    pa.k = a->k;
    pa.cs = a->cs;
    pa.tape = a->tape;
    if (ion) {
        PushArgs pi = pa;
        pi.electrons = false;
        pi.x = a->xi;
        pi.vx = a->vxi;
        pi.vy = a->vyi;
        pi.vz = a->vzi;
        for (int w = 0; w < 4; ++w) (&pi.rng0)[w] = a->rng_i[w];
        pi.death = a->death_i;
        pi.n_slots = a->n_slots_i;
        pi.acc = a->acc_i;
        pi.traj = store ? a->traj_i + (size_t)((t - seg0) / a->n_sub) * a->cap_i : nullptr;
        pi.c = a->k.factor_i;
        pi.dt = a->k.dt_i;
        pi.stot = a->stot_i;
        pi.cap = a->cap_i;
        pi.cap_i = a->cap_i;
        pi.acc_i = a->acc_i;
        k_adj_push<<<a->grid, 256, lds, st>>>(pi);
        PIC_LAUNCHED("k_adj_push");
    }
    PushArgs pe = pa;
    pe.electrons = true;
    pe.x = a->xe;
    pe.vx = a->vxe;
    pe.vy = a->vye;
    pe.vz = a->vze;
    for (int w = 0; w < 4; ++w) (&pe.rng0)[w] = a->rng_e[w];
    pe.death = a->death_e;
    pe.n_slots = a->n_slots_e;
    pe.acc = a->acc_e;
    pe.traj = store ? a->traj_e + (size_t)(t - seg0) * a->cap_e : nullptr;
    pe.c = a->k.factor_e;
    pe.dt = a->k.dt_e;
    pe.stot = a->stot_e;
    pe.ix = a->xi;
    pe.ivx = a->vxi;
    pe.ivy = a->vyi;
    pe.ivz = a->vzi;
    pe.i_death = a->death_i;
    pe.acc_i = a->acc_i;
    pe.cap = a->cap_e;
    pe.cap_i = a->cap_i;
    k_adj_push<<<a->grid, 256, lds, st>>>(pe);
    PIC_LAUNCHED("k_adj_push");
    if (!record) {
        ReplayArgs r{};
        r.t = t;
        r.tape = a->tape;
        r.xe = a->xe;
        r.vxe = a->vxe;
        r.xi = a->xi;
        r.vxi = a->vxi;
        r.acc_e = a->acc_e;
        r.acc_i = a->acc_i;
        r.inv_dx = a->k.inv_dx;
        r.n_g = n_g;
        k_adj_replay_tape<<<64, 256, 0, st>>>(r, ion);
        PIC_LAUNCHED("k_adj_replay_tape");
    }
    return PIC_OK;
}

// Reset the particle state to the uploaded one and deposit it.
int restart(pic_adjoint *a) {
    hipStream_t st = a->stream;
    PIC_HIP(hipMemcpyAsync(a->xe, a->x0e, 4ull * a->cap_e, hipMemcpyDeviceToDevice, st));
    PIC_HIP(hipMemcpyAsync(a->vxe, a->vx0e, 4ull * a->cap_e, hipMemcpyDeviceToDevice, st));
    PIC_HIP(hipMemcpyAsync(a->xi, a->x0i, 4ull * a->cap_i, hipMemcpyDeviceToDevice, st));
    PIC_HIP(hipMemcpyAsync(a->vxi, a->vx0i, 4ull * a->cap_i, hipMemcpyDeviceToDevice, st));
    PIC_HIP(hipMemsetAsync(a->acc_e, 0, 8ull * a->p.n_g, st));
    PIC_HIP(hipMemsetAsync(a->acc_i, 0, 8ull * a->p.n_g, st));
    PIC_HIP(hipMemsetAsync(a->nbar, 0, 8ull * a->p.n_g, st));
    return PIC_OK;
}

int deposit_at(pic_adjoint *a, uint32_t t) {
    const int n_g = a->p.n_g;
    k_adj_deposit<<<a->grid, 256, n_g * 8, a->stream>>>(a->xe, a->death_e, a->n_slots_e, t, a->cap_e, a->acc_e,
                                                        a->k.inv_dx, n_g);
    k_adj_deposit<<<a->grid, 256, n_g * 8, a->stream>>>(a->xi, a->death_i, a->n_slots_i, t, a->cap_i, a->acc_i,
                                                        a->k.inv_dx, n_g);
    PIC_LAUNCHED("k_adj_deposit");
    return PIC_OK;
}

int checkpoint(pic_adjoint *a, uint32_t seg) {
    hipStream_t st = a->stream;
    const size_t oe = (size_t)seg * a->cap_e, oi = (size_t)seg * a->cap_i;
    PIC_HIP(hipMemcpyAsync(a->ck_xe + oe, a->xe, 4ull * a->cap_e, hipMemcpyDeviceToDevice, st));
    PIC_HIP(hipMemcpyAsync(a->ck_vxe + oe, a->vxe, 4ull * a->cap_e, hipMemcpyDeviceToDevice, st));
    PIC_HIP(hipMemcpyAsync(a->ck_xi + oi, a->xi, 4ull * a->cap_i, hipMemcpyDeviceToDevice, st));
    PIC_HIP(hipMemcpyAsync(a->ck_vxi + oi, a->vxi, 4ull * a->cap_i, hipMemcpyDeviceToDevice, st));
    return PIC_OK;
}

// A whole-window forward: the record pass or a replay.  Checkpoints every K
// steps; J from the window-averaged ion density.
int forward(pic_adjoint *a, bool record, double *J) {
    hipStream_t st = a->stream;
    TRY(restart(a));
    if (record) {
        uint32_t c0[AC_WORDS] = {a->n0e, a->n0i, 0, 0, 0, 0, 0, 0};
        PIC_HIP(hipMemcpyAsync(a->cnt, c0, sizeof(c0), hipMemcpyHostToDevice, st));
        PIC_HIP(hipMemcpyAsync(a->n_slots_e, &a->n0e, 4, hipMemcpyHostToDevice, st));
        PIC_HIP(hipMemcpyAsync(a->n_slots_i, &a->n0i, 4, hipMemcpyHostToDevice, st));
        PIC_HIP(hipMemcpyAsync(a->death_e, a->death0_e, 4ull * a->cap_e, hipMemcpyDeviceToDevice, st));
        PIC_HIP(hipMemcpyAsync(a->death_i, a->death0_i, 4ull * a->cap_i, hipMemcpyDeviceToDevice, st));
    }
    TRY(deposit_at(a, 0));
    for (uint32_t t = 0; t < a->T; ++t) {
        if (t % a->K == 0) TRY(checkpoint(a, t / a->K));
        TRY(fwd_step(a, t, record, true, 0, false));
    }
    if (record) {
        // closing stamps (offsets and watermarks after the last step)
        uint32_t c[AC_WORDS];
        PIC_HIP(hipMemcpyAsync(c, a->cnt, sizeof(c), hipMemcpyDeviceToHost, st));
        PIC_HIP(hipStreamSynchronize(st));
        if (c[AC_OVERFLOW]) {
            set_error("pic_adj_record: particle or tape capacity exceeded");
            return PIC_ERR_CAPACITY;
        }
        const uint32_t T = a->T;
        PIC_HIP(hipMemcpy(a->n_slots_e + T, &c[AC_SLOTS_E], 4, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(a->n_slots_i + T, &c[AC_SLOTS_I], 4, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(a->tape.off_e + T, &c[AC_COLL_E], 4, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(a->tape.off_i + T, &c[AC_COLL_I], 4, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(a->tape.off_ion + T, &c[AC_ION], 4, hipMemcpyHostToDevice));
        a->recorded = true;
    }
    // the record pass leaves every slot's taped death step (step + 1 of its
    // absorption, or never) in death_e / death_i: replays only read them
    a->nbar_h.assign(a->p.n_g, 0.0);
    PIC_HIP(hipMemcpyAsync(a->nbar_h.data(), a->nbar, 8ull * a->p.n_g, hipMemcpyDeviceToHost, st));
    PIC_HIP(hipStreamSynchronize(st));
    const double M = (double)((a->T + a->n_sub - 1) / a->n_sub);
    double num = 0.0, den = 0.0;
    for (int q = 0; q < a->p.n_g; ++q) {
        a->nbar_h[q] /= M;
        const double d = a->nbar_h[q] - a->target[q];
        num += d * d;
        den += a->target[q] * a->target[q];
    }
    if (J) *J = den > 0.0 ? num / den : num;
    a->have_forward = true;
    return PIC_OK;
}

int backward(pic_adjoint *a, double *grad_cos, double *grad_sin, double *dJdV_out) {
    hipStream_t st = a->stream;
    const int n_g = a->p.n_g;
    const uint32_t T = a->T, K = a->K;
    // objective adjoint per ion step: dJ/dn_i(t) = 2 (nbar - n*) / (M sum n*^2)
    {
        const double M = (double)((T + a->n_sub - 1) / a->n_sub);
        double den = 0.0;
        for (int q = 0; q < n_g; ++q) den += a->target[q] * a->target[q];
        if (den <= 0.0) den = 1.0;
        std::vector<double> lo(n_g);
        for (int q = 0; q < n_g; ++q) lo[q] = 2.0 * (a->nbar_h[q] - a->target[q]) / (M * den);
        PIC_HIP(hipMemcpyAsync(a->lobj, lo.data(), 8ull * n_g, hipMemcpyHostToDevice, st));
    }
    PIC_HIP(hipMemsetAsync(a->lxe, 0, 4ull * a->cap_e, st));
    PIC_HIP(hipMemsetAsync(a->lve, 0, 4ull * a->cap_e, st));
    PIC_HIP(hipMemsetAsync(a->lxi, 0, 4ull * a->cap_i, st));
    PIC_HIP(hipMemsetAsync(a->lvi, 0, 4ull * a->cap_i, st));
    PIC_HIP(hipMemsetAsync(a->lE, 0, 8ull * n_g, st));
    PIC_HIP(hipMemsetAsync(a->ln_held, 0, 8ull * n_g, st));
    const uint32_t nseg = (T + K - 1) / K;
    for (int seg = (int)nseg - 1; seg >= 0; --seg) {
        const uint32_t s = (uint32_t)seg * K, e = std::min(T, s + K);
        // recompute the segment from its checkpoint, keeping x(t) and E(t)
        const size_t oe = (size_t)seg * a->cap_e, oi = (size_t)seg * a->cap_i;
        PIC_HIP(hipMemcpyAsync(a->xe, a->ck_xe + oe, 4ull * a->cap_e, hipMemcpyDeviceToDevice, st));
        PIC_HIP(hipMemcpyAsync(a->vxe, a->ck_vxe + oe, 4ull * a->cap_e, hipMemcpyDeviceToDevice, st));
        PIC_HIP(hipMemcpyAsync(a->xi, a->ck_xi + oi, 4ull * a->cap_i, hipMemcpyDeviceToDevice, st));
        PIC_HIP(hipMemcpyAsync(a->vxi, a->ck_vxi + oi, 4ull * a->cap_i, hipMemcpyDeviceToDevice, st));
        PIC_HIP(hipMemsetAsync(a->acc_e, 0, 8ull * n_g, st));
        PIC_HIP(hipMemsetAsync(a->acc_i, 0, 8ull * n_g, st));
        TRY(deposit_at(a, s));
        for (uint32_t t = s; t < e; ++t) TRY(fwd_step(a, t, false, false, s, true));
        // adjoint steps in reverse
        for (int ti = (int)e - 1; ti >= (int)s; --ti) {
            const uint32_t t = (uint32_t)ti;
            const bool ion = ion_step(a, t);
            BackTapeArgs bt{t, ion, a->tape, a->lxe, a->lve, a->lxi, a->lvi};
            k_adj_tape_back<<<64, 256, 0, st>>>(bt);
            PIC_LAUNCHED("k_adj_tape_back");
            const float *eh = a->ehist + (size_t)(t - s) * n_g;
            BackPushArgs bp{};
            bp.t = t;
            bp.ehist = eh;
            bp.lE = a->lE;
            bp.inv_dx = a->k.inv_dx;
            bp.n_g = n_g;
            if (ion) {
                BackPushArgs bi = bp;
                bi.traj = a->traj_i + (size_t)((t - s) / a->n_sub) * a->cap_i;
                bi.lx = a->lxi;
                bi.lv = a->lvi;
                bi.death = a->death_i;
                bi.n_slots = a->n_slots_i;
                bi.cap = a->cap_i;
                bi.c = a->k.factor_i;
                bi.dt = a->k.dt_i;
                k_adj_push_back<<<a->grid, 256, n_g * sizeof(double), st>>>(bi);
            }
            BackPushArgs be = bp;
            be.traj = a->traj_e + (size_t)(t - s) * a->cap_e;
            be.lx = a->lxe;
            be.lv = a->lve;
            be.death = a->death_e;
            be.n_slots = a->n_slots_e;
            be.cap = a->cap_e;
            be.c = a->k.factor_e;
            be.dt = a->k.dt_e;
            k_adj_push_back<<<a->grid, 256, n_g * sizeof(double), st>>>(be);
            PIC_LAUNCHED("k_adj_push_back");
            BackFieldArgs bf{t, ion, n_g, a->p.dx, (double)a->k.wf, a->lE, a->ln_held, a->lobj, a->ge, a->gi, a->dJdV};
            k_adj_field_back<<<1, 256, (3 * n_g + 512) * sizeof(double), st>>>(bf);
            PIC_LAUNCHED("k_adj_field_back");
            BackDepArgs bd{t, be.traj, a->lxe, a->death_e, a->n_slots_e, a->cap_e, a->ge, a->k.inv_dx, n_g};
            k_adj_deposit_back<<<a->grid, 256, n_g * 4, st>>>(bd);
            if (ion) {
                BackDepArgs bdi{t, a->traj_i + (size_t)((t - s) / a->n_sub) * a->cap_i, a->lxi, a->death_i,
                                a->n_slots_i, a->cap_i, a->gi, a->k.inv_dx, n_g};
                k_adj_deposit_back<<<a->grid, 256, n_g * 4, st>>>(bdi);
            }
            PIC_LAUNCHED("k_adj_deposit_back");
        }
    }
    std::vector<double> g(T);
    PIC_HIP(hipMemcpyAsync(g.data(), a->dJdV, 8ull * T, hipMemcpyDeviceToHost, st));
    PIC_HIP(hipStreamSynchronize(st));
    for (int h = 0; h < PIC_MAX_HARMONICS; ++h) {
        double gc = 0.0, gs = 0.0;
        for (uint32_t t = 0; t < T; ++t) {
            const double ph = 2.0 * M_PI * (double)((a->t0 + t) % (uint32_t)a->p.n_t) / (double)a->p.n_t;
            gc += g[t] * std::cos((h + 1) * ph);
            gs += g[t] * std::sin((h + 1) * ph);
        }
        if (grad_cos) grad_cos[h] = gc;
        if (grad_sin) grad_sin[h] = gs;
    }
    if (dJdV_out) std::copy(g.begin(), g.end(), dJdV_out);
    return PIC_OK;
}

void free_all(pic_adjoint *a) {
    void *bufs[] = {a->xe, a->vxe, a->vye, a->vze, a->xi, a->vxi, a->vyi, a->vzi, a->x0e, a->vx0e, a->x0i,
                    a->vx0i, a->death_e, a->death_i, a->death0_e, a->death0_i, a->n_slots_e, a->n_slots_i,
                    a->cnt, a->tape.off_e, a->tape.off_i, a->tape.off_ion, a->tape.slot_e, a->tape.slot_i,
                    a->tape.v_e, a->tape.v_i, a->tape.par, a->tape.es, a->tape.is, a->tape.ve, a->tape.vi,
                    a->acc_e, a->acc_i, a->n_e, a->n_i, a->nbar, a->efield, a->ehist, a->traj_e, a->traj_i,
                    a->ck_xe, a->ck_vxe, a->ck_xi, a->ck_vxi, a->lxe, a->lve, a->lxi, a->lvi, a->lE,
                    a->ln_held, a->lobj, a->ge, a->gi, a->dJdV, a->cs, a->stot_e, a->stot_i};
    for (void *b : bufs) (void)hipFree(b);
    for (int w = 0; w < 4; ++w) {
        (void)hipFree(a->rng_e[w]);
        (void)hipFree(a->rng_i[w]);
    }
    if (a->stream) (void)hipStreamDestroy(a->stream);
}

}  // namespace

extern "C" {

int pic_adj_create(const pic_adj_config_t *c, pic_adjoint_t **out) {
    if (!c || !out || c->capacity_e == 0 || c->capacity_i == 0 || c->steps == 0 || c->params.dt_e <= 0.0) {
        set_error("pic_adj_create: capacities, steps and derived params are required");
        return PIC_ERR_ARG;
    }
    const uint32_t n_sub = (uint32_t)c->params.n_sub;
    const uint32_t K = c->checkpoint ? c->checkpoint : std::min<uint32_t>(c->steps, 200u);
    if (c->steps % n_sub || K % n_sub || c->t0 % n_sub || c->params.n_g < 3 || c->params.n_g > 2048) {
        set_error("pic_adj_create: steps, checkpoint interval and t0 must be multiples of n_sub, n_g at most 2048");
        return PIC_ERR_ARG;
    }
    pic_adjoint *a = new pic_adjoint();
    a->p = c->params;
    a->k = make_kphys(c->params);
    a->T = c->steps;
    a->K = K;
    a->t0 = c->t0;
    a->n_sub = n_sub;
    a->cap_e = c->capacity_e;
    a->cap_i = c->capacity_i;
    a->grid = device_cus() * 4;
    memcpy(a->wave_cos, c->params.wave_cos, sizeof(a->wave_cos));
    memcpy(a->wave_sin, c->params.wave_sin, sizeof(a->wave_sin));
    a->target.assign(a->p.n_g, 0.0);
    const uint64_t tape_cap = c->tape_capacity ? c->tape_capacity
                                               : std::max<uint64_t>(1u << 16, (uint64_t)c->steps * c->capacity_e / 64);
    if (tape_cap > 0xffffffffull) {
        delete a;
        set_error("pic_adj_create: tape capacity above 2^32 events");
        return PIC_ERR_ARG;
    }
    a->tape.cap = (uint32_t)tape_cap;
    auto fail = [&](int r) {
        free_all(a);
        delete a;
        return r;
    };
    int rc;
    if (hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(hip_fail(hipGetLastError(), "hipStreamCreate"));
    const size_t ce = a->cap_e, ci = a->cap_i, n_g = a->p.n_g, T = a->T, nseg = (a->T + K - 1) / K;
    float **fe[] = {&a->xe, &a->vxe, &a->vye, &a->vze, &a->x0e, &a->vx0e, &a->lxe, &a->lve};
    float **fi[] = {&a->xi, &a->vxi, &a->vyi, &a->vzi, &a->x0i, &a->vx0i, &a->lxi, &a->lvi};
    for (float **b : fe)
        if ((rc = dalloc(b, ce))) return fail(rc);
    for (float **b : fi)
        if ((rc = dalloc(b, ci))) return fail(rc);
    for (int w = 0; w < 4; ++w) {
        if ((rc = dalloc(&a->rng_e[w], ce))) return fail(rc);
        if ((rc = dalloc(&a->rng_i[w], ci))) return fail(rc);
    }
    if ((rc = dalloc(&a->death_e, ce)) || (rc = dalloc(&a->death0_e, ce)) || (rc = dalloc(&a->death_i, ci)) ||
        (rc = dalloc(&a->death0_i, ci)) || (rc = dalloc(&a->n_slots_e, T + 1)) || (rc = dalloc(&a->n_slots_i, T + 1)) ||
        (rc = dalloc(&a->cnt, AC_WORDS)) || (rc = dalloc(&a->tape.off_e, T + 1)) ||
        (rc = dalloc(&a->tape.off_i, T + 1)) || (rc = dalloc(&a->tape.off_ion, T + 1)))
        return fail(rc);
    if ((rc = dalloc(&a->tape.slot_e, tape_cap)) || (rc = dalloc(&a->tape.v_e, tape_cap)) ||
        (rc = dalloc(&a->tape.slot_i, tape_cap)) || (rc = dalloc(&a->tape.v_i, tape_cap)) ||
        (rc = dalloc(&a->tape.par, tape_cap)) || (rc = dalloc(&a->tape.es, tape_cap)) ||
        (rc = dalloc(&a->tape.is, tape_cap)) || (rc = dalloc(&a->tape.ve, tape_cap)) ||
        (rc = dalloc(&a->tape.vi, tape_cap)))
        return fail(rc);
    if ((rc = dalloc(&a->acc_e, n_g)) || (rc = dalloc(&a->acc_i, n_g)) || (rc = dalloc(&a->n_e, n_g)) ||
        (rc = dalloc(&a->n_i, n_g)) || (rc = dalloc(&a->nbar, n_g)) || (rc = dalloc(&a->efield, n_g)) ||
        (rc = dalloc(&a->ehist, (size_t)K * n_g)) || (rc = dalloc(&a->traj_e, (size_t)K * ce)) ||
        (rc = dalloc(&a->traj_i, (size_t)(K / n_sub + 1) * ci)) || (rc = dalloc(&a->ck_xe, nseg * ce)) ||
        (rc = dalloc(&a->ck_vxe, nseg * ce)) || (rc = dalloc(&a->ck_xi, nseg * ci)) ||
        (rc = dalloc(&a->ck_vxi, nseg * ci)) || (rc = dalloc(&a->lE, n_g)) || (rc = dalloc(&a->ln_held, n_g)) ||
        (rc = dalloc(&a->lobj, n_g)) || (rc = dalloc(&a->ge, n_g)) || (rc = dalloc(&a->gi, n_g)) ||
        (rc = dalloc(&a->dJdV, T)))
        return fail(rc);
    if ((rc = dalloc(&a->cs, (size_t)PIC_N_CS * PIC_CS_RANGES)) || (rc = dalloc(&a->stot_e, PIC_CS_RANGES)) ||
        (rc = dalloc(&a->stot_i, PIC_CS_RANGES)))
        return fail(rc);
    {
        std::vector<float> cs((size_t)PIC_N_CS * PIC_CS_RANGES), se(PIC_CS_RANGES), si(PIC_CS_RANGES);
        if ((rc = pic_cross_sections(a->p.gas_density, cs.data(), se.data(), si.data()))) return fail(rc);
        if (hipMemcpy(a->cs, cs.data(), cs.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(a->stot_e, se.data(), se.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(a->stot_i, si.data(), si.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return fail(hip_fail(hipGetLastError(), "upload tables"));
    }
    if ((rc = pic_seed_streams(c->seed_e, a->cap_e, a->rng_e[0], a->rng_e[1], a->rng_e[2], a->rng_e[3], a->stream)) ||
        (rc = pic_seed_streams(c->seed_i, a->cap_i, a->rng_i[0], a->rng_i[1], a->rng_i[2], a->rng_i[3], a->stream)))
        return fail(rc);
    if (hipStreamSynchronize(a->stream) != hipSuccess) return fail(hip_fail(hipGetLastError(), "seed"));
    *out = a;
    return PIC_OK;
}

int pic_adj_destroy(pic_adjoint_t *a) {
    if (!a) return PIC_OK;
    (void)hipStreamSynchronize(a->stream);
    free_all(a);
    delete a;
    return PIC_OK;
}

int pic_adj_upload(pic_adjoint_t *a, int species, const float *x, const float *vx, const float *vy,
                   const float *vz, uint32_t n) {
    if (!a || (species != PIC_ELECTRONS && species != PIC_IONS) || (n && (!x || !vx || !vy || !vz))) {
        set_error("pic_adj_upload: bad argument");
        return PIC_ERR_ARG;
    }
    const bool el = species == PIC_ELECTRONS;
    const uint32_t cap = el ? a->cap_e : a->cap_i;
    if (n > cap) {
        set_error("pic_adj_upload: more particles than the capacity");
        return PIC_ERR_ARG;
    }
    if (a->recorded) {
        set_error("pic_adj_upload: the event tape is already recorded (create a new object)");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(a->stream));
    float *dst[4] = {el ? a->x0e : a->x0i, el ? a->vx0e : a->vx0i, el ? a->vye : a->vyi, el ? a->vze : a->vzi};
    const float *src[4] = {x, vx, vy, vz};
    for (int w = 0; w < 4; ++w)
        if (n) PIC_HIP(hipMemcpy(dst[w], src[w], 4ull * n, hipMemcpyHostToDevice));
    std::vector<uint32_t> d(cap, kNever);
    for (uint32_t i = n; i < cap; ++i) d[i] = 0u;  // free slots: not alive
    PIC_HIP(hipMemcpy(el ? a->death0_e : a->death0_i, d.data(), 4ull * cap, hipMemcpyHostToDevice));
    (el ? a->n0e : a->n0i) = n;
    a->uploaded |= el ? 1u : 2u;
    return PIC_OK;
}

int pic_adj_set_target(pic_adjoint_t *a, const double *n_i_target) {
    if (!a || !n_i_target) return PIC_ERR_ARG;
    a->target.assign(n_i_target, n_i_target + a->p.n_g);
    return PIC_OK;
}

int pic_adj_record(pic_adjoint_t *a, double *J) {
    if (!a || a->uploaded != 3u || a->recorded) {
        set_error("pic_adj_record: upload both species once, then record once");
        return PIC_ERR_ARG;
    }
    return forward(a, true, J);
}

int pic_adj_forward(pic_adjoint_t *a, const double *wave_cos, const double *wave_sin, double *J,
                    double *nbar) {
    if (!a || !a->recorded) {
        set_error("pic_adj_forward: record the event tape first");
        return PIC_ERR_ARG;
    }
    if (wave_cos) memcpy(a->wave_cos, wave_cos, sizeof(a->wave_cos));
    if (wave_sin) memcpy(a->wave_sin, wave_sin, sizeof(a->wave_sin));
    TRY(forward(a, false, J));
    if (nbar) std::copy(a->nbar_h.begin(), a->nbar_h.end(), nbar);
    return PIC_OK;
}

int pic_adj_backward(pic_adjoint_t *a, double *grad_cos, double *grad_sin, double *dJ_dV) {
    if (!a || !a->have_forward || !a->recorded) {
        set_error("pic_adj_backward: run pic_adj_record or pic_adj_forward first");
        return PIC_ERR_ARG;
    }
    return backward(a, grad_cos, grad_sin, dJ_dV);
}

int pic_adj_tape_sizes(pic_adjoint_t *a, uint32_t *n_coll_e, uint32_t *n_coll_i, uint32_t *n_ionisations,
                       uint32_t *slots_e, uint32_t *slots_i) {
    if (!a || !a->recorded) return PIC_ERR_ARG;
    uint32_t c[AC_WORDS];
    PIC_HIP(hipStreamSynchronize(a->stream));
    PIC_HIP(hipMemcpy(c, a->cnt, sizeof(c), hipMemcpyDeviceToHost));
    if (n_coll_e) *n_coll_e = c[AC_COLL_E];
    if (n_coll_i) *n_coll_i = c[AC_COLL_I];
    if (n_ionisations) *n_ionisations = c[AC_ION];
    if (slots_e) *slots_e = c[AC_SLOTS_E];
    if (slots_i) *slots_i = c[AC_SLOTS_I];
    return PIC_OK;
}

int pic_adj_download_tape(pic_adjoint_t *a, uint32_t *off_e, uint32_t *slot_e, float *v_e, uint32_t *off_i,
                          uint32_t *slot_i, float *v_i, uint32_t *off_ion, uint32_t *parent, uint32_t *new_e,
                          uint32_t *new_i, float *v_new_e, float *v_new_i, uint32_t *slots_e, uint32_t *slots_i,
                          uint32_t *death_e, uint32_t *death_i) {
    if (!a || !a->recorded) return PIC_ERR_ARG;
    uint32_t ne = 0, ni = 0, nion = 0, se = 0, si = 0;
    TRY(pic_adj_tape_sizes(a, &ne, &ni, &nion, &se, &si));
    const size_t T1 = a->T + 1;
    struct {
        void *dst;
        const void *src;
        size_t bytes;
    } cp[] = {{off_e, a->tape.off_e, 4 * T1},     {slot_e, a->tape.slot_e, 4ull * ne}, {v_e, a->tape.v_e, 4ull * ne},
              {off_i, a->tape.off_i, 4 * T1},     {slot_i, a->tape.slot_i, 4ull * ni}, {v_i, a->tape.v_i, 4ull * ni},
              {off_ion, a->tape.off_ion, 4 * T1}, {parent, a->tape.par, 4ull * nion},  {new_e, a->tape.es, 4ull * nion},
              {new_i, a->tape.is, 4ull * nion},   {v_new_e, a->tape.ve, 4ull * nion},  {v_new_i, a->tape.vi, 4ull * nion},
              {slots_e, a->n_slots_e, 4 * T1},    {slots_i, a->n_slots_i, 4 * T1},     {death_e, a->death_e, 4ull * se},
              {death_i, a->death_i, 4ull * si}};
    for (auto &c : cp)
        if (c.dst && c.bytes) PIC_HIP(hipMemcpy(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost));
    return PIC_OK;
}

}  // extern "C"
