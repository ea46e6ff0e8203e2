// pic_internal.h -- launchers shared between the stage entry points
// (pic_kernels.hip) and the simulation object (pic_sim.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/cudapic_amd.h"
#include "pic_device.h"

namespace pic {

// thread-local error text behind pic_last_error()
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
#define PIC_HIP(call)                                      \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return ::pic::hip_fail(e_, #call); \
    } while (0)
#define PIC_LAUNCHED(what)                                         \
    do {                                                           \
        hipError_t e_ = hipGetLastError();                         \
        if (e_ != hipSuccess) return ::pic::hip_fail(e_, what);    \
    } while (0)

int device_cus();            // compute units of the current device (cached)
int device_lds_per_cu();     // LDS bytes per compute unit
int device_lds_per_block();  // LDS bytes one workgroup may allocate

KPhys make_kphys(const pic_params_t &p);
// Upper bound of an electron's collision probability in one step over the
// energies of the table (the speed at the top of each energy bin; the
// table's sigma_tot n_gas scaled by stot_scale): the threshold of the
// collision-candidate masks.  A particle above it (beyond the table's last
// bin) loads its RNG words on every step.
float coll_threshold(const float *stot_e, double stot_scale, const pic_params_t &p);

// Workgroup sizes the persistent particle kernels are built for (their
// launch bounds allow up to kMaxParticleBlock lanes; the default is 256).
constexpr int kMaxParticleBlock = 512;
inline bool particle_block_ok(int b) { return b == 128 || b == 256 || b == 512; }
// Persistent workgroups by default: 1024 lanes per compute unit (4 waves
// per SIMD, what the fused kernels' ~112 VGPRs allow resident).
inline int default_particle_grid(int cus, int block) { return cus * (1024 / block); }
// Default steps per run of deferred electron RNG write-backs: 8, or the
// largest power of two below it that divides 4*n_sub (a captured period
// must end with a write-back).
inline uint32_t default_rng_defer(int n_sub) {
    uint32_t d = 8;
    while (d > 1 && (4u * (uint32_t)n_sub) % d) d >>= 1;
    return d;
}

// Launch shape of the persistent particle kernels.
struct Shape {
    int block = 256;   // threads per workgroup (multiple of 64, <= 1024)
    int grid = 0;      // persistent workgroups (0 = 8 per compute unit)
    int ppt = 4;       // particles per thread per tile (compacting kernels)
    int ppt_inplace = 2;  // particles per thread per tile (in-place kernel)
    int replicas = 1;  // global density replicas the flush spreads over
    int rep_stride = 0;
    int lds_reps = 0;  // LDS histograms per workgroup at most (0 = one per wave as LDS allows)
};
Shape resolve(Shape s);

// Grid geometry and constants of one Poisson solve.
struct PoissonGeom {
    int n_g, per, reps, rep_stride;
    double e_charge, alpha, inv_dx, half_inv_dx, half_dx_over_eps;
};

// Poisson solve request (one workgroup of 1024 lanes).
struct PoissonJob {
    int precision = PIC_PREC_FP64;
    int n_g = 0;
    const void *e_dens = nullptr;  // float* (FP32/COMP) or double* (FP64)
    const void *i_dens = nullptr;
    int replicas = 1, rep_stride = 0;
    double pot0 = 0.0;                // used when time == nullptr
    const uint32_t *time = nullptr;   // device step index within the cycle
    // with time: pot0 of the step read from pot0_in when given (else computed
    // from the waveform), pot0 of the next step written to pot0_next
    const double *pot0_in = nullptr;
    double *pot0_next = nullptr;
    // with pot0_next: pot0 of every step t < n_t (launch_pot0_table), read
    // instead of evaluating the waveform inside the solve
    const double *pot0_tab = nullptr;
    // when given, the ion node densities (n_g doubles, summed over the
    // replicas of i_dens by an earlier solve or launch_sum_replicas) replace i_dens
    const double *i_sum = nullptr;
    // the summed node densities the solve used (n_g doubles each), optional
    double *dens_e_out = nullptr, *dens_i_out = nullptr;
    int n_t = 4000, n_sub = 20;
    double wave_cos[PIC_MAX_HARMONICS] = {}, wave_sin[PIC_MAX_HARMONICS] = {};
    double e_charge = 0, alpha = 0, inv_dx = 0, half_inv_dx = 0, half_dx_over_eps = 0;
    float *pot = nullptr, *efield = nullptr;
    // fused-path extras, all optional
    double *cumul_e = nullptr, *cumul_i = nullptr;  // measurement accumulation
    unsigned long long *meas_steps = nullptr;       // [2] electron / ion steps measured
    void *zero_a = nullptr; int zero_a_words = 0;   // 32-bit words to clear
    void *zero_b = nullptr; int zero_b_words = 0;   // cleared on ion steps only
    uint32_t *zero_cnt = nullptr; int zero_cnt_words = 0;          // cleared every step
    uint32_t *zero_cnt_ion = nullptr; int zero_cnt_ion_words = 0;  // cleared on ion steps
    const uint32_t *copy_cnt_src = nullptr; uint32_t *copy_cnt_dst = nullptr; int copy_cnt_words = 0;
    uint32_t *advance_time = nullptr;  // == time: *time <- (*time + 1) % n_t after the solve
    // ion steps only: *copy_cnt_ion_dst = *copy_cnt_ion_src as read before any clear
    const uint32_t *copy_cnt_ion_src = nullptr; uint32_t *copy_cnt_ion_dst = nullptr;
    // >= 0: the step index (instead of reading *time); with pot0_tab the
    // step's pot0 is pot0_tab[t_value] (pot0_in is ignored)
    int64_t t_value = -1;
    // XT maps (measurement steps, with cumul_e): pot, E, n_e, n_i of the step
    // added to [4][n_g][xt_n] at time bin t / xt_bin; xt_count[bin] += 1
    double *xt_grid = nullptr;
    unsigned long long *xt_count = nullptr;
    uint32_t xt_bin = 0;
    int xt_n = 0;
};
int launch_poisson(const PoissonJob &j, hipStream_t s);
// *out = pot0 of the step *time (the waveform the Poisson kernel evaluates)
int launch_pot0(const PoissonJob &j, double *out, hipStream_t s);
// tab[t] = pot0 of step t, t < j.n_t
int launch_pot0_table(const PoissonJob &j, double *tab, hipStream_t s);
// out[k] = node k of dens (j.precision's type, j.replicas / j.rep_stride)
// summed over the replicas exactly as the Poisson kernel sums it
int launch_sum_replicas(const PoissonJob &j, const void *dens, double *out, hipStream_t s);

// n_cap bounds the live count read from n_active (an overflowed counter can
// exceed the array it counts)
int launch_deposit(const float *x, const uint32_t *n_active, void *dens, bool f64,
                   const KPhys &k, const Shape &sh, hipStream_t s, uint32_t n_cap = 0xffffffffu);
int launch_fill(void *ptr, uint32_t value, uint64_t words, hipStream_t s);
// *dst = src ? min(*src, clamp) : 0
int launch_copy_u32(uint32_t *dst, const uint32_t *src, hipStream_t s, uint32_t clamp = 0xffffffffu);
// *d0 = min(*s0, clamp) and *d1 = min(*s1, clamp) in one launch
int launch_copy2_u32(uint32_t *d0, const uint32_t *s0, uint32_t *d1, const uint32_t *s1, hipStream_t s,
                     uint32_t clamp = 0xffffffffu);

struct MoveJob {
    const float *efield;
    const float *sx, *svx, *svy, *svz;
    float *dx, *dvx, *dvy, *dvz;
    uint32_t *alive;
    const uint32_t *n_active;
    float factor, dt, half_m_over_e;
    unsigned long long *abs_counts;  // [2] or null
    uint32_t *ifed;                  // [2*200] or null
    uint32_t n_cap;                  // bound on *n_active
};
int launch_move_compact(const MoveJob &j, const KPhys &k, const Shape &sh, hipStream_t s);

struct CollideEJob {
    const float *stot_e, *cs;
    const uint32_t *n_active;
    pic_species_t e, i;
    uint32_t *alive_e, *alive_i, *overflow;
    unsigned long long *n_coll;
};
int launch_collide_e(const CollideEJob &j, const KPhys &k, const Shape &sh, hipStream_t s);

struct CollideIJob {
    const float *stot_i, *cs;
    const uint32_t *n_active;
    pic_species_t i;
    unsigned long long *n_coll;
};
int launch_collide_i(const CollideIJob &j, const KPhys &k, const Shape &sh, hipStream_t s);

int launch_accum_density(const void *dens, bool f64, double *cumul, int n_g, hipStream_t s);
// XT maps of one species (C/eduPIC.cc:558-575, 603-614): for every live
// particle of a segmented array, with the mid-step velocity u = vx + 0.5
// factor E(x) and the CIC weights c of its two nodes, adds c, c u, c energy
// and (electrons) c sigma_ion(energy) |v| dt_e n_gas into xt[q][node][bin],
// bin = t / bin_steps for the step t the clock held before the Poisson solve
// of this step advanced it.  Electrons 4 quantities, ions 3.
struct XtJob {
    bool electrons;
    const float *x, *vx, *vy, *vz;
    const uint32_t *cnt;
    int n_seg;
    uint32_t seg_size;
    const float *efield;
    const float *sigma_ion;  // electrons: ionisation cross section [PIC_CS_RANGES]
    float gas_density;
    const uint32_t *clock;
    int n_t;
    uint32_t bin_steps;
    int n_xt;
    double *xt;
    float *stage;  // per-workgroup partial maps, xt_stage_floats(n_g) floats
};
// staging floats launch_xt needs for n_g nodes
size_t xt_stage_floats(int n_g);
int launch_xt(const XtJob &j, const KPhys &k, hipStream_t s);

int launch_eepf(const float *x, const float *vx, const float *vy, const float *vz,
                const float *efield, const uint32_t *n_active, uint32_t *eepf, const KPhys &k,
                const Shape &sh, hipStream_t s, uint32_t n_cap = 0xffffffffu);

// Fused particle step of the fast path (one kernel per species per step):
// push, survivor filter, wave64 compaction, MCC collision with the
// destination slot's RNG stream, and CIC deposit of the new positions into
// the next step's density (plus EEPF in measurement steps).  A species is
// stored as n_seg segments of seg_size slots, each dense with its own count.
constexpr int kMaxSegments = 64;
// Slots per species (and per ensemble) on the fused path: its kernels
// address the particle arrays by 32-bit byte offsets.
constexpr uint32_t kMaxSlots = 1u << 30;
// Measurement counters (pic_sim_read_measurements order) are kept in
// kCounterLanes copies of kCounterWords each; workgroup b of a particle
// kernel adds into copy b % kCounterLanes, so that thousands of workgroups
// ending together do not serialise on one address.  Readers sum the copies.
constexpr int kCounterWords = 8;
constexpr int kCounterLanes = 64;
struct FusedJob {
    bool electrons;
    bool inplace;                         // update in the occupied slots (sx.. are written)
    const float *efield;
    const float *sx, *svx, *svy, *svz;    // input segments
    const uint32_t *cnt_in;               // [n_seg]
    float *dx, *dvx, *dvy, *dvz;          // output segments
    uint32_t *cnt_out;                    // [n_seg], cleared before the launch
    int n_seg;
    uint32_t seg_size;
    uint32_t *rng0, *rng1, *rng2, *rng3;  // per slot of the output segments
    // electrons: ionisation products (new ions go to the current ion segments)
    float *ix, *ivx, *ivy, *ivz;
    uint32_t *ion_cnt;                    // [n_seg]
    uint32_t ion_seg_size;
    void *ion_dens;                       // next ion density (new ions deposit here)
    // densities
    void *dens;                           // next density of this species
    bool dens_f64;
    // tables and diagnostics
    const float *stot, *cs;
    uint32_t *overflow;
    unsigned long long *counters;         // [kCounterLanes][kCounterWords]
    uint32_t *eepf;                       // measurement steps only
    uint32_t *ifed;                       // ions only
    // 32-bit words the workgroups clear between them while they run (the
    // density buffer the step's Poisson solve has read; optional)
    uint32_t *zero;
    uint32_t zero_words;
    // Deferred RNG write-back (in-place electron steps only; 0 and 0 on
    // every other launch): on entry the RNG words of each live slot are
    // rng_lag draws behind its stream (holes and slots past the watermark
    // are exact); the launch leaves them rng_next_lag draws behind (0: every
    // slot exact).  A live particle that does not collide then writes no
    // RNG words; a colliding one and a newborn store their state rewound by
    // rng_next_lag draws.
    uint32_t rng_lag, rng_next_lag;
    // Collision-candidate masks of a deferred run (in-place electron steps,
    // not on measurement steps): one byte per slot, bit L set when the
    // slot's draw at entry lag L of the run may fall below coll_thr (an
    // upper bound of the collision probability; a step whose probability
    // exceeds it loads the words regardless).  mask_read: the run's masks
    // are valid, and a live slot whose bit for this step's lag is clear
    // neither loads nor advances its RNG words.  mask_write (rng_next_lag
    // == 0): every slot writes the masks of the next run from its exact
    // words.  Colliding particles rewrite their masks from their new state,
    // newborn electrons get all bits set.
    uint8_t *mask;
    bool mask_read, mask_write;
    float coll_thr;
};
int launch_fused(const FusedJob &j, const KPhys &k, const Shape &sh, hipStream_t s);
// Largest lag rng_lag / rng_next_lag may take.
constexpr uint32_t kMaxRngLag = 64;
// *out += number of x >= 0 in the first cnt[q] slots of each segment q
int launch_count_live(const float *x, const uint32_t *cnt, int n_seg, uint32_t seg_size, uint32_t *out,
                      hipStream_t s);

// ---------------------------------------------------------------- ensembles
// One independent discharge of an ensemble: its physics, its driving
// amplitude and the regions it owns in the pooled particle arrays.
// Instances share the grid, the time step and the masses, so they differ in
// KPhys only by the weight factor wf, the gas density (stot_scale, the
// summed cross-section table is shared) and the thermal velocity spread of
// the gas atoms (normal_range).
struct EnsInst {
    KPhys k;
    double v0;
    float coll_thr;         // collision-candidate mask threshold (the instance's largest probability)
    uint32_t e_off, e_cap;  // electrons (and their RNG slots)
    uint32_t i_off, i_cap;  // ions (and their RNG slots)
};

// per-instance device counters (uint32 words): the electron watermark of
// the in-place steps (two parities: the step reads one and appends newborns
// after a copy of it in the other), the ion watermark of the two ion
// buffers, the compaction pass's output count and the overflow flag
enum EnsCnt { EC_E0 = 0, EC_I0 = 2, EC_C = 4, EC_OVERFLOW = 5, EC_WORDS = 8 };

struct EnsPoissonJob {
    int precision;
    PoissonGeom g;               // reps / rep_stride of one instance's density buffer
    const EnsInst *inst;
    int n_inst;
    const void *e_dens, *i_dens; // [n_inst][inst_stride]
    int inst_stride;             // entries per instance
    const uint32_t *time;
    int n_t, n_sub;
    float *pot, *efield;         // [n_inst][n_g]
    double *cumul_e, *cumul_i;   // [n_inst][n_g] or null
    unsigned long long *counters;  // [n_inst][8]: [6],[7] measured steps
    void *zero_a, *zero_b;       // next accumulators [n_inst][inst_stride] (zero_b on ion steps)
    uint32_t *cnt;               // [n_inst][EC_WORDS]
    // counter words: c[copy_dst] = c[copy_src] (when copy_dst >= 0), then
    // c[zc0] = 0 and c[zc1] = 0 (each when >= 0)
    int copy_src, copy_dst, zc0, zc1;
    // optional: pot0 of the step per instance (read; the next step's written,
    // from trig, the cos / sin table of launch_ens_pot0_init); without it the
    // kernel evaluates the waveform
    double *pot0;
    const double *trig;
    // optional: summed ion densities [n_inst][n_g], read instead of i_dens
    // (steps without ions) or written (ion steps)
    const double *i_sum_in;
    double *i_sum_out;
};
int launch_poisson_ens(const EnsPoissonJob &j, hipStream_t s);
// trig = the phase table of n_t steps; pot0[b] = v0_b cos(2 pi t / n_t)
int launch_ens_pot0_init(const EnsInst *inst, int n_inst, int n_t, uint32_t t, double *trig, double *pot0,
                         hipStream_t s);
// out[b][k] = node k of instance b's density (replicas summed as the Poisson kernel sums them)
int launch_sum_replicas_ens(const PoissonGeom &g, int precision, const void *dens, int inst_stride, int n_inst,
                            double *out, hipStream_t s);

// In-place particle step of every instance in one launch (the ensemble
// counterpart of the in-place FusedJob): pooled arrays updated in their
// slots, lost particles leave holes, newborn electrons and ions appended
// after each instance's watermark.
struct EnsStepJob {
    bool electrons;
    const EnsInst *inst;
    int n_inst;
    KPhys k;                               // shared physics (instance 0)
    const float *efield;                   // [n_inst][n_g]
    float *x, *vx, *vy, *vz;               // pooled particles of this species
    uint32_t *rng0, *rng1, *rng2, *rng3;   // pooled per-slot streams of this species
    float *ix, *ivx, *ivy, *ivz;           // pooled ions (electron step: appends)
    uint32_t *cnt;                         // [n_inst][EC_WORDS]
    int c_in, c_out, c_ion;                // counter words (c_out, c_ion: electron step)
    void *dens, *ion_dens;                 // [n_inst][inst_stride]
    int inst_stride;
    bool dens_f64;
    const float *stot, *cs;                // stot: summed cross sections (scaled per instance)
    unsigned long long *counters;          // [n_inst][8]
    uint32_t *time;                        // advanced by one step (electron step)
    int n_t;
    uint32_t *zero;                        // cleared while the kernel runs (optional)
    uint32_t zero_words;
    // deferred RNG write-back of the electrons, as FusedJob (0 and 0 for ions)
    uint32_t rng_lag, rng_next_lag;
    // collision-candidate masks of the pooled electron slots (null: none),
    // read on the steps of a deferred run or written by its write-back step;
    // thresholds per instance (EnsInst::coll_thr)
    uint8_t *mask;
    bool mask_read, mask_write;
};
int launch_inplace_ens(const EnsStepJob &j, const Shape &sh, hipStream_t s);

// Compaction pass: the live particles (x >= 0) of each instance's first
// c[c_in] slots of src are packed into the same region of dst, counted in
// c[c_out] (cleared before the launch).  RNG streams stay with the slots.
struct EnsCompactJob {
    bool electrons;
    const EnsInst *inst;
    int n_inst;
    const float *src[4];
    float *dst[4];
    uint32_t *cnt;
    int c_in, c_out;
};
int launch_compact_ens(const EnsCompactJob &j, hipStream_t s);
// out[i] += live particles (x >= 0) in the first c[word] slots of instance i
int launch_count_live_ens(const float *x, const EnsInst *inst, int n_inst, const uint32_t *cnt, int word,
                          bool electrons, uint32_t *out, hipStream_t s);

}  // namespace pic
