// pic_sim.hip -- the simulation object: the device state the reference keeps
// in GpuSimState (Rust-CUDA/src/main.rs:132-350) and the RF step loop of
// main.rs:1723-1828, on one HIP stream, with no host synchronisation inside
// a step.
//
// Two step paths share the state layout:
//  * PIC_PATH_REFERENCE runs the reference's stage order and counter
//    protocol with one kernel per stage (deposit e, [deposit i], Poisson,
//    push+compact e, [push+compact i], collisions e, [collisions i]).
//  * PIC_PATH_FUSED runs Poisson (which also clears the next step's
//    accumulators and counters and advances the device clock) and one fused
//    particle kernel per species.  Its launch sequence repeats every
//    4*N_SUB steps with identical arguments, so it is captured once into a
//    hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pic_internal.h"

using namespace pic;

namespace {
// device counter block (uint32 words): live counts per [parity][segment]
enum Cnt {
    C_E = 0,                        // electrons [2][kMaxSegments]
    C_I = 2 * kMaxSegments,         // ions [2][kMaxSegments]
    C_ALIVE = 4 * kMaxSegments,     // reference path scratch counter
    C_OVERFLOW = C_ALIVE + 1,
    C_TIME = C_ALIVE + 2,           // step index within the cycle (fused path clock)
    C_P0 = C_ALIVE + 4,             // pot0 of that step (double, 8-byte aligned)
    C_WORDS = C_ALIVE + 16
};
// Default segment count: one per 128 Ki slots, at most 32.  Segments spread
// the slot reservations of millions of particles over separate counters; a
// compaction sends each tile's survivors to one segment, so a segment must
// hold its share of tiles plus one tile, which small capacities cannot
// afford with many segments.
constexpr int kMeasCounters = kCounterWords;
constexpr uint32_t kCkptMagic = 0x43504943u;  // "CIPC"
constexpr uint32_t kCkptVersion = 2;
}  // namespace

struct pic_sim {
    pic_sim_config_t cfg;
    pic_params_t p;
    KPhys k;
    Shape sh;
    hipStream_t stream = nullptr;
    int prec = PIC_PREC_FP32;
    bool f64 = false;
    uint32_t cap = 0;                 // MAX_PARTICLES per species
    int nseg = 1;                     // segments per species (1 on the reference path)
    uint32_t seg = 0;                 // slots per segment
    uint32_t slots = 0;               // nseg * seg
    // particles: [parity][x, vx, vy, vz], nseg segments of seg slots each
    float *e[2][4] = {}, *ion[2][4] = {};
    uint32_t *e_rng[4] = {}, *i_rng[4] = {};
    uint32_t *cnt = nullptr;
    // grid
    float *pot = nullptr, *efield = nullptr;
    double *grid_e = nullptr, *grid_i = nullptr;  // node densities of the last solve
    double *pot0_tab = nullptr;                     // pot0 of every step of the period (n_t)
    void *e_dens[2] = {}, *i_dens[2] = {};  // replicas * stride entries of 8 bytes
    int dens_words = 0;                      // entries per buffer (replicas * stride)
    // measurement
    double *cumul_e = nullptr, *cumul_i = nullptr;
    uint32_t *eepf = nullptr, *ifed = nullptr;
    unsigned long long *counters = nullptr;
    // XT maps (fused path, measurement steps): kXtPlanes raw planes of
    // [n_g][xt_n] -- pot, E, n_e, n_i of the grid; electron weight, weighted
    // mid-step velocity, energy and ionisation rate; ion weight, velocity,
    // energy -- and the measured steps of each time bin
    double *xt = nullptr;
    unsigned long long *xt_count = nullptr;
    float *xt_stage = nullptr;  // per-workgroup partial maps of the particle pass
    uint32_t xt_bin = 0;
    int xt_n = 0;
    // tables
    float *cs = nullptr, *stot_e = nullptr, *stot_i = nullptr;
    // bookkeeping
    int ce = 0, ci = 0;      // current particle-array parity (electrons, ions)
    int pe = 0;              // electron counter parity (flips every fused step)
    int cd = 0, hi = 0;      // fused: current e density, held ion density
    uint32_t t = 0;          // next step's index in the cycle (host mirror)
    unsigned long long steps = 0;
    bool need_prologue = true;
    unsigned long long ref_meas[2] = {0, 0};
    // fused path: steps between compactions (a multiple of 4*N_SUB)
    uint32_t compact_every = 800;
    // fused path: deferred electron RNG write-back -- steps per run (divides
    // 4*N_SUB) and the draws the stored words of live slots are behind at
    // the next step (0 between calls)
    uint32_t rng_defer = 8;
    uint32_t e_lag = 0;
    // collision-candidate masks of the deferred runs (a byte per electron
    // slot, null when off): mask_ok while the masks hold the current run's
    // candidates (written by the last write-back step, cleared by anything
    // else that changes the electrons or their tables)
    uint8_t *e_mask = nullptr;
    bool mask_ok = false;
    float coll_thr = 1.0f;
    // captured runs of steps inside one 4*N_SUB-step period (whole periods,
    // and the partial ones a timed window starts or ends with), keyed by the
    // measure flag, whether the run starts with a compaction, the roles at
    // its start, and its offset in the period and length; each remembers the
    // roles it leaves behind
    static constexpr int kGraphSlots = 16;
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        int key[10] = {};
        int after[5] = {};
        bool mask_after = false;
    } graphs[kGraphSlots];
    int graph_next = 0;  // slot the next capture replaces when all are taken
    hipEvent_t ev[2] = {nullptr, nullptr};
};

namespace {

template <typename T>
int dalloc(T **p, size_t n) {
    PIC_HIP(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    PIC_HIP(hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return PIC_OK;
}

#define TRY(x)                     \
    do {                           \
        int rc_ = (x);             \
        if (rc_ != PIC_OK) return rc_; \
    } while (0)

uint32_t *ecnt(pic_sim *s, int parity) { return s->cnt + C_E + parity * kMaxSegments; }
uint32_t *icnt(pic_sim *s, int parity) { return s->cnt + C_I + parity * kMaxSegments; }

pic_species_t species(pic_sim *s, int which, int parity) {
    pic_species_t sp;
    float **a = which == PIC_ELECTRONS ? s->e[parity] : s->ion[parity];
    uint32_t **r = which == PIC_ELECTRONS ? s->e_rng : s->i_rng;
    sp.x = a[0];
    sp.vx = a[1];
    sp.vy = a[2];
    sp.vz = a[3];
    sp.rng0 = r[0];
    sp.rng1 = r[1];
    sp.rng2 = r[2];
    sp.rng3 = r[3];
    sp.capacity = s->slots;
    return sp;
}

PoissonJob base_poisson(pic_sim *s) {
    PoissonJob j;
    const pic_params_t &p = s->p;
    const double E = 1.60217662e-19, EPS0 = 8.85418781e-12;
    j.precision = s->prec;
    j.n_g = p.n_g;
    j.n_t = p.n_t;
    j.n_sub = p.n_sub;
    j.e_charge = E;
    j.alpha = -p.dx * p.dx / EPS0;
    j.inv_dx = p.inv_dx;
    j.half_inv_dx = 0.5 * p.inv_dx;
    j.half_dx_over_eps = p.dx / (2.0 * EPS0);
    for (int h = 0; h < PIC_MAX_HARMONICS; ++h) {
        j.wave_cos[h] = p.wave_cos[h];
        j.wave_sin[h] = p.wave_sin[h];
    }
    j.pot = s->pot;
    j.efield = s->efield;
    j.dens_e_out = s->grid_e;
    j.dens_i_out = s->grid_i;
    j.replicas = s->sh.replicas;
    j.rep_stride = s->sh.rep_stride;
    return j;
}

// pot0 of step t as the host computes it in main.rs:1741, for any waveform.
double host_pot0(const pic_params_t &p, uint32_t t) {
    double v = 0.0;
    for (int h = 0; h < PIC_MAX_HARMONICS; ++h) {
        if (p.wave_cos[h] == 0.0 && p.wave_sin[h] == 0.0) continue;
        const double ph = ((double)t / (double)p.n_t) * 2.0 * 3.141592653589793 * (double)(h + 1);
        v += p.wave_cos[h] * std::cos(ph) + p.wave_sin[h] * std::sin(ph);
    }
    return v;
}

// ---------------------------------------------------------------- reference path
int step_reference(pic_sim *s, bool measure) {
    hipStream_t st = s->stream;
    const bool ion = (s->t % (uint32_t)s->p.n_sub) == 0;
    // the ion density is refreshed on ion steps; after an upload, a load or
    // a clock change it is deposited on the first step whatever its index
    const bool dep_i = ion || s->need_prologue;
    s->need_prologue = false;
    uint32_t *ne = ecnt(s, s->pe);
    uint32_t *ni = icnt(s, s->ci);
    uint32_t *alive = s->cnt + C_ALIVE;
    Shape one = s->sh;
    one.replicas = 1;
    // main.rs:1729-1739
    TRY(launch_fill(s->e_dens[0], 0u, (uint64_t)s->p.n_g * (s->f64 ? 2 : 1), st));
    TRY(launch_deposit(s->e[s->ce][0], ne, s->e_dens[0], s->f64, s->k, one, st, s->cap));
    if (dep_i) {
        TRY(launch_fill(s->i_dens[0], 0u, (uint64_t)s->p.n_g * (s->f64 ? 2 : 1), st));
        TRY(launch_deposit(s->ion[s->ci][0], ni, s->i_dens[0], s->f64, s->k, one, st, s->cap));
    }
    // main.rs:1741-1744
    PoissonJob pj = base_poisson(s);
    pj.replicas = 1;
    pj.rep_stride = s->p.n_g;
    pj.e_dens = s->e_dens[0];
    pj.i_dens = s->i_dens[0];
    const double p0 = host_pot0(s->p, s->t);
    pj.pot0 = s->prec == PIC_PREC_FP64 ? p0 : (double)(float)p0;
    // the alive counter of the electron compaction is cleared by the solve
    // (main.rs:1758), which runs just before it: one launch fewer per step
    pj.zero_cnt = alive;
    pj.zero_cnt_words = 1;
    TRY(launch_poisson(pj, st));
    // main.rs:1746-1756
    if (measure) {
        TRY(launch_accum_density(s->e_dens[0], s->f64, s->cumul_e, s->p.n_g, st));
        s->ref_meas[0]++;
        if (ion) {
            TRY(launch_accum_density(s->i_dens[0], s->f64, s->cumul_i, s->p.n_g, st));
            s->ref_meas[1]++;
        }
    }
    // main.rs:1758-1768: electrons (alive cleared by the solve above)
    MoveJob mj{s->efield, s->e[s->ce][0], s->e[s->ce][1], s->e[s->ce][2], s->e[s->ce][3],
               s->e[1 - s->ce][0], s->e[1 - s->ce][1], s->e[1 - s->ce][2], s->e[1 - s->ce][3],
               alive, ne, s->k.factor_e, s->k.dt_e, 0.0f, s->counters + 0, nullptr, s->cap};
    TRY(launch_move_compact(mj, s->k, s->sh, st));
    s->ce ^= 1;
    s->pe ^= 1;
    ne = ecnt(s, s->pe);
    TRY(launch_copy_u32(ne, alive, st));
    // main.rs:1770-1783: ions
    if (ion) {
        TRY(launch_copy_u32(alive, nullptr, st));
        MoveJob mi{s->efield, s->ion[s->ci][0], s->ion[s->ci][1], s->ion[s->ci][2], s->ion[s->ci][3],
                   s->ion[1 - s->ci][0], s->ion[1 - s->ci][1], s->ion[1 - s->ci][2],
                   s->ion[1 - s->ci][3], alive, ni, s->k.factor_i, s->k.dt_i, s->k.half_mi_over_e,
                   s->counters + 2, s->ifed, s->cap};
        TRY(launch_move_compact(mi, s->k, s->sh, st));
        s->ci ^= 1;
        ni = icnt(s, s->ci);
        TRY(launch_copy2_u32(ni, alive, alive, ne, st, s->cap));
    }
    // main.rs:1785-1792
    CollideEJob cj{s->stot_e, s->cs, ne, species(s, PIC_ELECTRONS, s->ce), species(s, PIC_IONS, s->ci),
                   alive, ni, s->cnt + C_OVERFLOW, s->counters + 4};
    TRY(launch_collide_e(cj, s->k, s->sh, st));
    // appends past the capacity were dropped (and flagged): clamp the counts
    TRY(launch_copy2_u32(ne, alive, ni, ni, st, s->cap));
    // main.rs:1794-1800
    if (ion) {
        CollideIJob ij{s->stot_i, s->cs, ni, species(s, PIC_IONS, s->ci), s->counters + 5};
        TRY(launch_collide_i(ij, s->k, s->sh, st));
    }
    // main.rs:1802-1807
    if (measure) {
        float **a = s->e[s->ce];
        TRY(launch_eepf(a[0], a[1], a[2], a[3], s->efield, ne, s->eepf, s->k, s->sh, st, s->cap));
    }
    return PIC_OK;
}

// ---------------------------------------------------------------- fused path
constexpr int kXtPlanes = 11;

uint64_t dens_words32(pic_sim *s) { return (uint64_t)s->dens_words * (s->f64 ? 2 : 1); }

// Deposit the current particles into the accumulators the next Poisson
// solve reads (after an upload, a checkpoint load or a path start).
int fused_prologue(pic_sim *s) {
    hipStream_t st = s->stream;
    Shape sh = s->sh;
    s->mask_ok = false;  // the particles changed under the masks
    TRY(launch_fill(s->e_dens[s->cd], 0u, dens_words32(s), st));
    TRY(launch_fill(s->e_dens[1 - s->cd], 0u, dens_words32(s), st));  // the first step deposits here
    TRY(launch_fill(s->i_dens[0], 0u, dens_words32(s), st));
    TRY(launch_fill(s->i_dens[1], 0u, dens_words32(s), st));
    for (int q = 0; q < s->nseg; ++q) {
        const size_t o = (size_t)q * s->seg;
        TRY(launch_deposit(s->e[s->ce][0] + o, ecnt(s, s->pe) + q, s->e_dens[s->cd], s->f64, s->k, sh, st, s->seg));
        TRY(launch_deposit(s->ion[s->ci][0] + o, icnt(s, s->ci) + q, s->i_dens[0], s->f64, s->k, sh, st, s->seg));
        TRY(launch_deposit(s->ion[s->ci][0] + o, icnt(s, s->ci) + q, s->i_dens[1], s->f64, s->k, sh, st, s->seg));
    }
    // the device clock follows the host's; the pot0 of its step is computed
    // by the device waveform (each Poisson solve then reads the next one from
    // the table of the period); the held ion density is summed over its
    // replicas into grid_i, which the solves of steps without ions read
    PIC_HIP(hipMemcpyAsync(s->cnt + C_TIME, &s->t, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    {
        PoissonJob pj = base_poisson(s);
        pj.time = s->cnt + C_TIME;
        TRY(launch_pot0(pj, (double *)(s->cnt + C_P0), st));
        TRY(launch_pot0_table(pj, s->pot0_tab, st));
        TRY(launch_sum_replicas(pj, s->i_dens[s->hi], s->grid_i, st));
    }
    PIC_HIP(hipStreamSynchronize(st));
    s->need_prologue = false;
    return PIC_OK;
}

// One fused step.  Every compact_every steps the particle kernels compact (src ->
// dst arrays, holes squeezed out, survivors take the RNG stream of their new
// slot); the other steps update the particles in place.  The electron live
// counters are double-buffered every step (the in-place kernel appends
// newborns after a frozen copy of the watermark).  The electrons' RNG words
// are written back in runs of rng_defer steps aligned to multiples of
// rng_defer in the cycle; the step before a compaction, the last step of the
// cycle and the last step of a call (`last`) write them all back, so that
// every call and every captured period starts and ends with exact words.
int step_fused(pic_sim *s, bool measure, bool last, bool time_kernel, hipEvent_t e0, hipEvent_t e1) {
    hipStream_t st = s->stream;
    const uint32_t t = s->t;
    const bool ion = (t % (uint32_t)s->p.n_sub) == 0;
    const bool comp_e = (t % s->compact_every) == 0;
    const bool comp_i = ion && comp_e;
    const uint32_t lag = s->e_lag;
    const uint32_t lag_out = (last || comp_e || (t + 1u) % s->rng_defer == 0 || t + 1u == (uint32_t)s->p.n_t)
                                 ? 0u
                                 : lag + 1u;
    if (comp_e && lag) {
        set_error("pic_sim: internal: a compacting step starts behind on RNG words");
        return PIC_ERR_ARG;
    }
    // candidate masks: read inside a run whose masks are valid, written by a
    // write-back step whose next run has steps that can use them (the next
    // step neither compacts nor writes back)
    const bool masks_on = s->e_mask && !comp_e && !measure;
    const bool mask_read = masks_on && lag_out != 0u && s->mask_ok;
    bool mask_write = false;
    if (masks_on && lag_out == 0u) {
        const uint32_t t1 = (t + 1u) % (uint32_t)s->p.n_t;
        mask_write = t1 % s->compact_every != 0 && (t1 + 1u) % s->rng_defer != 0 && t1 + 1u != (uint32_t)s->p.n_t;
    }
    // Poisson of this step: e density accumulated by the previous step,
    // ion density = the held one, or on ion steps the accumulated one
    PoissonJob pj = base_poisson(s);
    pj.e_dens = s->e_dens[s->cd];
    const int i_read = ion ? 1 - s->hi : s->hi;
    pj.i_dens = s->i_dens[i_read];
    pj.time = s->cnt + C_TIME;
    pj.advance_time = s->cnt + C_TIME;
    pj.pot0_in = (const double *)(s->cnt + C_P0);
    pj.pot0_next = (double *)(s->cnt + C_P0);
    pj.pot0_tab = s->pot0_tab;
    // between ion steps the held ion density is the one the last ion step's
    // solve summed into grid_i
    if (!ion) pj.i_sum = s->grid_i;
    if (ion) {
        pj.zero_b = s->i_dens[s->hi];  // old held buffer becomes the accumulator
        pj.zero_b_words = (int)dens_words32(s);
    }
    if (comp_e) {
        pj.zero_cnt = ecnt(s, 1 - s->pe);
        pj.zero_cnt_words = s->nseg;
    } else {
        pj.copy_cnt_src = ecnt(s, s->pe);
        pj.copy_cnt_dst = ecnt(s, 1 - s->pe);
        pj.copy_cnt_words = s->nseg;
    }
    if (comp_i) {
        pj.zero_cnt_ion = icnt(s, 1 - s->ci);
        pj.zero_cnt_ion_words = s->nseg;
    }
    if (measure) {
        pj.cumul_e = s->cumul_e;
        pj.cumul_i = s->cumul_i;
        pj.meas_steps = s->counters + 6;
        if (s->xt) {
            pj.xt_grid = s->xt;
            pj.xt_count = s->xt_count;
            pj.xt_bin = s->xt_bin;
            pj.xt_n = s->xt_n;
        }
    }
    TRY(launch_poisson(pj, st));
    if (measure && s->xt) {
        // particle XT quantities from the positions and velocities before
        // this step's push, with this step's field (C/eduPIC.cc:558-575, 603-614)
        const size_t plane = (size_t)s->p.n_g * s->xt_n;
        XtJob xj;
        memset(&xj, 0, sizeof(xj));
        xj.n_seg = s->nseg;
        xj.seg_size = s->seg;
        xj.efield = s->efield;
        xj.sigma_ion = s->cs + 2 * (size_t)PIC_CS_RANGES;
        xj.gas_density = (float)s->p.gas_density;
        xj.clock = s->cnt + C_TIME;
        xj.n_t = s->p.n_t;
        xj.bin_steps = s->xt_bin;
        xj.n_xt = s->xt_n;
        xj.stage = s->xt_stage;
        if (ion) {
            XtJob xi = xj;
            xi.electrons = false;
            xi.x = s->ion[s->ci][0];
            xi.vx = s->ion[s->ci][1];
            xi.vy = s->ion[s->ci][2];
            xi.vz = s->ion[s->ci][3];
            xi.cnt = icnt(s, s->ci);
            xi.xt = s->xt + 8 * plane;
            TRY(launch_xt(xi, s->k, st));
        }
        xj.electrons = true;
        xj.x = s->e[s->ce][0];
        xj.vx = s->e[s->ce][1];
        xj.vy = s->e[s->ce][2];
        xj.vz = s->e[s->ce][3];
        xj.cnt = ecnt(s, s->pe);
        xj.xt = s->xt + 4 * plane;
        TRY(launch_xt(xj, s->k, st));
    }
    const int i_acc = ion ? s->hi : 1 - s->hi;  // accumulator for the next ion step
    FusedJob base;
    memset(&base, 0, sizeof(base));
    base.efield = s->efield;
    base.n_seg = s->nseg;
    base.seg_size = s->seg;
    base.dens_f64 = s->f64;
    base.cs = s->cs;
    base.overflow = s->cnt + C_OVERFLOW;
    base.counters = s->counters;
    if (ion) {
        FusedJob fi = base;
        fi.electrons = false;
        fi.inplace = !comp_i;
        fi.sx = s->ion[s->ci][0];
        fi.svx = s->ion[s->ci][1];
        fi.svy = s->ion[s->ci][2];
        fi.svz = s->ion[s->ci][3];
        fi.cnt_in = icnt(s, s->ci);
        fi.dx = s->ion[1 - s->ci][0];
        fi.dvx = s->ion[1 - s->ci][1];
        fi.dvy = s->ion[1 - s->ci][2];
        fi.dvz = s->ion[1 - s->ci][3];
        fi.cnt_out = icnt(s, 1 - s->ci);
        fi.rng0 = s->i_rng[0];
        fi.rng1 = s->i_rng[1];
        fi.rng2 = s->i_rng[2];
        fi.rng3 = s->i_rng[3];
        fi.dens = s->i_dens[i_acc];
        fi.stot = s->stot_i;
        fi.ifed = s->ifed;
        TRY(launch_fused(fi, s->k, s->sh, st));
        if (comp_i) s->ci ^= 1;
    }
    FusedJob fe = base;
    fe.electrons = true;
    fe.inplace = !comp_e;
    fe.sx = s->e[s->ce][0];
    fe.svx = s->e[s->ce][1];
    fe.svy = s->e[s->ce][2];
    fe.svz = s->e[s->ce][3];
    fe.cnt_in = ecnt(s, s->pe);
    fe.dx = s->e[1 - s->ce][0];
    fe.dvx = s->e[1 - s->ce][1];
    fe.dvy = s->e[1 - s->ce][2];
    fe.dvz = s->e[1 - s->ce][3];
    fe.cnt_out = ecnt(s, 1 - s->pe);
    fe.rng0 = s->e_rng[0];
    fe.rng1 = s->e_rng[1];
    fe.rng2 = s->e_rng[2];
    fe.rng3 = s->e_rng[3];
    fe.ix = s->ion[s->ci][0];
    fe.ivx = s->ion[s->ci][1];
    fe.ivy = s->ion[s->ci][2];
    fe.ivz = s->ion[s->ci][3];
    fe.ion_cnt = icnt(s, s->ci);
    fe.ion_seg_size = s->seg;
    fe.ion_dens = s->i_dens[i_acc];
    fe.dens = s->e_dens[1 - s->cd];
    fe.stot = s->stot_e;
    fe.eepf = measure ? s->eepf : nullptr;
    // the density this step's solve read is the next step's accumulator
    fe.zero = (uint32_t *)s->e_dens[s->cd];
    fe.zero_words = (uint32_t)dens_words32(s);
    fe.rng_lag = lag;
    fe.rng_next_lag = lag_out;
    fe.mask = s->e_mask;
    fe.mask_read = mask_read;
    fe.mask_write = mask_write;
    fe.coll_thr = s->coll_thr;
    if (time_kernel) PIC_HIP(hipEventRecord(e0, st));
    TRY(launch_fused(fe, s->k, s->sh, st));
    if (time_kernel) PIC_HIP(hipEventRecord(e1, st));
    s->e_lag = lag_out;
    // a step that leaves a run's masks stale (a compaction, a measured step,
    // a step inside an unmasked run) leaves them off until the next write
    s->mask_ok = lag_out == 0u ? mask_write : mask_read;
    if (comp_e) s->ce ^= 1;
    s->pe ^= 1;
    s->cd ^= 1;
    if (ion) s->hi ^= 1;
    return PIC_OK;
}

void invalidate_graphs(pic_sim *s) {
    for (auto &g : s->graphs)
        if (g.exec) {
            (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
        }
}

int one_step(pic_sim *s, bool measure, bool last) {
    int rc = s->cfg.path == PIC_PATH_FUSED ? step_fused(s, measure, last, false, nullptr, nullptr)
                                           : step_reference(s, measure);
    if (rc != PIC_OK) return rc;
    s->t = (s->t + 1) % (uint32_t)s->p.n_t;
    s->steps++;
    return PIC_OK;
}

void get_roles(const pic_sim *s, int *r) {
    r[0] = s->ce;
    r[1] = s->ci;
    r[2] = s->cd;
    r[3] = s->hi;
    r[4] = s->pe;
}

void set_roles(pic_sim *s, const int *r) {
    s->ce = r[0];
    s->ci = r[1];
    s->cd = r[2];
    s->hi = r[3];
    s->pe = r[4];
}

// The graph of `len` fused steps from now, inside one 4*N_SUB-step period
// ((t % 4*N_SUB) + len <= 4*N_SUB): found in the cache by (measure,
// compaction at the first step, roles, offset in the period, len) or, when
// `capture` is set, captured (otherwise *out = nullptr).  A period holds at
// most one compaction (at its first step), so over a whole period the
// density, counter and ion-hold parities return to their start values and
// only the array parities can differ at its end.  A run starts with exact
// RNG words (it starts a call or a period) and its last step writes them
// back (it ends a period, where the write-back runs are aligned, or ends
// the call).
int graph_for(pic_sim *s, bool measure, uint32_t len, bool capture, pic_sim::GraphSlot **out) {
    const uint32_t period = 4u * (uint32_t)s->p.n_sub;
    int key[10];
    key[0] = measure ? 1 : 0;
    key[1] = (s->t % s->compact_every) == 0 ? 1 : 0;
    get_roles(s, key + 2);
    key[7] = (int)(s->t % period);
    key[8] = (int)len;
    key[9] = s->mask_ok ? 1 : 0;  // the run's first steps read the masks
    *out = nullptr;
    for (auto &g : s->graphs)
        if (g.exec && memcmp(g.key, key, sizeof(key)) == 0) {
            *out = &g;
            return PIC_OK;
        }
    if (!capture) return PIC_OK;
    const uint32_t t0 = s->t;
    const unsigned long long steps0 = s->steps;
    if (s->e_lag) {
        set_error("pic_sim: internal: a captured run starts behind on RNG words");
        return PIC_ERR_ARG;
    }
    hipGraph_t g = nullptr;
    PIC_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = PIC_OK;
    for (uint32_t i = 0; i < len && rc == PIC_OK; ++i) rc = one_step(s, measure, i + 1 == len);
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    // capture does not execute: remember the roles the run leaves, then
    // restore the host mirror (a run starts and ends with exact RNG words)
    int after[5];
    get_roles(s, after);
    const bool mask_after = s->mask_ok;
    s->t = t0;
    s->steps = steps0;
    s->e_lag = 0;
    s->mask_ok = key[9] != 0;
    set_roles(s, key + 2);
    if (rc != PIC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    pic_sim::GraphSlot *slot = nullptr;
    for (auto &gs : s->graphs)
        if (!gs.exec) {
            slot = &gs;
            break;
        }
    if (!slot) {
        slot = &s->graphs[s->graph_next];
        s->graph_next = (s->graph_next + 1) % pic_sim::kGraphSlots;
        (void)hipGraphExecDestroy(slot->exec);
    }
    slot->exec = ex;
    memcpy(slot->key, key, sizeof(key));
    memcpy(slot->after, after, sizeof(after));
    slot->mask_after = mask_after;
    *out = slot;
    return PIC_OK;
}

// n steps in runs that end at period boundaries: whole 4*N_SUB-step periods
// from the graph cache (captured on first use), partial runs from the cache
// when a timed window captured them, the rest eagerly.  dry: only capture
// the graphs of every run of the n steps (partial ones included), then
// restore the host state, so that a timed window replays graphs only and
// does not include their capture.
int run_steps(pic_sim *s, uint32_t n, bool measure, bool dry = false) {
    if (s->cfg.path == PIC_PATH_FUSED && s->need_prologue) TRY(fused_prologue(s));
    const uint32_t period = 4u * (uint32_t)s->p.n_sub;
    const bool graphs = s->cfg.use_graph && s->cfg.path == PIC_PATH_FUSED &&
                        (uint32_t)s->p.n_t % period == 0;
    int roles0[5];
    get_roles(s, roles0);
    const uint32_t t0 = s->t;
    const unsigned long long steps0 = s->steps;
    const bool mask0 = s->mask_ok;
    int rc = PIC_OK;
    while (n > 0 && rc == PIC_OK) {
        if (graphs && s->e_lag == 0) {
            const uint32_t len = std::min(n, period - s->t % period);
            pic_sim::GraphSlot *g = nullptr;
            if ((rc = graph_for(s, measure, len, dry || len == period, &g)) != PIC_OK) break;
            if (g) {
                if (!dry) {
                    const hipError_t e = hipGraphLaunch(g->exec, s->stream);
                    if (e != hipSuccess) {
                        rc = hip_fail(e, "hipGraphLaunch");
                        break;
                    }
                }
                set_roles(s, g->after);
                s->mask_ok = g->mask_after;
                s->t = (s->t + len) % (uint32_t)s->p.n_t;
                s->steps += len;
                n -= len;
                continue;
            }
        }
        if (dry) break;
        rc = one_step(s, measure, n == 1);
        --n;
    }
    if (dry) {
        set_roles(s, roles0);
        s->t = t0;
        s->steps = steps0;
        s->e_lag = 0;
        s->mask_ok = mask0;
    }
    return rc;
}

int free_all(pic_sim *s) {
    invalidate_graphs(s);
    for (int q = 0; q < 2; ++q)
        for (int c = 0; c < 4; ++c) {
            (void)hipFree(s->e[q][c]);
            (void)hipFree(s->ion[q][c]);
        }
    for (int c = 0; c < 4; ++c) {
        (void)hipFree(s->e_rng[c]);
        (void)hipFree(s->i_rng[c]);
    }
    (void)hipFree(s->xt);
    (void)hipFree(s->xt_count);
    (void)hipFree(s->xt_stage);
    (void)hipFree(s->e_mask);
    void *bufs[] = {s->cnt, s->pot, s->efield, s->grid_e, s->grid_i, s->pot0_tab, s->e_dens[0], s->e_dens[1], s->i_dens[0], s->i_dens[1],
                    s->cumul_e, s->cumul_i, s->eepf, s->ifed, s->counters, s->cs, s->stot_e, s->stot_i};
    for (void *b : bufs) (void)hipFree(b);
    for (int i = 0; i < 2; ++i)
        if (s->ev[i]) (void)hipEventDestroy(s->ev[i]);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    return PIC_OK;
}

}  // namespace

extern "C" {

void pic_sim_config_default(pic_sim_config_t *c) {
    memset(c, 0, sizeof(*c));
    pic_params_default(&c->params);
    pic_params_derive(&c->params);
    c->capacity = 120000;  // main.rs:79
    const uint32_t se[4] = {0x12345678u, 0x11112222u, 0x22223333u, 0x33334444u};  // main.rs:1667
    const uint32_t si[4] = {0x44445555u, 0x55556666u, 0x66667777u, 0x77778888u};  // main.rs:1671
    memcpy(c->seed_e, se, sizeof(se));
    memcpy(c->seed_i, si, sizeof(si));
    c->path = PIC_PATH_FUSED;
    c->use_graph = 1;
    c->measure = 1;
}

int pic_sim_create(const pic_sim_config_t *c, pic_sim_t **out) {
    if (!c || !out || c->capacity == 0) {
        set_error("pic_sim_create: bad argument");
        return PIC_ERR_ARG;
    }
    if (c->capacity > kMaxSlots - 64u * kMaxSegments) {
        set_error("pic_sim_create: capacity above 2^30 - 4096 slots per species");
        return PIC_ERR_ARG;
    }
    if (c->params.dt_e <= 0.0 || c->params.n_g < 3 || c->params.n_g > 4096) {
        set_error("pic_sim_create: params not derived (call pic_params_derive) or n_g out of [3, 4096]");
        return PIC_ERR_ARG;
    }
    pic_sim *s = new pic_sim();
    s->cfg = *c;
    s->p = c->params;
    s->k = make_kphys(s->p);
    s->prec = s->p.precision;
    if (s->prec < PIC_PREC_FP32 || s->prec > PIC_PREC_FP64) {
        delete s;
        set_error("pic_sim_create: unknown precision");
        return PIC_ERR_ARG;
    }
    s->f64 = s->prec == PIC_PREC_FP64;
    {
        const uint32_t period = 4u * (uint32_t)s->p.n_sub;
        s->compact_every = c->compact_every ? c->compact_every : std::max(period, 800u / period * period);
        if (s->compact_every % period) {
            delete s;
            set_error("pic_sim_create: compact_every must be a multiple of 4*n_sub");
            return PIC_ERR_ARG;
        }
        s->rng_defer = c->rng_defer ? c->rng_defer : default_rng_defer(s->p.n_sub);
        if (s->rng_defer > kMaxRngLag || period % s->rng_defer) {
            delete s;
            set_error("pic_sim_create: rng_defer must divide 4*n_sub and be at most 64");
            return PIC_ERR_ARG;
        }
        if (c->cand_masks < -1 || c->cand_masks > 1 || !(c->cand_threshold >= 0.0f)) {
            delete s;
            set_error("pic_sim_create: cand_masks must be -1 (off), 0 or 1 (on), cand_threshold >= 0");
            return PIC_ERR_ARG;
        }
    }
    s->cap = c->capacity;
    // the fused path may keep each species in segments with their own
    // counters; the reference path in one (its stage kernels see a plain
    // array).  One by default: the in-place kernels map each tile to its
    // segment by a search over the segments' tile prefix, which costs more
    // (step 40.6 us with 32 segments, 38.6 us with one, at 3.8 M electrons)
    // than the spread of the compaction's slot reservations over several
    // counters saves once every compact_every steps
    s->nseg = c->path == PIC_PATH_FUSED ? (c->segments > 0 ? (int)c->segments : 1) : 1;
    if (s->nseg > kMaxSegments) {
        delete s;
        set_error("pic_sim_create: at most 64 segments");
        return PIC_ERR_ARG;
    }
    if (s->nseg > 1 && (uint64_t)s->nseg * 64 > s->cap) s->nseg = std::max<int>(1, (int)(s->cap / 64));
    s->seg = (uint32_t)(((uint64_t)s->cap + s->nseg - 1) / s->nseg);
    if (s->nseg > 1) s->seg = (s->seg + 63) / 64 * 64;
    s->slots = s->seg * (uint32_t)s->nseg;
    Shape sh;
    sh.block = c->block_size > 0 ? c->block_size : 256;
    if (!particle_block_ok(sh.block) || (c->path != PIC_PATH_FUSED && sh.block != 256)) {
        delete s;
        set_error("pic_sim_create: block_size must be 0 (256), 128, 256 or 512, and 0 or 256 on the reference "
                  "path (its stage kernels have fixed shapes)");
        return PIC_ERR_ARG;
    }
    // the fused kernels hold ~110 VGPRs (4 waves per SIMD): a persistent grid
    // of 1024 lanes per compute unit is fully resident
    sh.grid = c->grid_blocks > 0 ? c->grid_blocks
                                 : (c->path == PIC_PATH_FUSED ? default_particle_grid(device_cus(), sh.block) : 0);
    // particles per lane of a compacting tile (on the reference path a tile
    // of 1024 lanes reserves its output with one atomic on the single live
    // counter, and 4 keeps several tiles per workgroup to overlap them)
    sh.ppt = c->ppt > 0 ? c->ppt : 4;
    sh.ppt_inplace = c->ppt_inplace > 0 ? c->ppt_inplace : 2;
    sh.lds_reps = std::max(0, c->lds_replicas);
    // global density replicas: 1, or a multiple of 8 (the Poisson kernel
    // reads them in groups of eight)
    sh.replicas = c->replicas > 0 ? c->replicas : (c->path == PIC_PATH_FUSED ? 8 : 1);
    if (sh.replicas > 1) sh.replicas = std::min(64, (sh.replicas + 7) / 8 * 8);
    // replica r of a density buffer starts at element r * rep_stride (n_g
    // rounded up to 32 elements), so the same node of different replicas
    // never shares a cache line
    sh.rep_stride = ((s->p.n_g + 31) / 32) * 32;
    s->sh = resolve(sh);
    s->dens_words = s->sh.replicas * s->sh.rep_stride;
    int rc = PIC_OK;
    auto fail = [&](int r) {
        free_all(s);
        delete s;
        return r;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(hip_fail(hipGetLastError(), "hipStreamCreate"));
    for (int q = 0; q < 2; ++q)
        for (int k = 0; k < 4; ++k) {
            if ((rc = dalloc(&s->e[q][k], s->slots))) return fail(rc);
            if ((rc = dalloc(&s->ion[q][k], s->slots))) return fail(rc);
        }
    for (int k = 0; k < 4; ++k) {
        if ((rc = dalloc(&s->e_rng[k], s->slots))) return fail(rc);
        if ((rc = dalloc(&s->i_rng[k], s->slots))) return fail(rc);
    }
    if ((rc = dalloc(&s->cnt, C_WORDS))) return fail(rc);
    if ((rc = dalloc(&s->pot, s->p.n_g))) return fail(rc);
    if ((rc = dalloc(&s->efield, s->p.n_g))) return fail(rc);
    if ((rc = dalloc(&s->grid_e, s->p.n_g))) return fail(rc);
    if ((rc = dalloc(&s->grid_i, s->p.n_g))) return fail(rc);
    if ((rc = dalloc(&s->pot0_tab, (size_t)s->p.n_t))) return fail(rc);
    for (int q = 0; q < 2; ++q) {
        if ((rc = dalloc((double **)&s->e_dens[q], s->dens_words))) return fail(rc);
        if ((rc = dalloc((double **)&s->i_dens[q], s->dens_words))) return fail(rc);
    }
    if ((rc = dalloc(&s->cumul_e, s->p.n_g))) return fail(rc);
    if ((rc = dalloc(&s->cumul_i, s->p.n_g))) return fail(rc);
    if ((rc = dalloc(&s->eepf, PIC_N_EEPF))) return fail(rc);
    if ((rc = dalloc(&s->ifed, 2 * PIC_N_IFED))) return fail(rc);
    if ((rc = dalloc(&s->counters, (size_t)kCounterLanes * kMeasCounters))) return fail(rc);
    if ((rc = dalloc(&s->cs, (size_t)PIC_N_CS * PIC_CS_RANGES))) return fail(rc);
    if ((rc = dalloc(&s->stot_e, PIC_CS_RANGES))) return fail(rc);
    if ((rc = dalloc(&s->stot_i, PIC_CS_RANGES))) return fail(rc);
    if (c->path == PIC_PATH_FUSED && c->cand_masks >= 0 && (rc = dalloc(&s->e_mask, s->slots))) return fail(rc);
    for (int i = 0; i < 2; ++i)
        if (hipEventCreate(&s->ev[i]) != hipSuccess) return fail(hip_fail(hipGetLastError(), "hipEventCreate"));
    // tables (main.rs:1649, 1663) and RNG streams (main.rs:1666-1675)
    {
        std::vector<float> cs((size_t)PIC_N_CS * PIC_CS_RANGES), se(PIC_CS_RANGES), si(PIC_CS_RANGES);
        if ((rc = pic_cross_sections(s->p.gas_density, cs.data(), se.data(), si.data()))) return fail(rc);
        if ((rc = pic_sim_set_tables(s, cs.data(), se.data(), si.data()))) return fail(rc);
    }
    if ((rc = pic_seed_streams(c->seed_e, s->slots, s->e_rng[0], s->e_rng[1], s->e_rng[2], s->e_rng[3], s->stream)))
        return fail(rc);
    if ((rc = pic_seed_streams(c->seed_i, s->slots, s->i_rng[0], s->i_rng[1], s->i_rng[2], s->i_rng[3], s->stream)))
        return fail(rc);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(hip_fail(hipGetLastError(), "seed"));
    *out = s;
    return PIC_OK;
}

int pic_sim_destroy(pic_sim_t *s) {
    if (!s) return PIC_OK;
    (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return PIC_OK;
}

int pic_sim_set_tables(pic_sim_t *s, const float *cs, const float *stot_e, const float *stot_i) {
    if (!s || !cs || !stot_e || !stot_i) return PIC_ERR_ARG;
    PIC_HIP(hipMemcpy(s->cs, cs, sizeof(float) * PIC_N_CS * PIC_CS_RANGES, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(s->stot_e, stot_e, sizeof(float) * PIC_CS_RANGES, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(s->stot_i, stot_i, sizeof(float) * PIC_CS_RANGES, hipMemcpyHostToDevice));
    s->coll_thr = s->cfg.cand_threshold > 0.0f ? s->cfg.cand_threshold : coll_threshold(stot_e, 1.0, s->p);
    s->mask_ok = false;
    return PIC_OK;
}

int pic_sim_cand_threshold(pic_sim_t *s, float *threshold, int *masks_on) {
    if (!s) return PIC_ERR_ARG;
    if (threshold) *threshold = s->coll_thr;
    if (masks_on) *masks_on = s->e_mask != nullptr;
    return PIC_OK;
}

int pic_sim_upload(pic_sim_t *s, int species_id, const float *x, const float *vx, const float *vy,
                   const float *vz, uint32_t n) {
    if (!s || (species_id != PIC_ELECTRONS && species_id != PIC_IONS) || n > s->cap ||
        (n && (!x || !vx || !vy || !vz))) {
        set_error("pic_sim_upload: bad argument (n above capacity?)");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(s->stream));
    float **a = species_id == PIC_ELECTRONS ? s->e[s->ce] : s->ion[s->ci];
    uint32_t *cnt = species_id == PIC_ELECTRONS ? ecnt(s, s->pe) : icnt(s, s->ci);
    const float *src[4] = {x, vx, vy, vz};
    // consecutive runs of the host arrays go to consecutive segments
    std::vector<uint32_t> c(kMaxSegments, 0u);
    uint32_t done = 0;
    for (int q = 0; q < s->nseg; ++q) {
        const uint32_t m = n / s->nseg + ((uint32_t)q < n % s->nseg ? 1u : 0u);
        c[q] = m;
        for (int w = 0; w < 4; ++w)
            if (m) PIC_HIP(hipMemcpy(a[w] + (size_t)q * s->seg, src[w] + done, 4ull * m, hipMemcpyHostToDevice));
        done += m;
    }
    PIC_HIP(hipMemcpy(cnt, c.data(), 4ull * kMaxSegments, hipMemcpyHostToDevice));
    s->need_prologue = true;
    invalidate_graphs(s);
    return PIC_OK;
}

int pic_sim_segments(pic_sim_t *s, int species_id, uint32_t *n_seg, uint32_t *seg_size, uint32_t *counts) {
    if (!s || (species_id != PIC_ELECTRONS && species_id != PIC_IONS)) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    if (n_seg) *n_seg = (uint32_t)s->nseg;
    if (seg_size) *seg_size = s->seg;
    if (counts) {
        uint32_t *cnt = species_id == PIC_ELECTRONS ? ecnt(s, s->pe) : icnt(s, s->ci);
        PIC_HIP(hipMemcpy(counts, cnt, 4ull * s->nseg, hipMemcpyDeviceToHost));
        for (int q = 0; q < s->nseg; ++q) counts[q] = std::min(counts[q], s->seg);
    }
    return PIC_OK;
}

// Live particles of a species in segment order, holes (x < 0) skipped, with
// the slot of each when `slots` is given.
static int gather_live(pic_sim *s, int species_id, float *x, float *vx, float *vy, float *vz,
                       uint32_t *slots, uint32_t cap, uint32_t *n) {
    std::vector<uint32_t> c(kMaxSegments, 0u);
    TRY(pic_sim_segments(s, species_id, nullptr, nullptr, c.data()));
    float **a = species_id == PIC_ELECTRONS ? s->e[s->ce] : s->ion[s->ci];
    float *dst[4] = {x, vx, vy, vz};
    std::vector<float> buf[4];
    uint32_t total = 0;
    for (int q = 0; q < s->nseg; ++q) {
        const uint32_t m = c[q];
        if (!m) continue;
        for (int w = 0; w < 4; ++w) {
            buf[w].resize(m);
            if (w == 0 || dst[w])
                PIC_HIP(hipMemcpy(buf[w].data(), a[w] + (size_t)q * s->seg, 4ull * m, hipMemcpyDeviceToHost));
        }
        for (uint32_t i = 0; i < m; ++i) {
            if (!(buf[0][i] >= 0.0f)) continue;
            if (total < cap) {
                for (int w = 0; w < 4; ++w)
                    if (dst[w]) dst[w][total] = buf[w][i];
                if (slots) slots[total] = q * s->seg + i;
            }
            ++total;
        }
    }
    if (n) *n = total;
    return PIC_OK;
}

int pic_sim_download(pic_sim_t *s, int species_id, float *x, float *vx, float *vy, float *vz,
                     uint32_t cap, uint32_t *n) {
    if (!s || (species_id != PIC_ELECTRONS && species_id != PIC_IONS)) return PIC_ERR_ARG;
    return gather_live(s, species_id, x, vx, vy, vz, nullptr, cap, n);
}

int pic_sim_download_slots(pic_sim_t *s, int species_id, uint32_t *slots, uint32_t cap, uint32_t *n) {
    if (!s || (species_id != PIC_ELECTRONS && species_id != PIC_IONS)) return PIC_ERR_ARG;
    return gather_live(s, species_id, nullptr, nullptr, nullptr, nullptr, slots, cap, n);
}

int pic_sim_rng_upload(pic_sim_t *s, int species_id, const uint32_t *w0, const uint32_t *w1,
                       const uint32_t *w2, const uint32_t *w3, uint32_t n) {
    if (!s || n > s->slots || !w0 || !w1 || !w2 || !w3) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    uint32_t **r = species_id == PIC_ELECTRONS ? s->e_rng : s->i_rng;
    const uint32_t *src[4] = {w0, w1, w2, w3};
    for (int q = 0; q < 4; ++q) PIC_HIP(hipMemcpy(r[q], src[q], 4ull * n, hipMemcpyHostToDevice));
    if (species_id == PIC_ELECTRONS) s->mask_ok = false;
    return PIC_OK;
}

int pic_sim_rng_download(pic_sim_t *s, int species_id, uint32_t *w0, uint32_t *w1, uint32_t *w2,
                         uint32_t *w3, uint32_t n) {
    if (!s || n > s->slots) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    uint32_t **r = species_id == PIC_ELECTRONS ? s->e_rng : s->i_rng;
    uint32_t *dst[4] = {w0, w1, w2, w3};
    for (int q = 0; q < 4; ++q)
        if (dst[q]) PIC_HIP(hipMemcpy(dst[q], r[q], 4ull * n, hipMemcpyDeviceToHost));
    return PIC_OK;
}

int pic_sim_counts(pic_sim_t *s, uint32_t *n_e, uint32_t *n_i) {
    if (!s) return PIC_ERR_ARG;
    uint32_t tot[2] = {0u, 0u};
    if (s->cfg.path == PIC_PATH_FUSED) {
        // in-place steps leave holes: count the live slots on the device
        uint32_t *d = s->cnt + C_ALIVE;
        for (int sp = 0; sp < 2; ++sp) {
            TRY(launch_copy_u32(d, nullptr, s->stream));
            float *x = sp == 0 ? s->e[s->ce][0] : s->ion[s->ci][0];
            uint32_t *c = sp == 0 ? ecnt(s, s->pe) : icnt(s, s->ci);
            TRY(launch_count_live(x, c, s->nseg, s->seg, d, s->stream));
            PIC_HIP(hipMemcpyAsync(&tot[sp], d, 4, hipMemcpyDeviceToHost, s->stream));
            PIC_HIP(hipStreamSynchronize(s->stream));
        }
    } else {
        std::vector<uint32_t> c(kMaxSegments);
        for (int sp = 0; sp < 2; ++sp) {
            TRY(pic_sim_segments(s, sp, nullptr, nullptr, c.data()));
            for (int q = 0; q < s->nseg; ++q) tot[sp] += c[q];
        }
    }
    if (n_e) *n_e = tot[0];
    if (n_i) *n_i = tot[1];
    return PIC_OK;
}

int pic_sim_time(pic_sim_t *s, uint32_t *t_in_cycle, unsigned long long *steps_done) {
    if (!s) return PIC_ERR_ARG;
    if (t_in_cycle) *t_in_cycle = s->t;
    if (steps_done) *steps_done = s->steps;
    return PIC_OK;
}

int pic_sim_set_time(pic_sim_t *s, uint32_t t) {
    if (!s || t >= (uint32_t)s->p.n_t) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    s->t = t;
    s->need_prologue = true;
    invalidate_graphs(s);
    return PIC_OK;
}

int pic_sim_run(pic_sim_t *s, uint32_t n_steps, int measure) {
    if (!s) return PIC_ERR_ARG;
    return run_steps(s, n_steps, measure != 0);
}

int pic_sim_synchronize(pic_sim_t *s) {
    if (!s) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    return PIC_OK;
}

int pic_sim_check(pic_sim_t *s) {
    if (!s) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    uint32_t ov = 0;
    PIC_HIP(hipMemcpy(&ov, s->cnt + C_OVERFLOW, 4, hipMemcpyDeviceToHost));
    if (ov) {
        set_error("particle array capacity exceeded (particles were dropped)");
        return PIC_ERR_CAPACITY;
    }
    return PIC_OK;
}

int pic_sim_read_grid(pic_sim_t *s, float *pot, float *efield, double *e_dens, double *i_dens) {
    if (!s) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    const int n_g = s->p.n_g;
    if (pot) PIC_HIP(hipMemcpy(pot, s->pot, sizeof(float) * n_g, hipMemcpyDeviceToHost));
    if (efield) PIC_HIP(hipMemcpy(efield, s->efield, sizeof(float) * n_g, hipMemcpyDeviceToHost));
    // the summed node densities the last Poisson solve wrote out
    if (e_dens) PIC_HIP(hipMemcpy(e_dens, s->grid_e, sizeof(double) * n_g, hipMemcpyDeviceToHost));
    if (i_dens) PIC_HIP(hipMemcpy(i_dens, s->grid_i, sizeof(double) * n_g, hipMemcpyDeviceToHost));
    return PIC_OK;
}

int pic_sim_read_measurements(pic_sim_t *s, double *cumul_e, double *cumul_i, uint32_t *eepf,
                              uint32_t *ifed, unsigned long long *counters) {
    if (!s) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    const int n_g = s->p.n_g;
    if (cumul_e) PIC_HIP(hipMemcpy(cumul_e, s->cumul_e, sizeof(double) * n_g, hipMemcpyDeviceToHost));
    if (cumul_i) PIC_HIP(hipMemcpy(cumul_i, s->cumul_i, sizeof(double) * n_g, hipMemcpyDeviceToHost));
    if (eepf) PIC_HIP(hipMemcpy(eepf, s->eepf, sizeof(uint32_t) * PIC_N_EEPF, hipMemcpyDeviceToHost));
    if (ifed) PIC_HIP(hipMemcpy(ifed, s->ifed, sizeof(uint32_t) * 2 * PIC_N_IFED, hipMemcpyDeviceToHost));
    if (counters) {
        std::vector<unsigned long long> lanes((size_t)kCounterLanes * kMeasCounters);
        PIC_HIP(hipMemcpy(lanes.data(), s->counters, sizeof(unsigned long long) * lanes.size(),
                          hipMemcpyDeviceToHost));
        for (int c = 0; c < kMeasCounters; ++c) {
            counters[c] = 0;
            for (int l = 0; l < kCounterLanes; ++l) counters[c] += lanes[(size_t)l * kMeasCounters + c];
        }
        counters[6] += s->ref_meas[0];
        counters[7] += s->ref_meas[1];
    }
    return PIC_OK;
}

int pic_sim_reset_measurements(pic_sim_t *s) {
    if (!s) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    PIC_HIP(hipMemset(s->cumul_e, 0, sizeof(double) * s->p.n_g));
    PIC_HIP(hipMemset(s->cumul_i, 0, sizeof(double) * s->p.n_g));
    PIC_HIP(hipMemset(s->eepf, 0, sizeof(uint32_t) * PIC_N_EEPF));
    PIC_HIP(hipMemset(s->ifed, 0, sizeof(uint32_t) * 2 * PIC_N_IFED));
    PIC_HIP(hipMemset(s->counters, 0, sizeof(unsigned long long) * kCounterLanes * kMeasCounters));
    s->ref_meas[0] = s->ref_meas[1] = 0;
    if (s->xt) {
        PIC_HIP(hipMemset(s->xt, 0, sizeof(double) * kXtPlanes * s->p.n_g * s->xt_n));
        PIC_HIP(hipMemset(s->xt_count, 0, sizeof(unsigned long long) * s->xt_n));
    }
    return PIC_OK;
}

int pic_sim_set_xt(pic_sim_t *s, uint32_t bin_steps) {
    if (!s) return PIC_ERR_ARG;
    if (bin_steps && (s->cfg.path != PIC_PATH_FUSED || (uint32_t)s->p.n_t % bin_steps != 0)) {
        set_error("pic_sim_set_xt: XT maps are measured on the fused path, with bin_steps dividing n_t");
        return PIC_ERR_ARG;
    }
    if (bin_steps && (size_t)4 * s->p.n_g * sizeof(unsigned long long) > (size_t)device_lds_per_block()) {
        set_error("pic_sim_set_xt: the per-workgroup XT accumulators (32 B per node) exceed the LDS of a workgroup");
        return PIC_ERR_UNSUPPORTED;
    }
    PIC_HIP(hipStreamSynchronize(s->stream));
    invalidate_graphs(s);  // measurement graphs gain or lose the XT launches
    (void)hipFree(s->xt);
    (void)hipFree(s->xt_count);
    (void)hipFree(s->xt_stage);
    s->xt = nullptr;
    s->xt_count = nullptr;
    s->xt_stage = nullptr;
    s->xt_bin = 0;
    s->xt_n = 0;
    if (!bin_steps) return PIC_OK;
    const int n = (int)((uint32_t)s->p.n_t / bin_steps);
    int rc = dalloc(&s->xt, (size_t)kXtPlanes * s->p.n_g * n);
    if (rc == PIC_OK) rc = dalloc(&s->xt_count, (size_t)n);
    if (rc == PIC_OK) rc = dalloc(&s->xt_stage, xt_stage_floats(s->p.n_g));
    if (rc != PIC_OK) {  // leave the maps off, not half allocated
        (void)hipFree(s->xt);
        (void)hipFree(s->xt_count);
        (void)hipFree(s->xt_stage);
        s->xt = nullptr;
        s->xt_count = nullptr;
        s->xt_stage = nullptr;
        return rc;
    }
    s->xt_bin = bin_steps;
    s->xt_n = n;
    return PIC_OK;
}

int pic_sim_read_xt(pic_sim_t *s, double *out, uint32_t *n_xt) {
    if (!s) return PIC_ERR_ARG;
    if (n_xt) *n_xt = (uint32_t)s->xt_n;
    if (!out) return PIC_OK;
    if (!s->xt) {
        set_error("pic_sim_read_xt: XT maps are not enabled (pic_sim_set_xt)");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(s->stream));
    const int n_g = s->p.n_g, nx = s->xt_n;
    const size_t plane = (size_t)n_g * nx;
    std::vector<double> raw(kXtPlanes * plane);
    std::vector<unsigned long long> cnt(nx);
    PIC_HIP(hipMemcpy(raw.data(), s->xt, sizeof(double) * raw.size(), hipMemcpyDeviceToHost));
    PIC_HIP(hipMemcpy(cnt.data(), s->xt_count, sizeof(unsigned long long) * nx, hipMemcpyDeviceToHost));
    // norm_all_xt (C/eduPIC.cc:864-903), with the measured steps of each
    // bin in place of cycles * N_BIN
    const double e = 1.60217662e-19;
    const double rate_norm = s->p.weight / (s->p.area * s->p.dx) / s->p.dt_e;
    enum { POT, EF, NE, NI, CE, UE, EE, IZ, CI, UI, EI };
    enum { O_POT, O_EF, O_NE, O_NI, O_JE, O_JI, O_PE, O_PI, O_EE, O_EI, O_IZ };
    for (int i = 0; i < n_g; ++i)
        for (int jb = 0; jb < nx; ++jb) {
            const size_t o = (size_t)i * nx + jb;
            auto r = [&](int q) { return raw[q * plane + o]; };
            auto w = [&](int q, double v) { out[q * plane + o] = v; };
            const double f1 = cnt[jb] ? 1.0 / (double)cnt[jb] : 0.0;
            const double f2 = cnt[jb] ? rate_norm / (double)cnt[jb] : 0.0;
            const double pot = r(POT) * f1, ef = r(EF) * f1, ne = r(NE) * f1, ni = r(NI) * f1;
            double je = 0.0, ee = 0.0, iz = 0.0, ji = 0.0, ei = 0.0;
            if (r(CE) > 0.0) {
                const double ue = r(UE) / r(CE);
                je = -ue * ne * e;
                ee = r(EE) / r(CE);
                iz = r(IZ) * f2;
            }
            if (r(CI) > 0.0) {
                const double ui = r(UI) / r(CI);
                ji = ui * ni * e;
                ei = r(EI) / r(CI);
            }
            w(O_POT, pot);
            w(O_EF, ef);
            w(O_NE, ne);
            w(O_NI, ni);
            w(O_JE, je);
            w(O_JI, ji);
            w(O_PE, je * ef);
            w(O_PI, ji * ef);
            w(O_EE, ee);
            w(O_EI, ei);
            w(O_IZ, iz);
        }
    return PIC_OK;
}

int pic_sim_time_steps(pic_sim_t *s, uint32_t n_steps, int measure, float *ms_total, float *ms_kernel,
                       uint32_t *kernel_launches, unsigned long long *particle_steps) {
    if (!s) return PIC_ERR_ARG;
    const bool fused = s->cfg.path == PIC_PATH_FUSED;
    if (fused && s->need_prologue) TRY(fused_prologue(s));
    PIC_HIP(hipStreamSynchronize(s->stream));
    // particle-steps: electrons every step, ions on ion steps (counted from
    // the live counts before each step, on the device-side counters)
    uint32_t ne0 = 0, ni0 = 0;
    TRY(pic_sim_counts(s, &ne0, &ni0));
    uint32_t ion_steps = 0;
    for (uint32_t i = 0, tt = s->t; i < n_steps; ++i, tt = (tt + 1) % (uint32_t)s->p.n_t)
        ion_steps += (tt % (uint32_t)s->p.n_sub) == 0;
    float total = 0.0f, kern = 0.0f;
    uint32_t launches = 0;
    unsigned long long psteps = 0;
    if (ms_kernel && fused) {
        // eager steps, the electron particle kernel bracketed by events; the
        // chunks are synchronised to read the events back
        const uint32_t chunk = 64;
        std::vector<hipEvent_t> ev(2 * chunk);
        for (auto &e : ev) PIC_HIP(hipEventCreate(&e));
        hipEvent_t w0, w1;
        PIC_HIP(hipEventCreate(&w0));
        PIC_HIP(hipEventCreate(&w1));
        uint32_t done = 0;
        while (done < n_steps) {
            const uint32_t c = std::min(chunk, n_steps - done);
            PIC_HIP(hipEventRecord(w0, s->stream));
            for (uint32_t i = 0; i < c; ++i) {
                TRY(step_fused(s, measure != 0, done + i + 1 == n_steps, true, ev[2 * i], ev[2 * i + 1]));
                s->t = (s->t + 1) % (uint32_t)s->p.n_t;
                s->steps++;
            }
            PIC_HIP(hipEventRecord(w1, s->stream));
            PIC_HIP(hipEventSynchronize(w1));
            float ms = 0.0f;
            PIC_HIP(hipEventElapsedTime(&ms, w0, w1));
            total += ms;
            for (uint32_t i = 0; i < c; ++i) {
                PIC_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
                kern += ms;
            }
            launches += c;
            done += c;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        (void)hipEventDestroy(w0);
        (void)hipEventDestroy(w1);
    } else {
        TRY(run_steps(s, n_steps, measure != 0, true));
        PIC_HIP(hipEventRecord(s->ev[0], s->stream));
        TRY(run_steps(s, n_steps, measure != 0));
        PIC_HIP(hipEventRecord(s->ev[1], s->stream));
        PIC_HIP(hipEventSynchronize(s->ev[1]));
        PIC_HIP(hipEventElapsedTime(&total, s->ev[0], s->ev[1]));
    }
    // particle-steps over the window from the counts at both ends (the
    // live count changes by well under a percent per RF cycle at steady state)
    uint32_t ne1 = 0, ni1 = 0;
    TRY(pic_sim_counts(s, &ne1, &ni1));
    const double ne_avg = 0.5 * ((double)ne0 + (double)ne1);
    const double ni_avg = 0.5 * ((double)ni0 + (double)ni1);
    psteps = (unsigned long long)(ne_avg * n_steps + ni_avg * ion_steps);
    if (ms_total) *ms_total = total;
    if (ms_kernel) *ms_kernel = kern;
    if (kernel_launches) *kernel_launches = launches;
    if (particle_steps) *particle_steps = psteps;
    return PIC_OK;
}

// Checkpoint file (version 2): header u32[8] = magic, version, n_g, t,
// n_e, n_i, 0, 0; steps done (u64); per species the n live particles (x,
// vx, vy, vz as f32[n] each) and the four RNG words of each particle's slot
// in the same order (u32[n] each); then the measurements (cumul_e, cumul_i
// f64[n_g], EEPF u32[2000], IFED u32[400], counters u64[8]).  A load puts
// the particles into the first slots of its segments and gives each one
// the RNG stream it had, so every particle continues its own sequence of
// draws.  The continued run is the same discharge but not the same bits:
// slots freed by lost particles and the streams newborns take differ, and
// the next Poisson solve deposits every ion present (ions born since the
// last ion step included), where an uninterrupted run adds those at the
// next ion step.
int pic_sim_save(pic_sim_t *s, const char *path) {
    if (!s || !path) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    uint32_t ne = 0, ni = 0;
    int rc = pic_sim_counts(s, &ne, &ni);
    if (rc) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error(std::string("cannot open ") + path);
        return PIC_ERR_IO;
    }
    auto fail = [&](int r) {
        fclose(f);
        return r;
    };
    uint32_t hdr[8] = {kCkptMagic, kCkptVersion, (uint32_t)s->p.n_g, s->t, ne, ni, 0, 0};
    fwrite(hdr, sizeof(hdr), 1, f);
    fwrite(&s->steps, sizeof(s->steps), 1, f);
    std::vector<uint32_t> w[4];
    for (auto &v : w) v.resize(s->slots);
    for (int sp = 0; sp < 2; ++sp) {
        const uint32_t n = sp == 0 ? ne : ni;
        std::vector<float> a[4];
        for (auto &v : a) v.resize(std::max<uint32_t>(n, 1));
        std::vector<uint32_t> slots(std::max<uint32_t>(n, 1)), g(std::max<uint32_t>(n, 1));
        uint32_t got = 0, got_s = 0;
        if ((rc = pic_sim_download(s, sp, a[0].data(), a[1].data(), a[2].data(), a[3].data(), n, &got)) ||
            (rc = pic_sim_download_slots(s, sp, slots.data(), n, &got_s)) ||
            (rc = pic_sim_rng_download(s, sp, w[0].data(), w[1].data(), w[2].data(), w[3].data(), s->slots)))
            return fail(rc);
        if (got != n || got_s != n) {
            set_error("checkpoint: the live count changed while saving");
            return fail(PIC_ERR_IO);
        }
        for (auto &v : a) fwrite(v.data(), sizeof(float), n, f);
        for (int q = 0; q < 4; ++q) {
            for (uint32_t j = 0; j < n; ++j) g[j] = w[q][slots[j]];
            fwrite(g.data(), 4, n, f);
        }
    }
    std::vector<double> ce(s->p.n_g), cii(s->p.n_g);
    std::vector<uint32_t> ee(PIC_N_EEPF), ifd(2 * PIC_N_IFED);
    unsigned long long cnt[8];
    if ((rc = pic_sim_read_measurements(s, ce.data(), cii.data(), ee.data(), ifd.data(), cnt))) return fail(rc);
    fwrite(ce.data(), 8, s->p.n_g, f);
    fwrite(cii.data(), 8, s->p.n_g, f);
    fwrite(ee.data(), 4, PIC_N_EEPF, f);
    fwrite(ifd.data(), 4, 2 * PIC_N_IFED, f);
    fwrite(cnt, 8, 8, f);
    bool ok = ferror(f) == 0;
    fclose(f);
    if (!ok) {
        set_error("checkpoint write failed");
        return PIC_ERR_IO;
    }
    return PIC_OK;
}

int pic_sim_load(pic_sim_t *s, const char *path) {
    if (!s || !path) return PIC_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error(std::string("cannot open ") + path);
        return PIC_ERR_IO;
    }
    auto bad = [&](const char *why) {
        fclose(f);
        set_error(std::string("checkpoint ") + path + ": " + why);
        return PIC_ERR_IO;
    };
    uint32_t hdr[8];
    unsigned long long steps = 0;
    if (fread(hdr, sizeof(hdr), 1, f) != 1 || hdr[0] != kCkptMagic || hdr[1] != kCkptVersion)
        return bad("not a checkpoint of this version");
    if ((int)hdr[2] != s->p.n_g) return bad("grid size differs");
    if (hdr[4] > s->cap || hdr[5] > s->cap) return bad("particle count above capacity");
    // the header's counts must account for the whole file before anything
    // is allocated from them
    if (fseek(f, 0, SEEK_END) != 0) return bad("cannot seek");
    const long size = ftell(f);
    const uint64_t want = sizeof(hdr) + sizeof(steps) + 32ull * ((uint64_t)hdr[4] + hdr[5]) +
                          16ull * s->p.n_g + 4ull * (PIC_N_EEPF + 2 * PIC_N_IFED) + 64ull;
    if (size < 0 || (uint64_t)size != want) return bad("file size does not match its header");
    if (fseek(f, sizeof(hdr), SEEK_SET) != 0 || fread(&steps, sizeof(steps), 1, f) != 1) return bad("truncated");
    std::vector<uint32_t> w[4];
    for (auto &v : w) v.resize(s->slots);
    for (int sp = 0; sp < 2; ++sp) {
        const uint32_t n = sp == 0 ? hdr[4] : hdr[5];
        std::vector<float> a[4];
        for (auto &v : a) {
            v.resize(std::max<uint32_t>(n, 1));
            if (fread(v.data(), sizeof(float), n, f) != n) return bad("truncated particles");
        }
        std::vector<uint32_t> g[4];
        for (auto &v : g) {
            v.resize(std::max<uint32_t>(n, 1));
            if (fread(v.data(), 4, n, f) != n) return bad("truncated rng");
        }
        int rc = pic_sim_upload(s, sp, a[0].data(), a[1].data(), a[2].data(), a[3].data(), n);
        std::vector<uint32_t> slots(std::max<uint32_t>(n, 1));
        uint32_t got = 0;
        if (!rc) rc = pic_sim_download_slots(s, sp, slots.data(), n, &got);
        if (!rc) rc = pic_sim_rng_download(s, sp, w[0].data(), w[1].data(), w[2].data(), w[3].data(), s->slots);
        if (!rc && got != n) {
            set_error("checkpoint: the upload did not place every particle");
            rc = PIC_ERR_IO;
        }
        if (rc) {
            fclose(f);
            return rc;
        }
        // each particle gets the stream it had; free slots keep theirs
        for (int q = 0; q < 4; ++q)
            for (uint32_t j = 0; j < n; ++j) w[q][slots[j]] = g[q][j];
        if ((rc = pic_sim_rng_upload(s, sp, w[0].data(), w[1].data(), w[2].data(), w[3].data(), s->slots))) {
            fclose(f);
            return rc;
        }
    }
    std::vector<double> ce(s->p.n_g), cii(s->p.n_g);
    std::vector<uint32_t> ee(PIC_N_EEPF), ifd(2 * PIC_N_IFED);
    unsigned long long cnt[8];
    if (fread(ce.data(), 8, s->p.n_g, f) != (size_t)s->p.n_g || fread(cii.data(), 8, s->p.n_g, f) != (size_t)s->p.n_g ||
        fread(ee.data(), 4, PIC_N_EEPF, f) != PIC_N_EEPF || fread(ifd.data(), 4, 2 * PIC_N_IFED, f) != 2 * PIC_N_IFED ||
        fread(cnt, 8, 8, f) != 8)
        return bad("truncated measurements");
    fclose(f);
    PIC_HIP(hipMemcpy(s->cumul_e, ce.data(), 8ull * s->p.n_g, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(s->cumul_i, cii.data(), 8ull * s->p.n_g, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(s->eepf, ee.data(), 4ull * PIC_N_EEPF, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(s->ifed, ifd.data(), 4ull * 2 * PIC_N_IFED, hipMemcpyHostToDevice));
    PIC_HIP(hipMemset(s->counters, 0, sizeof(unsigned long long) * kCounterLanes * kMeasCounters));
    PIC_HIP(hipMemcpy(s->counters, cnt, 8ull * kMeasCounters, hipMemcpyHostToDevice));
    s->ref_meas[0] = s->ref_meas[1] = 0;
    s->t = hdr[3] % (uint32_t)s->p.n_t;
    s->steps = steps;
    s->need_prologue = true;
    invalidate_graphs(s);
    return PIC_OK;
}

}  // extern "C"
