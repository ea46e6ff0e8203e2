// pic_ensemble.hip -- ensembles of independent discharges on one GPU: every
// stage of every instance in one launch (one Poisson workgroup per
// instance, one persistent particle kernel per species over all instances),
// the 4*N_SUB-step launch pattern captured into a hipGraph.  On the graph
// path the instances are run in groups sized to the last-level cache (the
// MI355X's 256 MB Infinity Cache): a group runs a block of periods while
// its particles stay resident, then the next group runs the same steps.
//
// Instances share the grid (N_G, gap), the time step (frequency, N_T,
// N_SUB) and the precision; each has its own driving amplitude, gas
// pressure and temperature, superparticle weight, capacity, particles,
// RNG streams, counters, densities and measurements.  The per-instance
// total cross section is the shared summed table times the instance's gas
// density (the reference tabulates sigma_tot * n_gas, main.rs:1439-1440).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pic_internal.h"

using namespace pic;

struct pic_ensemble {
    int E = 0;
    pic_params_t p0;  // shared geometry / time step / precision
    std::vector<pic_params_t> params;
    std::vector<EnsInst> inst_h;
    EnsInst *inst_d = nullptr;
    uint32_t e_total = 0, i_total = 0;
    float *e[2][4] = {}, *ion[2][4] = {};
    uint32_t *e_rng[4] = {}, *i_rng[4] = {};
    uint32_t *cnt = nullptr;   // [E][EC_WORDS]
    uint32_t *clock = nullptr; // [E] device step index within the cycle; group [g0, g1) uses clock[g0]
    uint32_t *live = nullptr;  // [2][E] live counts for the host
    void *e_dens[2] = {}, *i_dens[2] = {};
    int inst_stride = 0;       // density entries per instance
    float *pot = nullptr, *efield = nullptr;
    double *cumul_e = nullptr, *cumul_i = nullptr;
    double *trig = nullptr;   // [n_t][2] cos, sin of the drive's phase
    double *pot0 = nullptr;   // [E] pot0 of the step each instance solves next
    double *i_sum = nullptr;  // [E][n_g] held ion densities summed over the replicas
    unsigned long long *counters = nullptr;  // [E][8]
    float *cs = nullptr, *stot_e = nullptr, *stot_i = nullptr;
    Shape sh;
    bool f64 = true;
    int prec = PIC_PREC_FP64;
    hipStream_t stream = nullptr;
    // roles: ce / ci the electron / ion array of the pair in use, cd / hi the
    // density buffers, pe the electron watermark parity, ac set after a
    // compaction step (the watermark is then the compaction's output count)
    int ce = 0, ci = 0, cd = 0, hi = 0, pe = 0, ac = 0;
    uint32_t t = 0;
    unsigned long long steps = 0;
    bool need_prologue = true;
    bool use_graph = true;
    uint32_t compact_every = 800;  // steps between compaction passes (a multiple of 4*N_SUB)
    // deferred electron RNG write-back: steps per run (divides 4*N_SUB) and
    // the draws the stored words of live slots are behind at the next step
    // (0 between calls and at every period start)
    uint32_t rng_defer = 8;
    uint32_t e_lag = 0;
    // collision-candidate masks of the deferred runs (a byte per pooled
    // electron slot, null when off): mask_ok while they hold the current
    // run's candidates (written by the last write-back step; a compaction,
    // a measured step, an upload or a new clock leaves them stale)
    uint8_t *e_mask = nullptr;
    bool mask_ok = false;
    // instance groups [grp[g], grp[g+1]) and the periods a group runs at a time
    std::vector<int> grp;
    uint32_t group_periods = 5;
    // streams the groups are dealt to (work[0] is the ensemble's stream;
    // the others fork from it at a block's start and join it at its end)
    std::vector<hipStream_t> work;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    // captured 4*N_SUB-step graphs of one group, keyed by the group, the
    // measure flag, whether the period starts with a compaction, and the
    // roles at its start (the mask state among them); each remembers the
    // roles it leaves behind
    static constexpr int kRoles = 7;
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        int key[kRoles + 3] = {};
        int after[kRoles] = {};
    };
    std::vector<GraphSlot> graphs;
    int graph_next = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
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

uint64_t dens_words32(pic_ensemble *s) { return (uint64_t)s->E * s->inst_stride * (s->f64 ? 2 : 1); }

// counter word holding the electron watermark the next step reads
int e_word(const pic_ensemble *s) { return s->ac ? EC_C : EC_E0 + s->pe; }
int i_word(const pic_ensemble *s) { return EC_I0 + s->ci; }

EnsPoissonJob poisson_job(pic_ensemble *s, int g0, int g1) {
    const double E = 1.60217662e-19, EPS0 = 8.85418781e-12;
    EnsPoissonJob j;
    memset(&j, 0, sizeof(j));
    j.precision = s->prec;
    j.g.n_g = s->p0.n_g;
    j.g.reps = s->sh.replicas;
    j.g.rep_stride = s->sh.rep_stride;
    j.g.e_charge = E;
    j.g.alpha = -s->p0.dx * s->p0.dx / EPS0;
    j.g.inv_dx = s->p0.inv_dx;
    j.g.half_inv_dx = 0.5 * s->p0.inv_dx;
    j.g.half_dx_over_eps = s->p0.dx / (2.0 * EPS0);
    j.inst = s->inst_d + g0;
    j.n_inst = g1 - g0;
    j.inst_stride = s->inst_stride;
    j.time = s->clock + g0;
    j.n_t = s->p0.n_t;
    j.n_sub = s->p0.n_sub;
    j.pot = s->pot + (size_t)g0 * s->p0.n_g;
    j.efield = s->efield + (size_t)g0 * s->p0.n_g;
    j.counters = s->counters + (size_t)g0 * 8;
    j.cnt = s->cnt + (size_t)g0 * EC_WORDS;
    j.copy_src = j.copy_dst = j.zc0 = j.zc1 = -1;
    return j;
}

// density buffer b shifted to instance g0
void *dens_at(const pic_ensemble *s, void *b, int g0) {
    return (char *)b + (size_t)g0 * s->inst_stride * (s->f64 ? 8 : 4);
}

int prologue(pic_ensemble *s) {
    hipStream_t st = s->stream;
    s->mask_ok = false;
    TRY(launch_fill(s->e_dens[s->cd], 0u, dens_words32(s), st));
    TRY(launch_fill(s->e_dens[1 - s->cd], 0u, dens_words32(s), st));
    TRY(launch_fill(s->i_dens[0], 0u, dens_words32(s), st));
    TRY(launch_fill(s->i_dens[1], 0u, dens_words32(s), st));
    Shape sh = s->sh;
    const size_t es = s->f64 ? 8 : 4;
    for (int i = 0; i < s->E; ++i) {
        const EnsInst &in = s->inst_h[i];
        uint32_t *c = s->cnt + (size_t)i * EC_WORDS;
        char *ed = (char *)s->e_dens[s->cd] + (size_t)i * s->inst_stride * es;
        TRY(launch_deposit(s->e[s->ce][0] + in.e_off, c + e_word(s), ed, s->f64, in.k, sh, st, in.e_cap));
        for (int b = 0; b < 2; ++b) {
            char *id = (char *)s->i_dens[b] + (size_t)i * s->inst_stride * es;
            TRY(launch_deposit(s->ion[s->ci][0] + in.i_off, c + i_word(s), id, s->f64, in.k, sh, st, in.i_cap));
        }
    }
    const std::vector<uint32_t> clk(s->E, s->t);
    PIC_HIP(hipMemcpyAsync(s->clock, clk.data(), 4ull * s->E, hipMemcpyHostToDevice, st));
    // the Poisson kernels read pot0 and, between ion steps, the summed held
    // ion density; each writes the next step's pot0 and, on ion steps, the sums
    TRY(launch_ens_pot0_init(s->inst_d, s->E, s->p0.n_t, s->t, s->trig, s->pot0, st));
    {
        const EnsPoissonJob pj = poisson_job(s, 0, s->E);
        TRY(launch_sum_replicas_ens(pj.g, s->prec, s->i_dens[s->hi], s->inst_stride, s->E, s->i_sum, st));
    }
    PIC_HIP(hipStreamSynchronize(st));
    s->need_prologue = false;
    return PIC_OK;
}

EnsCompactJob compact_job(pic_ensemble *s, int g0, int g1, bool electrons, float *const *src, float *const *dst,
                          int c_in, int c_out) {
    EnsCompactJob j;
    memset(&j, 0, sizeof(j));
    j.electrons = electrons;
    j.inst = s->inst_d + g0;
    j.n_inst = g1 - g0;
    for (int q = 0; q < 4; ++q) {
        j.src[q] = src[q];
        j.dst[q] = dst[q];
    }
    j.cnt = s->cnt + (size_t)g0 * EC_WORDS;
    j.c_in = c_in;
    j.c_out = c_out;
    return j;
}

// The electrons' RNG write-back lag after the step at s->t, given the lag
// s->e_lag it starts with: runs of rng_defer steps aligned to multiples of
// rng_defer in the cycle; a compacting step (the compaction pass moves
// particles to slots whose streams they continue), the last step of the
// cycle and the last step of a call write the words back.
uint32_t next_lag(const pic_ensemble *s, bool last) {
    const uint32_t t = s->t;
    if (last || t % s->compact_every == 0 || (t + 1u) % s->rng_defer == 0 || t + 1u == (uint32_t)s->p0.n_t) return 0;
    return s->e_lag + 1u;
}

// Candidate masks of the step at s->t that ends lag_out behind: read inside
// a deferred run whose masks are valid, written by a write-back step whose
// next run has steps that can use them (the next step neither compacts nor
// writes back).  No masks on compacting or measured steps.
void mask_mode(const pic_ensemble *s, bool measure, uint32_t lag_out, bool *read, bool *write) {
    const uint32_t t = s->t;
    const bool on = s->e_mask && t % s->compact_every != 0 && !measure;
    *read = on && lag_out != 0u && s->mask_ok;
    *write = false;
    if (on && lag_out == 0u) {
        const uint32_t t1 = (t + 1u) % (uint32_t)s->p0.n_t;
        *write = t1 % s->compact_every != 0 && (t1 + 1u) % s->rng_defer != 0 && t1 + 1u != (uint32_t)s->p0.n_t;
    }
}

// One ensemble step of the instances [g0, g1).  Particles are updated in
// place (lost ones leave holes, newborn electrons and ions are appended
// after the watermarks); every compact_every steps a compaction pass after
// the step packs each species into the other array of its pair.  Electron
// watermarks alternate between two counter words every step: the step
// reads one, the Poisson kernel copies it into the other, newborns are
// appended there.  The electrons' RNG words start s->e_lag draws behind and
// end lag_out behind.  The roles, the lag and the host clock are advanced
// by end_step once every group has had the step.
int step_range(pic_ensemble *s, bool measure, int g0, int g1, uint32_t lag_out) {
    hipStream_t st = s->stream;
    const uint32_t t = s->t;
    const bool ion = (t % (uint32_t)s->p0.n_sub) == 0;
    const bool comp_e = (t % s->compact_every) == 0;
    const bool comp_i = ion && comp_e;
    const int in_e = e_word(s), out_e = EC_E0 + (1 - s->pe);
    int ci = s->ci;
    EnsPoissonJob pj = poisson_job(s, g0, g1);
    pj.e_dens = dens_at(s, s->e_dens[s->cd], g0);
    pj.i_dens = dens_at(s, s->i_dens[ion ? 1 - s->hi : s->hi], g0);
    // the next electron accumulator is cleared by this step's electron
    // kernel (the buffer the solve reads now); the ion one here
    pj.zero_b = dens_at(s, s->i_dens[s->hi], g0);
    pj.pot0 = s->pot0 + g0;
    pj.trig = s->trig;
    if (ion) pj.i_sum_out = s->i_sum + (size_t)g0 * s->p0.n_g;
    else pj.i_sum_in = s->i_sum + (size_t)g0 * s->p0.n_g;
    pj.copy_src = in_e;
    pj.copy_dst = out_e;
    if (comp_e) pj.zc0 = EC_C;
    if (comp_i) pj.zc1 = EC_I0 + (1 - ci);
    if (measure) {
        pj.cumul_e = s->cumul_e + (size_t)g0 * s->p0.n_g;
        pj.cumul_i = s->cumul_i + (size_t)g0 * s->p0.n_g;
    }
    TRY(launch_poisson_ens(pj, st));
    const int i_acc = ion ? s->hi : 1 - s->hi;
    EnsStepJob base;
    memset(&base, 0, sizeof(base));
    base.inst = s->inst_d + g0;
    base.n_inst = g1 - g0;
    base.k = s->inst_h[0].k;
    base.efield = s->efield + (size_t)g0 * s->p0.n_g;
    base.cnt = s->cnt + (size_t)g0 * EC_WORDS;
    base.inst_stride = s->inst_stride;
    base.dens_f64 = s->f64;
    base.cs = s->cs;
    base.counters = s->counters + (size_t)g0 * 8;
    base.n_t = s->p0.n_t;
    if (ion) {
        EnsStepJob fi = base;
        fi.electrons = false;
        fi.x = s->ion[ci][0];
        fi.vx = s->ion[ci][1];
        fi.vy = s->ion[ci][2];
        fi.vz = s->ion[ci][3];
        fi.rng0 = s->i_rng[0];
        fi.rng1 = s->i_rng[1];
        fi.rng2 = s->i_rng[2];
        fi.rng3 = s->i_rng[3];
        fi.c_in = EC_I0 + ci;
        fi.dens = dens_at(s, s->i_dens[i_acc], g0);
        fi.stot = s->stot_i;
        TRY(launch_inplace_ens(fi, s->sh, st));
        if (comp_i) {
            TRY(launch_compact_ens(
                compact_job(s, g0, g1, false, s->ion[ci], s->ion[1 - ci], EC_I0 + ci, EC_I0 + (1 - ci)), st));
            ci ^= 1;
        }
    }
    EnsStepJob fe = base;
    fe.electrons = true;
    fe.x = s->e[s->ce][0];
    fe.vx = s->e[s->ce][1];
    fe.vy = s->e[s->ce][2];
    fe.vz = s->e[s->ce][3];
    fe.rng0 = s->e_rng[0];
    fe.rng1 = s->e_rng[1];
    fe.rng2 = s->e_rng[2];
    fe.rng3 = s->e_rng[3];
    fe.ix = s->ion[ci][0];
    fe.ivx = s->ion[ci][1];
    fe.ivy = s->ion[ci][2];
    fe.ivz = s->ion[ci][3];
    fe.c_in = in_e;
    fe.c_out = out_e;
    fe.c_ion = EC_I0 + ci;
    fe.dens = dens_at(s, s->e_dens[1 - s->cd], g0);
    fe.ion_dens = dens_at(s, s->i_dens[i_acc], g0);
    fe.stot = s->stot_e;
    fe.time = s->clock + g0;
    fe.zero = (uint32_t *)dens_at(s, s->e_dens[s->cd], g0);
    fe.zero_words = (uint32_t)((size_t)(g1 - g0) * s->inst_stride * (s->f64 ? 2 : 1));
    fe.rng_lag = s->e_lag;
    fe.rng_next_lag = lag_out;
    fe.mask = s->e_mask;
    mask_mode(s, measure, lag_out, &fe.mask_read, &fe.mask_write);
    TRY(launch_inplace_ens(fe, s->sh, st));
    if (comp_e) TRY(launch_compact_ens(compact_job(s, g0, g1, true, s->e[s->ce], s->e[1 - s->ce], out_e, EC_C), st));
    return PIC_OK;
}

// the roles, the RNG write-back lag, the mask state and the host clock
// after a step
void end_step(pic_ensemble *s, bool measure, uint32_t lag_out) {
    bool mread, mwrite;
    mask_mode(s, measure, lag_out, &mread, &mwrite);
    s->mask_ok = lag_out == 0u ? mwrite : mread;
    s->e_lag = lag_out;
    const uint32_t t = s->t;
    const bool ion = (t % (uint32_t)s->p0.n_sub) == 0;
    const bool comp_e = (t % s->compact_every) == 0;
    if (ion && comp_e) s->ci ^= 1;
    if (comp_e) s->ce ^= 1;
    s->ac = comp_e ? 1 : 0;
    s->pe ^= 1;
    s->cd ^= 1;
    if (ion) s->hi ^= 1;
    s->t = (s->t + 1) % (uint32_t)s->p0.n_t;
    s->steps++;
}

// one step of every instance, group by group (each group's kernels advance
// its own device clock); `last`: the call's last step
int step(pic_ensemble *s, bool measure, bool last) {
    const uint32_t lag_out = next_lag(s, last);
    for (size_t g = 0; g + 1 < s->grp.size(); ++g) TRY(step_range(s, measure, s->grp[g], s->grp[g + 1], lag_out));
    end_step(s, measure, lag_out);
    return PIC_OK;
}

void invalidate_graphs(pic_ensemble *s) {
    for (auto &g : s->graphs)
        if (g.exec) {
            (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
        }
}

void get_roles(const pic_ensemble *s, int *r) {
    r[0] = s->ce;
    r[1] = s->ci;
    r[2] = s->cd;
    r[3] = s->hi;
    r[4] = s->pe;
    r[5] = s->ac;
    r[6] = s->mask_ok ? 1 : 0;
}

void set_roles(pic_ensemble *s, const int *r) {
    s->ce = r[0];
    s->ci = r[1];
    s->cd = r[2];
    s->hi = r[3];
    s->pe = r[4];
    s->ac = r[5];
    s->mask_ok = r[6] != 0;
}

// The graph of group g's 4*N_SUB steps starting now (t a multiple of
// 4*N_SUB), from the cache or captured; a period holds at most one
// compaction (at its first step).
int graph_for(pic_ensemble *s, bool measure, int g, pic_ensemble::GraphSlot **out) {
    int key[pic_ensemble::kRoles + 3];
    key[0] = measure ? 1 : 0;
    key[1] = (s->t % s->compact_every) == 0 ? 1 : 0;
    get_roles(s, key + 2);
    key[pic_ensemble::kRoles + 2] = g;
    for (auto &gs : s->graphs)
        if (gs.exec && memcmp(gs.key, key, sizeof(key)) == 0) {
            *out = &gs;
            return PIC_OK;
        }
    const int period = 4 * s->p0.n_sub;
    const uint32_t t0 = s->t;
    const unsigned long long st0 = s->steps;
    if (s->e_lag) {
        set_error("pic_ens: internal: a period starts behind on RNG words");
        return PIC_ERR_ARG;
    }
    hipGraph_t gr = nullptr;
    PIC_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = PIC_OK;
    for (int i = 0; i < period && rc == PIC_OK; ++i) {
        // the period's last step writes the words back (rng_defer divides it)
        const uint32_t lag_out = next_lag(s, false);
        rc = step_range(s, measure, s->grp[g], s->grp[g + 1], lag_out);
        end_step(s, measure, lag_out);
    }
    hipError_t e = hipStreamEndCapture(s->stream, &gr);
    // capture does not execute: remember the roles the period leaves, then
    // restore the host mirror
    int after[pic_ensemble::kRoles];
    get_roles(s, after);
    s->t = t0;
    s->steps = st0;
    s->e_lag = 0;
    set_roles(s, key + 2);
    if (rc != PIC_OK) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    pic_ensemble::GraphSlot *slot = nullptr;
    for (auto &gs : s->graphs)
        if (!gs.exec) {
            slot = &gs;
            break;
        }
    if (!slot) {
        slot = &s->graphs[s->graph_next];
        s->graph_next = (s->graph_next + 1) % (int)s->graphs.size();
        (void)hipGraphExecDestroy(slot->exec);
    }
    slot->exec = ex;
    memcpy(slot->key, key, sizeof(key));
    memcpy(slot->after, after, sizeof(after));
    *out = slot;
    return PIC_OK;
}

// Graph path: blocks of up to group_periods periods; each group runs the
// whole block (its graphs back to back on the stream) before the next
// group starts it from the same time and roles.  dry: only capture the
// graphs the blocks will replay (up to the first eager step), then restore
// the host state, so that a timed run does not include graph capture.
int run(pic_ensemble *s, uint32_t n, bool measure, bool dry = false) {
    if (s->need_prologue) TRY(prologue(s));
    const uint32_t period = 4u * (uint32_t)s->p0.n_sub;
    const bool graphs = s->use_graph && (uint32_t)s->p0.n_t % period == 0;
    if (dry) {
        int roles0[pic_ensemble::kRoles];
        get_roles(s, roles0);
        const uint32_t t0 = s->t;
        int rc = PIC_OK;
        for (int g = 0; g + 1 < (int)s->grp.size() && rc == PIC_OK && graphs; ++g) {
            set_roles(s, roles0);
            s->t = t0;
            for (uint32_t m = n; m >= period && s->t % period == 0 && rc == PIC_OK; m -= period) {
                pic_ensemble::GraphSlot *gs = nullptr;
                if ((rc = graph_for(s, measure, g, &gs)) != PIC_OK) break;
                set_roles(s, gs->after);
                s->t = (s->t + period) % (uint32_t)s->p0.n_t;
            }
        }
        set_roles(s, roles0);
        s->t = t0;
        return rc;
    }
    while (n > 0) {
        if (graphs && s->t % period == 0 && n >= period) {
            const uint32_t block = std::min(s->group_periods, n / period);
            const int n_groups = (int)s->grp.size() - 1;
            const int ns = std::min((int)s->work.size(), n_groups);
            int roles0[pic_ensemble::kRoles], roles1[pic_ensemble::kRoles];
            get_roles(s, roles0);
            const uint32_t t0 = s->t;
            const unsigned long long st0 = s->steps;
            if (ns > 1) {
                PIC_HIP(hipEventRecord(s->fork, s->stream));
                for (int q = 1; q < ns; ++q) PIC_HIP(hipStreamWaitEvent(s->work[q], s->fork, 0));
            }
            int rc = PIC_OK;
            for (int g = 0; g < n_groups && rc == PIC_OK; ++g) {
                set_roles(s, roles0);
                s->t = t0;
                for (uint32_t b = 0; b < block && rc == PIC_OK; ++b) {
                    pic_ensemble::GraphSlot *gs = nullptr;
                    rc = graph_for(s, measure, g, &gs);
                    if (rc != PIC_OK) break;
                    const hipError_t e = hipGraphLaunch(gs->exec, s->work[g % ns]);
                    if (e != hipSuccess) {
                        rc = hip_fail(e, "hipGraphLaunch");
                        break;
                    }
                    set_roles(s, gs->after);
                    s->t = (s->t + period) % (uint32_t)s->p0.n_t;
                }
                get_roles(s, roles1);
            }
            // the join is recorded whatever happened, so that the ensemble's
            // stream orders everything launched on the others
            for (int q = 1; q < ns; ++q) {
                PIC_HIP(hipEventRecord(s->join[q], s->work[q]));
                PIC_HIP(hipStreamWaitEvent(s->stream, s->join[q], 0));
            }
            if (rc != PIC_OK) return rc;
            set_roles(s, roles1);
            s->steps = st0 + (unsigned long long)block * period;
            n -= block * period;
            continue;
        }
        TRY(step(s, measure, n == 1));
        --n;
    }
    return PIC_OK;
}

// Live particles of one instance in slot order (and their slots relative
// to the instance's region); *n = the live count.
int gather_live(pic_ensemble *s, int inst, int species, float *x, float *vx, float *vy, float *vz,
                uint32_t *slots, uint32_t cap, uint32_t *n) {
    PIC_HIP(hipStreamSynchronize(s->stream));
    const bool el = species == PIC_ELECTRONS;
    uint32_t c = 0;
    PIC_HIP(hipMemcpy(&c, s->cnt + (size_t)inst * EC_WORDS + (el ? e_word(s) : i_word(s)), 4,
                      hipMemcpyDeviceToHost));
    const EnsInst &in = s->inst_h[inst];
    const uint32_t m = std::min(c, el ? in.e_cap : in.i_cap), off = el ? in.e_off : in.i_off;
    float **a = el ? s->e[s->ce] : s->ion[s->ci];
    float *dst[4] = {x, vx, vy, vz};
    std::vector<float> buf[4];
    for (int w = 0; w < 4; ++w)
        if (m && (w == 0 || dst[w])) {
            buf[w].resize(m);
            PIC_HIP(hipMemcpy(buf[w].data(), a[w] + off, 4ull * m, hipMemcpyDeviceToHost));
        }
    uint32_t total = 0;
    for (uint32_t i = 0; i < m; ++i) {
        if (!(buf[0][i] >= 0.0f)) continue;  // a hole
        if (total < cap) {
            for (int w = 0; w < 4; ++w)
                if (dst[w]) dst[w][total] = buf[w][i];
            if (slots) slots[total] = i;
        }
        ++total;
    }
    if (n) *n = total;
    return PIC_OK;
}

void free_all(pic_ensemble *s) {
    for (size_t q = 1; q < s->work.size(); ++q) (void)hipStreamSynchronize(s->work[q]);
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
    (void)hipFree(s->e_mask);
    void *bufs[] = {s->inst_d, s->cnt, s->clock, s->live, s->e_dens[0], s->e_dens[1], s->i_dens[0], s->i_dens[1],
                    s->pot, s->efield, s->cumul_e, s->cumul_i, s->i_sum, s->pot0, s->trig,
                    s->counters, s->cs, s->stot_e, s->stot_i};
    for (void *b : bufs) (void)hipFree(b);
    for (int i = 0; i < 2; ++i)
        if (s->ev[i]) (void)hipEventDestroy(s->ev[i]);
    for (size_t q = 1; q < s->work.size(); ++q) {
        (void)hipStreamDestroy(s->work[q]);
        if (s->join[q]) (void)hipEventDestroy(s->join[q]);
    }
    if (s->fork) (void)hipEventDestroy(s->fork);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

bool same_shared(const pic_params_t &a, const pic_params_t &b) {
    return a.n_g == b.n_g && a.n_t == b.n_t && a.n_sub == b.n_sub && a.freq == b.freq && a.gap == b.gap &&
           a.precision == b.precision;
}

}  // namespace

extern "C" {

int pic_ens_create(const pic_ensemble_config_t *c, pic_ensemble_t **out) {
    if (!c || !out || !c->params || !c->capacity || c->n_instances <= 0 || c->n_instances > 256) {
        set_error("pic_ens_create: 1..256 instances with params and capacities are required");
        return PIC_ERR_ARG;
    }
    for (int i = 0; i < c->n_instances; ++i)
        if (c->params[i].dt_e <= 0.0 || !same_shared(c->params[0], c->params[i]) || c->capacity[i] == 0) {
            set_error("pic_ens_create: instances must share n_g, n_t, n_sub, frequency, gap and precision, "
                      "have derived params and a capacity");
            return PIC_ERR_ARG;
        }
    pic_ensemble *s = new pic_ensemble();
    s->E = c->n_instances;
    s->p0 = c->params[0];
    s->params.assign(c->params, c->params + s->E);
    s->prec = s->p0.precision;
    s->f64 = s->prec == PIC_PREC_FP64;
    s->use_graph = c->use_graph != 0;
    Shape sh;
    sh.block = c->block_size > 0 ? c->block_size : 256;
    if (!particle_block_ok(sh.block)) {
        delete s;
        set_error("pic_ens_create: block_size must be 0 (256), 128, 256 or 512");
        return PIC_ERR_ARG;
    }
    sh.grid = c->grid_blocks > 0 ? c->grid_blocks : default_particle_grid(device_cus(), sh.block);
    sh.ppt_inplace = c->ppt > 0 ? c->ppt : 2;
    sh.lds_reps = std::max(0, c->lds_replicas);
    {
        const uint32_t period = 4u * (uint32_t)s->p0.n_sub;
        s->compact_every = c->compact_every ? c->compact_every : std::max(period, 800u / period * period);
        if (s->compact_every % period) {
            delete s;
            set_error("pic_ens_create: compact_every must be a multiple of 4*n_sub");
            return PIC_ERR_ARG;
        }
        if (c->cand_masks < -1 || c->cand_masks > 1) {
            delete s;
            set_error("pic_ens_create: cand_masks must be -1 (off), 0 or 1 (on)");
            return PIC_ERR_ARG;
        }
        s->rng_defer = c->rng_defer ? c->rng_defer : default_rng_defer(s->p0.n_sub);
        if (s->rng_defer > kMaxRngLag || period % s->rng_defer) {
            delete s;
            set_error("pic_ens_create: rng_defer must divide 4*n_sub and be at most 64");
            return PIC_ERR_ARG;
        }
    }
    // global density replicas: 1 or a multiple of 8 (read in groups of eight)
    sh.replicas = c->replicas > 0 ? c->replicas : 8;
    if (sh.replicas > 1) sh.replicas = std::min(64, (sh.replicas + 7) / 8 * 8);
    sh.rep_stride = ((s->p0.n_g + 31) / 32) * 32;
    s->sh = resolve(sh);
    s->inst_stride = s->sh.replicas * s->sh.rep_stride;
    s->inst_h.resize(s->E);
    uint64_t tot = 0;
    const KPhys k0 = make_kphys(s->params[0]);
    for (int i = 0; i < s->E; ++i) {
        EnsInst &in = s->inst_h[i];
        in.k = make_kphys(s->params[i]);
        in.k.stot_scale = (float)s->params[i].gas_density;
        in.v0 = s->params[i].v0;
        // the particle kernels take everything else from instance 0
        KPhys kc = in.k;
        kc.wf = k0.wf;
        kc.normal_range = k0.normal_range;
        kc.stot_scale = k0.stot_scale;
        if (memcmp(&kc, &k0, sizeof(KPhys)) != 0) {
            delete s;
            set_error("pic_ens_create: instances may differ in v0, pressure, gas temperature and weight only");
            return PIC_ERR_ARG;
        }
        // regions start on 256-slot boundaries (whole cache lines per tile)
        const uint32_t cap = c->capacity[i];
        in.e_off = in.i_off = (uint32_t)tot;
        in.e_cap = in.i_cap = cap;
        tot += ((uint64_t)cap + 255u) / 256u * 256u;
    }
    if (tot > kMaxSlots) {
        delete s;
        set_error("pic_ens_create: total capacity above 2^30 slots per species");
        return PIC_ERR_ARG;
    }
    s->e_total = s->i_total = (uint32_t)tot;
    {
        // groups of consecutive instances: as many as a greedy cut at
        // group_slots electron slots needs, with the cuts then placed where
        // the capacities are shared out most evenly (no oversized group and
        // small remainder)
        const uint64_t budget = c->group_slots ? c->group_slots : 9000000u;
        s->group_periods = c->group_periods ? c->group_periods : 5u;
        std::vector<uint64_t> pre(s->E + 1, 0);
        int n_groups = 1;
        uint64_t acc = 0;
        for (int i = 0; i < s->E; ++i) {
            const uint64_t cap = s->inst_h[i].e_cap;
            pre[i + 1] = pre[i] + cap;
            if (acc > 0 && acc + cap > budget) {
                ++n_groups;
                acc = 0;
            }
            acc += cap;
        }
        s->grp.push_back(0);
        for (int g = 1; g < n_groups; ++g) {
            const double target = (double)pre[s->E] * g / n_groups;
            int best = s->grp.back() + 1;
            for (int i = best; i <= s->E - (n_groups - g); ++i)
                if (std::fabs((double)pre[i] - target) < std::fabs((double)pre[best] - target)) best = i;
            s->grp.push_back(best);
        }
        s->grp.push_back(s->E);
        s->graphs.resize(8 * (s->grp.size() - 1));
    }
    const int n_streams = (int)std::min<uint32_t>(c->group_streams ? c->group_streams : 1u, 8u);
    auto fail = [&](int r) {
        free_all(s);
        delete s;
        return r;
    };
    int rc;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(hip_fail(hipGetLastError(), "hipStreamCreate"));
    s->work.assign(1, s->stream);
    s->join.assign(1, nullptr);
    for (int q = 1; q < std::min(n_streams, (int)s->grp.size() - 1); ++q) {
        hipStream_t w = nullptr;
        hipEvent_t ev = nullptr;
        if (hipStreamCreateWithFlags(&w, hipStreamNonBlocking) != hipSuccess)
            return fail(hip_fail(hipGetLastError(), "hipStreamCreate"));
        s->work.push_back(w);
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
            return fail(hip_fail(hipGetLastError(), "hipEventCreate"));
        s->join.push_back(ev);
    }
    if (s->work.size() > 1 && hipEventCreateWithFlags(&s->fork, hipEventDisableTiming) != hipSuccess)
        return fail(hip_fail(hipGetLastError(), "hipEventCreate"));
    for (int q = 0; q < 2; ++q)
        for (int k = 0; k < 4; ++k) {
            if ((rc = dalloc(&s->e[q][k], s->e_total))) return fail(rc);
            if ((rc = dalloc(&s->ion[q][k], s->i_total))) return fail(rc);
        }
    for (int k = 0; k < 4; ++k) {
        if ((rc = dalloc(&s->e_rng[k], s->e_total))) return fail(rc);
        if ((rc = dalloc(&s->i_rng[k], s->i_total))) return fail(rc);
    }
    if (c->cand_masks >= 0 && (rc = dalloc(&s->e_mask, s->e_total))) return fail(rc);
    if ((rc = dalloc(&s->inst_d, s->E))) return fail(rc);
    if ((rc = dalloc(&s->cnt, (size_t)s->E * EC_WORDS))) return fail(rc);
    if ((rc = dalloc(&s->clock, (size_t)s->E))) return fail(rc);
    if ((rc = dalloc(&s->live, 2 * (size_t)s->E))) return fail(rc);
    for (int q = 0; q < 2; ++q) {
        if ((rc = dalloc((double **)&s->e_dens[q], (size_t)s->E * s->inst_stride))) return fail(rc);
        if ((rc = dalloc((double **)&s->i_dens[q], (size_t)s->E * s->inst_stride))) return fail(rc);
    }
    const size_t gridw = (size_t)s->E * s->p0.n_g;
    if ((rc = dalloc(&s->pot, gridw))) return fail(rc);
    if ((rc = dalloc(&s->efield, gridw))) return fail(rc);
    if ((rc = dalloc(&s->cumul_e, gridw))) return fail(rc);
    if ((rc = dalloc(&s->cumul_i, gridw))) return fail(rc);
    if ((rc = dalloc(&s->i_sum, gridw))) return fail(rc);
    if ((rc = dalloc(&s->pot0, (size_t)s->E))) return fail(rc);
    if ((rc = dalloc(&s->trig, 2 * (size_t)s->p0.n_t))) return fail(rc);
    if ((rc = dalloc(&s->counters, (size_t)s->E * 8))) return fail(rc);
    if ((rc = dalloc(&s->cs, (size_t)PIC_N_CS * PIC_CS_RANGES))) return fail(rc);
    if ((rc = dalloc(&s->stot_e, PIC_CS_RANGES))) return fail(rc);
    if ((rc = dalloc(&s->stot_i, PIC_CS_RANGES))) return fail(rc);
    for (int i = 0; i < 2; ++i)
        if (hipEventCreate(&s->ev[i]) != hipSuccess) return fail(hip_fail(hipGetLastError(), "hipEventCreate"));
    {
        // summed cross sections (gas density 1): scaled per instance; each
        // instance's mask threshold is its largest collision probability
        std::vector<float> cs((size_t)PIC_N_CS * PIC_CS_RANGES), se(PIC_CS_RANGES), si(PIC_CS_RANGES);
        if ((rc = pic_cross_sections(1.0, cs.data(), se.data(), si.data()))) return fail(rc);
        if (hipMemcpy(s->cs, cs.data(), cs.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->stot_e, se.data(), se.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->stot_i, si.data(), si.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return fail(hip_fail(hipGetLastError(), "upload tables"));
        for (int i = 0; i < s->E; ++i)
            s->inst_h[i].coll_thr = coll_threshold(se.data(), (double)s->inst_h[i].k.stot_scale, s->params[i]);
    }
    if (hipMemcpy(s->inst_d, s->inst_h.data(), sizeof(EnsInst) * s->E, hipMemcpyHostToDevice) != hipSuccess)
        return fail(hip_fail(hipGetLastError(), "upload instances"));
    if ((rc = pic_seed_streams(c->seed_e, s->e_total, s->e_rng[0], s->e_rng[1], s->e_rng[2], s->e_rng[3],
                               s->stream)))
        return fail(rc);
    if ((rc = pic_seed_streams(c->seed_i, s->i_total, s->i_rng[0], s->i_rng[1], s->i_rng[2], s->i_rng[3],
                               s->stream)))
        return fail(rc);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(hip_fail(hipGetLastError(), "seed"));
    *out = s;
    return PIC_OK;
}

int pic_ens_destroy(pic_ensemble_t *s) {
    if (!s) return PIC_OK;
    (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return PIC_OK;
}

int pic_ens_upload(pic_ensemble_t *s, int inst, int species, const float *x, const float *vx,
                   const float *vy, const float *vz, uint32_t n) {
    if (!s || inst < 0 || inst >= s->E || (species != PIC_ELECTRONS && species != PIC_IONS) ||
        (n && (!x || !vx || !vy || !vz))) {
        set_error("pic_ens_upload: bad argument");
        return PIC_ERR_ARG;
    }
    const EnsInst &in = s->inst_h[inst];
    const uint32_t cap = species == PIC_ELECTRONS ? in.e_cap : in.i_cap;
    if (n > cap) {
        set_error("pic_ens_upload: more particles than the instance capacity");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(s->stream));
    float **a = species == PIC_ELECTRONS ? s->e[s->ce] : s->ion[s->ci];
    const uint32_t off = species == PIC_ELECTRONS ? in.e_off : in.i_off;
    const float *src[4] = {x, vx, vy, vz};
    for (int q = 0; q < 4; ++q)
        if (n) PIC_HIP(hipMemcpy(a[q] + off, src[q], 4ull * n, hipMemcpyHostToDevice));
    uint32_t *c = s->cnt + (size_t)inst * EC_WORDS;
    PIC_HIP(hipMemcpy(c + (species == PIC_ELECTRONS ? e_word(s) : i_word(s)), &n, 4, hipMemcpyHostToDevice));
    s->need_prologue = true;
    invalidate_graphs(s);
    return PIC_OK;
}

int pic_ens_download(pic_ensemble_t *s, int inst, int species, float *x, float *vx, float *vy, float *vz,
                     uint32_t cap, uint32_t *n) {
    if (!s || inst < 0 || inst >= s->E || (species != PIC_ELECTRONS && species != PIC_IONS)) return PIC_ERR_ARG;
    return gather_live(s, inst, species, x, vx, vy, vz, nullptr, cap, n);
}

int pic_ens_download_slots(pic_ensemble_t *s, int inst, int species, uint32_t *slots, uint32_t cap, uint32_t *n) {
    if (!s || inst < 0 || inst >= s->E || (species != PIC_ELECTRONS && species != PIC_IONS)) return PIC_ERR_ARG;
    return gather_live(s, inst, species, nullptr, nullptr, nullptr, nullptr, slots, cap, n);
}

int pic_ens_set_time(pic_ensemble_t *s, uint32_t t) {
    if (!s || t >= (uint32_t)s->p0.n_t) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    s->t = t;
    s->need_prologue = true;
    invalidate_graphs(s);
    return PIC_OK;
}

int pic_ens_rng_download(pic_ensemble_t *s, int inst, int species, uint32_t *w0, uint32_t *w1,
                         uint32_t *w2, uint32_t *w3, uint32_t n) {
    if (!s || inst < 0 || inst >= s->E || (species != PIC_ELECTRONS && species != PIC_IONS)) return PIC_ERR_ARG;
    const EnsInst &in = s->inst_h[inst];
    const uint32_t off = species == PIC_ELECTRONS ? in.e_off : in.i_off;
    const uint32_t cap = species == PIC_ELECTRONS ? in.e_cap : in.i_cap;
    if (n > cap) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    uint32_t **r = species == PIC_ELECTRONS ? s->e_rng : s->i_rng;
    uint32_t *dst[4] = {w0, w1, w2, w3};
    for (int q = 0; q < 4; ++q)
        if (dst[q] && n) PIC_HIP(hipMemcpy(dst[q], r[q] + off, 4ull * n, hipMemcpyDeviceToHost));
    return PIC_OK;
}

int pic_ens_counts(pic_ensemble_t *s, uint32_t *n_e, uint32_t *n_i) {
    if (!s) return PIC_ERR_ARG;
    // in-place steps leave holes: count the live slots on the device
    TRY(launch_fill(s->live, 0u, 2ull * s->E, s->stream));
    TRY(launch_count_live_ens(s->e[s->ce][0], s->inst_d, s->E, s->cnt, e_word(s), true, s->live, s->stream));
    TRY(launch_count_live_ens(s->ion[s->ci][0], s->inst_d, s->E, s->cnt, i_word(s), false, s->live + s->E,
                              s->stream));
    std::vector<uint32_t> c(2 * (size_t)s->E);
    PIC_HIP(hipMemcpyAsync(c.data(), s->live, 4 * c.size(), hipMemcpyDeviceToHost, s->stream));
    PIC_HIP(hipStreamSynchronize(s->stream));
    for (int i = 0; i < s->E; ++i) {
        if (n_e) n_e[i] = c[i];
        if (n_i) n_i[i] = c[s->E + i];
    }
    return PIC_OK;
}

int pic_ens_run(pic_ensemble_t *s, uint32_t n_steps, int measure) {
    if (!s) return PIC_ERR_ARG;
    return run(s, n_steps, measure != 0);
}

int pic_ens_synchronize(pic_ensemble_t *s) {
    if (!s) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    return PIC_OK;
}

int pic_ens_check(pic_ensemble_t *s) {
    if (!s) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    std::vector<uint32_t> c((size_t)s->E * EC_WORDS);
    PIC_HIP(hipMemcpy(c.data(), s->cnt, 4 * c.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < s->E; ++i)
        if (c[(size_t)i * EC_WORDS + EC_OVERFLOW]) {
            set_error("ensemble instance " + std::to_string(i) + ": capacity exceeded (particles dropped)");
            return PIC_ERR_CAPACITY;
        }
    return PIC_OK;
}

int pic_ens_time_steps(pic_ensemble_t *s, uint32_t n_steps, int measure, float *ms_total,
                       unsigned long long *particle_steps) {
    if (!s) return PIC_ERR_ARG;
    if (s->need_prologue) TRY(prologue(s));
    std::vector<uint32_t> ne0(s->E), ni0(s->E), ne1(s->E), ni1(s->E);
    TRY(pic_ens_counts(s, ne0.data(), ni0.data()));
    uint32_t ion_steps = 0;
    for (uint32_t i = 0, tt = s->t; i < n_steps; ++i, tt = (tt + 1) % (uint32_t)s->p0.n_t)
        ion_steps += (tt % (uint32_t)s->p0.n_sub) == 0;
    TRY(run(s, n_steps, measure != 0, true));
    PIC_HIP(hipEventRecord(s->ev[0], s->stream));
    TRY(run(s, n_steps, measure != 0));
    PIC_HIP(hipEventRecord(s->ev[1], s->stream));
    PIC_HIP(hipEventSynchronize(s->ev[1]));
    float ms = 0.0f;
    PIC_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
    TRY(pic_ens_counts(s, ne1.data(), ni1.data()));
    double ps = 0.0;
    for (int i = 0; i < s->E; ++i)
        ps += 0.5 * ((double)ne0[i] + ne1[i]) * n_steps + 0.5 * ((double)ni0[i] + ni1[i]) * ion_steps;
    if (ms_total) *ms_total = ms;
    if (particle_steps) *particle_steps = (unsigned long long)ps;
    return PIC_OK;
}

int pic_ens_read_grid(pic_ensemble_t *s, int inst, float *pot, float *efield) {
    if (!s || inst < 0 || inst >= s->E) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    const size_t n = s->p0.n_g, off = (size_t)inst * n;
    if (pot) PIC_HIP(hipMemcpy(pot, s->pot + off, 4 * n, hipMemcpyDeviceToHost));
    if (efield) PIC_HIP(hipMemcpy(efield, s->efield + off, 4 * n, hipMemcpyDeviceToHost));
    return PIC_OK;
}

int pic_ens_read_measurements(pic_ensemble_t *s, int inst, double *cumul_e, double *cumul_i,
                              unsigned long long *counters) {
    if (!s || inst < 0 || inst >= s->E) return PIC_ERR_ARG;
    PIC_HIP(hipStreamSynchronize(s->stream));
    const size_t n = s->p0.n_g, off = (size_t)inst * n;
    if (cumul_e) PIC_HIP(hipMemcpy(cumul_e, s->cumul_e + off, 8 * n, hipMemcpyDeviceToHost));
    if (cumul_i) PIC_HIP(hipMemcpy(cumul_i, s->cumul_i + off, 8 * n, hipMemcpyDeviceToHost));
    if (counters) PIC_HIP(hipMemcpy(counters, s->counters + (size_t)inst * 8, 64, hipMemcpyDeviceToHost));
    return PIC_OK;
}

}  // extern "C"
