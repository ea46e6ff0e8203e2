// pic_step.hip -- the RF step on caller-owned buffers: the reference's
// GpuSimState (Rust-CUDA/src/main.rs:132-185) stays the caller's, and one
// call runs the loop body of main.rs:1725-1808 on the fused kernels (one
// Poisson solve and one particle kernel per species per step).
//
// What the fused step keeps between steps and the reference's state has no
// field for -- the next step's density accumulators (replicas, two
// parities per species), the frozen watermarks of the in-place kernels, the
// device clock and the pot0 of the period, the overflow flag and the
// diagnostic counters -- lives in a workspace.  pic_init allocates it (with
// a cache of captured period graphs); pic_step_noalloc carves it out of a
// fixed device array instead, for callers that cannot allocate.
//
// Counter protocol (one segment, the caller's arrays): *n_electrons is the
// watermark the electron kernel appends newborns after; the Poisson solve of
// each step copies it to the workspace word the kernel reads as its loop
// bound (and clears it on compacting steps, where the compaction re-counts
// the survivors into it).  Ions likewise on compacting ion steps.  Which
// density buffer a step reads and which it fills follow from the step index
// t alone (n_t a multiple of 2*n_sub), so no host bookkeeping is needed
// beyond the caller's t.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "pic_internal.h"

using namespace pic;

namespace {

enum StepCnt { SC_E = 0, SC_I = 1, SC_OVERFLOW = 2, SC_TIME = 3, SC_P0 = 4 /* double: words 4-5 */, SC_WORDS = 64 };
constexpr int kStepReplicas = 8;
constexpr int kMaxNt = 32768;
constexpr size_t kAlign = 256;

// Offsets (bytes) of the workspace arrays for a discharge.
struct StepLayout {
    size_t cnt, dens[4], grid_e, grid_i, tab, counters, ifed, total;
    int rep_stride, dens_words;
};

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

StepLayout step_layout(const pic_params_t &p) {
    StepLayout L;
    L.rep_stride = (p.n_g + 31) / 32 * 32;
    L.dens_words = kStepReplicas * L.rep_stride;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes);
        return at;
    };
    L.cnt = take(SC_WORDS * sizeof(uint32_t));
    for (int b = 0; b < 4; ++b) L.dens[b] = take((size_t)L.dens_words * sizeof(double));
    L.grid_e = take((size_t)p.n_g * sizeof(double));
    L.grid_i = take((size_t)p.n_g * sizeof(double));
    L.tab = take((size_t)p.n_t * sizeof(double));
    L.counters = take((size_t)kCounterLanes * kCounterWords * sizeof(unsigned long long));
    L.ifed = take(2 * PIC_N_IFED * sizeof(uint32_t));
    L.total = o;
    return L;
}

// Everything a step launches with, derived from the parameters and a
// workspace base address (plain data).
struct StepCtx {
    pic_params_t p;
    KPhys k;
    Shape sh;
    int prec = PIC_PREC_FP64;
    bool f64 = true;
    uint32_t compact_every = 800;
    uint32_t rng_defer = 8;  // steps per deferred RNG write-back run
    uint32_t *cnt = nullptr;
    void *e_dens[2] = {}, *i_dens[2] = {};
    double *grid_e = nullptr, *grid_i = nullptr, *tab = nullptr;
    unsigned long long *counters = nullptr;  // null: no diagnostics
    uint32_t *ifed = nullptr;
    int dens_words32 = 0;
};

int check_params(const pic_params_t *p) {
    if (!p || p->dt_e <= 0.0 || p->n_g < 3 || p->n_g > 4096) {
        set_error("pic_step: params not derived (pic_params_derive) or n_g outside [3, 4096]");
        return PIC_ERR_ARG;
    }
    if (p->n_sub < 1 || p->n_t < 2 || p->n_t > kMaxNt || p->n_t % (2 * p->n_sub) != 0) {
        set_error("pic_step: n_t must be a multiple of 2*n_sub and at most 32768");
        return PIC_ERR_ARG;
    }
    if (p->precision < PIC_PREC_FP32 || p->precision > PIC_PREC_FP64) {
        set_error("pic_step: unknown precision");
        return PIC_ERR_ARG;
    }
    return PIC_OK;
}

uint32_t compact_every(const pic_params_t &p, const pic_step_config_t *c) {
    const uint32_t period = 4u * (uint32_t)p.n_sub;
    return c && c->compact_every ? c->compact_every : std::max(period, 800u / period * period);
}

int check_config(const pic_params_t &p, const pic_step_config_t *c) {
    if (compact_every(p, c) % (4u * (uint32_t)p.n_sub)) {
        set_error("pic_step: compact_every must be a multiple of 4*n_sub");
        return PIC_ERR_ARG;
    }
    const uint32_t period = 4u * (uint32_t)p.n_sub;
    if (c && c->rng_defer && (c->rng_defer > kMaxRngLag || period % c->rng_defer)) {
        set_error("pic_step: rng_defer must divide 4*n_sub and be at most 64");
        return PIC_ERR_ARG;
    }
    return PIC_OK;
}

int make_ctx(const pic_params_t &p, const pic_step_config_t *c, uint8_t *ws, bool diagnostics, StepCtx *x) {
    const int rc = check_config(p, c);
    if (rc != PIC_OK) return rc;
    const StepLayout L = step_layout(p);
    x->p = p;
    x->k = make_kphys(p);
    x->prec = p.precision;
    x->f64 = p.precision == PIC_PREC_FP64;
    x->compact_every = compact_every(p, c);
    x->rng_defer = c && c->rng_defer ? c->rng_defer : default_rng_defer(p.n_sub);
    Shape sh;
    sh.block = 256;
    sh.grid = c && c->grid_blocks > 0 ? c->grid_blocks : default_particle_grid(device_cus(), sh.block);
    sh.ppt = c && c->ppt > 0 ? c->ppt : 4;
    sh.ppt_inplace = c && c->ppt_inplace > 0 ? c->ppt_inplace : 2;
    sh.lds_reps = c ? std::max(0, c->lds_replicas) : 0;
    sh.replicas = kStepReplicas;
    sh.rep_stride = L.rep_stride;
    x->sh = resolve(sh);
    x->cnt = (uint32_t *)(ws + L.cnt);
    x->e_dens[0] = ws + L.dens[0];
    x->e_dens[1] = ws + L.dens[1];
    x->i_dens[0] = ws + L.dens[2];
    x->i_dens[1] = ws + L.dens[3];
    x->grid_e = (double *)(ws + L.grid_e);
    x->grid_i = (double *)(ws + L.grid_i);
    x->tab = (double *)(ws + L.tab);
    x->counters = diagnostics ? (unsigned long long *)(ws + L.counters) : nullptr;
    x->ifed = diagnostics ? (uint32_t *)(ws + L.ifed) : nullptr;
    x->dens_words32 = L.dens_words * (x->f64 ? 2 : 1);
    return PIC_OK;
}

int check_state(const pic_state_t *s, int flags) {
    if (!s || !s->e_x || !s->e_vx || !s->e_vy || !s->e_vz || !s->i_x || !s->i_vx || !s->i_vy || !s->i_vz ||
        !s->n_electrons || !s->n_ions || !s->efield || !s->pot || !s->cs || !s->sigma_tot_e || !s->sigma_tot_i ||
        !s->rng_e0 || !s->rng_e1 || !s->rng_e2 || !s->rng_e3 || !s->rng_i0 || !s->rng_i1 || !s->rng_i2 ||
        !s->rng_i3 || !s->tmp_x || !s->tmp_vx || !s->tmp_vy || !s->tmp_vz || !s->tmp_i_x || !s->tmp_i_vx ||
        !s->tmp_i_vy || !s->tmp_i_vz || s->capacity == 0) {
        set_error("pic_step: a required buffer of the state is null or its capacity is 0");
        return PIC_ERR_ARG;
    }
    if (s->capacity > kMaxSlots) {
        set_error("pic_step: capacity above 2^30 slots per species");
        return PIC_ERR_ARG;
    }
    if ((s->e_density == nullptr) != (s->i_density == nullptr)) {
        set_error("pic_step: give both e_density and i_density or neither");
        return PIC_ERR_ARG;
    }
    if ((flags & PIC_STEP_MEASURE) && (!s->cumul_e_density || !s->cumul_i_density || !s->eepf_counts)) {
        set_error("pic_step: PIC_STEP_MEASURE needs cumul_e_density, cumul_i_density and eepf_counts");
        return PIC_ERR_ARG;
    }
    return PIC_OK;
}

// density buffer parities of step t: the electron density step t reads
// (the previous step filled it) and the held ion buffer (flips at every ion
// step); both return to their start after n_t steps when n_t is a multiple
// of 2*n_sub
int e_parity(uint32_t t) { return (int)(t & 1u); }
int i_parity(const StepCtx &x, uint32_t t) {
    const uint32_t n = (uint32_t)x.p.n_sub;
    return (int)(((t + n - 1u) / n) & 1u);
}

double *grid_e(const StepCtx &x, const pic_state_t &s) { return s.e_density ? s.e_density : x.grid_e; }
double *grid_i(const StepCtx &x, const pic_state_t &s) { return s.i_density ? s.i_density : x.grid_i; }

PoissonJob poisson_job(const StepCtx &x, const pic_state_t &s) {
    PoissonJob j;
    const pic_params_t &p = x.p;
    const double E = 1.60217662e-19, EPS0 = 8.85418781e-12;
    j.precision = x.prec;
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
    j.pot = s.pot;
    j.efield = s.efield;
    j.dens_e_out = grid_e(x, s);
    j.dens_i_out = grid_i(x, s);
    j.replicas = x.sh.replicas;
    j.rep_stride = x.sh.rep_stride;
    j.time = x.cnt + SC_TIME;
    return j;
}

// Deposit the caller's particles into the accumulators step t reads, set
// the clock to t and tabulate pot0 over the period (the fused path's
// prologue, on the caller's counts).
int prologue(const StepCtx &x, const pic_state_t &s, uint32_t t, hipStream_t st) {
    const int cd = e_parity(t), hi = i_parity(x, t);
    int rc = PIC_OK;
    for (int b = 0; b < 2 && rc == PIC_OK; ++b) {
        rc = launch_fill(x.e_dens[b], 0u, (uint64_t)x.dens_words32, st);
        if (rc == PIC_OK) rc = launch_fill(x.i_dens[b], 0u, (uint64_t)x.dens_words32, st);
    }
    if (rc == PIC_OK) rc = launch_fill(x.cnt + SC_TIME, t, 1, st);
    if (rc == PIC_OK)
        rc = launch_deposit(s.e_x, s.n_electrons, x.e_dens[cd], x.f64, x.k, x.sh, st, s.capacity);
    for (int b = 0; b < 2 && rc == PIC_OK; ++b)
        rc = launch_deposit(s.i_x, s.n_ions, x.i_dens[b], x.f64, x.k, x.sh, st, s.capacity);
    if (rc != PIC_OK) return rc;
    PoissonJob pj = poisson_job(x, s);
    rc = launch_pot0_table(pj, x.tab, st);
    if (rc == PIC_OK) rc = launch_pot0(pj, (double *)(x.cnt + SC_P0), st);
    if (rc == PIC_OK) rc = launch_sum_replicas(pj, x.i_dens[hi], grid_i(x, s), st);
    return rc;
}

void swap4(float *&a0, float *&a1, float *&a2, float *&a3, float *&b0, float *&b1, float *&b2, float *&b3) {
    std::swap(a0, b0);
    std::swap(a1, b1);
    std::swap(a2, b2);
    std::swap(a3, b3);
}

// Deferred RNG write-back of the electrons: the lag (draws the stored RNG
// words of live slots are behind their streams) after step t, given the lag
// of step t.  Runs of rng_defer steps, aligned to multiples of rng_defer in
// the cycle; the step before a compaction, the last step of the cycle and
// the last step of a call write the words back (lag 0), so every call
// starts and ends with exact words and every captured period (t a multiple
// of 4*n_sub, which rng_defer divides) starts and ends at lag 0.
uint32_t next_lag(const StepCtx &x, uint32_t t, uint32_t lag, bool last) {
    if (last || t % x.compact_every == 0 || (t + 1u) % x.rng_defer == 0 || t + 1u == (uint32_t)x.p.n_t) return 0;
    return lag + 1u;
}

// One RF step at cycle index t (main.rs:1725-1808): Poisson from the
// accumulated densities, then the ion kernel (ion steps) and the electron
// kernel, in place or compacting (t % compact_every == 0, both species).
// by_value: the solve takes t and pot0_tab[t] from the arguments (eager
// calls); otherwise from the device clock (captured periods, replayed at
// any t of the same phase).  Both advance the clock.  lag / lag_out: the
// electrons' deferred RNG write-back on entry and exit (0 on compacting
// steps).
int step_one(const StepCtx &x, pic_state_t &s, uint32_t t, bool measure, bool by_value, hipStream_t st,
             uint32_t lag = 0, uint32_t lag_out = 0) {
    const bool ion = (t % (uint32_t)x.p.n_sub) == 0;
    const bool comp_e = (t % x.compact_every) == 0;
    const bool comp_i = ion && comp_e;
    const int cd = e_parity(t), hi = i_parity(x, t);
    const int i_read = ion ? 1 - hi : hi;
    const int i_acc = ion ? hi : 1 - hi;
    uint32_t *const w_e = x.cnt + SC_E, *const w_i = x.cnt + SC_I;
    PoissonJob pj = poisson_job(x, s);
    pj.e_dens = x.e_dens[cd];
    pj.i_dens = x.i_dens[i_read];
    pj.advance_time = x.cnt + SC_TIME;
    pj.pot0_in = (const double *)(x.cnt + SC_P0);
    pj.pot0_next = (double *)(x.cnt + SC_P0);
    pj.pot0_tab = x.tab;
    if (by_value) pj.t_value = t;
    if (!ion) pj.i_sum = grid_i(x, s);
    if (ion) {
        pj.zero_b = x.i_dens[hi];
        pj.zero_b_words = x.dens_words32;
    }
    // the electron kernel's loop bound is a frozen copy of the watermark;
    // a compaction re-counts the survivors into the watermark
    pj.copy_cnt_src = s.n_electrons;
    pj.copy_cnt_dst = w_e;
    pj.copy_cnt_words = 1;
    if (comp_e) {
        pj.zero_cnt = s.n_electrons;
        pj.zero_cnt_words = 1;
    }
    if (comp_i) {
        pj.copy_cnt_ion_src = s.n_ions;
        pj.copy_cnt_ion_dst = w_i;
        pj.zero_cnt_ion = s.n_ions;
        pj.zero_cnt_ion_words = 1;
    }
    if (measure) {
        pj.cumul_e = s.cumul_e_density;
        pj.cumul_i = s.cumul_i_density;
        pj.meas_steps = x.counters ? x.counters + 6 : nullptr;
    }
    int rc = launch_poisson(pj, st);
    if (rc != PIC_OK) return rc;
    FusedJob base;
    memset(&base, 0, sizeof(base));
    base.efield = s.efield;
    base.n_seg = 1;
    base.seg_size = s.capacity;
    base.dens_f64 = x.f64;
    base.cs = s.cs;
    base.overflow = x.cnt + SC_OVERFLOW;
    base.counters = x.counters;
    if (ion) {
        FusedJob fi = base;
        fi.electrons = false;
        fi.inplace = !comp_i;
        fi.sx = s.i_x;
        fi.svx = s.i_vx;
        fi.svy = s.i_vy;
        fi.svz = s.i_vz;
        fi.cnt_in = comp_i ? w_i : s.n_ions;
        fi.dx = s.tmp_i_x;
        fi.dvx = s.tmp_i_vx;
        fi.dvy = s.tmp_i_vy;
        fi.dvz = s.tmp_i_vz;
        fi.cnt_out = s.n_ions;
        fi.rng0 = s.rng_i0;
        fi.rng1 = s.rng_i1;
        fi.rng2 = s.rng_i2;
        fi.rng3 = s.rng_i3;
        fi.dens = x.i_dens[i_acc];
        fi.stot = s.sigma_tot_i;
        fi.ifed = x.ifed;
        if ((rc = launch_fused(fi, x.k, x.sh, st)) != PIC_OK) return rc;
        // main.rs:1780-1782: the compacted ions live in the other arrays
        if (comp_i) swap4(s.i_x, s.i_vx, s.i_vy, s.i_vz, s.tmp_i_x, s.tmp_i_vx, s.tmp_i_vy, s.tmp_i_vz);
    }
    FusedJob fe = base;
    fe.electrons = true;
    fe.inplace = !comp_e;
    fe.sx = s.e_x;
    fe.svx = s.e_vx;
    fe.svy = s.e_vy;
    fe.svz = s.e_vz;
    fe.cnt_in = w_e;
    fe.dx = s.tmp_x;
    fe.dvx = s.tmp_vx;
    fe.dvy = s.tmp_vy;
    fe.dvz = s.tmp_vz;
    fe.cnt_out = s.n_electrons;
    fe.rng0 = s.rng_e0;
    fe.rng1 = s.rng_e1;
    fe.rng2 = s.rng_e2;
    fe.rng3 = s.rng_e3;
    fe.ix = s.i_x;
    fe.ivx = s.i_vx;
    fe.ivy = s.i_vy;
    fe.ivz = s.i_vz;
    fe.ion_cnt = s.n_ions;
    fe.ion_seg_size = s.capacity;
    fe.ion_dens = x.i_dens[i_acc];
    fe.dens = x.e_dens[1 - cd];
    fe.stot = s.sigma_tot_e;
    fe.eepf = measure ? s.eepf_counts : nullptr;
    fe.zero = (uint32_t *)x.e_dens[cd];
    fe.zero_words = (uint32_t)x.dens_words32;
    if (comp_e && (lag || lag_out)) {
        set_error("pic_step: a compacting step starts and ends with exact RNG words");
        return PIC_ERR_ARG;
    }
    fe.rng_lag = lag;
    fe.rng_next_lag = lag_out;
    if ((rc = launch_fused(fe, x.k, x.sh, st)) != PIC_OK) return rc;
    // main.rs:1764-1767
    if (comp_e) swap4(s.e_x, s.e_vx, s.e_vy, s.e_vz, s.tmp_x, s.tmp_vx, s.tmp_vy, s.tmp_vz);
    return PIC_OK;
}

int steps_eager(const StepCtx &x, pic_state_t *s, uint32_t t, uint32_t n, int flags, hipStream_t st) {
    const bool measure = (flags & PIC_STEP_MEASURE) != 0;
    if (flags & PIC_STEP_RESET) {
        const int rc = prologue(x, *s, t, st);
        if (rc != PIC_OK) return rc;
    }
    uint32_t lag = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t lag_out = next_lag(x, t, lag, i + 1 == n);
        const int rc = step_one(x, *s, t, measure, true, st, lag, lag_out);
        if (rc != PIC_OK) return rc;
        lag = lag_out;
        t = (t + 1u) % (uint32_t)x.p.n_t;
    }
    return PIC_OK;
}

constexpr size_t kScratchBytes = 1536u * 1024u;
constexpr int kMaxDevices = 64;

}  // namespace

// ------------------------------------------------------------ no-allocation workspace
// The fixed device array behind pic_step_noalloc (each device holds its own
// copy until the process ends).  Its name is a host-side stand-in, not a
// device address: host code uses only the per-device address
// hipGetSymbolAddress returns.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

int scratch_ptr(uint8_t **out) {
    static std::mutex mu;
    static uint8_t *per_dev[kMaxDevices] = {};
    int dev = 0;
    PIC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("pic_step_noalloc: device index beyond the table of workspaces");
        return PIC_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!per_dev[dev]) {
        void *p = nullptr;
        PIC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        per_dev[dev] = (uint8_t *)p;
    }
    *out = per_dev[dev];
    return PIC_OK;
}

}  // namespace

struct pic_ctx {
    StepCtx x;
    pic_params_t p;
    uint8_t *ws = nullptr;
    size_t ws_bytes = 0;
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        pic_state_t key{};
        int measure = 0, comp = 0;
    } graphs[8];
    int graph_next = 0;
};

namespace {

void drop_graphs(pic_ctx *c) {
    for (auto &g : c->graphs)
        if (g.exec) {
            (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
        }
}

// The captured 4*n_sub steps from t (a multiple of 4*n_sub) for this state
// and flags: the density parities at the start of such a period are always
// the same, so a graph depends on the buffers, the measure flag and whether
// the period starts with a compaction.
int period_graph(pic_ctx *c, const pic_state_t &s, uint32_t t, bool measure, hipStream_t st, hipGraphExec_t *out) {
    const int comp = (t % c->x.compact_every) == 0;
    for (auto &g : c->graphs)
        if (g.exec && g.measure == (int)measure && g.comp == comp && memcmp(&g.key, &s, sizeof(s)) == 0) {
            *out = g.exec;
            return PIC_OK;
        }
    const uint32_t period = 4u * (uint32_t)c->p.n_sub;
    pic_state_t tmp = s;
    hipGraph_t g = nullptr;
    PIC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = PIC_OK;
    uint32_t lag = 0;
    for (uint32_t i = 0; i < period && rc == PIC_OK; ++i) {
        const uint32_t lag_out = next_lag(c->x, t + i, lag, false);
        rc = step_one(c->x, tmp, t + i, measure, false, st, lag, lag_out);
        lag = lag_out;
    }
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != PIC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) return hip_fail(e2, "hipGraphInstantiate");
    pic_ctx::GraphSlot *slot = nullptr;
    for (auto &gs : c->graphs)
        if (!gs.exec) {
            slot = &gs;
            break;
        }
    if (!slot) {
        slot = &c->graphs[c->graph_next];
        c->graph_next = (c->graph_next + 1) % 8;
        (void)hipGraphExecDestroy(slot->exec);
    }
    slot->exec = ex;
    slot->key = s;
    slot->measure = measure;
    slot->comp = comp;
    *out = ex;
    return PIC_OK;
}

}  // namespace

extern "C" {

int pic_init(const pic_params_t *p, const pic_step_config_t *cfg, pic_ctx_t **out) {
    if (!out) {
        set_error("pic_init: null output");
        return PIC_ERR_ARG;
    }
    int rc = check_params(p);
    if (rc == PIC_OK) rc = check_config(*p, cfg);
    if (rc != PIC_OK) return rc;
    pic_ctx *c = new pic_ctx();
    c->p = *p;
    const StepLayout L = step_layout(*p);
    c->ws_bytes = L.total;
    if (hipMalloc((void **)&c->ws, L.total) != hipSuccess) {
        delete c;
        return hip_fail(hipGetLastError(), "pic_init: hipMalloc");
    }
    if (hipMemset(c->ws, 0, L.total) != hipSuccess || (rc = make_ctx(*p, cfg, c->ws, true, &c->x)) != PIC_OK) {
        (void)hipFree(c->ws);
        delete c;
        return rc != PIC_OK ? rc : hip_fail(hipGetLastError(), "pic_init: hipMemset");
    }
    *out = c;
    return PIC_OK;
}

int pic_destroy(pic_ctx_t *c) {
    if (!c) return PIC_OK;
    drop_graphs(c);
    (void)hipFree(c->ws);
    delete c;
    return PIC_OK;
}

int pic_step(pic_ctx_t *c, pic_state_t *s, uint32_t t, int flags, pic_stream_t stream) {
    if (!c) {
        set_error("pic_step: null context");
        return PIC_ERR_ARG;
    }
    int rc = check_state(s, flags);
    if (rc != PIC_OK) return rc;
    if (t >= (uint32_t)c->p.n_t) {
        set_error("pic_step: t beyond the RF cycle");
        return PIC_ERR_ARG;
    }
    return steps_eager(c->x, s, t, 1, flags, (hipStream_t)stream);
}

int pic_step_graph(pic_ctx_t *c, pic_state_t *s, uint32_t t, uint32_t n_steps, int flags, pic_stream_t stream) {
    if (!c) {
        set_error("pic_step_graph: null context");
        return PIC_ERR_ARG;
    }
    int rc = check_state(s, flags);
    if (rc != PIC_OK) return rc;
    if (t >= (uint32_t)c->p.n_t) {
        set_error("pic_step_graph: t beyond the RF cycle");
        return PIC_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    const bool measure = (flags & PIC_STEP_MEASURE) != 0;
    const uint32_t period = 4u * (uint32_t)c->p.n_sub;
    if (flags & PIC_STEP_RESET) {
        if ((rc = prologue(c->x, *s, t, st)) != PIC_OK) return rc;
    } else if (st && t % period == 0 && n_steps >= period) {
        // a replayed period reads the clock and the pot0 word: set them
        // for t (an eager step leaves them set for the step after it)
        if ((rc = launch_fill(c->x.cnt + SC_TIME, t, 1, st)) != PIC_OK) return rc;
        PoissonJob pj = poisson_job(c->x, *s);
        if ((rc = launch_pot0(pj, (double *)(c->x.cnt + SC_P0), st)) != PIC_OK) return rc;
    }
    uint32_t lag = 0;  // deferred RNG write-back of the eager steps
    while (n_steps > 0) {
        if (st && t % period == 0 && n_steps >= period) {
            // the step before t wrote the RNG words back (4*n_sub is a
            // multiple of rng_defer), as the captured period assumes
            if (lag) {
                set_error("pic_step_graph: internal: a period starts behind on RNG words");
                return PIC_ERR_ARG;
            }
            hipGraphExec_t ex = nullptr;
            if ((rc = period_graph(c, *s, t, measure, st, &ex)) != PIC_OK) return rc;
            const hipError_t e = hipGraphLaunch(ex, st);
            if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
            if (t % c->x.compact_every == 0) {
                swap4(s->e_x, s->e_vx, s->e_vy, s->e_vz, s->tmp_x, s->tmp_vx, s->tmp_vy, s->tmp_vz);
                swap4(s->i_x, s->i_vx, s->i_vy, s->i_vz, s->tmp_i_x, s->tmp_i_vx, s->tmp_i_vy, s->tmp_i_vz);
            }
            t = (t + period) % (uint32_t)c->p.n_t;
            n_steps -= period;
            continue;
        }
        const uint32_t lag_out = next_lag(c->x, t, lag, n_steps == 1);
        if ((rc = step_one(c->x, *s, t, measure, true, st, lag, lag_out)) != PIC_OK) return rc;
        lag = lag_out;
        t = (t + 1u) % (uint32_t)c->p.n_t;
        --n_steps;
    }
    return PIC_OK;
}

int pic_step_lagged(pic_ctx_t *c, pic_state_t *s, uint32_t t, int flags, uint32_t rng_lag, uint32_t rng_lag_out,
                    pic_stream_t stream) {
    if (!c) {
        set_error("pic_step_lagged: null context");
        return PIC_ERR_ARG;
    }
    int rc = check_state(s, flags);
    if (rc != PIC_OK) return rc;
    if (t >= (uint32_t)c->p.n_t) {
        set_error("pic_step_lagged: t beyond the RF cycle");
        return PIC_ERR_ARG;
    }
    if (rng_lag >= kMaxRngLag || (rng_lag_out != 0 && rng_lag_out != rng_lag + 1)) {
        set_error("pic_step_lagged: lag below 64, lag_out 0 or lag + 1");
        return PIC_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    if ((flags & PIC_STEP_RESET) && (rc = prologue(c->x, *s, t, st)) != PIC_OK) return rc;
    return step_one(c->x, *s, t, (flags & PIC_STEP_MEASURE) != 0, true, st, rng_lag, rng_lag_out);
}

int pic_step_diagnostics(pic_ctx_t *c, unsigned long long counters[8], uint32_t *ifed, uint32_t *overflow,
                         pic_stream_t stream) {
    if (!c) {
        set_error("pic_step_diagnostics: null context");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (counters) {
        unsigned long long h[kCounterLanes * kCounterWords];
        PIC_HIP(hipMemcpy(h, c->x.counters, sizeof(h), hipMemcpyDeviceToHost));
        for (int w = 0; w < kCounterWords; ++w) {
            counters[w] = 0;
            for (int l = 0; l < kCounterLanes; ++l) counters[w] += h[l * kCounterWords + w];
        }
    }
    if (ifed) PIC_HIP(hipMemcpy(ifed, c->x.ifed, 2 * PIC_N_IFED * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (overflow) PIC_HIP(hipMemcpy(overflow, c->x.cnt + SC_OVERFLOW, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return PIC_OK;
}

int pic_step_noalloc(const pic_params_t *p, pic_state_t *s, uint32_t t, int flags, pic_stream_t stream) {
    int rc = check_params(p);
    if (rc != PIC_OK) return rc;
    if ((rc = check_state(s, flags)) != PIC_OK) return rc;
    if (t >= (uint32_t)p->n_t) {
        set_error("pic_step_noalloc: t beyond the RF cycle");
        return PIC_ERR_ARG;
    }
    if (step_layout(*p).total > kScratchBytes) {
        set_error("pic_step_noalloc: n_g / n_t need more than the 1.5 MiB no-allocation workspace");
        return PIC_ERR_ARG;
    }
    uint8_t *ws = nullptr;
    if ((rc = scratch_ptr(&ws)) != PIC_OK) return rc;
    StepCtx x;
    if ((rc = make_ctx(*p, nullptr, ws, false, &x)) != PIC_OK) return rc;
    return steps_eager(x, s, t, 1, flags, (hipStream_t)stream);
}

}  // extern "C"
