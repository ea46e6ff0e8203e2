/*
 * cudapic_amd.h -- C ABI of the MI355X (gfx950) PIC/MCC engine for the
 * 1d3v argon RF discharge of sekulas/cudaPIC.
 *
 * Two layers, both plain C (pointers, sizes, status codes; no HIP or torch
 * types -- a stream is passed as void*, i.e. a hipStream_t):
 *
 *  1. Stage entry points, one per kernel of the reference's
 *     `#[cuda_module] mod kernels` (Rust-CUDA/src/main.rs:352-1358).  They take
 *     raw device pointers and a stream, read live counts from device memory
 *     (never from the host), and are asynchronous and graph-capturable.
 *
 *  2. A simulation object that owns the device state the reference keeps in
 *     `GpuSimState` (main.rs:132-350) and runs the RF step loop of main.rs:
 *     1723-1828, either in the reference's stage order or on the fused
 *     MI355X path (one particle kernel per species per step).
 *
 * Every function returns PIC_OK (0) or an error code; none aborts.
 */
#ifndef CUDAPIC_AMD_H
#define CUDAPIC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *pic_stream_t; /* hipStream_t; NULL = the null stream */

enum pic_status {
    PIC_OK = 0,
    PIC_ERR_HIP = 1,       /* a HIP runtime call failed (see pic_last_error) */
    PIC_ERR_ARG = 2,       /* invalid argument (null pointer, size, mode) */
    PIC_ERR_CAPACITY = 3,  /* a particle array overflowed (device flag set) */
    PIC_ERR_UNSUPPORTED = 4,
    PIC_ERR_IO = 5
};

/* Precision of the density flush and of the Poisson prefix scans. */
enum pic_precision {
    PIC_PREC_FP32 = 0,  /* f32 density atomics, f32 scans (the reference) */
    PIC_PREC_COMP = 1,  /* f32 density atomics, compensated (float-float) scans */
    PIC_PREC_FP64 = 2   /* f64 density atomics, f64 scans */
};

enum pic_species_id { PIC_ELECTRONS = 0, PIC_IONS = 1 };

#define PIC_CS_RANGES 1000000 /* main.rs:69 */
#define PIC_N_CS 5            /* main.rs:61 */
#define PIC_N_EEPF 2000       /* main.rs:75 */
#define PIC_N_IFED 200        /* C/eduPIC.cc:124 */
#define PIC_MAX_HARMONICS 16

/* Discharge parameters and the f32 factors the reference pre-folds
 * (main.rs:27-110).  Set the first block, then call pic_params_derive(). */
typedef struct pic_params {
    int32_t n_g, n_t, n_sub, precision;
    double freq, v0, gap, pressure, t_gas, weight, area;
    /* derived by pic_params_derive() */
    double dt_e, dt_i, dx, inv_dx, gas_density;
    float inv_dx_f, factor_e, factor_i, weight_factor, dt_e_f, dt_i_f, gap_f;
    float normal_range, f1, f2, half_me_over_e, half_mu_over_e, half_mi_over_e;
    float e_charge_f, alpha_f, half_dx_over_eps_f, half_inv_dx_f;
    float min_x_f, max_x_f, de_eepf_f, de_cs_f, me_f, exc_th_f, ion_th_f, eepf_kick_f;
    /* driving waveform pot0(t) = sum_k a_k cos(2 pi k t/N_T) + b_k sin(2 pi k t/N_T);
     * pic_params_derive() sets a_1 = v0 and the rest to 0 (main.rs:1741). */
    double wave_cos[PIC_MAX_HARMONICS], wave_sin[PIC_MAX_HARMONICS];
} pic_params_t;

/* The reference case (main.rs:39-49): N_G=400, N_T=4000, N_SUB=20,
 * 13.56 MHz, 250 V, 25 mm, 10 Pa, 350 K, W=7e4, A=1e-4 m^2, FP32. */
void pic_params_default(pic_params_t *p);
/* Fills the derived block exactly as main.rs:52-110 computes it. */
int pic_params_derive(pic_params_t *p);

/* SoA particle arrays of one species (the reference's e_x/e_vx/e_vy/e_vz or
 * i_*, main.rs:134-143) with the per-slot xoshiro128+ state words
 * (rng_e0..rng_e3 / rng_i0..rng_i3, main.rs:163-170). */
typedef struct pic_species {
    float *x, *vx, *vy, *vz;
    uint32_t *rng0, *rng1, *rng2, *rng3;
    uint32_t capacity; /* MAX_PARTICLES */
} pic_species_t;

/* ------------------------------------------------------------------------
 * Host-side tables (main.rs:1378-1496)
 * ---------------------------------------------------------------------- */

/* Cross-section tables on 1e6 energy bins of 1 meV into host buffers:
 * cs[5*PIC_CS_RANGES] (e elastic, e excitation, e ionisation, i isotropic,
 * i backward) and the total macroscopic cross sections (main.rs:1378-1444). */
int pic_cross_sections(double gas_density, float *cs, float *stot_e, float *stot_i);

/* eduPIC's run report (C/eduPIC.cc:926-978, check_and_save_info; no cudaPIC
 * counterpart): plasma characteristics at the centre of the gap and the four
 * stability and accuracy conditions.  Inputs: the mean measured electron
 * density at node N_G/2 [m^-3], the mean electron energy in the EEPF window
 * [eV], the simulated time of the run [s], the collisions counted in it and
 * the live particle counts; stot_e / stot_i are pic_cross_sections' totals.
 * Host only.  conditions_ok is 0 if any condition is violated. */
typedef struct pic_report {
    double density_center;   /* m^-3 */
    double plasma_freq;      /* electron plasma frequency at the centre [rad/s] */
    double debye_length;     /* [m], from kT = 2/3 of the mean energy */
    double coll_freq_e;      /* collisions / time / N_e [1/s] */
    double coll_freq_i;
    double wp_dt_e;          /* plasma_freq * DT_E      (OK below 0.20) */
    double dx_over_debye;    /* DX / debye_length       (OK below 1.00) */
    double nu_max_e_dt_e;    /* max_E v sigma_tot_e * DT_E (OK below 0.05) */
    double nu_max_i_dt_i;    /* max_E g sigma_tot_i * DT_I (OK below 0.05) */
    double cfl_energy;       /* 0.5 m_e (DX/DT_E)^2 [eV] */
    int32_t conditions_ok;
    int32_t pad;
} pic_report_t;
int pic_report(const pic_params_t *p, const float *stot_e, const float *stot_i, double density_center,
               double mean_energy_center, double sim_time, unsigned long long n_e_coll,
               unsigned long long n_i_coll, uint32_t n_e, uint32_t n_i, pic_report_t *out);

/* ------------------------------------------------------------------------
 * Stage entry points (one per reference kernel)
 * ---------------------------------------------------------------------- */

/* get_density (main.rs:356-406): CIC deposit of x[0..*n_active) into
 * `density` (n_g values, f32 or f64 as `density_f64` says), which the
 * caller zeroes first (main.rs:1729).  Nodes 0 and n_g-1 are doubled. */
int pic_get_density(const float *x, const uint32_t *n_active, void *density,
                    int density_f64, const pic_params_t *p, pic_stream_t stream);

/* solve_poisson_scan_f32 (main.rs:408-494): potential and central-difference
 * field from the two densities (f32 arrays for PIC_PREC_FP32/COMP, f64 for
 * PIC_PREC_FP64) with pot[0] = pot0, pot[n_g-1] = 0.  One workgroup. */
int pic_solve_poisson(const void *e_density, const void *i_density, double pot0,
                      int precision, float *pot, float *efield, const pic_params_t *p,
                      pic_stream_t stream);

/* solve_poisson_dps_flexible (main.rs:496-564): psi'' = source with
 * Dirichlet ends, 3 <= n <= 4096 nodes of spacing dx. */
int pic_solve_poisson_flexible(const float *source, int n, float psi_left, float psi_right,
                               float dx, int precision, float *pot, pic_stream_t stream);

/* move_and_compact (main.rs:566-678): leapfrog push of src[0..*n_active)
 * in `efield`, keep 0 <= x' <= L, compact the survivors into dst and add
 * their number to *alive_counter (caller zeroes it, main.rs:1758).
 * Optional: abs_counts[2] += particles lost at x<0 / x>L; ifed[2*200] +=
 * energy histogram (1 eV bins) of the lost particles with mass factor
 * half_m_over_e (pass 0 and NULL to skip). */
int pic_move_and_compact(const float *efield, const float *src_x, const float *src_vx,
                         const float *src_vy, const float *src_vz, float *dst_x,
                         float *dst_vx, float *dst_vy, float *dst_vz,
                         uint32_t *alive_counter, const uint32_t *n_active, float factor,
                         float dt, float half_m_over_e, unsigned long long *abs_counts,
                         uint32_t *ifed, const pic_params_t *p, pic_stream_t stream);

/* check_collisions_e + collision_e (main.rs:971-1165): electron/Ar MCC over
 * slots [0, *n_active).  Ionisations append an electron at (*alive_e)++ and
 * an ion at (*alive_i)++; *alive_e must equal *n_active on entry
 * (main.rs:1782).  Appends beyond the capacities are dropped and set
 * *overflow = 1.  n_coll (optional) += number of collisions. */
int pic_check_collisions_e(const float *stot_e, const float *cs, const uint32_t *n_active,
                           const pic_species_t *electrons, const pic_species_t *ions,
                           uint32_t *alive_e, uint32_t *alive_i, uint32_t *overflow,
                           unsigned long long *n_coll, const pic_params_t *p,
                           pic_stream_t stream);

/* check_collisions_i + collision_i (main.rs:1167-1292): Ar+/Ar MCC with a
 * Maxwellian target atom over slots [0, *n_active). */
int pic_check_collisions_i(const float *stot_i, const float *cs, const uint32_t *n_active,
                           const pic_species_t *ions, unsigned long long *n_coll,
                           const pic_params_t *p, pic_stream_t stream);

/* accumulate_density (main.rs:1294-1299): cumul[k] += (f64) density[k]. */
int pic_accumulate_density(const void *density, int density_f64, double *cumul, int n_g,
                           pic_stream_t stream);

/* accumulate_eepf (main.rs:1301-1339): time-centred energy histogram
 * (2000 bins of 0.05 eV) of electrons in 0.45 L < x < 0.55 L. */
int pic_accumulate_eepf(const float *x, const float *vx, const float *vy, const float *vz,
                        const float *efield, const uint32_t *n_active, uint32_t *eepf,
                        const pic_params_t *p, pic_stream_t stream);

/* xoshiro128p_test_wrapper (main.rs:1341-1357), generalised: n successive
 * next() outputs from state[4] (host values) into out[n] (device), the final
 * state into state_out[4] (device). */
int pic_xoshiro_next_n(const uint32_t state[4], uint32_t n, uint32_t *out,
                       uint32_t *state_out, pic_stream_t stream);

/* xoshiro128_seed_streams (main.rs:1446-1496) on the device: slot k of the
 * four planar word arrays gets the master state jumped k times (k in
 * [0, n)); bit-identical to the sequential host seeding. */
int pic_seed_streams(const uint32_t master[4], uint32_t n, uint32_t *rng0, uint32_t *rng1,
                     uint32_t *rng2, uint32_t *rng3, pic_stream_t stream);

/* ------------------------------------------------------------------------
 * Simulation object (GpuSimState + the RF loop, main.rs:112-350, 1611-1878)
 * ---------------------------------------------------------------------- */

enum pic_path {
    PIC_PATH_REFERENCE = 0, /* the reference's stage order and counter protocol */
    PIC_PATH_FUSED = 1      /* push+compact+collide+deposit fused per species */
};

typedef struct pic_sim_config {
    pic_params_t params;       /* derived */
    uint32_t capacity;         /* MAX_PARTICLES per species (main.rs:79), at most
                                  2^30 - 4096 */
    uint32_t segments;         /* fused path: segments per species, each with its own
                                  live counter (0 = 1, at most 64); the reference
                                  path uses 1 */
    uint32_t seed_e[4];        /* master xoshiro states (main.rs:1667, 1671) */
    uint32_t seed_i[4];
    int32_t path;              /* enum pic_path */
    int32_t use_graph;         /* fused path: capture 4*N_SUB steps into a hipGraph */
    int32_t block_size;        /* lanes per workgroup of the fused particle
                                  kernels: 128, 256 or 512 (0 = 256); other
                                  values, and any but 0 / 256 on the reference
                                  path, are rejected */
    int32_t grid_blocks;       /* persistent grid: 0 = tuned default */
    int32_t ppt;               /* particles per thread per tile, compacting
                                  steps: 1, 2, 4 or 8; 0 = 4 */
    int32_t replicas;          /* global density replicas: 0 = default */
    int32_t measure;           /* allocate measurement buffers */
    int32_t ppt_inplace;       /* particles per thread per tile, in-place
                                  steps: 0 = default (2) */
    int32_t lds_replicas;      /* fused path: at most this many LDS density
                                  histograms per workgroup (shared by its
                                  waves): 0 = one per wave as LDS allows */
    uint32_t compact_every;    /* fused path: steps between the compactions that
                                  squeeze out the holes left by lost particles
                                  (a multiple of 4*n_sub; 0 = 800) */
    uint32_t rng_defer;        /* fused path: steps per run of deferred electron
                                  RNG write-backs (as pic_step_config_t; the
                                  words are exact whenever a call returns) */
    int32_t cand_masks;        /* fused path: collision-candidate masks of the
                                  deferred runs (a byte per electron slot: the
                                  steps of the run whose draw can fall below
                                  the largest collision probability; the other
                                  steps of a live slot read no RNG words):
                                  0 or 1 = on, -1 = off.  The draws and their
                                  outcomes are the same either way. */
    float cand_threshold;      /* the masks' probability threshold: 0 = the
                                  largest collision probability of the tables'
                                  energies; a lower value makes particles above
                                  it load their words on every step, a higher
                                  one makes more steps candidates */
} pic_sim_config_t;

typedef struct pic_sim pic_sim_t;

void pic_sim_config_default(pic_sim_config_t *c);
int pic_sim_create(const pic_sim_config_t *c, pic_sim_t **out);
int pic_sim_destroy(pic_sim_t *s);

/* Host SoA upload of n particles into a species (main.rs:250-278); resets
 * the step parity bookkeeping.  Download copies the live particles out
 * (main.rs:315-339); *n receives the live count (at most cap are copied). */
int pic_sim_upload(pic_sim_t *s, int species, const float *x, const float *vx,
                   const float *vy, const float *vz, uint32_t n);
int pic_sim_download(pic_sim_t *s, int species, float *x, float *vx, float *vy, float *vz,
                     uint32_t cap, uint32_t *n);
/* Storage of a species: number of segments, slots per segment and the
 * number of slots in use in each segment (counts[n_seg], may be NULL).
 * Slots in use q*seg_size + [0, counts[q]) hold the particles; on the fused
 * path a slot whose particle left the gap in an in-place step is a hole
 * (x < 0) until the next compacting step (every compact_every steps).
 * pic_sim_download returns the live particles segment by segment; the
 * fused path needs a few percent of headroom in `capacity` because segments
 * fill independently. */
int pic_sim_segments(pic_sim_t *s, int species, uint32_t *n_seg, uint32_t *seg_size,
                     uint32_t *counts);
/* Slot index of every particle pic_sim_download returns, in the same order. */
int pic_sim_download_slots(pic_sim_t *s, int species, uint32_t *slots, uint32_t cap, uint32_t *n);
/* Per-slot RNG words, planar, n slots (upload replaces, download copies). */
int pic_sim_rng_upload(pic_sim_t *s, int species, const uint32_t *w0, const uint32_t *w1,
                       const uint32_t *w2, const uint32_t *w3, uint32_t n);
int pic_sim_rng_download(pic_sim_t *s, int species, uint32_t *w0, uint32_t *w1,
                         uint32_t *w2, uint32_t *w3, uint32_t n);
/* Replace the cross-section tables (host arrays as pic_cross_sections). */
int pic_sim_set_tables(pic_sim_t *s, const float *cs, const float *stot_e, const float *stot_i);
/* The collision-candidate masks' probability threshold in use (the
 * configured one, or the largest collision probability of the current
 * tables' energies) and whether the masks are on (no counterpart in the
 * reference: its collision test reads every slot's words every step). */
int pic_sim_cand_threshold(pic_sim_t *s, float *threshold, int *masks_on);

/* Live counts (synchronises the sim stream; the reference's checkpoint read,
 * main.rs:1810-1816). */
int pic_sim_counts(pic_sim_t *s, uint32_t *n_e, uint32_t *n_i);
/* Step index within the RF cycle (0..N_T-1) of the next step, and total steps run. */
int pic_sim_time(pic_sim_t *s, uint32_t *t_in_cycle, unsigned long long *steps_done);
int pic_sim_set_time(pic_sim_t *s, uint32_t t_in_cycle);

/* Run n_steps RF time steps asynchronously on the sim stream.  measure != 0
 * accumulates densities and the EEPF (main.rs:1746-1756, 1802-1807). */
int pic_sim_run(pic_sim_t *s, uint32_t n_steps, int measure);
int pic_sim_synchronize(pic_sim_t *s);
/* PIC_ERR_CAPACITY if any particle array overflowed since creation. */
int pic_sim_check(pic_sim_t *s);

/* Grid state after the last Poisson solve: pot, efield (f32), and the
 * electron / ion densities that solve used (f64). Any pointer may be NULL. */
int pic_sim_read_grid(pic_sim_t *s, float *pot, float *efield, double *e_dens, double *i_dens);
/* Measurement accumulators: cumul_e/i (f64 n_g), eepf (u32 2000),
 * ifed (u32 2*200: powered, grounded), counters[8] = electrons absorbed at
 * the powered / grounded electrode, ions ditto, electron collisions, ion
 * collisions, electron and ion measurement steps.  NULL skips. */
int pic_sim_read_measurements(pic_sim_t *s, double *cumul_e, double *cumul_i, uint32_t *eepf,
                              uint32_t *ifed, unsigned long long *counters);
int pic_sim_reset_measurements(pic_sim_t *s);
/* Spatio-temporal (XT) maps of eduPIC's measurement mode (C/eduPIC.cc:126-145,
 * 558-575, 603-614, 705-715), fused path only: on every measured step the
 * grid quantities and the CIC-weighted mid-step particle quantities are
 * added into time bins of bin_steps steps (bin_steps divides n_t; eduPIC
 * uses 20).  0 switches them off.  No reference counterpart in cudaPIC,
 * which measures densities and the EEPF only. */
int pic_sim_set_xt(pic_sim_t *s, uint32_t bin_steps);
/* The 11 maps normalised as norm_all_xt (C/eduPIC.cc:864-903) with each
 * bin's measured steps: out[q][k][b], q in eduPIC's file order pot, efield,
 * ne, ni, je, ji, powere (heate), poweri (heati), meanee, meanei, ioniz;
 * k the node, b < *n_xt = n_t / bin_steps the time bin.  out NULL: *n_xt only. */
int pic_sim_read_xt(pic_sim_t *s, double *out, uint32_t *n_xt);

/* Timing: run n_steps and report wall milliseconds (HIP events on the sim
 * stream) and the summed duration of the dominant particle kernel
 * (electron particle kernel) with its launch count.  ms_kernel NULL: with
 * graphs (fused path, n_t a multiple of 4*N_SUB) the window replays
 * captured graphs only -- whole periods and the partial periods it starts
 * or ends with, captured before the window opens; otherwise eager steps
 * bracketed by events. */
int pic_sim_time_steps(pic_sim_t *s, uint32_t n_steps, int measure, float *ms_total,
                       float *ms_kernel, uint32_t *kernel_launches,
                       unsigned long long *particle_steps);

/* Binary checkpoint of the whole device state (particles, RNG words,
 * counters, held densities, time, measurements). */
int pic_sim_save(pic_sim_t *s, const char *path);
int pic_sim_load(pic_sim_t *s, const char *path);

/* ------------------------------------------------------------------------
 * The RF step on caller-owned buffers (the reference's GpuSimState and the
 * loop body of main.rs:1725-1808), on the fused kernels
 * ---------------------------------------------------------------------- */

/* The device buffers of the reference's GpuSimState (main.rs:132-185), owned
 * by the caller; every particle and RNG array has `capacity` slots.  A step
 * updates the particles in their slots; a particle that leaves the gap
 * becomes a hole (x = -1) below *n_electrons / *n_ions until the next
 * compacting step (t % compact_every == 0, an ion step: both species are
 * compacted into the tmp arrays and the struct's pointers swapped, as the
 * reference's host swaps its handles at main.rs:1764-1767, 1780-1782).  The
 * counts are therefore watermarks of the slots in use; readers skip x < 0.
 * The RNG words stay with their slots (main.rs:1764-1767 swaps x and v only). */
typedef struct pic_state {
    float *e_x, *e_vx, *e_vy, *e_vz;
    float *i_x, *i_vx, *i_vy, *i_vz;
    uint32_t *n_electrons, *n_ions;  /* device words */
    float *efield, *pot;             /* n_g each, written by every step */
    /* optional (both or neither): the f64 node densities each step's Poisson
     * solve used; state the next steps read, like the particles (NULL keeps
     * them in the context) */
    double *e_density, *i_density;
    const float *cs, *sigma_tot_e, *sigma_tot_i;  /* as pic_cross_sections */
    uint32_t *rng_e0, *rng_e1, *rng_e2, *rng_e3;
    uint32_t *rng_i0, *rng_i1, *rng_i2, *rng_i3;
    float *tmp_x, *tmp_vx, *tmp_vy, *tmp_vz;          /* electrons' compaction target */
    /* ions' compaction target: no reference counterpart (the reference
     * shares tmp_* between species; the fused step compacts both in one step) */
    float *tmp_i_x, *tmp_i_vx, *tmp_i_vy, *tmp_i_vz;
    double *cumul_e_density, *cumul_i_density;        /* PIC_STEP_MEASURE: n_g each */
    uint32_t *eepf_counts;                            /* PIC_STEP_MEASURE: PIC_N_EEPF */
    uint32_t capacity;                                /* MAX_PARTICLES, at most 2^30 */
    uint32_t pad;
} pic_state_t;

enum pic_step_flags {
    PIC_STEP_MEASURE = 1, /* accumulate densities and the EEPF (main.rs:1746-1756, 1802-1807) */
    PIC_STEP_RESET = 2    /* deposit the current particles first: required on the first step
                             and after the caller changes particles, counts or the step order */
};

/* Launch shape of the fused kernels (0 = default, as pic_sim_config_t);
 * compact_every is a multiple of 4*n_sub (0 = 800). */
typedef struct pic_step_config {
    uint32_t compact_every;
    int32_t grid_blocks, ppt, ppt_inplace, lds_replicas;
    uint32_t rng_defer;  /* electrons' deferred RNG write-back: inside one
                            call a live slot that does not collide stores no
                            RNG words for up to this many consecutive steps
                            (must divide 4*n_sub, at most 64; 1 = write back
                            every step; 0 = 8).  Every call returns with the
                            reference's exact words. */
} pic_step_config_t;

typedef struct pic_ctx pic_ctx_t;

/* The workspace of the fused step for one discharge (density accumulators,
 * frozen watermarks, device clock, pot0 table, overflow flag, diagnostic
 * counters) and a cache of captured period graphs.  params derived;
 * n_t a multiple of 2*n_sub, at most 32768; cfg may be NULL. */
int pic_init(const pic_params_t *p, const pic_step_config_t *cfg, pic_ctx_t **out);
int pic_destroy(pic_ctx_t *ctx);
/* One RF step at cycle index t (0 <= t < n_t; the reference's t,
 * main.rs:1741), asynchronous on `stream`; consecutive calls pass
 * consecutive t.  Updates *s when the step compacts. */
int pic_step(pic_ctx_t *ctx, pic_state_t *s, uint32_t t, int flags, pic_stream_t stream);
/* n_steps steps from t: whole 4*n_sub-step periods (from t a multiple of
 * 4*n_sub) replayed from captured hipGraphs, the rest as pic_step.  Graphs
 * need a non-null stream (NULL runs every step eagerly) and are keyed by the
 * state's buffers. */
int pic_step_graph(pic_ctx_t *ctx, pic_state_t *s, uint32_t t, uint32_t n_steps, int flags,
                   pic_stream_t stream);
/* Synchronises `stream`; then (each may be NULL) counters[8] as
 * pic_sim_read_measurements since pic_init, ifed[2*200] of the absorbed
 * ions, and the overflow flag (1: appends beyond capacity were dropped). */
int pic_step_diagnostics(pic_ctx_t *ctx, unsigned long long counters[8], uint32_t *ifed, uint32_t *overflow,
                         pic_stream_t stream);
/* pic_step without a context, for callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc).  No reference counterpart.  The workspace is a fixed 1.5 MiB
 * device array per GPU, enough for every discharge pic_init accepts
 * (n_g <= 4096, n_t <= 32768: 1.32 MiB; the reference case uses 0.15 MiB;
 * a larger need returns PIC_ERR_ARG); the call allocates nothing and uses
 * the default launch shape.  The array
 * keeps the last call's state: consecutive steps of one discharge pass
 * consecutive t, the first with PIC_STEP_RESET, and switching discharges on
 * a GPU needs PIC_STEP_RESET again; calls that overlap on one GPU race.  No
 * diagnostic counters or IFED are kept; a capacity overflow shows as a
 * count above capacity. */
int pic_step_noalloc(const pic_params_t *p, pic_state_t *s, uint32_t t, int flags, pic_stream_t stream);
/* One pic_step with the electrons' deferred RNG write-back made explicit,
 * for tests that check the deferral one step at a time.  No reference
 * counterpart.  On entry the RNG words of every live electron slot
 * (x >= 0, below *n_electrons) are rng_lag draws behind the slot's stream;
 * holes and slots past the watermark hold exact words.  On return the live
 * slots (newborns included) are rng_lag_out draws behind, holes and
 * unused slots exact.  rng_lag < 64; rng_lag_out 0 or rng_lag + 1; both 0
 * at a compacting step.  pic_step is pic_step_lagged(.., 0, 0, ..). */
int pic_step_lagged(pic_ctx_t *ctx, pic_state_t *s, uint32_t t, int flags, uint32_t rng_lag,
                    uint32_t rng_lag_out, pic_stream_t stream);

/* ------------------------------------------------------------------------
 * Ensembles of independent discharges (one launch per stage per group of instances)
 * ---------------------------------------------------------------------- */

/* Instances share n_g, n_t, n_sub, frequency, gap and precision (taken from
 * params[0]); each has its own v0, pressure, gas temperature, weight and
 * capacity.  No reference counterpart: the reference runs one discharge per
 * process; this is the batched mode of the same step.  Particles are
 * updated in the slots they occupy, as on the fused single-discharge path
 * (lost particles leave holes, a compaction pass every compact_every steps), so
 * a particle's RNG stream is its slot's until the next compaction. */
typedef struct pic_ensemble_config {
    int32_t n_instances;          /* 1..256 */
    const pic_params_t *params;   /* [n_instances], derived */
    const uint32_t *capacity;     /* [n_instances] slots per species (at most 2^30 in
                                     all, each rounded up to 256) */
    uint32_t seed_e[4], seed_i[4];
    /* use_graph: replay 4*N_SUB steps from a hipGraph; grid_blocks (0 = 1024
     * lanes per compute unit), ppt (particles per thread per tile, 1, 2 or 4; 0 =
     * 2) and replicas (global density replicas, 0 = 8) shape the particle
     * kernels; lds_replicas and compact_every as in pic_sim_config_t */
    int32_t use_graph, grid_blocks, ppt, replicas, lds_replicas;
    uint32_t compact_every;
    /* Temporal blocking of the graph path: consecutive instances whose
     * electron capacities sum to at most group_slots form a group (an
     * instance above it is a group alone; 0 = 9,000,000, 0xffffffff = one
     * group of all), and each group runs group_periods 4*N_SUB-step graphs
     * (0 = 5) before the next one starts, so that a group's particles stay
     * in the GPU's last-level cache from step to step.  The groups are
     * dealt round-robin to group_streams streams (0 = 1; more overlap one
     * group's Poisson solves with another's particle kernels, at the price
     * of two groups sharing the cache).  Every
     * instance still goes through the same steps; the order in which
     * instances' steps execute is all that changes. */
    uint32_t group_slots, group_periods, group_streams;
    /* lanes per workgroup of the particle kernel: 128, 256 or 512 (0 = 256) */
    int32_t block_size;
    /* steps per run of deferred electron RNG write-backs, as in
     * pic_sim_config_t (divides 4*n_sub, at most 64; 1 = every step; 0 = 8) */
    uint32_t rng_defer;
    /* collision-candidate masks of the deferred runs, as in pic_sim_config_t
     * (0 or 1 = on, -1 = off), each instance with its own threshold (its
     * largest collision probability); the draws and outcomes are the same
     * either way */
    int32_t cand_masks;
} pic_ensemble_config_t;

typedef struct pic_ensemble pic_ensemble_t;

int pic_ens_create(const pic_ensemble_config_t *c, pic_ensemble_t **out);
int pic_ens_destroy(pic_ensemble_t *e);
int pic_ens_upload(pic_ensemble_t *e, int inst, int species, const float *x, const float *vx,
                   const float *vy, const float *vz, uint32_t n);
/* The live particles of one instance in slot order (*n = live count). */
int pic_ens_download(pic_ensemble_t *e, int inst, int species, float *x, float *vx, float *vy,
                     float *vz, uint32_t cap, uint32_t *n);
/* Slot (relative to the instance's region) of each particle
 * pic_ens_download returns, in the same order. */
int pic_ens_download_slots(pic_ensemble_t *e, int inst, int species, uint32_t *slots, uint32_t cap,
                           uint32_t *n);
/* Step index within the RF cycle of the next step (as pic_sim_set_time). */
int pic_ens_set_time(pic_ensemble_t *e, uint32_t t_in_cycle);
/* RNG words of the first n slots of one instance's region. */
int pic_ens_rng_download(pic_ensemble_t *e, int inst, int species, uint32_t *w0, uint32_t *w1,
                         uint32_t *w2, uint32_t *w3, uint32_t n);
/* n_e[i], n_i[i] for every instance (synchronises). */
int pic_ens_counts(pic_ensemble_t *e, uint32_t *n_e, uint32_t *n_i);
int pic_ens_run(pic_ensemble_t *e, uint32_t n_steps, int measure);
int pic_ens_synchronize(pic_ensemble_t *e);
int pic_ens_check(pic_ensemble_t *e);
/* Wall time (HIP events) of n_steps over all instances and the particle-steps done. */
int pic_ens_time_steps(pic_ensemble_t *e, uint32_t n_steps, int measure, float *ms_total,
                       unsigned long long *particle_steps);
int pic_ens_read_grid(pic_ensemble_t *e, int inst, float *pot, float *efield);
/* cumul_e/i (f64 n_g) and counters[8] (as pic_sim_read_measurements) of one instance. */
int pic_ens_read_measurements(pic_ensemble_t *e, int inst, double *cumul_e, double *cumul_i,
                              unsigned long long *counters);

/* ------------------------------------------------------------------------
 * Adjoint: gradients of the window-averaged ion density misfit
 * ---------------------------------------------------------------------- */

/* No reference counterpart: the reference has no differentiation.  A record
 * pass runs the RF step (the reference's stage order) over a window of
 * `steps` steps and tapes its discrete events (collisions with their
 * post-collision v_x, absorptions, ionisations with the newborns' slots and
 * v_x); later passes replay the tape, so x and v_x are differentiable
 * functions of the waveform pot0(t) = sum_h wave_cos[h] cos(2 pi (h+1) t/N_T)
 * + wave_sin[h] sin(...).  Objective J = sum_k (nbar_k - target_k)^2 /
 * sum_k target_k^2, nbar the ion density averaged over the window's ion
 * steps.  The backward pass restores a checkpoint every `checkpoint` steps
 * and recomputes the segment before its adjoint steps (push, deposit, field
 * solve).  Densities are accumulated in 64-bit fixed point, so every pass is
 * deterministic. */
typedef struct pic_adj_config {
    pic_params_t params;        /* derived (n_g <= 2048); the waveform is wave_cos / wave_sin */
    uint32_t capacity_e, capacity_i;
    uint32_t steps;             /* window length (a multiple of n_sub) */
    uint32_t checkpoint;        /* steps between checkpoints (multiple of n_sub; 0 = 200) */
    uint32_t t0;                /* step of the RF cycle the window starts at (multiple of n_sub) */
    uint32_t tape_capacity;     /* events per kind (0 = steps * capacity_e / 64) */
    uint32_t seed_e[4], seed_i[4];
} pic_adj_config_t;

typedef struct pic_adjoint pic_adjoint_t;

int pic_adj_create(const pic_adj_config_t *c, pic_adjoint_t **out);
int pic_adj_destroy(pic_adjoint_t *a);
/* Starting state of the window (before pic_adj_record). */
int pic_adj_upload(pic_adjoint_t *a, int species, const float *x, const float *vx, const float *vy,
                   const float *vz, uint32_t n);
/* Target ion density profile (n_g values, m^-3). */
int pic_adj_set_target(pic_adjoint_t *a, const double *n_i_target);
/* The record pass (once): Monte Carlo collisions, event tape, *J. */
int pic_adj_record(pic_adjoint_t *a, double *J);
/* A frozen-event forward pass with another waveform (NULL keeps the
 * current coefficients); *J and nbar[n_g] (either may be NULL). */
int pic_adj_forward(pic_adjoint_t *a, const double *wave_cos, const double *wave_sin, double *J,
                    double *nbar);
/* dJ/dwave_cos[16], dJ/dwave_sin[16] and dJ/dpot0(t) for every window step
 * (any may be NULL) at the waveform of the last forward or record pass. */
int pic_adj_backward(pic_adjoint_t *a, double *grad_cos, double *grad_sin, double *dJ_dV);
/* Event counts of the tape and the slots in use at the window's end. */
int pic_adj_tape_sizes(pic_adjoint_t *a, uint32_t *n_coll_e, uint32_t *n_coll_i, uint32_t *n_ionisations,
                       uint32_t *slots_e, uint32_t *slots_i);
/* The tape: per-step offsets [steps+1] into the collision lists (slot, v_x)
 * of each species and into the ionisation list (parent, new electron slot,
 * new ion slot, their v_x), the slot watermarks at the start of each step
 * [steps+1] and every slot's death step (step + 1 of its absorption, or
 * 0xffffffff).  Any pointer may be NULL. */
int pic_adj_download_tape(pic_adjoint_t *a, uint32_t *off_e, uint32_t *slot_e, float *v_e, uint32_t *off_i,
                          uint32_t *slot_i, float *v_i, uint32_t *off_ion, uint32_t *parent, uint32_t *new_e,
                          uint32_t *new_i, float *v_new_e, float *v_new_i, uint32_t *slots_e, uint32_t *slots_i,
                          uint32_t *death_e, uint32_t *death_i);

/* Device memory helpers for callers without their own allocator. */
int pic_malloc(void **ptr, uint64_t bytes);
int pic_free(void *ptr);
int pic_memcpy_htod(void *dst, const void *src, uint64_t bytes);
int pic_memcpy_dtoh(void *dst, const void *src, uint64_t bytes);
int pic_memset(void *dst, int value, uint64_t bytes);
int pic_device_synchronize(void);
/* A non-blocking stream of the default priority (for pic_step_graph) and its release. */
int pic_stream_create(pic_stream_t *out);
int pic_stream_destroy(pic_stream_t stream);
int pic_stream_synchronize(pic_stream_t stream);
/* Kernel copy of `bytes` (16-byte aligned) with 16-byte vector accesses: the
 * reference point for HBM traffic counters and achievable bandwidth (no
 * reference counterpart). */
int pic_stream_copy(void *dst, const void *src, uint64_t bytes, pic_stream_t stream);
/* Read-only counterpart (16-byte loads; *sum += a digest of the data, so
 * the loads happen), and the runtime's own device-to-device copy: the two
 * cross-checks of the copy kernel's bandwidth (no reference counterpart). */
int pic_stream_read(const void *src, uint64_t bytes, unsigned long long *sum, pic_stream_t stream);
int pic_memcpy_dtod_async(void *dst, const void *src, uint64_t bytes, pic_stream_t stream);
/* Text of the last error on this thread. */
const char *pic_last_error(void);
/* Device properties used for launch shapes: compute units, wave size. */
int pic_device_info(int *compute_units, int *wave_size, char *arch, int arch_len);
/* sizeof(pic_params_t), sizeof(pic_species_t), sizeof(pic_sim_config_t),
 * sizeof(pic_ensemble_config_t), sizeof(pic_state_t), sizeof(pic_step_config_t),
 * so that a binding can verify its struct layout (no reference counterpart). */
int pic_abi_layout(uint64_t out[6]);

#ifdef __cplusplus
}
#endif
#endif
