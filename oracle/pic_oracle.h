/*
 * pic_oracle.h -- CPU oracle for the 1d3v PIC/MCC RF step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial CPU restatement of the
 * reference algorithm (sekulas/cudaPIC Rust-CUDA/src/main.rs, physics from
 * C/eduPIC.cc).  It is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py as the checker; the product library never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line it follows.  Where the
 * reference's own summation order or transcendental approximations cannot
 * be reproduced (cuda-oxide block_scan, PTX *.approx instructions), the
 * oracle uses correctly rounded libm and the tests compare within stated
 * tolerances; integer work (xoshiro128+, jump, compaction counts, indices)
 * is compared bit-exactly.
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Discharge description + the FP32 factors the reference pre-folds
 * (Rust-CUDA/src/main.rs:27-110).  Filled by or_params_derive(). */
typedef struct {
    int32_t n_g, n_t, n_sub;
    double freq, v0, gap, pressure, t_gas, weight, area;
    /* derived, f64 (main.rs:52-58) */
    double dt_e, dt_i, dx, inv_dx, gas_density;
    /* f32 casts (main.rs:80-109) */
    float inv_dx_f, factor_e, factor_i, weight_factor, dt_e_f, dt_i_f, gap_f;
    float normal_range, f1, f2, half_me_over_e, half_mu_over_e;
    float e_charge_f, alpha_f, half_dx_over_eps_f, half_inv_dx_f;
    float min_x_f, max_x_f, de_eepf_f, de_cs_f;
    float me_f, exc_th_f, ion_th_f;
} or_params;

void or_params_derive(or_params *p, int n_g, int n_t, int n_sub, double freq,
                      double v0, double gap, double pressure, double t_gas,
                      double weight, double area);

/* xoshiro128+ 1.0 (main.rs:860-885, xorshiro128c/xoshiro128p.c:36-55) */
uint32_t or_xo_next(uint32_t s[4]);
void     or_xo_jump(uint32_t s[4]);
void     or_xo_skip_soa(uint32_t *s0, uint32_t *s1, uint32_t *s2, uint32_t *s3, uint32_t n, const uint32_t *k);
float    or_u32_to_real(uint32_t x);
void     or_seed_streams(const uint32_t master[4], uint32_t n, uint32_t *s0,
                         uint32_t *s1, uint32_t *s2, uint32_t *s3);
void     or_three_normal(uint32_t s[4], float sigma, float out[3]);

/* cross sections (main.rs:1378-1444 == C/eduPIC.cc:173-229) */
void or_cross_sections(double gas_density, float *cs /* 5*1e6 */,
                       float *stot_e, float *stot_i);

/* deposition (main.rs:356-406): reference f32 weights, f64 sums */
void or_deposit(const float *x, uint32_t n, const or_params *p, double *dens);
/* deposition with f64 positions and weights (the all-FP64 reference) */
void or_deposit_f64(const double *x, uint32_t n, const or_params *p, double *dens);

/* Poisson + field */
void or_poisson_thomas(const double *rho, double pot0, const or_params *p,
                       double *pot, double *efield);          /* eduPIC.cc:461-494 */
void or_poisson_scan_f64(const double *e_d, const double *i_d, double pot0,
                         const or_params *p, double *pot, double *efield); /* main.rs:408-494 */
void or_poisson_scan_f32(const float *e_d, const float *i_d, float pot0,
                         const or_params *p, float *pot, float *efield);   /* main.rs:408-494 */
void or_poisson_flexible_f64(const double *src, int n, double psi_l, double psi_r,
                             double dx, double *pot);             /* main.rs:496-564 */

/* push + survivor filter + order-preserving compaction (main.rs:566-678).
 * Returns the survivor count.  abs_counts[0/1] += electrons lost at x<0 / x>L.
 * ifed (optional, 2*200 bins) collects absorbed-ion energies (eduPIC.cc:625-667). */
uint32_t or_move_and_compact(const float *efield, const float *sx, const float *svx,
                             const float *svy, const float *svz, uint32_t n,
                             float factor, float dt, float mass_over_2e,
                             const or_params *p, float *dx_, float *dvx, float *dvy,
                             float *dvz, uint64_t abs_counts[2], uint32_t *ifed);

/* electron collisions over slots [0, n_active) (main.rs:971-1165).
 * Spawned electrons go to *alive_e (post-increment), ions to *alive_i. */
void or_collide_e(const float *stot_e, const float *cs, uint32_t n_active,
                  float *x, float *vx, float *vy, float *vz,
                  uint32_t *r0, uint32_t *r1, uint32_t *r2, uint32_t *r3,
                  float *ix, float *ivx, float *ivy, float *ivz,
                  uint32_t *alive_e, uint32_t *alive_i, uint32_t cap_e,
                  uint32_t cap_i, const or_params *p, uint64_t *n_coll);
/* one electron slot (the body of check_collisions_e), used by replay tests */
int or_collide_e_one(const float *stot_e, const float *cs, float v[3],
                     uint32_t s[4], float spawn_e_v[3], float spawn_i_v[3],
                     const or_params *p);

/* ion collisions over slots [0, n_active) (main.rs:1167-1292) */
void or_collide_i(const float *stot_i, const float *cs, uint32_t n_active,
                  float *vx, float *vy, float *vz,
                  uint32_t *r0, uint32_t *r1, uint32_t *r2, uint32_t *r3,
                  const or_params *p, uint64_t *n_coll);
int or_collide_i_one(const float *stot_i, const float *cs, float v[3],
                     uint32_t s[4], const or_params *p);

/* Energies (bit patterns lo..hi, step stride) whose cross-section bin from
 * the reciprocal with one fused correction differs from the IEEE division's
 * (main.rs:1002-1005): the check of the GPU kernels' bin formula (corrected = 0:
 * without the correction). */
uint64_t or_bin_division_mismatches(float de_cs, uint32_t lo, uint32_t hi, uint32_t stride,
                                    int corrected);

/* EEPF histogram of centre electrons (main.rs:1301-1339) */
void or_accumulate_eepf(const float *x, const float *vx, const float *vy,
                        const float *vz, const float *efield, uint32_t n,
                        const or_params *p, uint32_t *eepf);

/* Whole RF step in the reference stage order and counter protocol
 * (main.rs:1725-1808), FP32 particles, FP64 density and Poisson.
 * State arrays are capacity-sized; t is the step index within the cycle. */
typedef struct {
    uint32_t cap;
    float *ex, *evx, *evy, *evz, *ix, *ivx, *ivy, *ivz;
    float *tx, *tvx, *tvy, *tvz;              /* compaction scratch */
    uint32_t *er[4], *ir[4];
    uint32_t n_e, n_i;
    double *e_dens, *i_dens;                  /* n_g each */
    float *pot, *efield;                      /* n_g each */
    const float *cs, *stot_e, *stot_i;
    uint64_t abs_e[2], abs_i[2], coll_e, coll_i;
    uint32_t *ifed;                           /* optional 2*200 */
} or_state;

void or_step(or_state *s, const or_params *p, int t, int collisions);

#ifdef __cplusplus
}
#endif
#endif
