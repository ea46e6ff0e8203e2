/*
 * pic_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see pic_oracle.h).
 *
 * Serial restatement of sekulas/cudaPIC's RF step (Rust-CUDA/src/main.rs)
 * with the FP64 physics of eduPIC (C/eduPIC.cc).  Compiled with
 * -ffp-contract=off so that every fused multiply-add is the explicit fmaf()
 * written below; the HIP kernels spell the same operations the same way,
 * which is what makes the push/compaction comparison bit-exact.
 */
#include "pic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.141592653589793
#define OR_TWO_PI (2.0 * OR_PI)
#define E_CHARGE 1.60217662e-19
#define E_MASS 9.10938356e-31
#define AR_MASS 6.63352090e-26
#define MU_ARAR (AR_MASS / 2.0)
#define K_BOLTZMANN 1.38064852e-23
#define EPSILON0 8.85418781e-12
#define CS_RANGES 1000000
#define DE_CS 0.001
#define E_EXC_TH 11.5
#define E_ION_TH 15.8
#define N_EEPF 2000
#define N_IFED 200

/* main.rs:27-110: every f32 factor is computed in f64 then cast once. */
void or_params_derive(or_params *p, int n_g, int n_t, int n_sub, double freq,
                      double v0, double gap, double pressure, double t_gas,
                      double weight, double area) {
    memset(p, 0, sizeof(*p));
    p->n_g = n_g;
    p->n_t = n_t;
    p->n_sub = n_sub;
    p->freq = freq;
    p->v0 = v0;
    p->gap = gap;
    p->pressure = pressure;
    p->t_gas = t_gas;
    p->weight = weight;
    p->area = area;
    double period = 1.0 / freq;
    p->dt_e = period / (double)n_t;
    p->dt_i = (double)n_sub * p->dt_e;
    p->dx = gap / (double)(n_g - 1);
    p->inv_dx = 1.0 / p->dx;
    p->gas_density = pressure / (K_BOLTZMANN * t_gas);
    p->inv_dx_f = (float)p->inv_dx;
    p->factor_e = (float)(p->dt_e / E_MASS * (-E_CHARGE));
    p->factor_i = (float)(p->dt_i / AR_MASS * E_CHARGE);
    p->weight_factor = (float)(weight / (area * p->dx));
    p->dt_e_f = (float)p->dt_e;
    p->dt_i_f = (float)p->dt_i;
    p->gap_f = (float)gap;
    p->normal_range = (float)sqrt(K_BOLTZMANN * t_gas / AR_MASS);
    p->f1 = (float)(E_MASS / (E_MASS + AR_MASS));
    p->f2 = (float)(AR_MASS / (E_MASS + AR_MASS));
    p->half_me_over_e = (float)(0.5 * E_MASS / E_CHARGE);
    p->half_mu_over_e = (float)(0.5 * MU_ARAR / E_CHARGE);
    p->e_charge_f = (float)E_CHARGE;
    p->alpha_f = (float)(-p->dx * p->dx / EPSILON0);
    p->half_dx_over_eps_f = (float)(p->dx / (2.0 * EPSILON0));
    p->half_inv_dx_f = (float)(0.5 * p->inv_dx);
    p->min_x_f = (float)(0.45 * gap);
    p->max_x_f = (float)(0.55 * gap);
    p->de_eepf_f = (float)0.05;
    p->de_cs_f = (float)DE_CS;
    p->me_f = (float)E_MASS;
    p->exc_th_f = (float)E_EXC_TH;
    p->ion_th_f = (float)E_ION_TH;
}

/* ---- xoshiro128+ (main.rs:860-885; xoshiro128p.c:36-55) ---------------- */
static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

uint32_t or_xo_next(uint32_t s[4]) {
    uint32_t result = s[0] + s[3];
    uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl32(s[3], 11);
    return result;
}

/* main.rs:1473-1487: XOR of the states reached at the set bits of the
 * jump polynomial, i.e. the state advanced by 2^64 calls of next(). */
void or_xo_jump(uint32_t s[4]) {
    static const uint32_t JUMP[4] = {0x8764000b, 0xf542d2d3, 0x6fa035c3, 0x77f2db5b};
    uint32_t acc[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; ++w)
        for (int b = 0; b < 32; ++b) {
            if (JUMP[w] & (1u << b)) {
                acc[0] ^= s[0];
                acc[1] ^= s[1];
                acc[2] ^= s[2];
                acc[3] ^= s[3];
            }
            (void)or_xo_next(s);
        }
    memcpy(s, acc, sizeof(acc));
}

/* Test support, no reference counterpart: the SoA streams s0..s3[j]
 * (j < n) each advanced by k[j] draws of or_xo_next. */
void or_xo_skip_soa(uint32_t *s0, uint32_t *s1, uint32_t *s2, uint32_t *s3, uint32_t n, const uint32_t *k) {
    for (uint32_t j = 0; j < n; ++j) {
        uint32_t st[4] = {s0[j], s1[j], s2[j], s3[j]};
        for (uint32_t r = 0; r < k[j]; ++r) (void)or_xo_next(st);
        s0[j] = st[0];
        s1[j] = st[1];
        s2[j] = st[2];
        s3[j] = st[3];
    }
}

/* main.rs:883-885: 23 high bits into the mantissa of 1.0, minus 1. */
float or_u32_to_real(uint32_t x) {
    union { uint32_t u; float f; } c;
    c.u = 0x3F800000u | (x >> 9);
    return c.f - 1.0f;
}

/* main.rs:1446-1496: stream k is the master state jumped k times. */
void or_seed_streams(const uint32_t master[4], uint32_t n, uint32_t *s0,
                     uint32_t *s1, uint32_t *s2, uint32_t *s3) {
    uint32_t st[4] = {master[0], master[1], master[2], master[3]};
    for (uint32_t k = 0; k < n; ++k) {
        s0[k] = st[0];
        s1[k] = st[1];
        s2[k] = st[2];
        s3[k] = st[3];
        or_xo_jump(st);
    }
}

static inline float draw(uint32_t s[4]) { return or_u32_to_real(or_xo_next(s)); }

/* main.rs:908-969: four uniforms clamped at 1e-9, two Box-Muller pairs,
 * three normals kept, scaled by sigma. */
void or_three_normal(uint32_t s[4], float sigma, float out[3]) {
    float u1 = fmaxf(draw(s), 1e-9f);
    float u2 = fmaxf(draw(s), 1e-9f);
    float u3 = fmaxf(draw(s), 1e-9f);
    float u4 = fmaxf(draw(s), 1e-9f);
    const float two_pi = (float)OR_TWO_PI;
    float rad1 = sqrtf(-2.0f * logf(u1));
    float a1 = two_pi * u2;
    float rad2 = sqrtf(-2.0f * logf(u3));
    float a2 = two_pi * u4;
    out[0] = rad1 * cosf(a1) * sigma;
    out[1] = rad1 * sinf(a1) * sigma;
    out[2] = rad2 * cosf(a2) * sigma;
}

/* ---- cross sections (main.rs:1378-1444; C/eduPIC.cc:173-229) ----------- */
void or_cross_sections(double gas_density, float *cs, float *stot_e, float *stot_i) {
    for (int i = 0; i < CS_RANGES; ++i) {
        double e = (i == 0) ? DE_CS : (double)i * DE_CS;
        /* Phelps & Petrovic 1999 electron cross sections */
        double qmel = fabs(6.0 / pow(1.0 + e / 0.1 + pow(e / 0.6, 2.0), 3.3) -
                           1.1 * pow(e, 1.4) / (1.0 + pow(e / 15.0, 1.2)) /
                               sqrt(1.0 + pow(e / 5.5, 2.5) + pow(e / 60.0, 4.1))) +
                      0.05 / pow(1.0 + e / 10.0, 2.0) + 0.01 * pow(e, 3.0) / (1.0 + pow(e / 12.0, 6.0));
        double qexc = 0.0, qion = 0.0;
        if (e > E_EXC_TH)
            qexc = 0.034 * pow(e - 11.5, 1.1) * (1.0 + pow(e / 15.0, 2.8)) / (1.0 + pow(e / 23.0, 5.5)) +
                   0.023 * (e - 11.5) / pow(1.0 + e / 80.0, 1.9);
        if (e > E_ION_TH)
            qion = 970.0 * (e - 15.8) / pow(70.0 + e, 2.0) + 0.06 * pow(e - 15.8, 2.0) * exp(-e / 9.0);
        /* Phelps 1994 ion cross sections at E_lab = 2 E_com */
        double el = 2.0 * e;
        double qmom = 1.15e-18 * pow(el, -0.1) * pow(1.0 + 0.015 / el, 0.6);
        double qiso = 2e-19 * pow(el, -0.5) / (1.0 + el) + 3e-19 * el / pow(1.0 + el / 3.0, 2.0);
        double qback = 0.5 * (qmom - qiso);
        cs[0 * CS_RANGES + i] = (float)(qmel * 1.0e-20);
        cs[1 * CS_RANGES + i] = (float)(qexc * 1.0e-20);
        cs[2 * CS_RANGES + i] = (float)(qion * 1.0e-20);
        cs[3 * CS_RANGES + i] = (float)qiso;
        cs[4 * CS_RANGES + i] = (float)qback;
    }
    /* main.rs:1432-1441: totals summed in f64 from the f32 table values */
    for (int i = 0; i < CS_RANGES; ++i) {
        double se = (double)cs[i] + (double)cs[CS_RANGES + i] + (double)cs[2 * CS_RANGES + i];
        double si = (double)cs[3 * CS_RANGES + i] + (double)cs[4 * CS_RANGES + i];
        stot_e[i] = (float)(se * gas_density);
        stot_i[i] = (float)(si * gas_density);
    }
}

/* ---- deposition (main.rs:356-406) -------------------------------------- */
/* Cell index is clamped to [0, n_g-2] so that x == L deposits onto the last
 * two nodes instead of past the end of the grid (main.rs:385-386 would index
 * N_G there). */
static inline int cell_of(float pos, int n_g) {
    int q = (int)pos;
    q = q < 0 ? 0 : q;
    return q > n_g - 2 ? n_g - 2 : q;
}

void or_deposit(const float *x, uint32_t n, const or_params *p, double *dens) {
    const int ng = p->n_g;
    const float wf = p->weight_factor;
    for (int k = 0; k < ng; ++k) dens[k] = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        float pos = x[i] * p->inv_dx_f;
        int q = cell_of(pos, ng);
        float rem = pos - (float)q;
        float c1 = (1.0f - rem) * wf;
        float c2 = rem * wf;
        dens[q] += (double)c1;
        dens[q + 1] += (double)c2;
    }
    dens[0] *= 2.0;
    dens[ng - 1] *= 2.0;
}

void or_deposit_f64(const double *x, uint32_t n, const or_params *p, double *dens) {
    const int ng = p->n_g;
    const double wf = p->weight / (p->area * p->dx);
    for (int k = 0; k < ng; ++k) dens[k] = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double c0 = x[i] * p->inv_dx;
        int q = (int)c0;
        if (q > ng - 2) q = ng - 2;
        dens[q] += (q + 1 - c0) * wf;
        dens[q + 1] += (c0 - q) * wf;
    }
    dens[0] *= 2.0;
    dens[ng - 1] *= 2.0;
}

/* ---- Poisson (C/eduPIC.cc:461-494, textbook Thomas algorithm) ---------- */
void or_poisson_thomas(const double *rho, double pot0, const or_params *p,
                       double *pot, double *efield) {
    const int ng = p->n_g;
    const double dx = p->dx, alpha = -dx * dx / EPSILON0, s = 1.0 / (2.0 * dx);
    double *g = (double *)malloc(sizeof(double) * ng);
    double *w = (double *)malloc(sizeof(double) * ng);
    double *f = (double *)calloc(ng, sizeof(double));
    pot[0] = pot0;
    pot[ng - 1] = 0.0;
    for (int i = 1; i <= ng - 2; ++i) f[i] = alpha * rho[i];
    f[1] -= pot[0];
    f[ng - 2] -= pot[ng - 1];
    w[1] = 1.0 / -2.0;
    g[1] = f[1] / -2.0;
    for (int i = 2; i <= ng - 2; ++i) {
        w[i] = 1.0 / (-2.0 - w[i - 1]);
        g[i] = (f[i] - g[i - 1]) / (-2.0 - w[i - 1]);
    }
    pot[ng - 2] = g[ng - 2];
    for (int i = ng - 3; i > 0; --i) pot[i] = g[i] - w[i] * pot[i + 1];
    for (int i = 1; i <= ng - 2; ++i) efield[i] = (pot[i - 1] - pot[i + 1]) * s;
    efield[0] = (pot[0] - pot[1]) * p->inv_dx - rho[0] * dx / (2.0 * EPSILON0);
    efield[ng - 1] = (pot[ng - 2] - pot[ng - 1]) * p->inv_dx + rho[ng - 1] * dx / (2.0 * EPSILON0);
    free(g);
    free(w);
    free(f);
}

/* main.rs:408-494: the Thomas recurrences for A=C=1, B=-2 unrolled into two
 * inclusive prefix sums: G_i = sum_{j<=i} j f_j, r_i = -G_i / (i (i+1)),
 * pot_i = i * sum_{j>=i} r_j  (1 <= i <= N_G-2). */
void or_poisson_scan_f64(const double *e_d, const double *i_d, double pot0,
                         const or_params *p, double *pot, double *efield) {
    const int ng = p->n_g;
    const double dx = p->dx, alpha = -dx * dx / EPSILON0;
    double *rho = (double *)malloc(sizeof(double) * ng);
    double *r = (double *)calloc(ng, sizeof(double));
    for (int i = 0; i < ng; ++i) rho[i] = E_CHARGE * (i_d[i] - e_d[i]);
    double G = 0.0;
    for (int i = 1; i <= ng - 2; ++i) {
        double f = alpha * rho[i] - (i == 1 ? pot0 : 0.0);
        G += (double)i * f;
        r[i] = -G / ((double)i * (double)(i + 1));
    }
    double tail = 0.0;
    pot[ng - 1] = 0.0;
    for (int i = ng - 2; i >= 1; --i) {
        tail += r[i];
        pot[i] = (double)i * tail;
    }
    pot[0] = pot0;
    for (int i = 1; i <= ng - 2; ++i) efield[i] = (pot[i - 1] - pot[i + 1]) * (0.5 * p->inv_dx);
    efield[0] = (pot[0] - pot[1]) * p->inv_dx - rho[0] * dx / (2.0 * EPSILON0);
    efield[ng - 1] = (pot[ng - 2] - pot[ng - 1]) * p->inv_dx + rho[ng - 1] * dx / (2.0 * EPSILON0);
    free(rho);
    free(r);
}

/* main.rs:408-494 in FP32 with serial prefix sums (R_total - R_i + r_i form) */
void or_poisson_scan_f32(const float *e_d, const float *i_d, float pot0,
                         const or_params *p, float *pot, float *efield) {
    const int ng = p->n_g;
    float *rho = (float *)malloc(sizeof(float) * ng);
    float *r = (float *)calloc(ng, sizeof(float));
    float *R = (float *)calloc(ng, sizeof(float));
    for (int i = 0; i < ng; ++i) rho[i] = p->e_charge_f * (i_d[i] - e_d[i]);
    float G = 0.0f, acc = 0.0f;
    for (int i = 1; i <= ng - 2; ++i) {
        float f = p->alpha_f * rho[i] - (i == 1 ? pot0 : 0.0f);
        G += (float)i * f;
        r[i] = (-G / ((float)i + 1.0f)) / (float)i;
        acc += r[i];
        R[i] = acc;
    }
    float rt = R[ng - 2];
    pot[0] = pot0;
    for (int i = 1; i <= ng - 2; ++i) pot[i] = (float)i * (rt - R[i] + r[i]);
    pot[ng - 1] = 0.0f;
    for (int i = 1; i <= ng - 2; ++i) efield[i] = (pot[i - 1] - pot[i + 1]) * p->half_inv_dx_f;
    efield[0] = (pot[0] - pot[1]) * p->inv_dx_f - rho[0] * p->half_dx_over_eps_f;
    efield[ng - 1] = (pot[ng - 2] - pot[ng - 1]) * p->inv_dx_f + rho[ng - 1] * p->half_dx_over_eps_f;
    free(rho);
    free(r);
    free(R);
}

/* main.rs:496-564: general source, Dirichlet values on both ends */
void or_poisson_flexible_f64(const double *src, int n, double psi_l, double psi_r,
                             double dx, double *pot) {
    double *r = (double *)calloc(n, sizeof(double));
    double G = 0.0;
    for (int i = 1; i <= n - 2; ++i) {
        double f = dx * dx * src[i];
        if (i == 1) f -= psi_l;
        if (i == n - 2) f -= psi_r;
        G += (double)i * f;
        r[i] = -G / ((double)i * (double)(i + 1));
    }
    double tail = 0.0;
    for (int i = n - 2; i >= 1; --i) {
        tail += r[i];
        pot[i] = (double)i * tail;
    }
    pot[0] = psi_l;
    pot[n - 1] = psi_r;
    free(r);
}

/* ---- push + compaction (main.rs:566-678) ------------------------------- */
/* The FP32 operation sequence shared with the HIP kernel:
 *   pos = x*inv_dx; p = min((int)pos, n_g-2); c = pos - p;
 *   E   = fmaf(c, E[p+1], (1-c)*E[p]);
 *   vx' = fmaf(factor, E, vx);   x' = fmaf(vx', dt, x);
 *   keep iff 0 <= x' <= L.                                                */
uint32_t or_move_and_compact(const float *efield, const float *sx, const float *svx,
                             const float *svy, const float *svz, uint32_t n,
                             float factor, float dt, float mass_over_2e,
                             const or_params *p, float *dx_, float *dvx, float *dvy,
                             float *dvz, uint64_t abs_counts[2], uint32_t *ifed) {
    uint32_t out = 0;
    for (uint32_t i = 0; i < n; ++i) {
        float x = sx[i];
        float pos = x * p->inv_dx_f;
        int q = cell_of(pos, p->n_g);
        float c = pos - (float)q;
        float e = fmaf(c, efield[q + 1], (1.0f - c) * efield[q]);
        float nvx = fmaf(factor, e, svx[i]);
        float nx = fmaf(nvx, dt, x);
        if (nx >= 0.0f && nx <= p->gap_f) {
            dx_[out] = nx;
            dvx[out] = nvx;
            dvy[out] = svy[i];
            dvz[out] = svz[i];
            ++out;
        } else {
            int side = nx < 0.0f ? 0 : 1;
            if (abs_counts) abs_counts[side]++;
            if (ifed) {
                float v2 = nvx * nvx + svy[i] * svy[i] + svz[i] * svz[i];
                float en = mass_over_2e * v2;
                int b = (int)en; /* DE_IFED = 1 eV */
                if (en >= 0.0f && b < N_IFED) ifed[side * N_IFED + b]++;
            }
        }
    }
    return out;
}

/* ---- collisions (main.rs:971-1292; C/eduPIC.cc:300-455) ---------------- */
static inline int energy_index(float energy, const or_params *p) {
    float c = energy / p->de_cs_f + 0.5f;
    uint32_t idx = (uint32_t)c;
    return idx > CS_RANGES - 1 ? CS_RANGES - 1 : (int)idx;
}

/* Checker of an identity the GPU kernels rely on (not part of the
 * reference): they form energy / DE_CS from the rounded reciprocal
 * y = RN(1 / DE_CS) with one fused correction (Markstein's method:
 * q = RN(e*y), r = fma(-q, DE_CS, e), q' = RN(q + r*y)), and clamp with a
 * comparison that sends a NaN quotient to the last bin.  Returns how many
 * energies with the bit patterns lo, lo + stride, ... <= hi get another bin
 * than energy_index's IEEE division (main.rs:1002-1005); corrected = 0 drops
 * the correction (the plain product e*y), to show that the check can fail. */
uint64_t or_bin_division_mismatches(float de_cs, uint32_t lo, uint32_t hi, uint32_t stride, int corrected) {
    or_params p;
    memset(&p, 0, sizeof(p));
    p.de_cs_f = de_cs;
    const float y = 1.0f / de_cs;
    uint64_t bad = 0;
    for (uint64_t u = lo; u <= hi; u += stride) {
        float e;
        const uint32_t bits = (uint32_t)u;
        memcpy(&e, &bits, sizeof(e));
        float q = e * y;
        if (corrected) {
            const float r = fmaf(-q, de_cs, e);
            q = fmaf(r, y, q);
        }
        const float c = q + 0.5f;
        const int fast = c < (float)(CS_RANGES - 1) ? (int)(uint32_t)c : CS_RANGES - 1;
        const float ce = e / de_cs + 0.5f;
        /* the reference's cast saturates (Rust `as usize`); the C cast of an
           infinity is undefined, so the exact bin is clamped first */
        const int exact = ce < 4294967040.0f ? energy_index(e, &p) : CS_RANGES - 1;
        bad += fast != exact;
    }
    return bad;
}

static inline void euler_angles(float gx, float gy, float gz, float *st, float *ct,
                                float *sp, float *cp) {
    const float half_pi = (float)(0.5 * OR_PI);
    float theta, phi;
    if (gx == 0.0f) theta = half_pi;
    else theta = atan2f(sqrtf(gy * gy + gz * gz), gx);
    if (gy == 0.0f) phi = gz > 0.0f ? half_pi : -half_pi;
    else phi = atan2f(gz, gy);
    *st = sinf(theta);
    *ct = cosf(theta);
    *sp = sinf(phi);
    *cp = cosf(phi);
}

static inline void rotate(float g, float chi, float eta, float st, float ct, float sp,
                          float cp, float out[3]) {
    float sc = sinf(chi), cc = cosf(chi), se = sinf(eta), ce = cosf(eta);
    out[0] = g * (ct * cc - st * sc * ce);
    out[1] = g * (st * cp * cc + ct * cp * sc * ce - sp * sc * se);
    out[2] = g * (st * sp * cc + ct * sp * sc * ce + cp * sc * se);
}

/* Returns 0: no collision, 1: elastic, 2: excitation, 3: ionisation.  v is
 * updated in place; for 3 the spawned electron/ion velocities are written. */
int or_collide_e_one(const float *stot_e, const float *cs, float v[3], uint32_t s[4],
                     float spawn_e_v[3], float spawn_i_v[3], const or_params *p) {
    const float two_pi = (float)OR_TWO_PI, pi = (float)OR_PI;
    float v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    float energy = p->half_me_over_e * v2;
    int idx = energy_index(energy, p);
    float velocity = sqrtf(v2);
    float nu = stot_e[idx] * velocity;
    float rv = draw(s);
    float pc = 1.0f - expf(-nu * p->dt_e_f);
    if (!(rv < pc)) return 0;

    /* collision_e, main.rs:1035-1165 */
    float gx = v[0], gy = v[1], gz = v[2];
    float g = sqrtf(gx * gx + gy * gy + gz * gz);
    float wx = p->f1 * v[0], wy = p->f1 * v[1], wz = p->f1 * v[2];
    float t0 = cs[0 * CS_RANGES + idx];
    float t1 = t0 + cs[1 * CS_RANGES + idx];
    float t2 = t1 + cs[2 * CS_RANGES + idx];
    float st, ct, sp, cp;
    euler_angles(gx, gy, gz, &st, &ct, &sp, &cp);
    float chi, eta;
    int kind;
    float rnd = draw(s);
    if (rnd < t0 / t2) {
        chi = acosf(1.0f - 2.0f * draw(s));
        eta = two_pi * draw(s);
        kind = 1;
    } else if (rnd < t1 / t2) {
        float en = 0.5f * p->me_f * g * g;
        en = fabsf(en - p->exc_th_f * p->e_charge_f);
        g = sqrtf(2.0f * en / p->me_f);
        chi = acosf(1.0f - 2.0f * draw(s));
        eta = two_pi * draw(s);
        kind = 2;
    } else {
        float en = 0.5f * p->me_f * g * g;
        en = fabsf(en - p->ion_th_f * p->e_charge_f);
        float r = draw(s);
        float e_new = 10.0f * tanf(r * atanf(en / p->e_charge_f / 20.0f)) * p->e_charge_f;
        float e_orig = fabsf(en - e_new);
        g = sqrtf(2.0f * e_orig / p->me_f);
        float g_new = sqrtf(2.0f * e_new / p->me_f);
        chi = acosf(sqrtf(e_orig / en));
        float chi_new = acosf(sqrtf(e_new / en));
        eta = two_pi * draw(s);
        float eta_new = eta + pi;
        float gn[3];
        rotate(g_new, chi_new, eta_new, st, ct, sp, cp, gn);
        spawn_e_v[0] = wx + p->f2 * gn[0];
        spawn_e_v[1] = wy + p->f2 * gn[1];
        spawn_e_v[2] = wz + p->f2 * gn[2];
        or_three_normal(s, p->normal_range, spawn_i_v);
        kind = 3;
    }
    float go[3];
    rotate(g, chi, eta, st, ct, sp, cp, go);
    v[0] = wx + p->f2 * go[0];
    v[1] = wy + p->f2 * go[1];
    v[2] = wz + p->f2 * go[2];
    return kind;
}

void or_collide_e(const float *stot_e, const float *cs, uint32_t n_active,
                  float *x, float *vx, float *vy, float *vz,
                  uint32_t *r0, uint32_t *r1, uint32_t *r2, uint32_t *r3,
                  float *ix, float *ivx, float *ivy, float *ivz,
                  uint32_t *alive_e, uint32_t *alive_i, uint32_t cap_e,
                  uint32_t cap_i, const or_params *p, uint64_t *n_coll) {
    for (uint32_t i = 0; i < n_active; ++i) {
        uint32_t s[4] = {r0[i], r1[i], r2[i], r3[i]};
        float v[3] = {vx[i], vy[i], vz[i]};
        float se[3], si[3];
        int kind = or_collide_e_one(stot_e, cs, v, s, se, si, p);
        r0[i] = s[0];
        r1[i] = s[1];
        r2[i] = s[2];
        r3[i] = s[3];
        if (kind == 0) continue;
        if (n_coll) ++*n_coll;
        vx[i] = v[0];
        vy[i] = v[1];
        vz[i] = v[2];
        if (kind == 3) {
            uint32_t je = (*alive_e)++;
            if (je < cap_e) {
                x[je] = x[i];
                vx[je] = se[0];
                vy[je] = se[1];
                vz[je] = se[2];
            }
            uint32_t ji = (*alive_i)++;
            if (ji < cap_i) {
                ix[ji] = x[i];
                ivx[ji] = si[0];
                ivy[ji] = si[1];
                ivz[ji] = si[2];
            }
        }
    }
}

int or_collide_i_one(const float *stot_i, const float *cs, float v[3], uint32_t s[4],
                     const or_params *p) {
    const float two_pi = (float)OR_TWO_PI, pi = (float)OR_PI;
    float a[3];
    or_three_normal(s, p->normal_range, a);
    float gx = v[0] - a[0], gy = v[1] - a[1], gz = v[2] - a[2];
    float g2 = gx * gx + gy * gy + gz * gz;
    float energy = p->half_mu_over_e * g2;
    int idx = energy_index(energy, p);
    float g = sqrtf(g2);
    float nu = stot_i[idx] * g;
    float rv = draw(s);
    float pc = 1.0f - expf(-nu * p->dt_i_f);
    if (!(rv < pc)) return 0;

    /* collision_i, main.rs:1221-1292 */
    float t0 = cs[3 * CS_RANGES + idx];
    float t1 = t0 + cs[4 * CS_RANGES + idx];
    float wx = 0.5f * (v[0] + a[0]), wy = 0.5f * (v[1] + a[1]), wz = 0.5f * (v[2] + a[2]);
    float st, ct, sp, cp;
    euler_angles(gx, gy, gz, &st, &ct, &sp, &cp);
    float chi;
    int kind;
    float rnd = draw(s);
    if (rnd < t0 / t1) {
        chi = acosf(1.0f - 2.0f * draw(s));
        kind = 1;
    } else {
        chi = pi;
        kind = 2;
    }
    float eta = two_pi * draw(s);
    float go[3];
    rotate(g, chi, eta, st, ct, sp, cp, go);
    v[0] = wx + 0.5f * go[0];
    v[1] = wy + 0.5f * go[1];
    v[2] = wz + 0.5f * go[2];
    return kind;
}

void or_collide_i(const float *stot_i, const float *cs, uint32_t n_active,
                  float *vx, float *vy, float *vz,
                  uint32_t *r0, uint32_t *r1, uint32_t *r2, uint32_t *r3,
                  const or_params *p, uint64_t *n_coll) {
    for (uint32_t i = 0; i < n_active; ++i) {
        uint32_t s[4] = {r0[i], r1[i], r2[i], r3[i]};
        float v[3] = {vx[i], vy[i], vz[i]};
        int kind = or_collide_i_one(stot_i, cs, v, s, p);
        r0[i] = s[0];
        r1[i] = s[1];
        r2[i] = s[2];
        r3[i] = s[3];
        if (kind) {
            if (n_coll) ++*n_coll;
            vx[i] = v[0];
            vy[i] = v[1];
            vz[i] = v[2];
        }
    }
}

/* ---- EEPF (main.rs:1301-1339) ------------------------------------------ */
void or_accumulate_eepf(const float *x, const float *vx, const float *vy,
                        const float *vz, const float *efield, uint32_t n,
                        const or_params *p, uint32_t *eepf) {
    const float half_kick = (float)(p->dt_e * E_CHARGE / E_MASS);
    for (uint32_t i = 0; i < n; ++i) {
        float xi = x[i];
        if (xi <= p->min_x_f || xi >= p->max_x_f) continue;
        float pos = xi * p->inv_dx_f;
        int q = cell_of(pos, p->n_g);
        float c = pos - (float)q;
        float e = fmaf(c, efield[q + 1], (1.0f - c) * efield[q]);
        float mvx = vx[i] - 0.5f * e * half_kick;
        float v2 = mvx * mvx + vy[i] * vy[i] + vz[i] * vz[i];
        float en = p->half_me_over_e * v2;
        uint32_t b = (uint32_t)(en / p->de_eepf_f);
        if (b < N_EEPF) eepf[b]++;
    }
}

/* ---- whole step (main.rs:1725-1808) ------------------------------------ */
static void copy_back(float *d, const float *s, uint32_t n) { memcpy(d, s, sizeof(float) * n); }

void or_step(or_state *s, const or_params *p, int t, int collisions) {
    const int ion_step = (t % p->n_sub) == 0;
    const int ng = p->n_g;
    or_deposit(s->ex, s->n_e, p, s->e_dens);
    if (ion_step) or_deposit(s->ix, s->n_i, p, s->i_dens);
    /* main.rs:1741: host f64 phase, cast to f32 */
    float pot0 = (float)(p->v0 * cos(((double)t / (double)p->n_t) * OR_TWO_PI));
    double *pot = (double *)malloc(sizeof(double) * ng);
    double *ef = (double *)malloc(sizeof(double) * ng);
    or_poisson_scan_f64(s->e_dens, s->i_dens, (double)pot0, p, pot, ef);
    for (int k = 0; k < ng; ++k) {
        s->pot[k] = (float)pot[k];
        s->efield[k] = (float)ef[k];
    }
    free(pot);
    free(ef);
    uint32_t ne = or_move_and_compact(s->efield, s->ex, s->evx, s->evy, s->evz, s->n_e,
                                      p->factor_e, p->dt_e_f, 0.0f, p, s->tx, s->tvx,
                                      s->tvy, s->tvz, s->abs_e, NULL);
    copy_back(s->ex, s->tx, ne);
    copy_back(s->evx, s->tvx, ne);
    copy_back(s->evy, s->tvy, ne);
    copy_back(s->evz, s->tvz, ne);
    s->n_e = ne;
    if (ion_step) {
        uint32_t ni = or_move_and_compact(s->efield, s->ix, s->ivx, s->ivy, s->ivz, s->n_i,
                                          p->factor_i, p->dt_i_f,
                                          (float)(0.5 * AR_MASS / E_CHARGE), p, s->tx,
                                          s->tvx, s->tvy, s->tvz, s->abs_i, s->ifed);
        copy_back(s->ix, s->tx, ni);
        copy_back(s->ivx, s->tvx, ni);
        copy_back(s->ivy, s->tvy, ni);
        copy_back(s->ivz, s->tvz, ni);
        s->n_i = ni;
    }
    if (collisions) {
        uint32_t alive = s->n_e;
        or_collide_e(s->stot_e, s->cs, s->n_e, s->ex, s->evx, s->evy, s->evz, s->er[0],
                     s->er[1], s->er[2], s->er[3], s->ix, s->ivx, s->ivy, s->ivz, &alive,
                     &s->n_i, s->cap, s->cap, p, &s->coll_e);
        s->n_e = alive > s->cap ? s->cap : alive;
        if (s->n_i > s->cap) s->n_i = s->cap;
        if (ion_step)
            or_collide_i(s->stot_i, s->cs, s->n_i, s->ivx, s->ivy, s->ivz, s->ir[0], s->ir[1],
                         s->ir[2], s->ir[3], p, &s->coll_i);
    }
}
