/*
 * mfc_oracle.c -- CPU FP64 restatement of MicroFC's simulation RHS.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Compile with
 * -ffp-contract=off so that every multiply and add rounds separately, as the
 * reference's Fortran statements do.  Every function cites the reference
 * file:line whose arithmetic it restates; evaluation order follows the
 * Fortran left-to-right rule for operators of equal precedence, and x**2d0 is
 * evaluated as x*x.
 */
#include "mfc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads for the pencil and update loops (OpenMP builds); the result does
 * not depend on the count. */
void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}
int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

#define SGM_EPS 1e-16 /* m_global_parameters.fpp:39 */

/* ------------------------------------------------------------------------ */
/* sizes (m_global_parameters.fpp:285-381)                                   */

int orc_num_dims(const orc_case *c) { return 1 + (c->n > 0) + (c->p > 0); }
int orc_sys_size(const orc_case *c) { return 2 * c->num_fluids + orc_num_dims(c) + 1; }
/* m_global_parameters.fpp:355-359: twice the WENO half-width plus 2 ghost
 * layers when any fluid has a Reynolds number, else the half-width plus 2. */
int orc_buff_size(const orc_case *c) {
    const int polyn = (c->weno_order - 1) / 2;
    int i, r, visc = 0;
    for (i = 0; i < c->num_fluids; i++)
        for (r = 0; r < 2; r++) visc = visc || c->Re[i][r] > 0.0;
    return visc ? 2 * polyn + 2 : polyn + 2;
}

typedef struct {
    int m, n, p, b, bj, bk;
    long nxg, nyg, nzg, ng;
    int D, S, nf;
    int contxe, momxb, E_idx, advxb; /* 0-based */
} geom;

static void mkgeom(const orc_case *c, geom *g) {
    g->m = c->m;
    g->n = c->n;
    g->p = c->p;
    g->b = orc_buff_size(c);
    g->bj = c->n > 0 ? g->b : 0;
    g->bk = c->p > 0 ? g->b : 0;
    g->nxg = c->m + 2 * g->b + 1;
    g->nyg = c->n + 2 * g->bj + 1;
    g->nzg = c->p + 2 * g->bk + 1;
    g->ng = g->nxg * g->nyg * g->nzg;
    g->D = orc_num_dims(c);
    g->nf = c->num_fluids;
    g->S = 2 * g->nf + g->D + 1;
    g->contxe = g->nf - 1;
    g->momxb = g->nf;
    g->E_idx = g->nf + g->D;
    g->advxb = g->E_idx + 1;
}

long orc_ghost_len(const orc_case *c) {
    geom g;
    mkgeom(c, &g);
    return g.ng;
}

static inline long gidx(const geom *g, int i, int j, int k) {
    return (long)(i + g->b) + g->nxg * ((long)(j + g->bj) + g->nyg * (long)(k + g->bk));
}

static inline long ridx(const geom *g, int i, int j, int k) {
    return (long)i + (long)(g->m + 1) * ((long)j + (long)(g->n + 1) * (long)k);
}

/* ------------------------------------------------------------------------ */
/* grid ghost buffers: m_start_up.fpp:517-655                                */

/* Cell centres with ghosts from the ghosted cb and ds (m_start_up.fpp:356-359
 * for the interior, :551 and :582 for the ghosts). */
static void cell_centres(int s, int b, const double *cb_, const double *ds_, double *cc_) {
    const double *cb = cb_ + 1 + b, *ds = ds_ + b;
    double *cc = cc_ + b;
    int i;
    for (i = 0; i <= s; i++) cc[i] = cb[i - 1] + ds[i] / 2.0;
    for (i = 1; i <= b; i++) cc[-i] = cc[1 - i] - (ds[1 - i] + ds[-i]) / 2.0;
    for (i = 1; i <= b; i++) cc[s + i] = cc[s + (i - 1)] + (ds[s + (i - 1)] + ds[s + i]) / 2.0;
}

int orc_grid_buffers(int s, int b, int bc_beg, int bc_end, double *cb_, double *ds_, double *cc_) {
    double *cb = cb_ + 1 + b; /* cb[i] = s_cb(i), i in [-1-b, s+b] */
    double *ds = ds_ + b;     /* ds[i] = ds(i),   i in [-b, s+b]   */
    double *cc = cc_ + b;
    int i;
    /* m_start_up.fpp:356-359 */
    for (i = 0; i <= s; i++) ds[i] = cb[i] - cb[i - 1];
    /* beginning: m_start_up.fpp:527-550 */
    if (bc_beg <= -3) {
        for (i = 1; i <= b; i++) ds[-i] = ds[0];
    } else if (bc_beg == -2) {
        for (i = 1; i <= b; i++) ds[-i] = ds[i - 1];
    } else if (bc_beg == -1) {
        for (i = 1; i <= b; i++) ds[-i] = ds[s - (i - 1)];
    } else {
        return -1; /* processor boundaries are not part of a single-domain run */
    }
    for (i = 1; i <= b; i++) cb[-1 - i] = cb[-i] - ds[-i];
    /* end: m_start_up.fpp:557-582 */
    if (bc_end <= -3) {
        for (i = 1; i <= b; i++) ds[s + i] = ds[s];
    } else if (bc_end == -2) {
        for (i = 1; i <= b; i++) ds[s + i] = ds[s - (i - 1)];
    } else if (bc_end == -1) {
        for (i = 1; i <= b; i++) ds[s + i] = ds[i - 1];
    } else {
        return -1;
    }
    for (i = 1; i <= b; i++) cb[s + i] = cb[s + (i - 1)] + ds[s + i];
    cell_centres(s, b, cb_, ds_, cc_);
    (void)cc;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* WENO coefficients: m_weno.fpp:168-363                                     */

static inline double sq(double x) { return x * x; }

int orc_weno_coefficients(int s, int b, int weno_order, const double *cb_, double *out) {
    const double *cb = cb_ + 1 + b;
    const long nc = s + 2 * b + 1;
    const int polyn = (weno_order - 1) / 2;
    int i;
    memset(out, 0, sizeof(double) * 27 * (size_t)nc);
    if (weno_order == 1) return 0;
#define C(plane, j) out[(long)(plane) * nc + (long)((j) + b)]
    if (weno_order == 3) {
        /* m_weno.fpp:194-217 */
        for (i = -b - 1 + polyn; i <= s + b - 1 - polyn; i++) {
            const double S0 = cb[i], S1 = cb[i + 1], S2 = cb[i + 2], Sm1 = cb[i - 1];
            const int j = i + 1;
            C(6 + 0, j) = (S0 - S1) / (S0 - S2);  /* cbR(0,0) */
            C(6 + 2, j) = (S0 - S1) / (Sm1 - S1); /* cbR(1,0) */
            C(0 + 0, j) = -C(6 + 0, j);
            C(0 + 2, j) = -C(6 + 2, j);
            C(15 + 0, j) = (Sm1 - S1) / (Sm1 - S2); /* d_cbR(0) */
            C(12 + 0, j) = (Sm1 - S0) / (Sm1 - S2); /* d_cbL(0) */
            C(15 + 1, j) = 1.0 - C(15 + 0, j);
            C(12 + 1, j) = 1.0 - C(12 + 0, j);
            C(18 + 0, j) = 4.0 * sq(S0 - S1) / sq(S0 - S2);  /* beta(0,0) */
            C(18 + 3, j) = 4.0 * sq(S0 - S1) / sq(Sm1 - S1); /* beta(1,0) */
        }
        return 0;
    }
    /* WENO5: m_weno.fpp:223-346 */
    for (i = -b - 1 + polyn; i <= s + b - 1 - polyn; i++) {
        const double S0 = cb[i], S1 = cb[i + 1], S2 = cb[i + 2], S3 = cb[i + 3];
        const double Sm1 = cb[i - 1], Sm2 = cb[i - 2];
        const int j = i + 1;
        const double h4 = 4.0 * sq(S0 - S1);
        /* poly_coef_cbR(j,0,0), (1,0), (1,1), (2,1) */
        C(6 + 0, j) = ((S0 - S1) * (S1 - S2)) / ((S0 - S3) * (S3 - S1));
        C(6 + 2, j) = ((Sm1 - S1) * (S1 - S0)) / ((Sm1 - S2) * (S2 - S0));
        C(6 + 3, j) = ((S0 - S1) * (S1 - S2)) / ((Sm1 - S1) * (Sm1 - S2));
        C(6 + 5, j) = ((S0 - S1) * (S1 - Sm1)) / ((Sm2 - S0) * (Sm2 - S1));
        /* poly_coef_cbL(j,0,0), (1,0), (1,1), (2,1) */
        C(0 + 0, j) = ((S1 - S0) * (S0 - S2)) / ((S0 - S3) * (S3 - S1));
        C(0 + 2, j) = ((S0 - Sm1) * (S0 - S1)) / ((Sm1 - S2) * (S0 - S2));
        C(0 + 3, j) = ((S1 - S0) * (S0 - S2)) / ((Sm1 - S1) * (Sm1 - S2));
        C(0 + 5, j) = ((Sm1 - S0) * (S0 - S1)) / ((Sm2 - S0) * (Sm2 - S1));
        /* poly_coef_cbR(j,0,1), (2,0); cbL(j,0,1), (2,0) */
        C(6 + 1, j) = ((S0 - S2) + (S1 - S3)) / ((S0 - S2) * (S0 - S3)) * ((S0 - S1));
        C(6 + 4, j) = ((Sm2 - S1) + (Sm1 - S1)) / ((Sm1 - S1) * (S1 - Sm2)) * ((S1 - S0));
        C(0 + 1, j) = ((S0 - S2) + (S0 - S3)) / ((S0 - S2) * (S0 - S3)) * ((S1 - S0));
        C(0 + 4, j) = ((Sm2 - S0) + (Sm1 - S1)) / ((Sm2 - S1) * (S1 - Sm1)) * ((S0 - S1));
        /* ideal weights d_cbR(0), d_cbR(2), d_cbL(0), d_cbL(2), then (1) */
        C(15 + 0, j) = ((Sm2 - S1) * (S1 - Sm1)) / ((Sm2 - S3) * (S3 - Sm1));
        C(15 + 2, j) = ((S1 - S2) * (S1 - S3)) / ((Sm2 - S2) * (Sm2 - S3));
        C(12 + 0, j) = ((Sm2 - S0) * (S0 - Sm1)) / ((Sm2 - S3) * (S3 - Sm1));
        C(12 + 2, j) = ((S0 - S2) * (S0 - S3)) / ((Sm2 - S2) * (Sm2 - S3));
        C(15 + 1, j) = 1.0 - C(15 + 0, j) - C(15 + 2, j);
        C(12 + 1, j) = 1.0 - C(12 + 0, j) - C(12 + 2, j);
        /* smoothness-indicator coefficients beta_coef(j,r,q) */
        C(18 + 0, j) = h4 * (10.0 * sq(S1 - S0) + (S1 - S0) * (S2 - S1) + sq(S2 - S1)) /
                       (sq(S0 - S3) * sq(S1 - S3));
        C(18 + 1, j) = h4 *
                       (19.0 * sq(S1 - S0) - (S1 - S0) * (S3 - S1) +
                        2.0 * (S2 - S0) * ((S2 - S0) + (S3 - S1))) /
                       ((S0 - S2) * sq(S0 - S3) * (S3 - S1));
        C(18 + 2, j) = h4 *
                       (10.0 * sq(S1 - S0) + (S1 - S0) * ((S2 - S0) + (S3 - S1)) +
                        sq((S2 - S0) + (S3 - S1))) /
                       (sq(S0 - S2) * sq(S0 - S3));
        C(18 + 3, j) = h4 * (10.0 * sq(S1 - S0) + sq(S0 - Sm1) + (S0 - Sm1) * (S1 - S0)) /
                       (sq(Sm1 - S2) * sq(S0 - S2));
        C(18 + 4, j) = h4 *
                       ((S0 - S1) * ((S0 - Sm1) + 20.0 * (S1 - S0)) +
                        (2.0 * (S0 - Sm1) + (S1 - S0)) * (S2 - S0)) /
                       ((S1 - Sm1) * sq(Sm1 - S2) * (S2 - S0));
        C(18 + 5, j) = h4 * (10.0 * sq(S1 - S0) + (S1 - S0) * (S2 - S1) + sq(S2 - S1)) /
                       (sq(Sm1 - S1) * sq(Sm1 - S2));
        C(18 + 6, j) = h4 *
                       (12.0 * sq(S1 - S0) + sq((S0 - Sm2) + (S0 - Sm1)) +
                        3.0 * ((S0 - Sm2) + (S0 - Sm1)) * (S1 - S0)) /
                       (sq(Sm2 - S1) * sq(Sm1 - S1));
        C(18 + 7, j) = h4 *
                       (19.0 * sq(S1 - S0) + ((S0 - Sm2) * (S0 - S1)) +
                        2.0 * (S1 - Sm1) * ((S0 - Sm2) + (S1 - Sm1))) /
                       ((Sm2 - S0) * sq(Sm2 - S1) * (S1 - Sm1));
        C(18 + 8, j) = h4 * (10.0 * sq(S1 - S0) + sq(S0 - Sm1) + (S0 - Sm1) * (S1 - S0)) /
                       (sq(Sm2 - S0) * sq(Sm2 - S1));
    }
#undef C
    return 0;
}

/* ------------------------------------------------------------------------ */
/* WENO reconstruction of one variable at one cell: m_weno.fpp:391-540       */

void orc_weno_cell(int weno_order, double weno_eps, const double *cf, const double *v, double *vL,
                   double *vR) {
    const double vj = v[2];
    if (weno_order == 1) { /* m_weno.fpp:391-414 */
        *vL = vj;
        *vR = vj;
        return;
    }
    if (weno_order == 3) { /* m_weno.fpp:416-465 */
        const double d0 = v[3] - v[2], dm1 = v[2] - v[1];
        double poly0, poly1, beta0, beta1, a0, a1, sum;
        poly0 = vj + cf[0] * d0;
        poly1 = vj + cf[2] * dm1;
        beta0 = cf[18] * d0 * d0 + weno_eps;
        beta1 = cf[21] * dm1 * dm1 + weno_eps;
        a0 = cf[12] / (beta0 * beta0);
        a1 = cf[13] / (beta1 * beta1);
        sum = a0 + a1;
        *vL = (a0 / sum) * poly0 + (a1 / sum) * poly1;
        poly0 = vj + cf[6] * d0;
        poly1 = vj + cf[8] * dm1;
        a0 = cf[15] / (beta0 * beta0);
        a1 = cf[16] / (beta1 * beta1);
        sum = a0 + a1;
        *vR = (a0 / sum) * poly0 + (a1 / sum) * poly1;
        return;
    }
    /* WENO5: m_weno.fpp:476-531 */
    {
        const double d1 = v[4] - v[3], d0 = v[3] - v[2], dm1 = v[2] - v[1], dm2 = v[1] - v[0];
        double p0, p1, p2, b0, b1, b2, a0, a1, a2, sum, w0, w1, w2;
        p0 = vj + cf[0] * d1 + cf[1] * d0;
        p1 = vj + cf[2] * d0 + cf[3] * dm1;
        p2 = vj + cf[4] * dm1 + cf[5] * dm2;
        b0 = cf[18] * d1 * d1 + cf[19] * d1 * d0 + cf[20] * d0 * d0 + weno_eps;
        b1 = cf[21] * d0 * d0 + cf[22] * d0 * dm1 + cf[23] * dm1 * dm1 + weno_eps;
        b2 = cf[24] * dm1 * dm1 + cf[25] * dm1 * dm2 + cf[26] * dm2 * dm2 + weno_eps;
        a0 = cf[12] / (b0 * b0);
        a1 = cf[13] / (b1 * b1);
        a2 = cf[14] / (b2 * b2);
        sum = a0 + a1 + a2;
        w0 = a0 / sum;
        w1 = a1 / sum;
        w2 = a2 / sum;
        *vL = w0 * p0 + w1 * p1 + w2 * p2;
        p0 = vj + cf[6] * d1 + cf[7] * d0;
        p1 = vj + cf[8] * d0 + cf[9] * dm1;
        p2 = vj + cf[10] * dm1 + cf[11] * dm2;
        a0 = cf[15] / (b0 * b0);
        a1 = cf[16] / (b1 * b1);
        a2 = cf[17] / (b2 * b2);
        sum = a0 + a1 + a2;
        w0 = a0 / sum;
        w1 = a1 / sum;
        w2 = a2 / sum;
        *vR = w0 * p0 + w1 * p1 + w2 * p2;
    }
}

/* ------------------------------------------------------------------------ */
/* Learned WENO5 weights.  No reference counterpart: the model defined in
 * include/mfc_rhs.h replaces w_k of m_weno.fpp:508-510,526-528 by
 * w_plain_k exp(z_k) / sum_m w_plain_m exp(z_m), z = MLP(stencil features),
 * and keeps the plain value where the learned one is not finite or breaks a
 * bound the plain one keeps.                                                */

static int g_learned = 0;
static double g_w1[5][32], g_b1[32], g_w2[32][32], g_b2[32], g_w3[32][3], g_b3[3];

/* FP64 -> FP32 -> BF16, round to nearest even (finite values). */
static double bf16r(double a) {
    float f = (float)a;
    unsigned u;
    memcpy(&u, &f, sizeof u);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    memcpy(&f, &u, sizeof f);
    return f;
}

void orc_set_learned(const float *params) {
    int k, n;
    g_learned = params != NULL;
    if (!params) return;
    for (k = 0; k < 5; k++)
        for (n = 0; n < 32; n++) g_w1[k][n] = bf16r(params[k * 32 + n]);
    for (k = 0; k < 32; k++)
        for (n = 0; n < 32; n++) g_w2[k][n] = bf16r(params[544 + k * 32 + n]);
    for (k = 0; k < 32; k++)
        for (n = 0; n < 3; n++) g_w3[k][n] = bf16r(params[1600 + k * 16 + n]);
    for (n = 0; n < 32; n++) {
        g_b1[n] = params[512 + n];
        g_b2[n] = params[1568 + n];
    }
    for (n = 0; n < 3; n++) g_b3[n] = params[2112 + n];
}

/* Learned weights of one face from its oriented stencil (face between v[2]
 * and v[3]) and the plain weights wp in the same orientation.  The features
 * are part of the model's definition: x_k = (v_k - v_2) * (1/s) with s the
 * largest neighbour difference, evaluated in FP32 on the FP32 stencil (the
 * learned form reconstructs in FP32), then rounded to BF16. */
static void learned_weights(const double *v, const double *wp, double *w) {
    double x[5], h1[32], h2[32], z[3], zm, u[3], su, g;
    float vf[5], sf = 0.0f, vmf = 0.0f, inv;
    int k, n;
    for (k = 0; k < 5; k++) vf[k] = (float)v[k];
    for (k = 0; k < 4; k++) sf = fmaxf(sf, fabsf(vf[k + 1] - vf[k]));
    for (k = 0; k < 5; k++) vmf = fmaxf(vmf, fabsf(vf[k]));
    /* a stencil whose variation is rounding noise keeps the plain weights
       (the rule that admits training pairs, on the FP32 stencil) */
    if (!(sf > 1e-6f * vmf)) {
        for (k = 0; k < 3; k++) w[k] = wp[k];
        return;
    }
    inv = sf > 0.0f ? 1.0f / sf : 0.0f;
    for (k = 0; k < 5; k++) x[k] = bf16r((double)((vf[k] - vf[2]) * inv));
    for (n = 0; n < 32; n++) {
        double acc = 0.0;
        for (k = 0; k < 5; k++) acc += x[k] * g_w1[k][n];
        acc += g_b1[n];
        h1[n] = bf16r(acc > 0.0 ? acc : 0.0);
    }
    for (n = 0; n < 32; n++) {
        double acc = 0.0;
        for (k = 0; k < 32; k++) acc += h1[k] * g_w2[k][n];
        acc += g_b2[n];
        h2[n] = bf16r(acc > 0.0 ? acc : 0.0);
    }
    for (n = 0; n < 3; n++) {
        double acc = 0.0;
        for (k = 0; k < 32; k++) acc += h2[k] * g_w3[k][n];
        z[n] = acc + g_b3[n];
    }
    /* gate: the logits fade out as the plain weights approach the ideal
       weights of a uniform grid for the oriented face, and the plain weights
       are kept exactly within 0.005 of them */
    g = fmin(1.0, fmax(0.0, fabs(wp[0] - 0.3) + fabs(wp[1] - 0.6) + fabs(wp[2] - 0.1) - 0.005) / 0.03);
    if (g == 0.0) {
        for (k = 0; k < 3; k++) w[k] = wp[k];
        return;
    }
    for (k = 0; k < 3; k++) z[k] *= g;
    zm = fmax(z[0], fmax(z[1], z[2]));
    su = 0.0;
    for (k = 0; k < 3; k++) {
        u[k] = wp[k] * exp(z[k] - zm);
        su += u[k];
    }
    for (k = 0; k < 3; k++) w[k] = u[k] / su;
}

static double learned_value(const double *v, const double *p, const double *wp, int kind) {
    double w[3], lv, pv;
    int ok;
    learned_weights(v, wp, w);
    lv = w[0] * p[0] + w[1] * p[1] + w[2] * p[2];
    pv = wp[0] * p[0] + wp[1] * p[1] + wp[2] * p[2];
    ok = isfinite(lv);
    if (kind & 1) ok = ok && !(lv < 0.0 && pv >= 0.0);
    if (kind & 2) ok = ok && !(lv > 1.0 && pv <= 1.0);
    if (kind & 4) ok = ok && !(lv <= 0.0 && pv > 0.0);
    return ok ? lv : pv;
}

void orc_weno5_learned_cell(double weno_eps, const double *cf, const double *v, int kind, double *vL,
                            double *vR) {
    const double vj = v[2];
    const double d1 = v[4] - v[3], d0 = v[3] - v[2], dm1 = v[2] - v[1], dm2 = v[1] - v[0];
    const double b0 = cf[18] * d1 * d1 + cf[19] * d1 * d0 + cf[20] * d0 * d0 + weno_eps;
    const double b1 = cf[21] * d0 * d0 + cf[22] * d0 * dm1 + cf[23] * dm1 * dm1 + weno_eps;
    const double b2 = cf[24] * dm1 * dm1 + cf[25] * dm1 * dm2 + cf[26] * dm2 * dm2 + weno_eps;
    const double bb[3] = {b0 * b0, b1 * b1, b2 * b2};
    double p[3], a[3], wp[3], po[3], wo[3], vr[5], sum;
    int k;
    /* face j+1/2: the stencil is already oriented */
    p[0] = vj + cf[6] * d1 + cf[7] * d0;
    p[1] = vj + cf[8] * d0 + cf[9] * dm1;
    p[2] = vj + cf[10] * dm1 + cf[11] * dm2;
    for (k = 0; k < 3; k++) a[k] = cf[15 + k] / bb[k];
    sum = a[0] + a[1] + a[2];
    for (k = 0; k < 3; k++) wp[k] = a[k] / sum;
    *vR = learned_value(v, p, wp, kind);
    /* face j-1/2: reversed stencil, candidate k becomes oriented 2-k */
    p[0] = vj + cf[0] * d1 + cf[1] * d0;
    p[1] = vj + cf[2] * d0 + cf[3] * dm1;
    p[2] = vj + cf[4] * dm1 + cf[5] * dm2;
    for (k = 0; k < 3; k++) a[k] = cf[12 + k] / bb[k];
    sum = a[0] + a[1] + a[2];
    for (k = 0; k < 3; k++) {
        po[2 - k] = p[k];
        wo[2 - k] = a[k] / sum;
    }
    for (k = 0; k < 5; k++) vr[k] = v[4 - k];
    *vL = learned_value(vr, po, wo, kind);
}

/* ------------------------------------------------------------------------ */
/* HLLC at one face: m_riemann_solvers.fpp:134-325                           */

static inline double fmin_ref(double a, double b) { return (a < b) ? a : b; }
static inline double fmax_ref(double a, double b) { return (a > b) ? a : b; }

/* HLLC at one face; vsrc (may be NULL) receives vel_src_rs(:, :, 1:num_dims),
 * the velocity the viscous energy flux multiplies (:313-323). */
static void hllc_face_full(const orc_case *c, int dir, const double *qL, const double *qR, double *flux,
                           double *ustar, double *vsrc) {
    const int nf = c->num_fluids, D = orc_num_dims(c);
    const int contxe = nf - 1, E = nf + D, advxb = nf + D + 1;
    double velL[3], velR[3];
    double vel_L_rms = 0.0, vel_R_rms = 0.0;
    double rho_L = 0.0, gamma_L = 0.0, pi_inf_L = 0.0;
    double rho_R = 0.0, gamma_R = 0.0, pi_inf_R = 0.0;
    double pres_L, pres_R, E_L, E_R, H_L, H_R, c_L, c_R;
    double s_L, s_R, s_S, s_M, s_P, xi_L, xi_R, xi_M, xi_P;
    double uL, uR;
    int i;
    for (i = 0; i < D; i++) { /* :139-145 */
        velL[i] = qL[contxe + 1 + i];
        velR[i] = qR[contxe + 1 + i];
        vel_L_rms = vel_L_rms + velL[i] * velL[i];
        vel_R_rms = vel_R_rms + velR[i] * velR[i];
    }
    pres_L = qL[E];
    pres_R = qR[E];
    for (i = 0; i < nf; i++) { /* :158-167 */
        rho_L = rho_L + qL[i];
        gamma_L = gamma_L + qL[E + 1 + i] * c->gammas[i];
        pi_inf_L = pi_inf_L + qL[E + 1 + i] * c->pi_infs[i];
        rho_R = rho_R + qR[i];
        gamma_R = gamma_R + qR[E + 1 + i] * c->gammas[i];
        pi_inf_R = pi_inf_R + qR[E + 1 + i] * c->pi_infs[i];
    }
    E_L = gamma_L * pres_L + pi_inf_L + 0.5 * rho_L * vel_L_rms; /* :202-204 */
    E_R = gamma_R * pres_R + pi_inf_R + 0.5 * rho_R * vel_R_rms;
    H_L = (E_L + pres_L) / rho_L;
    H_R = (E_R + pres_R) / rho_R;
    /* c_avg (:209-218) is computed by the reference but never used. */
    c_L = ((H_L - 0.5 * vel_L_rms) / gamma_L); /* :220-223 */
    c_R = ((H_R - 0.5 * vel_R_rms) / gamma_R);
    c_L = sqrt(c_L);
    c_R = sqrt(c_R);
    uL = velL[dir];
    uR = velR[dir];
    s_L = fmin_ref(uL - c_L, uR - c_R); /* :232-233 */
    s_R = fmax_ref(uR + c_R, uL + c_L);
    s_S = (pres_R - pres_L + rho_L * uL * (s_L - uL) - rho_R * uR * (s_R - uR)) /
          (rho_L * (s_L - uL) - rho_R * (s_R - uR)); /* :235-240 */
    s_M = fmin_ref(0.0, s_L);
    s_P = fmax_ref(0.0, s_R);
    xi_L = (s_L - uL) / (s_L - s_S);
    xi_R = (s_R - uR) / (s_R - s_S);
    xi_M = (0.5 + copysign(0.5, s_S)); /* sign(5d-1, s_S): -0.0 gives -0.5 */
    xi_P = (0.5 - copysign(0.5, s_S));
    for (i = 0; i <= contxe; i++) /* :258-264 */
        flux[i] = xi_M * qL[i] * (uL + s_M * (xi_L - 1.0)) + xi_P * qR[i] * (uR + s_P * (xi_R - 1.0));
    for (i = 0; i < D; i++) { /* :270-286 */
        const double flg = (i == dir) ? 1.0 : 0.0;
        flux[contxe + 1 + i] =
            xi_M * (rho_L * (uL * velL[i] + s_M * (xi_L * (flg * s_S + (1.0 - flg) * velL[i]) - velL[i])) +
                    flg * (pres_L)) +
            xi_P * (rho_R * (uR * velR[i] + s_P * (xi_R * (flg * s_S + (1.0 - flg) * velR[i]) - velR[i])) +
                    flg * (pres_R));
    }
    flux[E] = xi_M * (uL * (E_L + pres_L) + s_M * (xi_L * (E_L + (s_S - uL) * (rho_L * s_S + pres_L / (s_L - uL))) - E_L)) +
              xi_P * (uR * (E_R + pres_R) + s_P * (xi_R * (E_R + (s_S - uR) * (rho_R * s_S + pres_R / (s_R - uR))) - E_R));
    for (i = advxb; i < advxb + nf; i++) /* :304-310 */
        flux[i] = xi_M * qL[i] * (uL + s_M * (xi_L - 1.0)) + xi_P * qR[i] * (uR + s_P * (xi_R - 1.0));
    /* contact velocity, :314-325 (dir_flg = 1 for the normal component) */
    *ustar = xi_M * (uL + 1.0 * s_M * (xi_L - 1.0)) + xi_P * (uR + 1.0 * s_P * (xi_R - 1.0));
    if (vsrc)
        for (i = 0; i < D; i++) {
            const double flg = (i == dir) ? 1.0 : 0.0;
            vsrc[i] = xi_M * (velL[i] + flg * s_M * (xi_L - 1.0)) + xi_P * (velR[i] + flg * s_P * (xi_R - 1.0));
        }
}

void orc_hllc_face(const orc_case *c, int dir, const double *qL, const double *qR, double *flux,
                   double *ustar) {
    hllc_face_full(c, dir, qL, qR, flux, ustar, NULL);
}

/* ------------------------------------------------------------------------ */
/* cons -> prim: m_variables_conversion.fpp:313-375 (+ :187-227, :98-106)     */

static void cons_to_prim_cell(const orc_case *c, const geom *g, const double *qc, double *qp,
                              long o) {
    double rho_K = 0.0, gamma_K = 0.0, pi_inf_K = 0.0, dyn_pres_K = 0.0;
    int i;
    for (i = 0; i <= g->contxe; i++) qp[i * g->ng + o] = qc[i * g->ng + o];
    for (i = 0; i < g->nf; i++) {
        rho_K = rho_K + qc[i * g->ng + o];
        gamma_K = gamma_K + qc[(g->advxb + i) * g->ng + o] * c->gammas[i];
        pi_inf_K = pi_inf_K + qc[(g->advxb + i) * g->ng + o] * c->pi_infs[i];
    }
    rho_K = fmax_ref(rho_K, SGM_EPS);
    for (i = g->momxb; i < g->momxb + g->D; i++) {
        qp[i * g->ng + o] = qc[i * g->ng + o] / rho_K;
        dyn_pres_K = dyn_pres_K + 0.5 * qc[i * g->ng + o] * qp[i * g->ng + o];
    }
    qp[g->E_idx * g->ng + o] = (qc[g->E_idx * g->ng + o] - dyn_pres_K - pi_inf_K) / gamma_K;
    for (i = g->advxb; i < g->advxb + g->nf; i++) qp[i * g->ng + o] = qc[i * g->ng + o];
}

int orc_cons_to_prim(const orc_case *c, const double *q_cons, double *q_prim) {
    geom g;
    long o;
    mkgeom(c, &g);
    for (o = 0; o < g.ng; o++) cons_to_prim_cell(c, &g, q_cons, q_prim, o);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ghost fill: m_rhs.fpp:686-908 (x over interior rows, then y over all x,    */
/* then z over all x and y -- the z part is the 3D generalisation)           */

static int bc_src(int bc_beg, int bc_end, int s, int i, int *flip) {
    *flip = 0;
    if (i < 0) {
        if (bc_beg <= -3) return 0;
        if (bc_beg == -2) { *flip = 1; return -i - 1; }
        return s + 1 + i; /* periodic: q(-j) = q(s-(j-1)) */
    }
    if (bc_end <= -3) return s;
    if (bc_end == -2) { *flip = 1; return 2 * s + 1 - i; }
    return i - (s + 1); /* periodic: q(s+j) = q(j-1) */
}

int orc_bc_fill(const orc_case *c, double *q) {
    geom g;
    int i, j, k, v, flip, src;
    mkgeom(c, &g);
    for (i = 0; i < 3; i++) {
        if (i < g.D && (c->bc_beg[i] < -12 || c->bc_beg[i] > -1 || c->bc_end[i] < -12 || c->bc_end[i] > -1))
            return -1;
    }
    /* x: rows j in [0,n], k in [0,p] */
    for (k = 0; k <= g.p; k++)
        for (j = 0; j <= g.n; j++)
            for (i = -g.b; i <= g.m + g.b; i++) {
                if (i >= 0 && i <= g.m) continue;
                src = bc_src(c->bc_beg[0], c->bc_end[0], g.m, i, &flip);
                for (v = 0; v < g.S; v++) {
                    double val = q[v * g.ng + gidx(&g, src, j, k)];
                    q[v * g.ng + gidx(&g, i, j, k)] = (flip && v == g.momxb) ? -val : val;
                }
            }
    if (g.D < 2) return 0;
    /* y: all x in [-b, m+b], k in [0,p] */
    for (k = 0; k <= g.p; k++)
        for (j = -g.b; j <= g.n + g.b; j++) {
            if (j >= 0 && j <= g.n) continue;
            src = bc_src(c->bc_beg[1], c->bc_end[1], g.n, j, &flip);
            for (i = -g.b; i <= g.m + g.b; i++)
                for (v = 0; v < g.S; v++) {
                    double val = q[v * g.ng + gidx(&g, i, src, k)];
                    q[v * g.ng + gidx(&g, i, j, k)] = (flip && v == g.momxb + 1) ? -val : val;
                }
        }
    if (g.D < 3) return 0;
    /* z: all x and y including ghosts */
    for (k = -g.b; k <= g.p + g.b; k++) {
        if (k >= 0 && k <= g.p) continue;
        src = bc_src(c->bc_beg[2], c->bc_end[2], g.p, k, &flip);
        for (j = -g.b; j <= g.n + g.b; j++)
            for (i = -g.b; i <= g.m + g.b; i++)
                for (v = 0; v < g.S; v++) {
                    double val = q[v * g.ng + gidx(&g, i, j, src)];
                    q[v * g.ng + gidx(&g, i, j, k)] = (flip && v == g.momxb + 2) ? -val : val;
                }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Viscous terms: m_viscous.fpp:159-591 (velocity gradients, by finite
 * differences or, with weno_Re_flux, by the divergence theorem on WENO face
 * values), m_riemann_solvers.fpp:169-231 (face
 * Reynolds numbers), :683-902 (viscous source flux), m_rhs.fpp:591-603 and
 * :639-651 (its divergence).  1D and 2D, as in the reference.               */

typedef struct {
    int size[2];                       /* Re_size(1:2) */
    int idx[2][ORC_NUM_FLUIDS_MAX];    /* Re_idx, 0-based fluid index */
    double res[2][ORC_NUM_FLUIDS_MAX]; /* Res */
    /* Face velocity gradients, D components of ng doubles each, over the
     * ghosted grid: Lx1/Rx1 = dqL/dqR_prim_dx_n(1), Ly2/Ry2 =
     * dqL/dqR_prim_dy_n(2), Lx2/Rx2 = dq*_prim_dx_n(2) and Ly1/Ry1 =
     * dq*_prim_dy_n(1).  By finite differences: one-sided differences for
     * the first two, their four-point averages for the others; with
     * weno_Re_flux, WENO face values of the divergence-theorem gradients.
     * Entries the reference does not compute hold NaN. */
    double *Lx1, *Rx1, *Ly2, *Ry2, *Lx2, *Rx2, *Ly1, *Ry1;
} visc;

/* Re_size, Re_idx, Res: m_global_parameters.fpp:314-338, m_viscous.fpp:36-42. */
static int visc_setup(const orc_case *c, visc *w) {
    int i, r;
    memset(w, 0, sizeof *w);
    for (r = 0; r < 2; r++)
        for (i = 0; i < c->num_fluids; i++)
            if (c->Re[i][r] > 0) {
                w->idx[r][w->size[r]] = i;
                w->res[r][w->size[r]] = c->Re[i][r];
                w->size[r]++;
            }
    return w->size[0] > 0 || w->size[1] > 0;
}

static void visc_free(visc *w) {
    free(w->Lx1); free(w->Rx1); free(w->Ly2); free(w->Ry2);
    free(w->Lx2); free(w->Rx2); free(w->Ly1); free(w->Ry1);
}

/* s_get_viscous, finite-difference branch (m_viscous.fpp:231-355) with the
 * index ranges s_compute_rhs sets first (m_rhs.fpp:415-419): x in [-b, m+b],
 * y in [-b, n+b] (2D) or {0} (1D).  The cell-centred gradients the branch also
 * computes (s_compute_fd_gradient, :340-352) feed nothing when weno_Re_flux is
 * off and are left out. */
static int visc_gradients(const orc_case *c, const geom *g, const double *qp, const double *ccx_,
                          const double *ccy_, visc *w) {
    const long tot = (long)g->D * g->ng;
    const int b = g->b, m = g->m, n = g->n;
    const int y0 = g->D > 1 ? -b : 0, y1 = g->D > 1 ? n + b : 0;
    const double *ccx = ccx_ + b, *ccy = g->D > 1 ? ccy_ + b : NULL;
    double **arr[8] = {&w->Lx1, &w->Rx1, &w->Ly2, &w->Ry2, &w->Lx2, &w->Rx2, &w->Ly1, &w->Ry1};
    int a, i, x, y;
    long o;
    (void)c;
    for (a = 0; a < (g->D > 1 ? 8 : 2); a++) {
        *arr[a] = (double *)malloc(sizeof(double) * tot);
        if (!*arr[a]) return -2;
        for (o = 0; o < tot; o++) (*arr[a])[o] = NAN;
    }
#define U(x, y) qp[(long)(g->momxb + i) * g->ng + gidx(g, x, y, 0)]
#define AT(A, x, y) A[(long)i * g->ng + gidx(g, x, y, 0)]
    for (i = 0; i < g->D; i++) {
        for (y = y0; y <= y1; y++)
            for (x = -b + 1; x <= m + b; x++)
                AT(w->Lx1, x, y) = (U(x, y) - U(x - 1, y)) / (ccx[x] - ccx[x - 1]);
        for (y = y0; y <= y1; y++)
            for (x = -b; x <= m + b - 1; x++)
                AT(w->Rx1, x, y) = (U(x + 1, y) - U(x, y)) / (ccx[x + 1] - ccx[x]);
        if (g->D == 1) continue;
        for (y = -b + 1; y <= n + b; y++)
            for (x = -b; x <= m + b; x++)
                AT(w->Ly2, x, y) = (U(x, y) - U(x, y - 1)) / (ccy[y] - ccy[y - 1]);
        for (y = -b; y <= n + b - 1; y++)
            for (x = -b; x <= m + b; x++)
                AT(w->Ry2, x, y) = (U(x, y + 1) - U(x, y)) / (ccy[y + 1] - ccy[y]);
    }
    if (g->D > 1)
        for (i = 0; i < g->D; i++) {
            for (y = -b + 1; y <= n + b; y++)
                for (x = -b + 1; x <= m + b - 1; x++) {
                    double t = AT(w->Lx1, x, y) + AT(w->Rx1, x, y) + AT(w->Lx1, x, y - 1) + AT(w->Rx1, x, y - 1);
                    AT(w->Lx2, x, y) = 25e-2 * t;
                }
            for (y = -b; y <= n + b - 1; y++)
                for (x = -b + 1; x <= m + b - 1; x++) {
                    double t = AT(w->Lx1, x, y + 1) + AT(w->Rx1, x, y + 1) + AT(w->Lx1, x, y) + AT(w->Rx1, x, y);
                    AT(w->Rx2, x, y) = 25e-2 * t;
                }
            for (y = -b + 1; y <= n + b - 1; y++)
                for (x = -b + 1; x <= m + b; x++) {
                    double t = AT(w->Ly2, x, y) + AT(w->Ry2, x, y) + AT(w->Ly2, x - 1, y) + AT(w->Ry2, x - 1, y);
                    AT(w->Ly1, x, y) = 25e-2 * t;
                }
            for (y = -b + 1; y <= n + b - 1; y++)
                for (x = -b; x <= m + b - 1; x++) {
                    double t = AT(w->Ly2, x + 1, y) + AT(w->Ry2, x + 1, y) + AT(w->Ly2, x, y) + AT(w->Ry2, x, y);
                    AT(w->Ry1, x, y) = 25e-2 * t;
                }
        }
#undef U
#undef AT
    return 0;
}

/* Stencil of cell j along direction e (0 = x, 1 = y) of plane a (ng doubles
 * over the ghosted grid) at transverse index t, as pencil() forms it. */
static void stencil_of(const orc_case *c, const geom *g, const double *a, int e, int j, int t, double *st) {
    int r;
    for (r = -2; r <= 2; r++) {
        if ((c->weno_order == 1 && r != 0) || (c->weno_order == 3 && (r < -1 || r > 1))) {
            st[r + 2] = 0.0;
            continue;
        }
        st[r + 2] = e == 0 ? a[gidx(g, j + r, t, 0)] : a[gidx(g, t, j + r, 0)];
    }
}

/* Velocity gradients with weno_Re_flux = T.
 * 1. s_get_viscous (m_viscous.fpp:186-214): for each direction e, WENO
 *    reconstruction along e of the velocity components over cells
 *    [-b+polyn, s+b-polyn] and every transverse index
 *    (s_reconstruct_cell_boundary_values_visc, :400-469), then the scalar
 *    divergence theorem (s_apply_scalar_divergence_theorem, :372-398):
 *    dq_prim_de_qp(j) = 1/ds(j) * (vR(j) - vL(j)) over [-b+1, s+b-1].
 * 2. In each direction d of s_compute_rhs (m_rhs.fpp:514-528), WENO
 *    reconstruction along d of dq_prim_dx_qp and dq_prim_dy_qp
 *    (s_reconstruct_cell_boundary_values_visc_deriv, m_viscous.fpp:471-591)
 *    into dqL/dqR_prim_dx_n(d) and dqL/dqR_prim_dy_n(d): L is the value at
 *    the cell's lower face, R at its upper face.
 * The theorem's outermost cells -b+1..-b+polyn-1 (and their mirror) read
 * face values step 1 never writes, and cells -b, s+b are never written; they
 * hold NaN here.  Their reconstructions reach only cells whose faces the
 * fluxes at -1/2 .. s+1/2 do not use (a NaN reaching an output would show). */
static int visc_gradients_weno(const orc_case *c, const geom *g, const double *qp, const double *const cb[3],
                               const double *const ds[3], visc *w) {
    const long tot = (long)g->D * g->ng;
    const int b = g->b, D = g->D, polyn = (c->weno_order - 1) / 2;
    const int ext[2] = {g->m, g->n};
    double **arr[8] = {&w->Lx1, &w->Rx1, &w->Ly2, &w->Ry2, &w->Lx2, &w->Rx2, &w->Ly1, &w->Ry1};
    double *G[2] = {NULL, NULL}, *coef[2] = {NULL, NULL}, *vel = NULL;
    int a, e, d, i, j, t, rc = 0;
    long o;
    for (a = 0; a < (D > 1 ? 8 : 2); a++) {
        *arr[a] = (double *)malloc(sizeof(double) * tot);
        if (!*arr[a]) return -2;
        for (o = 0; o < tot; o++) (*arr[a])[o] = NAN;
    }
    vel = (double *)malloc(sizeof(double) * tot);
    for (e = 0; e < D; e++) {
        const long nc = ext[e] + 2 * b + 1;
        G[e] = (double *)malloc(sizeof(double) * tot);
        coef[e] = (double *)malloc(sizeof(double) * 27 * nc);
        if (!G[e] || !coef[e]) rc = -2;
        else {
            for (o = 0; o < tot; o++) G[e][o] = NAN;
            orc_weno_coefficients(ext[e], b, c->weno_order, cb[e], coef[e]);
        }
    }
    if (!vel || rc) goto done;
    /* the velocity components of the primitive field */
    for (i = 0; i < D; i++)
        for (o = 0; o < g->ng; o++) vel[i * g->ng + o] = qp[(long)(g->momxb + i) * g->ng + o];
    /* 1. reconstruction along e and the divergence theorem */
    for (e = 0; e < D; e++) {
        const int s = ext[e], tl = D > 1 ? -b : 0, th = D > 1 ? ext[1 - e] + b : 0;
        const long nc = s + 2 * b + 1;
        for (i = 0; i < D; i++)
            for (t = tl; t <= th; t++)
                for (j = -b + polyn; j <= s + b - polyn; j++) {
                    double cf[27], st[5], vL, vR;
                    int q;
                    if (j < -b + 1 || j > s + b - 1) continue;
                    for (q = 0; q < 27; q++) cf[q] = coef[e][q * nc + j + b];
                    stencil_of(c, g, vel + (long)i * g->ng, e, j, t, st);
                    orc_weno_cell(c->weno_order, c->weno_eps, cf, st, &vL, &vR);
                    G[e][(long)i * g->ng + (e == 0 ? gidx(g, j, t, 0) : gidx(g, t, j, 0))] =
                        1.0 / ds[e][j + b] * (vR - vL);
                }
    }
    /* 2. reconstruction of each gradient along each sweep direction d */
    for (d = 0; d < D; d++) {
        const int s = ext[d], tl = D > 1 ? -b : 0, th = D > 1 ? ext[1 - d] + b : 0;
        const long nc = s + 2 * b + 1;
        for (e = 0; e < D; e++) {
            /* dq*_prim_dx_n(1) = Lx1/Rx1, dq*_prim_dy_n(1) = Ly1/Ry1,
             * dq*_prim_dx_n(2) = Lx2/Rx2, dq*_prim_dy_n(2) = Ly2/Ry2 */
            double *L = d == 0 ? (e == 0 ? w->Lx1 : w->Ly1) : (e == 0 ? w->Lx2 : w->Ly2);
            double *R = d == 0 ? (e == 0 ? w->Rx1 : w->Ry1) : (e == 0 ? w->Rx2 : w->Ry2);
            for (i = 0; i < D; i++)
                for (t = tl; t <= th; t++)
                    for (j = -b + polyn; j <= s + b - polyn; j++) {
                        double cf[27], st[5];
                        const long at = (long)i * g->ng + (d == 0 ? gidx(g, j, t, 0) : gidx(g, t, j, 0));
                        int q;
                        for (q = 0; q < 27; q++) cf[q] = coef[d][q * nc + j + b];
                        stencil_of(c, g, G[e] + (long)i * g->ng, d, j, t, st);
                        orc_weno_cell(c->weno_order, c->weno_eps, cf, st, &L[at], &R[at]);
                    }
        }
    }
done:
    free(vel);
    for (e = 0; e < 2; e++) { free(G[e]); free(coef[e]); }
    return rc;
}

/* The gradient part of s_populate_riemann_states_variables_buffers for
 * bc = -4 (m_riemann_solvers.fpp:488-508, 524-544, 560-576, 592-608): with
 * the caller's swap (the solver's dqL_prim_d*_vf are dqR_prim_d*_n), the
 * upper-face value of cell -1 takes the lower-face value of cell 0 and the
 * lower-face value of cell s+1 the upper-face value of cell s, over the
 * interior transverse range.  (With finite-difference gradients both sides
 * already hold the same value.) */
static void visc_riemann_buffers(const orc_case *c, const geom *g, visc *w) {
    int i, t;
#define AT(A, xx, yy) A[(long)i * g->ng + gidx(g, xx, yy, 0)]
    for (i = 0; i < g->D; i++) {
        for (t = 0; t <= g->n; t++) {
            if (c->bc_beg[0] == -4) {
                AT(w->Rx1, -1, t) = AT(w->Lx1, 0, t);
                if (g->D > 1) AT(w->Ry1, -1, t) = AT(w->Ly1, 0, t);
            }
            if (c->bc_end[0] == -4) {
                AT(w->Lx1, g->m + 1, t) = AT(w->Rx1, g->m, t);
                if (g->D > 1) AT(w->Ly1, g->m + 1, t) = AT(w->Ry1, g->m, t);
            }
        }
        if (g->D == 1) continue;
        for (t = 0; t <= g->m; t++) {
            if (c->bc_beg[1] == -4) {
                AT(w->Rx2, t, -1) = AT(w->Lx2, t, 0);
                AT(w->Ry2, t, -1) = AT(w->Ly2, t, 0);
            }
            if (c->bc_end[1] == -4) {
                AT(w->Lx2, t, g->n + 1) = AT(w->Rx2, t, g->n);
                AT(w->Ly2, t, g->n + 1) = AT(w->Ry2, t, g->n);
            }
        }
    }
#undef AT
}

/* Harmonic face Reynolds numbers Re_avg(1:2) from the reconstructed volume
 * fractions of the two face states (m_riemann_solvers.fpp:169-231). */
static void visc_re_avg(const visc *w, const geom *g, const double *ql, const double *qr, double *re_avg) {
    int r, q;
    for (r = 0; r < 2; r++) {
        double reL, reR;
        if (w->size[r] == 0) { re_avg[r] = 0.0; continue; }
        reL = 0.0;
        reR = 0.0;
        for (q = 0; q < w->size[r]; q++) {
            reL = ql[g->advxb + w->idx[r][q]] / w->res[r][q] + reL;
            reR = qr[g->advxb + w->idx[r][q]] / w->res[r][q] + reR;
        }
        reL = 1.0 / fmax_ref(reL, SGM_EPS);
        reR = 1.0 / fmax_ref(reR, SGM_EPS);
        re_avg[r] = 2.0 / (1.0 / reL + 1.0 / reR);
    }
}

/* s_compute_viscous_source_flux at one face (m_riemann_solvers.fpp:683-902),
 * with the caller's argument swap (m_rhs.fpp:545-556: the solver's left
 * gradients are dqR_prim_d*_n, evaluated at the face's left cell, its right
 * ones dqL_prim_d*_n at the right cell).  (x, y) is the face's left cell;
 * fs[0..D-1] momentum, fs[D] energy. */
static void visc_face_flux(const visc *w, const geom *g, int dir, int x, int y, const double *vsrc,
                           const double *re_avg, double *fs) {
    int i;
    const int D = g->D;
#define AT(A, c, xx, yy) A[(long)(c) * g->ng + gidx(g, xx, yy, 0)]
    for (i = 0; i <= D; i++) fs[i] = 0.0;
    if (dir == 0) {
        if (w->size[0] > 0) {
            const double dx1 = 5e-1 * (AT(w->Rx1, 0, x, y) + AT(w->Lx1, 0, x + 1, y));
            const double t11 = (4.0 / 3.0) * dx1 / re_avg[0];
            fs[0] = fs[0] - t11;
            fs[D] = fs[D] - vsrc[0] * t11;
        }
        if (w->size[1] > 0) {
            const double dx1 = 5e-1 * (AT(w->Rx1, 0, x, y) + AT(w->Lx1, 0, x + 1, y));
            const double t11 = dx1 / re_avg[1];
            fs[0] = fs[0] - t11;
            fs[D] = fs[D] - vsrc[0] * t11;
        }
        if (D == 1) return;
        if (w->size[0] > 0) {
            double dy[2], dx2, tau[2];
            for (i = 0; i < 2; i++) dy[i] = 5e-1 * (AT(w->Ry1, i, x, y) + AT(w->Ly1, i, x + 1, y));
            dx2 = 5e-1 * (AT(w->Rx1, 1, x, y) + AT(w->Lx1, 1, x + 1, y));
            tau[0] = -((2.0 / 3.0) * dy[1] / re_avg[0]);
            tau[1] = (dy[0] + dx2) / re_avg[0];
            for (i = 0; i < 2; i++) {
                fs[i] = fs[i] - tau[i];
                fs[D] = fs[D] - vsrc[i] * tau[i];
            }
        }
        if (w->size[1] > 0) {
            const double dy2 = 5e-1 * (AT(w->Ry1, 1, x, y) + AT(w->Ly1, 1, x + 1, y));
            const double t11 = dy2 / re_avg[1];
            fs[0] = fs[0] - t11;
            fs[D] = fs[D] - vsrc[0] * t11;
        }
    } else {
        if (w->size[0] > 0) {
            double dx[2], dy[2], tau[2];
            for (i = 0; i < 2; i++) {
                dx[i] = 5e-1 * (AT(w->Rx2, i, x, y) + AT(w->Lx2, i, x, y + 1));
                dy[i] = 5e-1 * (AT(w->Ry2, i, x, y) + AT(w->Ly2, i, x, y + 1));
            }
            tau[0] = (dy[0] + dx[1]) / re_avg[0];
            tau[1] = (4.0 * dy[1] - 2.0 * dx[0]) / (3.0 * re_avg[0]);
            for (i = 0; i < 2; i++) {
                fs[i] = fs[i] - tau[i];
                fs[D] = fs[D] - vsrc[i] * tau[i];
            }
        }
        if (w->size[1] > 0) {
            const double dx1 = 5e-1 * (AT(w->Rx2, 0, x, y) + AT(w->Lx2, 0, x, y + 1));
            const double dy2 = 5e-1 * (AT(w->Ry2, 1, x, y) + AT(w->Ly2, 1, x, y + 1));
            const double t22 = (dx1 + dy2) / re_avg[1];
            fs[1] = fs[1] - t22;
            fs[D] = fs[D] - vsrc[1] * t22;
        }
    }
#undef AT
}

/* ------------------------------------------------------------------------ */
/* s_compute_rhs: m_rhs.fpp:405-679                                          */

/* One pencil of one direction of s_compute_rhs: WENO (m_weno.fpp:470-536)
 * for cells -1..s+1, the Riemann-state buffers for bc = -4
 * (m_riemann_solvers.fpp:480-550), HLLC at faces -1/2..s+1/2 with the right
 * face value of cell j as the solver's left state (m_rhs.fpp:545-556), and
 * the flux divergence and alpha*div(u) source (m_rhs.fpp:565-652).  The
 * reference also reconstructs further ghost cells and ghost rows whose values
 * never reach the interior fluxes. */
static void pencil(const orc_case *c, const geom *gp, int dir, int s, int k1, int k2, const double *coef,
                   const double *qc, const double *qp, const double *dsd, const visc *w, double *rhs) {
    const geom g = *gp;
    const long nc = s + 2 * g.b + 1;
    const long ns = s + 3; /* cells -1..s+1 */
    double *vL = (double *)malloc(sizeof(double) * g.S * ns);
    double *vR = (double *)malloc(sizeof(double) * g.S * ns);
    double *flux = (double *)malloc(sizeof(double) * g.S * (s + 2));
    double *ustar = (double *)malloc(sizeof(double) * (s + 2));
    double *fsv = w ? (double *)malloc(sizeof(double) * (gp->D + 1) * (s + 2)) : NULL;
    int j, q, v;
#define POS(jj, I, J, K)                                                   \
    do {                                                                   \
        if (dir == 0) { I = (jj); J = k1; K = k2; }                        \
        else if (dir == 1) { I = k1; J = (jj); K = k2; }                   \
        else { I = k1; J = k2; K = (jj); }                                 \
    } while (0)
    for (j = -1; j <= s + 1; j++) {
        double cf[27];
        for (q = 0; q < 27; q++) cf[q] = coef[q * nc + j + g.b];
        for (v = 0; v < g.S; v++) {
            double st[5];
            int r;
            for (r = -2; r <= 2; r++) {
                int I, J, K, jj = j + r;
                if (c->weno_order == 1 && r != 0) { st[r + 2] = 0.0; continue; }
                if (c->weno_order == 3 && (r < -1 || r > 1)) { st[r + 2] = 0.0; continue; }
                POS(jj, I, J, K);
                st[r + 2] = qp[v * g.ng + gidx(&g, I, J, K)];
            }
            if (g_learned && c->weno_order == 5) {
                const int kind = (v < g.nf || v >= g.advxb ? 1 : 0) | (v >= g.advxb ? 2 : 0) |
                                 (v == g.E_idx ? 4 : 0);
                orc_weno5_learned_cell(c->weno_eps, cf, st, kind, &vL[v * ns + (j + 1)], &vR[v * ns + (j + 1)]);
            } else
                orc_weno_cell(c->weno_order, c->weno_eps, cf, st, &vL[v * ns + (j + 1)], &vR[v * ns + (j + 1)]);
        }
    }
    if (c->bc_beg[dir] == -4)
        for (v = 0; v < g.S; v++) vR[v * ns + 0] = vL[v * ns + 1];
    if (c->bc_end[dir] == -4)
        for (v = 0; v < g.S; v++) vL[v * ns + (s + 2)] = vR[v * ns + (s + 1)];
    for (j = -1; j <= s; j++) {
        double ql[3 * ORC_NUM_FLUIDS_MAX + 4], qr[3 * ORC_NUM_FLUIDS_MAX + 4], f[3 * ORC_NUM_FLUIDS_MAX + 4];
        for (v = 0; v < g.S; v++) {
            ql[v] = vR[v * ns + (j + 1)];
            qr[v] = vL[v * ns + (j + 2)];
        }
        double vsrc[3], re_avg[2], fs[4];
        hllc_face_full(c, dir, ql, qr, f, &ustar[j + 1], vsrc);
        for (v = 0; v < g.S; v++) flux[v * (s + 2) + (j + 1)] = f[v];
        if (w) {
            int I, J, K;
            POS(j, I, J, K);
            (void)K;
            visc_re_avg(w, &g, ql, qr, re_avg);
            visc_face_flux(w, &g, dir, I, J, vsrc, re_avg, fs);
            for (v = 0; v <= g.D; v++) fsv[v * (s + 2) + (j + 1)] = fs[v];
        }
    }
    for (v = 0; v < g.S; v++)
        for (j = 0; j <= s; j++) {
            int I, J, K;
            long r;
            POS(j, I, J, K);
            r = (long)v * (g.m + 1) * (g.n + 1) * (g.p + 1) + ridx(&g, I, J, K);
            if (dir == 0)
                rhs[r] = 1.0 / dsd[j + g.b] * (flux[v * (s + 2) + j] - flux[v * (s + 2) + j + 1]);
            else
                rhs[r] = rhs[r] + 1.0 / dsd[j + g.b] * (flux[v * (s + 2) + j] - flux[v * (s + 2) + j + 1]);
        }
    for (v = g.advxb; v < g.advxb + g.nf; v++)
        for (j = 0; j <= s; j++) {
            int I, J, K;
            long r;
            POS(j, I, J, K);
            r = (long)v * (g.m + 1) * (g.n + 1) * (g.p + 1) + ridx(&g, I, J, K);
            rhs[r] = rhs[r] + 1.0 / dsd[j + g.b] * qc[v * g.ng + gidx(&g, I, J, K)] * (ustar[j + 1] - ustar[j]);
        }
    if (w) /* m_rhs.fpp:591-603 (x), :639-651 (y): momentum and energy */
        for (v = 0; v <= g.D; v++)
            for (j = 0; j <= s; j++) {
                int I, J, K;
                long r;
                POS(j, I, J, K);
                r = (long)(g.momxb + v) * (g.m + 1) * (g.n + 1) * (g.p + 1) + ridx(&g, I, J, K);
                rhs[r] = rhs[r] + 1.0 / dsd[j + g.b] * (fsv[v * (s + 2) + j] - fsv[v * (s + 2) + j + 1]);
            }
#undef POS
    free(fsv);
    free(vL);
    free(vR);
    free(flux);
    free(ustar);
}

int orc_compute_rhs(const orc_case *c, const double *const cb[3], const double *const ds[3],
                    const double *q_cons, double *q_prim, double *rhs, int t_step) {
    geom g;
    double *qc, *qp, *coef;
    int dir, viscous;
    visc w;
    mkgeom(c, &g);
    if (c->weno_order != 1 && c->weno_order != 3 && c->weno_order != 5) return -1;
    qc = (double *)malloc(sizeof(double) * g.S * g.ng);
    qp = (double *)malloc(sizeof(double) * g.S * g.ng);
    if (!qc || !qp) { free(qc); free(qp); return -2; }
    /* :425-432 private copy, :435 ghost fill, :445-447 conversion */
    memcpy(qc, q_cons, sizeof(double) * g.S * g.ng);
    if (orc_bc_fill(c, qc) != 0) { free(qc); free(qp); return -1; }
    orc_cons_to_prim(c, qc, qp);
    if (t_step == c->t_step_stop) { free(qc); free(qp); return 1; } /* :452 */
    viscous = visc_setup(c, &w);
    if (viscous) { /* :455-464 */
        double *cc[2] = {NULL, NULL};
        int rc = 0;
        if (g.D > 2) rc = -1;
        for (dir = 0; dir < g.D && rc == 0; dir++) {
            const int s = dir == 0 ? g.m : g.n;
            cc[dir] = (double *)malloc(sizeof(double) * (s + 2 * g.b + 1));
            cell_centres(s, g.b, cb[dir], ds[dir], cc[dir]);
        }
        if (rc == 0)
            rc = c->weno_Re_flux ? visc_gradients_weno(c, &g, qp, cb, ds, &w)
                                 : visc_gradients(c, &g, qp, cc[0], cc[1], &w);
        if (rc == 0) visc_riemann_buffers(c, &g, &w);
        free(cc[0]);
        free(cc[1]);
        if (rc != 0) { visc_free(&w); free(qc); free(qp); return rc; }
    }

    for (dir = 0; dir < g.D; dir++) {
        /* Extents along the sweep direction (s) and the two transverse ones. */
        const int s = dir == 0 ? g.m : (dir == 1 ? g.n : g.p);
        const int t1 = dir == 0 ? g.n : g.m;                  /* first transverse  */
        const int t2 = dir == 2 ? g.n : g.p;                  /* second transverse */
        const long nc = s + 2 * g.b + 1;
        const long npencil = (long)(t1 + 1) * (t2 + 1);
        long pc;
        coef = (double *)malloc(sizeof(double) * 27 * nc);
        orc_weno_coefficients(s, g.b, c->weno_order, cb[dir], coef);
        /* Pencils are independent (each writes its own rhs cells), so they
         * may run on several threads; the arithmetic of each is unchanged. */
#pragma omp parallel for schedule(dynamic, 4)
        for (pc = 0; pc < npencil; pc++)
            pencil(c, &g, dir, s, (int)(pc % (t1 + 1)), (int)(pc / (t1 + 1)), coef, qc, qp, ds[dir],
                   viscous ? &w : NULL, rhs);
        free(coef);
    }
    if (viscous) visc_free(&w);
    if (q_prim) memcpy(q_prim, qp, sizeof(double) * g.S * g.ng); /* :659-675 */
    free(qc);
    free(qp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* TVD RK: m_time_steppers.fpp:129-362                                       */

int orc_tvd_rk_step(const orc_case *c, const double *const cb[3], const double *const ds[3],
                    double dt, double *q1, double *q2, double *q_prim, double *rhs, int t_step) {
    geom g;
    int v, i, j, k, rc;
    mkgeom(c, &g);
    const long nr = (long)(g.m + 1) * (g.n + 1) * (g.p + 1);
#define LOOP _Pragma("omp parallel for private(k, j, i)") for (v = 0; v < g.S; v++) for (k = 0; k <= g.p; k++) for (j = 0; j <= g.n; j++) for (i = 0; i <= g.m; i++)
#define Q(a) a[v * g.ng + gidx(&g, i, j, k)]
#define R rhs[v * nr + ridx(&g, i, j, k)]
    rc = orc_compute_rhs(c, cb, ds, q1, q_prim, rhs, t_step);
    if (rc < 0) return rc;
    if (t_step == c->t_step_stop) return 1;
    if (c->time_stepper == 1) {
        LOOP Q(q1) = Q(q1) + dt * R;
        return 0;
    }
    LOOP Q(q2) = Q(q1) + dt * R;
    rc = orc_compute_rhs(c, cb, ds, q2, q_prim, rhs, t_step);
    if (rc < 0) return rc;
    if (c->time_stepper == 2) {
        LOOP Q(q1) = (Q(q1) + Q(q2) + dt * R) / 2.0;
        return 0;
    }
    LOOP Q(q2) = (3.0 * Q(q1) + Q(q2) + dt * R) / 4.0;
    rc = orc_compute_rhs(c, cb, ds, q2, q_prim, rhs, t_step);
    if (rc < 0) return rc;
    LOOP Q(q1) = (Q(q1) + 2.0 * Q(q2) + 2.0 * dt * R) / 3.0;
#undef LOOP
#undef Q
#undef R
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ICFL: m_data_output.fpp:197-251                                           */

int orc_icfl_max(const orc_case *c, const double *const ds[3], const double *q_prim, double dt,
                 double *icfl_max) {
    geom g;
    int i, j, k, f;
    double mx = 0.0;
    int any = 0, nan = 0;
    mkgeom(c, &g);
    for (k = 0; k <= g.p; k++)
        for (j = 0; j <= g.n; j++)
            for (i = 0; i <= g.m; i++) {
                const long o = gidx(&g, i, j, k);
                double rho = 0.0, gamma = 0.0, pi_inf = 0.0, vel[3], pres, cc, icfl;
                for (f = 0; f < g.nf; f++) {
                    rho = rho + q_prim[f * g.ng + o];
                    gamma = gamma + q_prim[(g.advxb + f) * g.ng + o] * c->gammas[f];
                    pi_inf = pi_inf + q_prim[(g.advxb + f) * g.ng + o] * c->pi_infs[f];
                }
                for (f = 0; f < g.D; f++) vel[f] = q_prim[(g.momxb + f) * g.ng + o];
                pres = q_prim[g.E_idx * g.ng + o];
                cc = (((gamma + 1.0) * pres + pi_inf) / (gamma * rho));
                cc = sqrt(cc);
                if (g.D == 1) {
                    icfl = (dt / ds[0][i + g.b]) * (fabs(vel[0]) + cc);
                } else if (g.D == 2) {
                    icfl = dt / fmin_ref(ds[0][i + g.b] / (fabs(vel[0]) + cc),
                                         ds[1][j + g.b] / (fabs(vel[1]) + cc));
                } else {
                    icfl = dt / fmin_ref(fmin_ref(ds[0][i + g.b] / (fabs(vel[0]) + cc),
                                                  ds[1][j + g.b] / (fabs(vel[1]) + cc)),
                                         ds[2][k + g.b] / (fabs(vel[2]) + cc));
                }
                if (icfl != icfl) nan = 1;
                else if (!any || icfl > mx) { mx = icfl; any = 1; }
            }
    *icfl_max = nan ? NAN : mx;
    return 0;
}

/* VCFL maximum and Rc minimum of a viscous case (m_data_output.fpp:214-263)
 * with the mixture Reynolds numbers of m_variables_conversion.fpp:210-223:
 * Re_K(i) = 1/max(sum_j alpha(Re_idx(i,j))/Res(i,j), sgm_eps), dflt_real
 * before the sum when fluid i has no Reynolds number of that kind.
 * out[0] = ICFL max, out[1] = VCFL max, out[2] = Rc min (0, 0 when inviscid);
 * NaN anywhere gives NaN. */
int orc_stability(const orc_case *c, const double *const ds[3], const double *q_prim, double dt, double out[3]) {
    geom g;
    int i, j, f, r, visc = 0;
    double vmax = 0.0, rmin = 0.0;
    int any = 0, nan = 0;
    mkgeom(c, &g);
    orc_icfl_max(c, ds, q_prim, dt, &out[0]);
    for (f = 0; f < g.nf; f++)
        for (r = 0; r < 2; r++) visc = visc || c->Re[f][r] > 0.0;
    out[1] = 0.0;
    out[2] = 0.0;
    if (!visc) return 0;
    if (g.D > 2) return -1;
    for (j = 0; j <= g.n; j++)
        for (i = 0; i <= g.m; i++) {
            const long o = gidx(&g, i, j, 0);
            double rho = 0.0, gamma = 0.0, pi_inf = 0.0, Re[2], cc, vcfl, rc, dtre, inv;
            for (f = 0; f < g.nf; f++) {
                rho = rho + q_prim[f * g.ng + o];
                gamma = gamma + q_prim[(g.advxb + f) * g.ng + o] * c->gammas[f];
                pi_inf = pi_inf + q_prim[(g.advxb + f) * g.ng + o] * c->pi_infs[f];
            }
            for (r = 0; r < 2; r++) {
                int size = 0;
                double v = -1e6;
                for (f = 0; f < g.nf; f++) size += c->Re[f][r] > 0.0;
                if (size > 0) v = 0.0;
                for (f = 0; f < g.nf; f++)
                    if (c->Re[f][r] > 0.0) v = q_prim[(g.advxb + f) * g.ng + o] / c->Re[f][r] + v;
                Re[r] = 1.0 / fmax_ref(v, 1e-16);
            }
            cc = (((gamma + 1.0) * q_prim[g.E_idx * g.ng + o] + pi_inf) / (gamma * rho));
            cc = sqrt(cc);
            dtre = fmax_ref(dt / Re[0], dt / Re[1]);
            inv = fmax_ref(1.0 / Re[0], 1.0 / Re[1]);
            if (g.D == 1) {
                const double h = ds[0][i + g.b];
                vcfl = dtre / (h * h);
                rc = h * (fabs(q_prim[g.momxb * g.ng + o]) + cc) / inv;
            } else {
                const double h = fmin_ref(ds[0][i + g.b], ds[1][j + g.b]);
                vcfl = dtre / (h * h);
                rc = fmin_ref(ds[0][i + g.b] * (fabs(q_prim[g.momxb * g.ng + o]) + cc),
                              ds[1][j + g.b] * (fabs(q_prim[(g.momxb + 1) * g.ng + o]) + cc)) / inv;
            }
            if (vcfl != vcfl || rc != rc) nan = 1;
            else if (!any) { vmax = vcfl; rmin = rc; any = 1; }
            else {
                if (vcfl > vmax) vmax = vcfl;
                if (rc < rmin) rmin = rc;
            }
        }
    out[1] = nan ? NAN : vmax;
    out[2] = nan ? NAN : rmin;
    return 0;
}
