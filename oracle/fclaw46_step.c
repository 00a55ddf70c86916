/*
 * Oracle (test infrastructure only): plain-C restatement of the reference's
 * Clawpack-4.6 single-patch step.  Never linked into the product.
 *
 *   step2_wrap : src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_step2_wrap.f:1-127
 *   step2      : src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_step2.f:5-239
 *   flux2      : src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_flux2.f:4-228
 *   limiter    : src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_inlinelimiter.f:4-118
 *   Riemann    : applications/clawpack/{advection/2d/all,shallow/2d/rp,euler/2d/rp}/clawpack46_rp*.f
 *
 * Floating-point operations are written in the Fortran's evaluation order
 * (left to right, explicit parentheses kept) and the file is compiled with
 * -ffp-contract=off so that results are bit-identical to the Fortran.
 */
#include "fclaw46_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 1D slice arrays a(1-mbc:maxm+mbc, m) with 1-based m. */
#define S1(a, i, m) (a)[((m) - 1) * N + (i) + mbc - 1]
/* wave(1-mbc:maxm+mbc, meqn, mwaves) */
#define WV(i, m, mw) wave[((mw) - 1) * meqn * N + ((m) - 1) * N + (i) + mbc - 1]
#define SP(i, mw) s[((mw) - 1) * N + (i) + mbc - 1]

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* Roe averages shared between rpn2 and rpt2 (the reference's common /comroe/). */
typedef struct roe_ctx {
    int N, mbc;
    double *u, *v, *a, *h;                 /* shallow water */
    double *u2v2, *enth, *g1a2, *euv;      /* Euler (u, v, a shared) */
} roe_ctx_t;

#define R(arr, i) (arr)[(i) + mbc - 1]

/* ---------------- shallow water: clawpack46_rpn2.f / clawpack46_rpt2.f ---------------- */

static void rpn2_shallow(int ixy, int maxm, int meqn, int mwaves, int mbc, int mx,
                         const double *ql, const double *qr, double grav,
                         double *wave, double *s, double *amdq, double *apdq,
                         roe_ctx_t *c)
{
    const int N = maxm + 2 * mbc;
    int mu, mv, i, m, mw;
    double delta[4];
    (void) mwaves;
    if (ixy == 1) { mu = 2; mv = 3; } else { mu = 3; mv = 2; }

    /* clawpack46_rpn2.f:91-99 */
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        double hsqrtl, hsqrtr, hsq2;
        R(c->h, i) = (S1(qr, i - 1, 1) + S1(ql, i, 1)) / 2.0;
        hsqrtl = sqrt(S1(qr, i - 1, 1));
        hsqrtr = sqrt(S1(ql, i, 1));
        hsq2 = hsqrtl + hsqrtr;
        R(c->u, i) = (S1(qr, i - 1, mu) / hsqrtl + S1(ql, i, mu) / hsqrtr) / hsq2;
        R(c->v, i) = (S1(qr, i - 1, mv) / hsqrtl + S1(ql, i, mv) / hsqrtr) / hsq2;
        R(c->a, i) = sqrt(grav * R(c->h, i));
    }
    /* clawpack46_rpn2.f:105-129 */
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        const double u = R(c->u, i), v = R(c->v, i), a = R(c->a, i);
        double a1, a2, a3;
        delta[1] = S1(ql, i, 1) - S1(qr, i - 1, 1);
        delta[2] = S1(ql, i, mu) - S1(qr, i - 1, mu);
        delta[3] = S1(ql, i, mv) - S1(qr, i - 1, mv);
        a1 = ((u + a) * delta[1] - delta[2]) / (2.0 * a);
        a2 = -(v * delta[1]) + delta[3];
        a3 = (-((u - a) * delta[1]) + delta[2]) / (2.0 * a);

        WV(i, 1, 1) = a1;
        WV(i, mu, 1) = a1 * (u - a);
        WV(i, mv, 1) = a1 * v;
        SP(i, 1) = u - a;

        WV(i, 1, 2) = 0.0;
        WV(i, mu, 2) = 0.0;
        WV(i, mv, 2) = a2;
        SP(i, 2) = u;

        WV(i, 1, 3) = a3;
        WV(i, mu, 3) = a3 * (u + a);
        WV(i, mv, 3) = a3 * v;
        SP(i, 3) = u + a;
    }
    /* entropy fix, clawpack46_rpn2.f:170-229 */
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        double him1, s0, h1, hu1, s1, sfract, hi, s03, h3, hu3, s3;
        him1 = S1(qr, i - 1, 1);
        s0 = S1(qr, i - 1, mu) / him1 - sqrt(grav * him1);
        if (s0 > 0.0 && SP(i, 1) > 0.0) {
            for (m = 1; m <= 3; m++) S1(amdq, i, m) = 0.0;
            continue;
        }
        h1 = S1(qr, i - 1, 1) + WV(i, 1, 1);
        hu1 = S1(qr, i - 1, mu) + WV(i, mu, 1);
        s1 = hu1 / h1 - sqrt(grav * h1);
        if (s0 < 0.0 && s1 > 0.0)
            sfract = s0 * ((s1 - SP(i, 1)) / (s1 - s0));
        else if (SP(i, 1) < 0.0)
            sfract = SP(i, 1);
        else
            sfract = 0.0;
        for (m = 1; m <= 3; m++) S1(amdq, i, m) = sfract * WV(i, m, 1);

        if (SP(i, 2) > 0.0) continue;
        for (m = 1; m <= 3; m++) S1(amdq, i, m) = S1(amdq, i, m) + SP(i, 2) * WV(i, m, 2);

        hi = S1(ql, i, 1);
        s03 = S1(ql, i, mu) / hi + sqrt(grav * hi);
        h3 = S1(ql, i, 1) - WV(i, 1, 3);
        hu3 = S1(ql, i, mu) - WV(i, mu, 3);
        s3 = hu3 / h3 + sqrt(grav * h3);
        if (s3 < 0.0 && s03 > 0.0)
            sfract = s3 * ((s03 - SP(i, 3)) / (s03 - s3));
        else if (SP(i, 3) < 0.0)
            sfract = SP(i, 3);
        else
            continue;
        for (m = 1; m <= 3; m++) S1(amdq, i, m) = S1(amdq, i, m) + sfract * WV(i, m, 3);
    }
    /* clawpack46_rpn2.f:234-242 */
    for (m = 1; m <= 3; m++)
        for (i = 2 - mbc; i <= mx + mbc; i++) {
            double df = 0.0;
            for (mw = 1; mw <= mwaves; mw++) df = df + SP(i, mw) * WV(i, m, mw);
            S1(apdq, i, m) = df - S1(amdq, i, m);
        }
}

static void rpt2_shallow(int ixy, int maxm, int meqn, int mwaves, int mbc, int mx,
                         const double *asdq, double *bmasdq, double *bpasdq,
                         const roe_ctx_t *c)
{
    const int N = maxm + 2 * mbc;
    int mu, mv, i, m, mw;
    double waveb[4][4], sb[4];
    if (ixy == 1) { mu = 2; mv = 3; } else { mu = 3; mv = 2; }
    /* clawpack46_rpt2.f:287-319 */
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        const double u = R(c->u, i), v = R(c->v, i), a = R(c->a, i);
        double a1, a2, a3;
        a1 = ((v + a) * S1(asdq, i, 1) - S1(asdq, i, mv)) / (2.0 * a);
        a2 = S1(asdq, i, mu) - u * S1(asdq, i, 1);
        a3 = (-((v - a) * S1(asdq, i, 1)) + S1(asdq, i, mv)) / (2.0 * a);

        waveb[1][1] = a1;  waveb[mu][1] = a1 * u;  waveb[mv][1] = a1 * (v - a);  sb[1] = v - a;
        waveb[1][2] = 0.0; waveb[mu][2] = a2;      waveb[mv][2] = 0.0;          sb[2] = v;
        waveb[1][3] = a3;  waveb[mu][3] = a3 * u;  waveb[mv][3] = a3 * (v + a);  sb[3] = v + a;

        for (m = 1; m <= meqn; m++) {
            double bm = 0.0, bp = 0.0;
            for (mw = 1; mw <= mwaves; mw++) {
                bm = bm + dmin(sb[mw], 0.0) * waveb[m][mw];
                bp = bp + dmax(sb[mw], 0.0) * waveb[m][mw];
            }
            S1(bmasdq, i, m) = bm;
            S1(bpasdq, i, m) = bp;
        }
    }
}

/* ---------------- Euler: clawpack46_rpn2_euler4.f / clawpack46_rpt2_euler4.f ---------------- */

static void rpn2_euler(int ixy, int maxm, int meqn, int mwaves, int mbc, int mx,
                       const double *ql, const double *qr, double gamma,
                       double *wave, double *s, double *amdq, double *apdq,
                       roe_ctx_t *c)
{
    const int N = maxm + 2 * mbc;
    const double gamma1 = gamma - 1.0;
    int mu, mv, i, m, mw;
    double delta[5];
    if (ixy == 1) { mu = 2; mv = 3; } else { mu = 3; mv = 2; }

    /* clawpack46_rpn2_euler4.f:83-100 */
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        double rhsqrtl, rhsqrtr, pl, pr, rhsq2, a2;
        rhsqrtl = sqrt(S1(qr, i - 1, 1));
        rhsqrtr = sqrt(S1(ql, i, 1));
        pl = gamma1 * (S1(qr, i - 1, 4) - 0.5 * (S1(qr, i - 1, 2) * S1(qr, i - 1, 2) +
                                                 S1(qr, i - 1, 3) * S1(qr, i - 1, 3)) / S1(qr, i - 1, 1));
        pr = gamma1 * (S1(ql, i, 4) - 0.5 * (S1(ql, i, 2) * S1(ql, i, 2) +
                                             S1(ql, i, 3) * S1(ql, i, 3)) / S1(ql, i, 1));
        rhsq2 = rhsqrtl + rhsqrtr;
        R(c->u, i) = (S1(qr, i - 1, mu) / rhsqrtl + S1(ql, i, mu) / rhsqrtr) / rhsq2;
        R(c->v, i) = (S1(qr, i - 1, mv) / rhsqrtl + S1(ql, i, mv) / rhsqrtr) / rhsq2;
        R(c->enth, i) = (((S1(qr, i - 1, 4) + pl) / rhsqrtl + (S1(ql, i, 4) + pr) / rhsqrtr)) / rhsq2;
        R(c->u2v2, i) = R(c->u, i) * R(c->u, i) + R(c->v, i) * R(c->v, i);
        a2 = gamma1 * (R(c->enth, i) - .5 * R(c->u2v2, i));
        R(c->a, i) = sqrt(a2);
        R(c->g1a2, i) = gamma1 / a2;
        R(c->euv, i) = R(c->enth, i) - R(c->u2v2, i);
    }
    /* clawpack46_rpn2_euler4.f:106-149 */
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        const double u = R(c->u, i), v = R(c->v, i), a = R(c->a, i);
        const double enth = R(c->enth, i), u2v2 = R(c->u2v2, i);
        double a1, a2, a3, a4;
        delta[1] = S1(ql, i, 1) - S1(qr, i - 1, 1);
        delta[2] = S1(ql, i, mu) - S1(qr, i - 1, mu);
        delta[3] = S1(ql, i, mv) - S1(qr, i - 1, mv);
        delta[4] = S1(ql, i, 4) - S1(qr, i - 1, 4);
        a3 = R(c->g1a2, i) * (R(c->euv, i) * delta[1] + u * delta[2] + v * delta[3] - delta[4]);
        a2 = delta[3] - v * delta[1];
        a4 = (delta[2] + (a - u) * delta[1] - a * a3) / (2.0 * a);
        a1 = delta[1] - a3 - a4;

        WV(i, 1, 1) = a1;  WV(i, mu, 1) = a1 * (u - a);  WV(i, mv, 1) = a1 * v;
        WV(i, 4, 1) = a1 * (enth - u * a);  SP(i, 1) = u - a;

        WV(i, 1, 2) = 0.0; WV(i, mu, 2) = 0.0; WV(i, mv, 2) = a2;
        WV(i, 4, 2) = a2 * v;  SP(i, 2) = u;

        WV(i, 1, 3) = a3;  WV(i, mu, 3) = a3 * u;  WV(i, mv, 3) = a3 * v;
        WV(i, 4, 3) = a3 * 0.5 * u2v2;  SP(i, 3) = u;

        WV(i, 1, 4) = a4;  WV(i, mu, 4) = a4 * (u + a);  WV(i, mv, 4) = a4 * v;
        WV(i, 4, 4) = a4 * (enth + u * a);  SP(i, 4) = u + a;
    }
    /* entropy fix, clawpack46_rpn2_euler4.f:191-272 */
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        double rhoim1, pim1, cim1, s0, rho1, rhou1, rhov1, en1, p1, c1, s1, sfract;
        double rhoi, pi, ci, s3, rho2, rhou2, rhov2, en2, p2, c2, s2;
        rhoim1 = S1(qr, i - 1, 1);
        pim1 = gamma1 * (S1(qr, i - 1, 4) - 0.5 * (S1(qr, i - 1, mu) * S1(qr, i - 1, mu) +
                                                  S1(qr, i - 1, mv) * S1(qr, i - 1, mv)) / rhoim1);
        cim1 = sqrt(gamma * pim1 / rhoim1);
        s0 = S1(qr, i - 1, mu) / rhoim1 - cim1;
        if (s0 >= 0.0 && SP(i, 1) > 0.0) {
            for (m = 1; m <= meqn; m++) S1(amdq, i, m) = 0.0;
            continue;
        }
        rho1 = S1(qr, i - 1, 1) + WV(i, 1, 1);
        rhou1 = S1(qr, i - 1, mu) + WV(i, mu, 1);
        rhov1 = S1(qr, i - 1, mv) + WV(i, mv, 1);
        en1 = S1(qr, i - 1, 4) + WV(i, 4, 1);
        p1 = gamma1 * (en1 - 0.5 * (rhou1 * rhou1 + rhov1 * rhov1) / rho1);
        c1 = sqrt(gamma * p1 / rho1);
        s1 = rhou1 / rho1 - c1;
        if (s0 < 0.0 && s1 > 0.0)
            sfract = s0 * (s1 - SP(i, 1)) / (s1 - s0);
        else if (SP(i, 1) < 0.0)
            sfract = SP(i, 1);
        else
            sfract = 0.0;
        for (m = 1; m <= meqn; m++) S1(amdq, i, m) = sfract * WV(i, m, 1);

        if (SP(i, 2) >= 0.0) continue;
        for (m = 1; m <= meqn; m++) {
            S1(amdq, i, m) = S1(amdq, i, m) + SP(i, 2) * WV(i, m, 2);
            S1(amdq, i, m) = S1(amdq, i, m) + SP(i, 3) * WV(i, m, 3);
        }

        rhoi = S1(ql, i, 1);
        pi = gamma1 * (S1(ql, i, 4) - 0.5 * (S1(ql, i, mu) * S1(ql, i, mu) +
                                             S1(ql, i, mv) * S1(ql, i, mv)) / rhoi);
        ci = sqrt(gamma * pi / rhoi);
        s3 = S1(ql, i, mu) / rhoi + ci;
        rho2 = S1(ql, i, 1) - WV(i, 1, 4);
        rhou2 = S1(ql, i, mu) - WV(i, mu, 4);
        rhov2 = S1(ql, i, mv) - WV(i, mv, 4);
        en2 = S1(ql, i, 4) - WV(i, 4, 4);
        p2 = gamma1 * (en2 - 0.5 * (rhou2 * rhou2 + rhov2 * rhov2) / rho2);
        c2 = sqrt(gamma * p2 / rho2);
        s2 = rhou2 / rho2 + c2;
        if (s2 < 0.0 && s3 > 0.0)
            sfract = s2 * (s3 - SP(i, 4)) / (s3 - s2);
        else if (SP(i, 4) < 0.0)
            sfract = SP(i, 4);
        else
            continue;
        for (m = 1; m <= meqn; m++) S1(amdq, i, m) = S1(amdq, i, m) + sfract * WV(i, m, 4);
    }
    /* clawpack46_rpn2_euler4.f:277-285 */
    for (m = 1; m <= meqn; m++)
        for (i = 2 - mbc; i <= mx + mbc; i++) {
            double df = 0.0;
            for (mw = 1; mw <= mwaves; mw++) df = df + SP(i, mw) * WV(i, m, mw);
            S1(apdq, i, m) = df - S1(amdq, i, m);
        }
}

static void rpt2_euler(int ixy, int maxm, int meqn, int mwaves, int mbc, int mx,
                       const double *asdq, double *bmasdq, double *bpasdq,
                       const roe_ctx_t *c)
{
    const int N = maxm + 2 * mbc;
    int mu, mv, i, m, mw;
    double waveb[5][4], sb[4];
    (void) mwaves;
    if (ixy == 1) { mu = 2; mv = 3; } else { mu = 3; mv = 2; }
    /* clawpack46_rpt2_euler4.f:345-383 */
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        const double u = R(c->u, i), v = R(c->v, i), a = R(c->a, i);
        const double enth = R(c->enth, i), u2v2 = R(c->u2v2, i);
        double a1, a2, a3, a4;
        a3 = R(c->g1a2, i) * (R(c->euv, i) * S1(asdq, i, 1) + u * S1(asdq, i, mu) +
                              v * S1(asdq, i, mv) - S1(asdq, i, 4));
        a2 = S1(asdq, i, mu) - u * S1(asdq, i, 1);
        a4 = (S1(asdq, i, mv) + (a - v) * S1(asdq, i, 1) - a * a3) / (2.0 * a);
        a1 = S1(asdq, i, 1) - a3 - a4;

        waveb[1][1] = a1; waveb[mu][1] = a1 * u; waveb[mv][1] = a1 * (v - a);
        waveb[4][1] = a1 * (enth - v * a); sb[1] = v - a;

        waveb[1][2] = a3; waveb[mu][2] = a3 * u + a2; waveb[mv][2] = a3 * v;
        waveb[4][2] = a3 * 0.5 * u2v2 + a2 * u; sb[2] = v;

        waveb[1][3] = a4; waveb[mu][3] = a4 * u; waveb[mv][3] = a4 * (v + a);
        waveb[4][3] = a4 * (enth + v * a); sb[3] = v + a;

        for (m = 1; m <= meqn; m++) {
            double bm = 0.0, bp = 0.0;
            for (mw = 1; mw <= 3; mw++) {
                bm = bm + dmin(sb[mw], 0.0) * waveb[m][mw];
                bp = bp + dmax(sb[mw], 0.0) * waveb[m][mw];
            }
            S1(bmasdq, i, m) = bm;
            S1(bpasdq, i, m) = bp;
        }
    }
}

/* ---------------- advection: clawpack46_rpn2adv.f / clawpack46_rpt2adv.f ---------------- */

static void rpn2_adv(int ixy, int maxm, int meqn, int mwaves, int mbc, int mx,
                     const double *ql, const double *qr, const double *auxl,
                     double *wave, double *s, double *amdq, double *apdq)
{
    const int N = maxm + 2 * mbc;
    int i, mq, mw;
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        for (mq = 1; mq <= meqn; mq++)
            for (mw = 1; mw <= mwaves; mw++) WV(i, mq, mw) = 0.0;
        for (mq = 1; mq <= meqn; mq++) {
            double delta = S1(ql, i, mq) - S1(qr, i - 1, mq);
            for (mw = 1; mw <= mwaves; mw++) WV(i, mq, mw) = delta;
        }
        for (mw = 1; mw <= mwaves; mw++) SP(i, mw) = S1(auxl, i, ixy);
        for (mq = 1; mq <= meqn; mq++) {
            double am = 0.0, ap = 0.0;
            for (mw = 1; mw <= mwaves; mw++) {
                am = am + dmin(SP(i, mw), 0.0) * WV(i, mq, mw);
                ap = ap + dmax(SP(i, mw), 0.0) * WV(i, mq, mw);
            }
            S1(amdq, i, mq) = am;
            S1(apdq, i, mq) = ap;
        }
    }
}

static void rpt2_adv(int ixy, int maxm, int meqn, int mbc, int mx,
                     const double *aux2, const double *aux3, int imp,
                     const double *asdq, double *bmasdq, double *bpasdq)
{
    const int N = maxm + 2 * mbc;
    const int kv = 3 - ixy;
    int i, mq;
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        const int i1 = i - 2 + imp;
        for (mq = 1; mq <= meqn; mq++) {
            S1(bmasdq, i, mq) = dmin(S1(aux2, i1, kv), 0.0) * S1(asdq, i, mq);
            S1(bpasdq, i, mq) = dmax(S1(aux3, i1, kv), 0.0) * S1(asdq, i, mq);
        }
    }
}

/* ---------------- acoustics: clawpack46_rpn2_acoustics.f / _rpt2_acoustics.f ---------------- */

/* applications/clawpack/acoustics/2d/rp/clawpack46_rpn2_acoustics.f:39-75:
 * two waves, speeds -cc and cc, impedance zz */
static void rpn2_acoustics(int ixy, int maxm, int meqn, int mwaves, int mbc, int mx,
                           const double *ql, const double *qr, double cc, double zz,
                           double *wave, double *s, double *amdq, double *apdq)
{
    const int N = maxm + 2 * mbc;
    const int mu = ixy == 1 ? 2 : 3, mv = ixy == 1 ? 3 : 2;
    int i, m;
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        const double d1 = S1(ql, i, 1) - S1(qr, i - 1, 1);
        const double d2 = S1(ql, i, mu) - S1(qr, i - 1, mu);
        const double a1 = (-d1 + zz * d2) / (2.0 * zz);
        const double a2 = (d1 + zz * d2) / (2.0 * zz);
        WV(i, 1, 1) = -a1 * zz;
        WV(i, mu, 1) = a1;
        WV(i, mv, 1) = 0.0;
        SP(i, 1) = -cc;
        WV(i, 1, 2) = a2 * zz;
        WV(i, mu, 2) = a2;
        WV(i, mv, 2) = 0.0;
        SP(i, 2) = cc;
    }
    for (m = 1; m <= meqn; m++)
        for (i = 2 - mbc; i <= mx + mbc; i++) {
            S1(amdq, i, m) = SP(i, 1) * WV(i, m, 1);
            S1(apdq, i, m) = SP(i, 2) * WV(i, m, 2);
        }
}

/* applications/clawpack/acoustics/2d/rp/clawpack46_rpt2_acoustics.f:36-61 */
static void rpt2_acoustics(int ixy, int maxm, int mbc, int mx, double cc, double zz,
                           const double *asdq, double *bmasdq, double *bpasdq)
{
    const int N = maxm + 2 * mbc;
    const int mu = ixy == 1 ? 2 : 3, mv = ixy == 1 ? 3 : 2;
    int i;
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        const double a1 = (-S1(asdq, i, 1) + zz * S1(asdq, i, mv)) / (2.0 * zz);
        const double a2 = (S1(asdq, i, 1) + zz * S1(asdq, i, mv)) / (2.0 * zz);
        S1(bmasdq, i, 1) = cc * a1 * zz;
        S1(bmasdq, i, mu) = 0.0;
        S1(bmasdq, i, mv) = -cc * a1;
        S1(bpasdq, i, 1) = cc * a2 * zz;
        S1(bpasdq, i, mu) = 0.0;
        S1(bpasdq, i, mv) = cc * a2;
    }
}

/* ---------------- tracers: the advection solvers per component ---------------- */

static double tracer_scale(int n, int k) { return (double) (n - k) / (double) n; }

/* wave mw carries component mw only, at speed c_mw u_n; the fluctuations sum
 * over every wave as rpn2adv does */
static void rpn2_tracers(int ixy, int maxm, int meqn, int mwaves, int mbc, int mx,
                         const double *ql, const double *qr, const double *auxl,
                         double *wave, double *s, double *amdq, double *apdq)
{
    const int N = maxm + 2 * mbc;
    int i, mq, mw;
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        for (mw = 1; mw <= mwaves; mw++) {
            for (mq = 1; mq <= meqn; mq++)
                WV(i, mq, mw) = mq == mw ? S1(ql, i, mq) - S1(qr, i - 1, mq) : 0.0;
            SP(i, mw) = tracer_scale(mwaves, mw - 1) * S1(auxl, i, ixy);
        }
        for (mq = 1; mq <= meqn; mq++) {
            double am = 0.0, ap = 0.0;
            for (mw = 1; mw <= mwaves; mw++) {
                am = am + dmin(SP(i, mw), 0.0) * WV(i, mq, mw);
                ap = ap + dmax(SP(i, mw), 0.0) * WV(i, mq, mw);
            }
            S1(amdq, i, mq) = am;
            S1(apdq, i, mq) = ap;
        }
    }
}

/* rpt2adv per component with the component's transverse velocity */
static void rpt2_tracers(int ixy, int maxm, int meqn, int mbc, int mx,
                         const double *aux2, const double *aux3, int imp,
                         const double *asdq, double *bmasdq, double *bpasdq)
{
    const int N = maxm + 2 * mbc;
    const int kv = 3 - ixy;
    int i, mq;
    for (i = 2 - mbc; i <= mx + mbc; i++) {
        const int i1 = i - 2 + imp;
        for (mq = 1; mq <= meqn; mq++) {
            const double c = tracer_scale(meqn, mq - 1);
            S1(bmasdq, i, mq) = dmin(c * S1(aux2, i1, kv), 0.0) * S1(asdq, i, mq);
            S1(bpasdq, i, mq) = dmax(c * S1(aux3, i1, kv), 0.0) * S1(asdq, i, mq);
        }
    }
}

/* ---------------- limiter: clawpack46_inlinelimiter.f:39-115 ---------------- */

void oracle_inlinelimiter(int maxm, int meqn, int mwaves, int mbc, int mx,
                          double *wave, const double *s, const int *mthlim)
{
    const int N = maxm + 2 * mbc;
    int mw, i, m;
    for (mw = 1; mw <= mwaves; mw++) {
        double dotr;
        if (mthlim[mw - 1] == 0) continue;
        dotr = 0.0;
        for (i = 0; i <= mx + 1; i++) {
            double wnorm2 = 0.0, dotl = dotr, r, wlimitr, c;
            dotr = 0.0;
            for (m = 1; m <= meqn; m++) {
                wnorm2 = wnorm2 + WV(i, m, mw) * WV(i, m, mw);
                dotr = dotr + WV(i, m, mw) * WV(i + 1, m, mw);
            }
            if (i == 0) continue;
            if (wnorm2 == 0.0) continue;
            if (SP(i, mw) > 0.0)
                r = dotl / wnorm2;
            else
                r = dotr / wnorm2;
            switch (mthlim[mw - 1]) {
            case 1: /* minmod */
                wlimitr = dmax(0.0, dmin(1.0, r));
                break;
            case 2: /* superbee */
                wlimitr = dmax(dmax(0.0, dmin(1.0, 2.0 * r)), dmin(2.0, r));
                break;
            case 3: /* van Leer */
                wlimitr = (r + fabs(r)) / (1.0 + fabs(r));
                break;
            case 4: /* monotonized centered */
                c = (1.0 + r) / 2.0;
                wlimitr = dmax(0.0, dmin(dmin(c, 2.0), 2.0 * r));
                break;
            case 5: /* Beam-Warming */
                wlimitr = r;
                break;
            case 6: { /* generalized minmod; th is a REAL*4 literal in the reference */
                const double th = (double) 1.3f;
                wlimitr = dmax(0.0, dmin(dmin(th * r, (1.0 + r) / 2.0), th));
                break;
            }
            default: /* computed GOTO out of range falls through to label 10 (minmod) */
                wlimitr = dmax(0.0, dmin(1.0, r));
                break;
            }
            for (m = 1; m <= meqn; m++) WV(i, m, mw) = wlimitr * WV(i, m, mw);
        }
    }
}

/* ---------------- flux2: clawpack46_flux2.f:4-228 ---------------- */

typedef struct slice_work {
    double *wave, *s, *amdq, *apdq, *cqxx, *bmasdq, *bpasdq;
} slice_work_t;

static void flux2(const oracle_claw_t *claw, int ixy, int maxm, int meqn, int maux,
                  int mbc, int mx, const double *q1d, const double *dtdx1d,
                  const double *aux1, const double *aux2, const double *aux3,
                  double *faddm, double *faddp, double *gaddm, double *gaddp,
                  double *cfl1d, slice_work_t *w, roe_ctx_t *roe)
{
    const int N = maxm + 2 * mbc;
    const int mwaves = claw->mwaves;
    const int *method = claw->method;
    double *wave = w->wave, *s = w->s, *amdq = w->amdq, *apdq = w->apdq;
    double *cqxx = w->cqxx, *bmasdq = w->bmasdq, *bpasdq = w->bpasdq;
    int limit = 0, mw, m, i, jside;
#define D1(i) dtdx1d[(i) + mbc - 1]
#define G2(a, i, m, js) (a)[((js) - 1) * meqn * N + ((m) - 1) * N + (i) + mbc - 1]
    (void) maux;

    for (mw = 1; mw <= mwaves; mw++)
        if (claw->mthlim[mw - 1] > 0) limit = 1;

    for (jside = 1; jside <= 2; jside++)
        for (m = 1; m <= meqn; m++)
            for (i = 1 - mbc; i <= mx + mbc; i++) {
                S1(faddm, i, m) = 0.0;
                S1(faddp, i, m) = 0.0;
                G2(gaddm, i, m, jside) = 0.0;
                G2(gaddp, i, m, jside) = 0.0;
            }

    /* rpn2 is called with ql = qr = q1d and auxl = auxr = aux2 */
    switch (claw->solver) {
    case ORACLE_RP_ADVECTION:
        rpn2_adv(ixy, maxm, meqn, mwaves, mbc, mx, q1d, q1d, aux2, wave, s, amdq, apdq);
        break;
    case ORACLE_RP_SHALLOW:
        rpn2_shallow(ixy, maxm, meqn, mwaves, mbc, mx, q1d, q1d, claw->grav, wave, s, amdq, apdq, roe);
        break;
    case ORACLE_RP_ACOUSTICS:
        rpn2_acoustics(ixy, maxm, meqn, mwaves, mbc, mx, q1d, q1d, claw->p[2], claw->p[3], wave, s, amdq, apdq);
        break;
    case ORACLE_RP_TRACERS:
        rpn2_tracers(ixy, maxm, meqn, mwaves, mbc, mx, q1d, q1d, aux2, wave, s, amdq, apdq);
        break;
    default:
        rpn2_euler(ixy, maxm, meqn, mwaves, mbc, mx, q1d, q1d, claw->gamma, wave, s, amdq, apdq, roe);
        break;
    }

    for (i = 2 - mbc; i <= mx + mbc - 1; i++)
        for (m = 1; m <= meqn; m++) {
            S1(faddp, i, m) = S1(faddp, i, m) - S1(apdq, i, m);
            S1(faddm, i, m) = S1(faddm, i, m) + S1(amdq, i, m);
        }

    *cfl1d = 0.0;
    for (mw = 1; mw <= mwaves; mw++)
        for (i = 1; i <= mx + 1; i++)
            *cfl1d = dmax(dmax(*cfl1d, D1(i) * SP(i, mw)), -D1(i - 1) * SP(i, mw));

    if (method[1] == 2) {
        if (limit) oracle_inlinelimiter(maxm, meqn, mwaves, mbc, mx, wave, s, claw->mthlim);
        for (i = 2 - mbc; i <= mx + mbc; i++) {
            const double dtdxave = 0.5 * (D1(i - 1) + D1(i));
            for (m = 1; m <= meqn; m++) {
                S1(cqxx, i, m) = 0.0;
                for (mw = 1; mw <= mwaves; mw++)
                    S1(cqxx, i, m) = S1(cqxx, i, m) + fabs(SP(i, mw)) *
                        (1.0 - fabs(SP(i, mw)) * dtdxave) * WV(i, m, mw);
                S1(faddm, i, m) = S1(faddm, i, m) + 0.5 * S1(cqxx, i, m);
                S1(faddp, i, m) = S1(faddp, i, m) + 0.5 * S1(cqxx, i, m);
            }
        }
    }

    if (method[2] > 0) {
        if (method[1] > 1 && method[2] == 2)
            for (i = 2 - mbc; i <= mx + mbc - 1; i++)
                for (m = 1; m <= meqn; m++) {
                    S1(amdq, i, m) = S1(amdq, i, m) + S1(cqxx, i, m);
                    S1(apdq, i, m) = S1(apdq, i, m) - S1(cqxx, i, m);
                }

        /* ilr = 1: split the left-going fluctuation */
        switch (claw->solver) {
        case ORACLE_RP_ADVECTION:
            rpt2_adv(ixy, maxm, meqn, mbc, mx, aux2, aux3, 1, amdq, bmasdq, bpasdq);
            break;
        case ORACLE_RP_SHALLOW:
            rpt2_shallow(ixy, maxm, meqn, mwaves, mbc, mx, amdq, bmasdq, bpasdq, roe);
            break;
        case ORACLE_RP_ACOUSTICS:
            rpt2_acoustics(ixy, maxm, mbc, mx, claw->p[2], claw->p[3], amdq, bmasdq, bpasdq);
            break;
        case ORACLE_RP_TRACERS:
            rpt2_tracers(ixy, maxm, meqn, mbc, mx, aux2, aux3, 1, amdq, bmasdq, bpasdq);
            break;
        default:
            rpt2_euler(ixy, maxm, meqn, mwaves, mbc, mx, amdq, bmasdq, bpasdq, roe);
            break;
        }
        for (m = 1; m <= meqn; m++)
            for (i = 2 - mbc; i <= mx + mbc; i++) {
                double gupdate = 0.5 * D1(i - 1) * S1(bmasdq, i, m);
                G2(gaddm, i - 1, m, 1) = G2(gaddm, i - 1, m, 1) - gupdate;
                G2(gaddp, i - 1, m, 1) = G2(gaddp, i - 1, m, 1) - gupdate;
                gupdate = 0.5 * D1(i - 1) * S1(bpasdq, i, m);
                G2(gaddm, i - 1, m, 2) = G2(gaddm, i - 1, m, 2) - gupdate;
                G2(gaddp, i - 1, m, 2) = G2(gaddp, i - 1, m, 2) - gupdate;
            }

        /* ilr = 2: split the right-going fluctuation */
        switch (claw->solver) {
        case ORACLE_RP_ADVECTION:
            rpt2_adv(ixy, maxm, meqn, mbc, mx, aux2, aux3, 2, apdq, bmasdq, bpasdq);
            break;
        case ORACLE_RP_SHALLOW:
            rpt2_shallow(ixy, maxm, meqn, mwaves, mbc, mx, apdq, bmasdq, bpasdq, roe);
            break;
        case ORACLE_RP_ACOUSTICS:
            rpt2_acoustics(ixy, maxm, mbc, mx, claw->p[2], claw->p[3], apdq, bmasdq, bpasdq);
            break;
        case ORACLE_RP_TRACERS:
            rpt2_tracers(ixy, maxm, meqn, mbc, mx, aux2, aux3, 2, apdq, bmasdq, bpasdq);
            break;
        default:
            rpt2_euler(ixy, maxm, meqn, mwaves, mbc, mx, apdq, bmasdq, bpasdq, roe);
            break;
        }
        for (m = 1; m <= meqn; m++)
            for (i = 2 - mbc; i <= mx + mbc; i++) {
                double gupdate = 0.5 * D1(i) * S1(bmasdq, i, m);
                G2(gaddm, i, m, 1) = G2(gaddm, i, m, 1) - gupdate;
                G2(gaddp, i, m, 1) = G2(gaddp, i, m, 1) - gupdate;
                gupdate = 0.5 * D1(i) * S1(bpasdq, i, m);
                G2(gaddm, i, m, 2) = G2(gaddm, i, m, 2) - gupdate;
                G2(gaddp, i, m, 2) = G2(gaddp, i, m, 2) - gupdate;
            }
    }
#undef D1
#undef G2
}

void oracle_rpn2(const oracle_claw_t *claw, int ixy, int maxm, int mbc, int mx,
                 const double *q1d, const double *aux1d,
                 double *wave, double *s, double *amdq, double *apdq)
{
    const int N = maxm + 2 * mbc;
    roe_ctx_t roe;
    double *buf = (double *) calloc((size_t) 7 * N, sizeof(double));
    roe.N = N; roe.mbc = mbc;
    roe.u = buf; roe.v = buf + N; roe.a = buf + 2 * N; roe.h = buf + 3 * N;
    roe.u2v2 = buf + 3 * N; roe.enth = buf + 4 * N; roe.g1a2 = buf + 5 * N; roe.euv = buf + 6 * N;
    switch (claw->solver) {
    case ORACLE_RP_ADVECTION:
        rpn2_adv(ixy, maxm, claw->meqn, claw->mwaves, mbc, mx, q1d, q1d, aux1d, wave, s, amdq, apdq);
        break;
    case ORACLE_RP_SHALLOW:
        rpn2_shallow(ixy, maxm, claw->meqn, claw->mwaves, mbc, mx, q1d, q1d, claw->grav, wave, s, amdq, apdq, &roe);
        break;
    case ORACLE_RP_ACOUSTICS:
        rpn2_acoustics(ixy, maxm, claw->meqn, claw->mwaves, mbc, mx, q1d, q1d, claw->p[2], claw->p[3], wave, s, amdq, apdq);
        break;
    case ORACLE_RP_TRACERS:
        rpn2_tracers(ixy, maxm, claw->meqn, claw->mwaves, mbc, mx, q1d, q1d, aux1d, wave, s, amdq, apdq);
        break;
    default:
        rpn2_euler(ixy, maxm, claw->meqn, claw->mwaves, mbc, mx, q1d, q1d, claw->gamma, wave, s, amdq, apdq, &roe);
        break;
    }
    free(buf);
}

/* ---------------- step2 + step2_wrap ---------------- */

#define P2(a, i, j, m) (a)[((size_t) ((m) - 1) * (my + 2 * mbc) + (size_t) ((j) + mbc - 1)) * (mx + 2 * mbc) + (i) + mbc - 1]

int oracle_step2_patch(const oracle_claw_t *claw, int mx, int my, int mbc,
                       double *qold, const double *aux,
                       double dx, double dy, double dt, double *cflgrid,
                       double *fm_out, double *fp_out, double *gm_out, double *gp_out)
{
    const int meqn = claw->meqn, maux = claw->maux, mwaves = claw->mwaves;
    const int maxm = mx > my ? mx : my;
    const int N = maxm + 2 * mbc;
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    const double dtdx = dt / dx, dtdy = dt / dy;
    double *fm, *fp, *gm, *gp, *faddm, *faddp, *gaddm, *gaddp, *q1d, *dtdx1d, *dtdy1d;
    double *aux1, *aux2, *aux3, *buf, cfl1d;
    slice_work_t w;
    roe_ctx_t roe;
    int i, j, m, ma;
    size_t off = 0, total;

    total = 4 * psize + (size_t) N * (2 * meqn + 4 * meqn + meqn + 2 + 3 * (maux > 0 ? maux : 1))
          + (size_t) N * (meqn * mwaves + mwaves + 5 * meqn) + (size_t) 7 * N;
    buf = (double *) calloc(total, sizeof(double));
    if (!buf) return 1;
#define TAKE(p, n) do { p = buf + off; off += (n); } while (0)
    TAKE(fm, psize); TAKE(fp, psize); TAKE(gm, psize); TAKE(gp, psize);
    TAKE(faddm, (size_t) N * meqn); TAKE(faddp, (size_t) N * meqn);
    TAKE(gaddm, (size_t) 2 * N * meqn); TAKE(gaddp, (size_t) 2 * N * meqn);
    TAKE(q1d, (size_t) N * meqn); TAKE(dtdx1d, N); TAKE(dtdy1d, N);
    TAKE(aux1, (size_t) N * (maux > 0 ? maux : 1)); TAKE(aux2, (size_t) N * (maux > 0 ? maux : 1));
    TAKE(aux3, (size_t) N * (maux > 0 ? maux : 1));
    TAKE(w.wave, (size_t) N * meqn * mwaves); TAKE(w.s, (size_t) N * mwaves);
    TAKE(w.amdq, (size_t) N * meqn); TAKE(w.apdq, (size_t) N * meqn); TAKE(w.cqxx, (size_t) N * meqn);
    TAKE(w.bmasdq, (size_t) N * meqn); TAKE(w.bpasdq, (size_t) N * meqn);
    roe.N = N; roe.mbc = mbc;
    TAKE(roe.u, N); TAKE(roe.v, N); TAKE(roe.a, N); TAKE(roe.h, N);
    TAKE(roe.u2v2, N); TAKE(roe.enth, N); TAKE(roe.g1a2, N);
    roe.euv = roe.h; /* h (shallow) and euv (Euler) never coexist */
#undef TAKE

    *cflgrid = 0.0;
    for (i = 1 - mbc; i <= maxm + mbc; i++) {
        dtdx1d[i + mbc - 1] = dtdx;
        dtdy1d[i + mbc - 1] = dtdy;
    }

    /* x-sweeps: clawpack46_step2.f:113-167 */
    for (j = 2 - mbc; j <= my + mbc - 1; j++) {
        for (m = 1; m <= meqn; m++)
            for (i = 1 - mbc; i <= mx + mbc; i++) S1(q1d, i, m) = P2(qold, i, j, m);
        for (ma = 1; ma <= maux; ma++)
            for (i = 1 - mbc; i <= mx + mbc; i++) {
                S1(aux1, i, ma) = P2(aux, i, j - 1, ma);
                S1(aux2, i, ma) = P2(aux, i, j, ma);
                S1(aux3, i, ma) = P2(aux, i, j + 1, ma);
            }
        flux2(claw, 1, maxm, meqn, maux, mbc, mx, q1d, dtdx1d, aux1, aux2, aux3,
              faddm, faddp, gaddm, gaddp, &cfl1d, &w, &roe);
        *cflgrid = dmax(*cflgrid, cfl1d);
        for (m = 1; m <= meqn; m++)
            for (i = 2 - mbc; i <= mx + mbc; i++) {
                P2(fm, i, j, m) = P2(fm, i, j, m) + S1(faddm, i, m);
                P2(fp, i, j, m) = P2(fp, i, j, m) + S1(faddp, i, m);
                P2(gm, i, j, m) = P2(gm, i, j, m) + gaddm[(m - 1) * N + i + mbc - 1];
                P2(gp, i, j, m) = P2(gp, i, j, m) + gaddp[(m - 1) * N + i + mbc - 1];
                P2(gm, i, j + 1, m) = P2(gm, i, j + 1, m) + gaddm[meqn * N + (m - 1) * N + i + mbc - 1];
                P2(gp, i, j + 1, m) = P2(gp, i, j + 1, m) + gaddp[meqn * N + (m - 1) * N + i + mbc - 1];
            }
    }

    /* y-sweeps: clawpack46_step2.f:182-236 */
    for (i = 2 - mbc; i <= mx + mbc - 1; i++) {
        for (m = 1; m <= meqn; m++)
            for (j = 1 - mbc; j <= my + mbc; j++) S1(q1d, j, m) = P2(qold, i, j, m);
        for (ma = 1; ma <= maux; ma++)
            for (j = 1 - mbc; j <= my + mbc; j++) {
                S1(aux1, j, ma) = P2(aux, i - 1, j, ma);
                S1(aux2, j, ma) = P2(aux, i, j, ma);
                S1(aux3, j, ma) = P2(aux, i + 1, j, ma);
            }
        flux2(claw, 2, maxm, meqn, maux, mbc, my, q1d, dtdy1d, aux1, aux2, aux3,
              faddm, faddp, gaddm, gaddp, &cfl1d, &w, &roe);
        *cflgrid = dmax(*cflgrid, cfl1d);
        for (m = 1; m <= meqn; m++)
            for (j = 2 - mbc; j <= my + mbc; j++) {
                P2(gm, i, j, m) = P2(gm, i, j, m) + S1(faddm, j, m);
                P2(gp, i, j, m) = P2(gp, i, j, m) + S1(faddp, j, m);
                P2(fm, i, j, m) = P2(fm, i, j, m) + gaddm[(m - 1) * N + j + mbc - 1];
                P2(fp, i, j, m) = P2(fp, i, j, m) + gaddp[(m - 1) * N + j + mbc - 1];
                P2(fm, i + 1, j, m) = P2(fm, i + 1, j, m) + gaddm[meqn * N + (m - 1) * N + j + mbc - 1];
                P2(fp, i + 1, j, m) = P2(fp, i + 1, j, m) + gaddp[meqn * N + (m - 1) * N + j + mbc - 1];
            }
    }

    /* q update: clawpack46_step2_wrap.f:106-125 (mcapa = 0) */
    for (m = 1; m <= meqn; m++)
        for (i = 1 - mbc; i <= mx + mbc - 1; i++)
            for (j = 1 - mbc; j <= my + mbc - 1; j++)
                P2(qold, i, j, m) = P2(qold, i, j, m)
                    - dtdx * (P2(fm, i + 1, j, m) - P2(fp, i, j, m))
                    - dtdy * (P2(gm, i, j + 1, m) - P2(gp, i, j, m));

    if (fm_out) memcpy(fm_out, fm, psize * sizeof(double));
    if (fp_out) memcpy(fp_out, fp, psize * sizeof(double));
    if (gm_out) memcpy(gm_out, gm, psize * sizeof(double));
    if (gp_out) memcpy(gp_out, gp, psize * sizeof(double));
    free(buf);
    return 0;
}
