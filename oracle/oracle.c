/*
 * ORACLE -- test infrastructure, not product code.  See oracle.h.
 *
 * CPU restatement of the reference MBX MB-pol 2B/3B path.  Built with
 * -ffp-contract=off so that every a*b+c rounds twice, as the reference's
 * -O2 -std=c++17 build does on x86-64.  Reference paths are relative to src/.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------------ */
/* Polynomials                                                                                */
/* ------------------------------------------------------------------------------------------ */

/* Direct monomial summation of E = sum_k a_k sum_{m in S_k} x^m and its gradient.
 * Restates poly_2b_v6x::eval / poly_3b_v2x::eval (potential/2b/poly-2b-v6x.cpp:12936,
 * potential/3b/poly-3b-v2x.cpp:26982) from the extracted monomial table. */
double orc_poly_eval(const orc_poly* p, const double* a, const double* x, double* g) {
    if (!a) a = p->coef;
    double e = 0.0;
    if (g)
        for (int j = 0; j < p->nvars; ++j) g[j] = 0.0;
    for (int m = 0; m < p->nmono; ++m) {
        const int d = p->mono_deg[m];
        const int* id = p->mono_idx + p->stride * m;
        const double c = a[p->mono_k[m]];
        double pre[9], suf[9];
        pre[0] = 1.0;
        for (int q = 0; q < d; ++q) pre[q + 1] = pre[q] * x[id[q]];
        e += c * pre[d];
        if (g) {
            suf[d] = 1.0;
            for (int q = d - 1; q >= 0; --q) suf[q] = suf[q + 1] * x[id[q]];
            for (int q = 0; q < d; ++q) g[id[q]] += c * (pre[q] * suf[q + 1]);
        }
    }
    return e;
}

void orc_poly_basis(const orc_poly* p, const double* x, double* basis) {
    for (int k = 0; k < p->ncoef; ++k) basis[k] = 0.0;
    for (int m = 0; m < p->nmono; ++m) {
        const int* id = p->mono_idx + p->stride * m;
        double v = 1.0;
        for (int q = 0; q < p->mono_deg[m]; ++q) v *= x[id[q]];
        basis[p->mono_k[m]] += v;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* MB-pol 2B holder (potential/2b/x2b-v9x.cpp)                                                */
/* ------------------------------------------------------------------------------------------ */

/* Nonlinear parameters, x2b-v9x.cpp:1330-1345. */
static const double k_HH_intra = -6.480884773303821e-01;
static const double k_OH_intra = 1.674518993682975e+00;
static const double k_HH_coul = 1.148231864355956e+00;
static const double k_OH_coul = 1.205989761123099e+00;
static const double k_OO_coul = 1.395357065790959e+00;
static const double k_XH_main = 7.347036852042255e-01;
static const double k_XO_main = 7.998249864422826e-01;
static const double k_XX_main = 7.960663960630585e-01;
static const double in_plane_gamma = -9.721486914088159e-02;
static const double out_of_plane_gamma = 9.859272078406150e-02;
static const double r2i = 4.500000000000000e+00;
static const double r2f = 6.500000000000000e+00;

/* variable::v_exp / v_coul (x2b-v9x.cpp:54-91): value and g = d(value)/d(p1 - p2). */
static double v_exp2(double r0, double k, const double* xc, int o1, int o2, double* g) {
    g[0] = xc[o1] - xc[o2];
    g[1] = xc[o1 + 1] - xc[o2 + 1];
    g[2] = xc[o1 + 2] - xc[o2 + 2];
    const double r = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    const double e1 = exp(k * (r0 - r));
    const double gg = -k * e1 / r;
    g[0] *= gg;
    g[1] *= gg;
    g[2] *= gg;
    return e1;
}

static double v_coul2(double r0, double k, const double* xc, int o1, int o2, double* g) {
    g[0] = xc[o1] - xc[o2];
    g[1] = xc[o1 + 1] - xc[o2 + 1];
    g[2] = xc[o1 + 2] - xc[o2 + 2];
    const double rsq = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    const double r = sqrt(rsq);
    const double e1 = exp(k * (r0 - r));
    const double rinv = 1.0 / r;
    const double val = e1 * rinv;
    const double gg = -(k + rinv) * val * rinv;
    g[0] *= gg;
    g[1] *= gg;
    g[2] *= gg;
    return val;
}

/* monomer::setup (x2b-v9x.cpp:117-135): the two M-sites (X1, X2) of a water from the in-plane and
 * out-of-plane gammas; the MB-nrg CH4-H2O holder builds its Z sites the same way with MB-pol's
 * values (x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp:119-160, 866-872). */
static void msite_setup_w(const double* ohh, double ipg, double oopg, double* oh1, double* oh2, double* x1,
                          double* x2) {
    for (int i = 0; i < 3; ++i) {
        oh1[i] = ohh[i + 3] - ohh[i];
        oh2[i] = ohh[i + 6] - ohh[i];
    }
    const double v[3] = {oh1[1] * oh2[2] - oh1[2] * oh2[1], oh1[2] * oh2[0] - oh1[0] * oh2[2],
                         oh1[0] * oh2[1] - oh1[1] * oh2[0]};
    for (int i = 0; i < 3; ++i) {
        const double in_plane = ohh[i] + 0.5 * ipg * (oh1[i] + oh2[i]);
        const double out_of_plane = oopg * v[i];
        x1[i] = in_plane + out_of_plane;
        x2[i] = in_plane - out_of_plane;
    }
}

/* monomer::grads (x2b-v9x.cpp:139-160): X-site gradients back onto O, H1, H2. */
static void msite_grads_w(const double* oh1, const double* oh2, double ipg, double oopg, const double* g1,
                          const double* g2, double* grd) {
    const double gm[3] = {g1[0] - g2[0], g1[1] - g2[1], g1[2] - g2[2]};
    const double t1[3] = {oh2[1] * gm[2] - oh2[2] * gm[1], oh2[2] * gm[0] - oh2[0] * gm[2],
                          oh2[0] * gm[1] - oh2[1] * gm[0]};
    const double t2[3] = {oh1[1] * gm[2] - oh1[2] * gm[1], oh1[2] * gm[0] - oh1[0] * gm[2],
                          oh1[0] * gm[1] - oh1[1] * gm[0]};
    for (int i = 0; i < 3; ++i) {
        const double gsum = g1[i] + g2[i];
        const double in_plane = 0.5 * ipg * gsum;
        const double gh1 = in_plane + oopg * t1[i];
        const double gh2 = in_plane - oopg * t2[i];
        grd[i + 0] += gsum - (gh1 + gh2);
        grd[i + 3] += gh1;
        grd[i + 6] += gh2;
    }
}

/* f_switch (x2b-v9x.cpp:1349-1362). */
static double switch2(double r, double* g) {
    if (r > r2f) {
        *g = 0.0;
        return 0.0;
    } else if (r > r2i) {
        const double t1 = M_PI / (r2f - r2i);
        const double x = (r - r2i) * t1;
        *g = -sin(x) * t1 / 2.0;
        return (1.0 + cos(x)) / 2.0;
    }
    *g = 0.0;
    return 1.0;
}

/* Site offsets inside the 30-double coordinate block (x2b-v9x.cpp:1614-1627). */
enum { Oa = 0, Ha1 = 3, Ha2 = 6, Ob = 9, Hb1 = 12, Hb2 = 15, Xa1 = 18, Xa2 = 21, Xb1 = 24, Xb2 = 27 };

/* The 31 variable definitions of x2b-v9x.cpp:1640-1675: (kind, k, o1, o2); kind 0 = v_exp
 * intra (r0 = 1), 1 = v_coul inter (r0 = 4), 2 = v_exp inter (r0 = 4). */
typedef struct {
    int kind;
    int k;
    int o1, o2;
} var2_def;
enum { KHHI, KOHI, KHHC, KOHC, KOOC, KXH, KXO, KXX };
static const var2_def V2[31] = {
    {0, KHHI, Ha1, Ha2}, {0, KHHI, Hb1, Hb2}, {0, KOHI, Oa, Ha1},  {0, KOHI, Oa, Ha2},  {0, KOHI, Ob, Hb1},
    {0, KOHI, Ob, Hb2},  {1, KHHC, Ha1, Hb1}, {1, KHHC, Ha1, Hb2}, {1, KHHC, Ha2, Hb1}, {1, KHHC, Ha2, Hb2},
    {1, KOHC, Oa, Hb1},  {1, KOHC, Oa, Hb2},  {1, KOHC, Ob, Ha1},  {1, KOHC, Ob, Ha2},  {1, KOOC, Oa, Ob},
    {2, KXH, Xa1, Hb1},  {2, KXH, Xa1, Hb2},  {2, KXH, Xa2, Hb1},  {2, KXH, Xa2, Hb2},  {2, KXH, Xb1, Ha1},
    {2, KXH, Xb1, Ha2},  {2, KXH, Xb2, Ha1},  {2, KXH, Xb2, Ha2},  {2, KXO, Oa, Xb1},   {2, KXO, Oa, Xb2},
    {2, KXO, Ob, Xa1},   {2, KXO, Ob, Xa2},   {2, KXX, Xa1, Xb1},  {2, KXX, Xa1, Xb2},  {2, KXX, Xa2, Xb1},
    {2, KXX, Xa2, Xb2}};

/* orc_mbpol2b with the ten nonlinear parameters given (theta, in the order k_HH_intra k_OH_intra
 * k_HH_coul k_OH_coul k_OO_coul k_XH_main k_XO_main k_XX_main in_plane_gamma out_of_plane_gamma;
 * NULL = the reference's values). */
double orc_mbpol2b_t(const orc_poly* p, const double* a, const double* theta, const double* w1, const double* w2,
                     double* g1, double* g2, double* vir, int* evaluated) {
    static const double th0[10] = {k_HH_intra, k_OH_intra, k_HH_coul,      k_OH_coul,     k_OO_coul,
                                   k_XH_main,  k_XO_main,  k_XX_main, in_plane_gamma, out_of_plane_gamma};
    const double* th = theta ? theta : th0;
    const double* kv = th;
    /* r_OO and the cutoff test (x2b-v9x.cpp:1589-1600) */
    const double dOO[3] = {w1[0] - w2[0], w1[1] - w2[1], w1[2] - w2[2]};
    const double rOO = sqrt(dOO[0] * dOO[0] + dOO[1] * dOO[1] + dOO[2] * dOO[2]);
    if (evaluated) *evaluated = 0;
    if (rOO > r2f) return 0.0;
    if (evaluated) *evaluated = 1;

    double xc[30];
    memcpy(xc, w1, 9 * sizeof(double));
    memcpy(xc + 9, w2, 9 * sizeof(double));
    double oh1a[3], oh2a[3], oh1b[3], oh2b[3];
    msite_setup_w(xc + Oa, th[8], th[9], oh1a, oh2a, xc + Xa1, xc + Xa2);
    msite_setup_w(xc + Ob, th[8], th[9], oh1b, oh2b, xc + Xb1, xc + Xb2);

    double v[31], vg[31][3];
    for (int j = 0; j < 31; ++j) {
        const var2_def* d = &V2[j];
        if (d->kind == 0)
            v[j] = v_exp2(1.0, kv[d->k], xc, d->o1, d->o2, vg[j]);
        else if (d->kind == 1)
            v[j] = v_coul2(4.0, kv[d->k], xc, d->o1, d->o2, vg[j]);
        else
            v[j] = v_exp2(4.0, kv[d->k], xc, d->o1, d->o2, vg[j]);
    }
    double gp[31];
    const double e2b = orc_poly_eval(p, a, v, g1 ? gp : NULL);

    double gsw;
    const double sw = switch2(rOO, &gsw);
    if (g1) {
        /* variable::grads (x2b-v9x.cpp:95-103), then monomer::grads and the switch
         * (x2b-v9x.cpp:1718-1759) */
        double xg[30];
        memset(xg, 0, sizeof(xg));
        for (int j = 0; j < 31; ++j) {
            for (int i = 0; i < 3; ++i) {
                const double dd = gp[j] * vg[j][i];
                xg[V2[j].o1 + i] += dd;
                xg[V2[j].o2 + i] -= dd;
            }
        }
        msite_grads_w(oh1a, oh2a, th[8], th[9], xg + Xa1, xg + Xa2, xg + Oa);
        msite_grads_w(oh1b, oh2b, th[8], th[9], xg + Xb1, xg + Xb2, xg + Ob);
        double gd1[9], gd2[9];
        for (int j = 0; j < 9; ++j) {
            gd1[j] = sw * xg[j];
            gd2[j] = sw * xg[j + 9];
        }
        gsw *= e2b / rOO;
        for (int j = 0; j < 3; ++j) {
            const double dd = gsw * dOO[j];
            gd1[j] += dd;
            gd2[j] -= dd;
        }
        for (int j = 0; j < 9; ++j) {
            g1[j] += gd1[j];
            g2[j] += gd2[j];
        }
        if (vir) {
            /* virial of this dimer (x2b-v9x.cpp:1762-1800): -sum w_a g_b over its 6 atoms */
            static const int ab[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
            static const int slot[6] = {0, 1, 2, 4, 5, 8};
            for (int q = 0; q < 6; ++q) {
                const int ca = ab[q][0], cb = ab[q][1];
                vir[slot[q]] -= w1[ca] * gd1[cb] + w1[3 + ca] * gd1[3 + cb] + w1[6 + ca] * gd1[6 + cb] +
                                w2[ca] * gd2[cb] + w2[3 + ca] * gd2[3 + cb] + w2[6 + ca] * gd2[6 + cb];
            }
            vir[3] = vir[1];
            vir[6] = vir[2];
            vir[7] = vir[5];
        }
    }
    return sw * e2b;
}

double orc_mbpol2b(const orc_poly* p, const double* a, const double* w1, const double* w2, double* g1, double* g2,
                   double* vir, int* evaluated) {
    return orc_mbpol2b_t(p, a, NULL, w1, w2, g1, g2, vir, evaluated);
}

/* ------------------------------------------------------------------------------------------ */
/* MB-pol 3B holder (potential/3b/x3b-v2x.cpp)                                                */
/* ------------------------------------------------------------------------------------------ */

/* Parameters, x3b-v2x.cpp:1256-1271. */
static const double r3i = 0.000000000000000e+00;
static const double r3f = 4.500000000000000e+00;
static const double kHH_intra = -2.254303257872797e+00;
static const double kOH_intra = -2.564771901404151e-01;
static const double kHH = 4.247920544074333e-01;
static const double kOH = 8.128985941165371e-01;
static const double kOO = 3.749457984616480e-02;
static const double dHH_intra = 1.690594510379166e+00;
static const double dOH_intra = -2.999999868517452e+00;
static const double dHH = 3.499031358429095e+00;
static const double dOH = 4.854042963514281e+00;
static const double dOO = 4.816312044947604e-08;

/* var / g_var (x3b-v2x.cpp:52-76). */
static double var3(double k, double r0, const double* a1, const double* a2) {
    const double dx[3] = {a1[0] - a2[0], a1[1] - a2[1], a1[2] - a2[2]};
    const double dsq = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    const double d = sqrt(dsq);
    return exp(-k * (d - r0));
}

static void g_var3(double g, double k, double r0, const double* a1, const double* a2, double* g1, double* g2) {
    const double dx[3] = {a1[0] - a2[0], a1[1] - a2[1], a1[2] - a2[2]};
    const double dsq = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    const double d = sqrt(dsq);
    const double gg = -k * g * exp(-k * (d - r0)) / d;
    for (int i = 0; i < 3; ++i) {
        g1[i] += gg * dx[i];
        g2[i] -= gg * dx[i];
    }
}

/* f_switch (x3b-v2x.cpp:1275-1288). */
static double switch3(double r, double* g) {
    if (r > r3f) {
        *g = 0.0;
        return 0.0;
    } else if (r > r3i) {
        const double t1 = M_PI / (r3f - r3i);
        const double x = (r - r3i) * t1;
        *g = -sin(x) * t1 / 2.0;
        return (1.0 + cos(x)) / 2.0;
    }
    *g = 0.0;
    return 1.0;
}

/* Atom slots of a trimer: 0..8 = Oa Ha1 Ha2 Ob Hb1 Hb2 Oc Hc1 Hc2 (by 3-vector). */
enum { AO = 0, AH1, AH2, BO, BH1, BH2, CO, CH1, CH2 };
/* The 36 variables of x3b-v2x.cpp:1548-1584: (param set, atom1, atom2). */
typedef struct {
    int set;
    int a1, a2;
} var3_def;
enum { P_HHI, P_OHI, P_HH, P_OH, P_OO };
static const var3_def V3[36] = {
    {P_HHI, AH1, AH2}, {P_HHI, BH1, BH2}, {P_HHI, CH1, CH2}, {P_OHI, AO, AH1},  {P_OHI, AO, AH2},
    {P_OHI, BO, BH1},  {P_OHI, BO, BH2},  {P_OHI, CO, CH1},  {P_OHI, CO, CH2},  {P_HH, AH1, BH1},
    {P_HH, AH1, BH2},  {P_HH, AH1, CH1},  {P_HH, AH1, CH2},  {P_HH, AH2, BH1},  {P_HH, AH2, BH2},
    {P_HH, AH2, CH1},  {P_HH, AH2, CH2},  {P_HH, BH1, CH1},  {P_HH, BH1, CH2},  {P_HH, BH2, CH1},
    {P_HH, BH2, CH2},  {P_OH, AO, BH1},   {P_OH, AO, BH2},   {P_OH, AO, CH1},   {P_OH, AO, CH2},
    {P_OH, BO, AH1},   {P_OH, BO, AH2},   {P_OH, BO, CH1},   {P_OH, BO, CH2},   {P_OH, CO, AH1},
    {P_OH, CO, AH2},   {P_OH, CO, BH1},   {P_OH, CO, BH2},   {P_OO, AO, BO},    {P_OO, AO, CO},
    {P_OO, BO, CO}};

/* orc_mbpol3b with the ten nonlinear parameters given (theta = kHH_intra kOH_intra kHH kOH kOO
 * dHH_intra dOH_intra dHH dOH dOO; NULL = the reference's values). */
double orc_mbpol3b_t(const orc_poly* p, const double* a, const double* theta, const double* w1, const double* w2,
                     const double* w3, double* g1, double* g2, double* g3, double* vir, int* evaluated) {
    static const double th0[10] = {kHH_intra, kOH_intra, kHH, kOH, kOO, dHH_intra, dOH_intra, dHH, dOH, dOO};
    const double* th = theta ? theta : th0;
    const double* ks = th;
    const double* ds = th + 5;
    /* O-O distances and the skip rule (x3b-v2x.cpp:1495-1515) */
    double dab[3], dac[3], dbc[3];
    for (int j = 0; j < 3; ++j) {
        dab[j] = w1[j] - w2[j];
        dac[j] = w1[j] - w3[j];
        dbc[j] = w2[j] - w3[j];
    }
    double rabsq = 0.0, racsq = 0.0, rbcsq = 0.0;
    for (int j = 0; j < 3; ++j) {
        rabsq += dab[j] * dab[j];
        racsq += dac[j] * dac[j];
        rbcsq += dbc[j] * dbc[j];
    }
    const double rab = sqrt(rabsq), rac = sqrt(racsq), rbc = sqrt(rbcsq);
    const int skip = (rab > r3f && rac > r3f) || (rab > r3f && rbc > r3f) || (rac > r3f && rbc > r3f);
    if (evaluated) *evaluated = !skip;
    if (skip) return 0.0;

    const double* at[9] = {w1, w1 + 3, w1 + 6, w2, w2 + 3, w2 + 6, w3, w3 + 3, w3 + 6};
    double x[36];
    for (int j = 0; j < 36; ++j) x[j] = var3(ks[V3[j].set], ds[V3[j].set], at[V3[j].a1], at[V3[j].a2]);
    double gp[36];
    const double e3 = orc_poly_eval(p, a, x, g1 ? gp : NULL);

    double gab, gac, gbc;
    const double sab = switch3(rab, &gab);
    const double sac = switch3(rac, &gac);
    const double sbc = switch3(rbc, &gbc);
    const double s = sab * sac + sab * sbc + sac * sbc;
    if (g1) {
        /* x3b-v2x.cpp:1614-1686 */
        double gd[27];
        memset(gd, 0, sizeof(gd));
        for (int j = 0; j < 36; ++j) {
            const double gj = gp[j] * s;
            g_var3(gj, ks[V3[j].set], ds[V3[j].set], at[V3[j].a1], at[V3[j].a2], gd + 3 * V3[j].a1,
                   gd + 3 * V3[j].a2);
        }
        gab *= (sac + sbc) * e3 / rab;
        gac *= (sab + sbc) * e3 / rac;
        gbc *= (sab + sac) * e3 / rbc;
        for (int n = 0; n < 3; ++n) {
            gd[3 * AO + n] += gab * dab[n] + gac * dac[n];
            gd[3 * BO + n] += gbc * dbc[n] - gab * dab[n];
            gd[3 * CO + n] -= gac * dac[n] + gbc * dbc[n];
        }
        for (int j = 0; j < 9; ++j) {
            g1[j] += gd[j];
            g2[j] += gd[9 + j];
            g3[j] += gd[18 + j];
        }
        if (vir) {
            /* virial of this trimer (x3b-v2x.cpp:1687-1726) */
            static const int ab[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
            static const int slot[6] = {0, 1, 2, 4, 5, 8};
            const double* w[3] = {w1, w2, w3};
            for (int q = 0; q < 6; ++q) {
                const int ca = ab[q][0], cb = ab[q][1];
                double acc = vir[slot[q]];
                for (int m = 0; m < 3; ++m)
                    for (int t = 0; t < 3; ++t) acc += -w[m][3 * t + ca] * gd[9 * m + 3 * t + cb];
                vir[slot[q]] = acc;
            }
            vir[3] = vir[1];
            vir[6] = vir[2];
            vir[7] = vir[5];
        }
    }
    return s * e3;
}

double orc_mbpol3b(const orc_poly* p, const double* a, const double* w1, const double* w2, const double* w3,
                   double* g1, double* g2, double* g3, double* vir, int* evaluated) {
    return orc_mbpol3b_t(p, a, NULL, w1, w2, w3, g1, g2, g3, vir, evaluated);
}

/* Independent n-mers (order 2: [n][18] = O H H O H H; order 3: [n][27]), each evaluated as
 * given (no imaging): e[n] and, when g is not NULL, g[n][18 or 27].  Parallel over n-mers. */
void orc_mbpol_nmers(int order, const orc_poly* p, const double* a, const double* theta, long n, const double* xyz,
                     double* e, double* g) {
    const int w = 9 * order;
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < n; ++i) {
        const double* x = xyz + (size_t)i * w;
        double gl[27] = {0};
        if (order == 2)
            e[i] = orc_mbpol2b_t(p, a, theta, x, x + 9, g ? gl : NULL, g ? gl + 9 : NULL, NULL, NULL);
        else
            e[i] = orc_mbpol3b_t(p, a, theta, x, x + 9, x + 18, g ? gl : NULL, g ? gl + 9 : NULL,
                                 g ? gl + 18 : NULL, NULL, NULL);
        if (g)
            for (int j = 0; j < w; ++j) g[(size_t)i * w + j] = gl[j];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Geometry                                                                                   */
/* ------------------------------------------------------------------------------------------ */

int orc_invert_unit_cell(const double* box, double* bi) {
    /* tools/math_tools.cpp:37-64 */
    const double det = box[0] * (box[4] * box[8] - box[7] * box[5]) - box[1] * (box[3] * box[8] - box[5] * box[6]) +
                       box[2] * (box[3] * box[7] - box[4] * box[6]);
    if (fabs(det) < 1E-36) return -1;
    const double di = 1 / det;
    bi[0] = (box[4] * box[8] - box[7] * box[5]) * di;
    bi[1] = (box[2] * box[7] - box[1] * box[8]) * di;
    bi[2] = (box[1] * box[5] - box[2] * box[4]) * di;
    bi[3] = (box[5] * box[6] - box[3] * box[8]) * di;
    bi[4] = (box[0] * box[8] - box[2] * box[6]) * di;
    bi[5] = (box[3] * box[2] - box[0] * box[5]) * di;
    bi[6] = (box[3] * box[7] - box[6] * box[4]) * di;
    bi[7] = (box[6] * box[1] - box[0] * box[7]) * di;
    bi[8] = (box[0] * box[4] - box[3] * box[1]) * di;
    return 0;
}

static void to_frac(const double* bi, const double* p, double* r) {
    r[0] = bi[0] * p[0] + bi[3] * p[1] + bi[6] * p[2];
    r[1] = bi[1] * p[0] + bi[4] * p[1] + bi[7] * p[2];
    r[2] = bi[2] * p[0] + bi[5] * p[1] + bi[8] * p[2];
}

static void to_cart(const double* b, const double* r, double* p) {
    p[0] = b[0] * r[0] + b[3] * r[1] + b[6] * r[2];
    p[1] = b[1] * r[0] + b[4] * r[1] + b[7] * r[2];
    p[2] = b[2] * r[0] + b[5] * r[1] + b[8] * r[2];
}

void orc_fix_monomer_coordinates(double* xyz, const double* box, const double* bi, int nmon, const int* nat,
                                 const int* first_index) {
    /* bblock/sys_tools.cpp:287-319 */
    for (int i = 0; i < nmon; ++i) {
        double* x0 = xyz + 3 * first_index[i];
        double r0[3];
        to_frac(bi, x0, r0);
        for (int c = 0; c < 3; ++c) r0[c] -= floor(r0[c] + 0.5);
        to_cart(box, r0, x0);
        for (int j = 1; j < nat[i]; ++j) {
            double* xj = x0 + 3 * j;
            double r[3];
            to_frac(bi, xj, r);
            for (int c = 0; c < 3; ++c) r[c] -= floor(r[c] - r0[c] + 0.5);
            to_cart(box, r, xj);
        }
    }
}

void orc_close_image(const double* box, const double* bi, const double* xyz1, int nat2, double* xyz2) {
    /* bblock/sys_tools.cpp:330-360 */
    double r1[3], r2[3], d[3];
    to_frac(bi, xyz1, r1);
    to_frac(bi, xyz2, r2);
    for (int c = 0; c < 3; ++c) d[c] = floor(r2[c] - r1[c] + 0.5);
    for (int j = 0; j < nat2; ++j) {
        double r[3];
        to_frac(bi, xyz2 + 3 * j, r);
        for (int c = 0; c < 3; ++c) r[c] -= d[c];
        to_cart(box, r, xyz2 + 3 * j);
    }
}

/* GetFractionalCoordinates (kdtree/kdtree_utils.h:165-194). */
static void frac_coords(const double* box, const double* point, double* coord) {
    const double A = box[0];
    const double B = sqrt(box[3] * box[3] + box[4] * box[4]);
    const double C = sqrt(box[6] * box[6] + box[7] * box[7] + box[8] * box[8]);
    const double AdotB = box[0] * box[3];
    const double AdotC = box[0] * box[6];
    const double BdotC = box[3] * box[6] + box[4] * box[7];
    const double gamma = acos(AdotB / A / B);
    const double beta = acos(AdotC / A / C);
    const double alpha = acos(BdotC / B / C);
    const double ca = cos(alpha), cb = cos(beta), cg = cos(gamma), sg = sin(gamma);
    const double omega = A * B * C * sqrt(1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg);
    coord[0] = 1 / A * point[0] - cg / A / sg * point[1] + B * C * (ca * cg - cb) / omega / sg * point[2];
    coord[1] = 1 / B / sg * point[1] + A * C * (cb * cg - ca) / omega / sg * point[2];
    coord[2] = A * B * sg / omega * point[2];
}

typedef struct {
    double d;
    long idx;
} match_t;

static int match_cmp(const void* a, const void* b) {
    const match_t* x = (const match_t*)a;
    const match_t* y = (const match_t*)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

long orc_clusters(int nmax, double cutoff, int nmon, int use_pbc, const double* box, const double* bi,
                  const double* first_xyz, const int* islocal, const int* tag, int use_ghost, int* out, long cap) {
    /* monomers entering the cloud (sys_tools.cpp:501-517) */
    int* mon_index = (int*)malloc(sizeof(int) * (nmon > 0 ? nmon : 1));
    int* tag_index = (int*)malloc(sizeof(int) * (nmon > 0 ? nmon : 1));
    double* xyz = (double*)malloc(sizeof(double) * 3 * (nmon > 0 ? nmon : 1));
    int np = 0;
    for (int i = 0; i < nmon; ++i) {
        if (use_ghost || islocal[i]) {
            memcpy(xyz + 3 * np, first_xyz + 3 * i, 3 * sizeof(double));
            mon_index[np] = i;
            tag_index[np] = tag[i];
            ++np;
        }
    }
    long found = 0;
    if (np < 2 || (nmax > 2 && np < 3) || (nmax > 3 && np < 4)) goto done;
    {
        /* XyzToCloud (kdtree_utils.h:196-279): slot 0 = re-wrapped point, slots 1..7 images */
        const long size = use_pbc ? 8L * np : np;
        double* pts = (double*)calloc(3 * size, sizeof(double));
        static const int ind[3] = {0, -1, 1};
        for (int i = 0; i < np; ++i) {
            double* q = xyz + 3 * i;
            double uvw[3];
            if (use_pbc) {
                double r[3];
                to_frac(bi, q, r);
                for (int c = 0; c < 3; ++c) r[c] -= floor(r[c] + 0.5);
                to_cart(box, r, q);
                frac_coords(box, q, uvw);
            }
            memcpy(pts + 3 * i, q, 3 * sizeof(double));
            if (!use_pbc) continue;
            for (int m2 = 0; m2 < 3; ++m2) {
                const int m = ind[m2];
                if ((uvw[0] + m) > 1.0 || (uvw[0] + m) < -1.0) continue;
                for (int n2 = 0; n2 < 3; ++n2) {
                    const int n = ind[n2];
                    if ((uvw[1] + n) > 1.0 || (uvw[1] + n) < -1.0) continue;
                    for (int l2 = 0; l2 < 3; ++l2) {
                        if (m2 == 0 && n2 == 0 && l2 == 0) continue;
                        const int l = ind[l2];
                        if ((uvw[2] + l) > 1.0 || (uvw[2] + l) < -1.0) continue;
                        long shift;
                        if (m == 0 && n == 0)
                            shift = 1;
                        else if (m == 0 && l == 0)
                            shift = 2;
                        else if (n == 0 && l == 0)
                            shift = 3;
                        else if (m != 0 && n != 0 && l != 0)
                            shift = 4;
                        else if (m == 0)
                            shift = 5;
                        else if (n == 0)
                            shift = 6;
                        else
                            shift = 7;
                        shift *= np;
                        const double sx = (double)m * box[0] + (double)n * box[3] + (double)l * box[6];
                        const double sy = (double)m * box[1] + (double)n * box[4] + (double)l * box[7];
                        const double sz = (double)m * box[2] + (double)n * box[5] + (double)l * box[8];
                        pts[3 * (i + shift) + 0] = q[0] + sx;
                        pts[3 * (i + shift) + 1] = q[1] + sy;
                        pts[3 * (i + shift) + 2] = q[2] + sz;
                    }
                }
            }
        }
        const double c2 = cutoff * cutoff;
        match_t* mt = (match_t*)malloc(sizeof(match_t) * size);
        /* queries run over monomer indices i = 0..nmon-1 and address the cloud, tag_index and
         * mon_index with that same i (sys_tools.cpp:533-547); with every monomer local the cloud
         * index and the monomer index coincide. */
        for (int i = 0; i < nmon && i < np; ++i) {
            if (!islocal[i]) continue;
            const double* pt = pts + 3 * i;
            long nm = 0;
            /* radius search, strict dist < r^2 (nanoflann.hpp:166-171, kdtree_utils.h:63-69) */
            for (long s = 0; s < size; ++s) {
                const double d0 = pt[0] - pts[3 * s];
                const double d1 = pt[1] - pts[3 * s + 1];
                const double d2 = pt[2] - pts[3 * s + 2];
                const double d = d0 * d0 + d1 * d1 + d2 * d2;
                if (d < c2) {
                    mt[nm].d = d;
                    mt[nm].idx = s;
                    ++nm;
                }
            }
            qsort(mt, nm, sizeof(match_t), match_cmp);
            if (nmax == 2) {
                for (long j = 0; j < nm; ++j) {
                    const int idj = (int)(mt[j].idx % np);
                    if (tag_index[i] < tag_index[idj]) {
                        if (found < cap) {
                            out[2 * found] = mon_index[i];
                            out[2 * found + 1] = mon_index[idj];
                        }
                        ++found;
                    }
                }
            } else if (nmax == 3) {
                for (long j = 0; j < nm; ++j) {
                    for (long k = j + 1; k < nm; ++k) {
                        const double* pj = pts + 3 * mt[j].idx;
                        const double* pk = pts + 3 * mt[k].idx;
                        const double djk0 = pj[0] - pk[0], djk1 = pj[1] - pk[1], djk2 = pj[2] - pk[2];
                        const int idj = (int)(mt[j].idx % np);
                        const int idk = (int)(mt[k].idx % np);
                        if (idj == i || idk == i || idj == idk) continue;
                        if (djk0 * djk0 + djk1 * djk1 + djk2 * djk2 > c2 ||
                            (tag_index[i] < tag_index[idj] && tag_index[i] < tag_index[idk])) {
                            if (found < cap) {
                                out[3 * found] = mon_index[i];
                                out[3 * found + 1] = mon_index[idj];
                                out[3 * found + 2] = mon_index[idk];
                            }
                            ++found;
                        }
                    }
                }
            } else if (nmax == 4) {
                /* tetramers, the center-3-neighbour rule (sys_tools.cpp:594-634): no self-image check */
                for (long j = 0; j < nm; ++j)
                    for (long k = j + 1; k < nm; ++k)
                        for (long l = k + 1; l < nm; ++l) {
                            const double* pj = pts + 3 * mt[j].idx;
                            const double* pk = pts + 3 * mt[k].idx;
                            const double* pl = pts + 3 * mt[l].idx;
                            double a0 = pj[0] - pk[0], a1 = pj[1] - pk[1], a2 = pj[2] - pk[2];
                            const double d2jk = a0 * a0 + a1 * a1 + a2 * a2;
                            a0 = pj[0] - pl[0], a1 = pj[1] - pl[1], a2 = pj[2] - pl[2];
                            const double d2jl = a0 * a0 + a1 * a1 + a2 * a2;
                            a0 = pk[0] - pl[0], a1 = pk[1] - pl[1], a2 = pk[2] - pl[2];
                            const double d2kl = a0 * a0 + a1 * a1 + a2 * a2;
                            const int idj = (int)(mt[j].idx % np);
                            const int idk = (int)(mt[k].idx % np);
                            const int idl = (int)(mt[l].idx % np);
                            const int ti = tag_index[i];
                            if (!((ti >= tag_index[idj] && d2jk < c2 && d2jl < c2) ||
                                  (ti >= tag_index[idk] && d2jk < c2 && d2kl < c2) ||
                                  (ti >= tag_index[idl] && d2kl < c2 && d2jl < c2))) {
                                if (found < cap) {
                                    out[4 * found] = mon_index[i];
                                    out[4 * found + 1] = mon_index[idj];
                                    out[4 * found + 2] = mon_index[idk];
                                    out[4 * found + 3] = mon_index[idl];
                                }
                                ++found;
                            }
                        }
            }
        }
        free(mt);
        free(pts);
    }
done:
    free(mon_index);
    free(tag_index);
    free(xyz);
    return found;
}

/* ------------------------------------------------------------------------------------------ */
/* Pure-water system                                                                          */
/* ------------------------------------------------------------------------------------------ */

int orc_water_system(const orc_poly* p2, const orc_poly* p3, int nmon, double* xyz, const double* box,
                     double cutoff2b, double cutoff3b, double* e2b, double* e3b, double* grad, double* virial,
                     long* counts, int* dimers, long cap2, int* trimers, long cap3, double* e_dimer,
                     double* e_trimer) {
    double bi[9];
    const int pbc = box != NULL;
    int* nat = (int*)malloc(sizeof(int) * nmon);
    int* fi = (int*)malloc(sizeof(int) * nmon);
    int* loc = (int*)malloc(sizeof(int) * nmon);
    int* tag = (int*)malloc(sizeof(int) * nmon);
    double* first = (double*)malloc(sizeof(double) * 3 * nmon);
    for (int i = 0; i < nmon; ++i) {
        nat[i] = 3;
        fi[i] = 3 * i;
        loc[i] = 1;
        tag[i] = 3 * i;
    }
    if (pbc) {
        if (orc_invert_unit_cell(box, bi)) return -1;
        orc_fix_monomer_coordinates(xyz, box, bi, nmon, nat, fi);
    }
    for (int i = 0; i < nmon; ++i) memcpy(first + 3 * i, xyz + 9 * i, 3 * sizeof(double));
    memset(grad, 0, sizeof(double) * 9 * nmon);
    memset(virial, 0, sizeof(double) * 9);
    *e2b = 0.0;
    *e3b = 0.0;
    for (int order = 2; order <= 3; ++order) {
        const double cut = order == 2 ? cutoff2b : cutoff3b;
        const long n = orc_clusters(order, cut, nmon, pbc, box, bi, first, loc, tag, 0, NULL, 0);
        int* list = (int*)malloc(sizeof(int) * order * (n > 0 ? n : 1));
        orc_clusters(order, cut, nmon, pbc, box, bi, first, loc, tag, 0, list, n);
        int* out = order == 2 ? dimers : trimers;
        const long cap = order == 2 ? cap2 : cap3;
        if (out) memcpy(out, list, sizeof(int) * order * (n < cap ? n : cap));
        double* per = order == 2 ? e_dimer : e_trimer;
        long nev = 0;
        double etot = 0.0;
#pragma omp parallel reduction(+ : nev, etot)
        {
            double* gl = (double*)calloc(9 * (size_t)nmon, sizeof(double));
            double vl[9] = {0};
#pragma omp for schedule(dynamic, 64)
            for (long t = 0; t < n; ++t) {
                const int* id = list + order * t;
                double w[27], g[27], v[9] = {0};
                memcpy(w, xyz + 9 * id[0], 9 * sizeof(double));
                memcpy(w + 9, xyz + 9 * id[1], 9 * sizeof(double));
                if (order == 3) memcpy(w + 18, xyz + 9 * id[2], 9 * sizeof(double));
                if (pbc) {
                    orc_close_image(box, bi, w, 3, w + 9);
                    if (order == 3) orc_close_image(box, bi, w, 3, w + 18);
                }
                memset(g, 0, sizeof(g));
                int ev = 0;
                double e;
                if (order == 2)
                    e = orc_mbpol2b(p2, NULL, w, w + 9, g, g + 9, v, &ev);
                else
                    e = orc_mbpol3b(p3, NULL, w, w + 9, w + 18, g, g + 9, g + 18, v, &ev);
                if (per) per[t] = e;
                nev += ev;
                etot += e;
                for (int m = 0; m < order; ++m)
                    for (int j = 0; j < 9; ++j) gl[9 * id[m] + j] += g[9 * m + j];
                for (int k = 0; k < 9; ++k) vl[k] += v[k];
            }
#pragma omp critical(orc_reduce)
            {
                for (long j = 0; j < 9L * nmon; ++j) grad[j] += gl[j];
                for (int k = 0; k < 9; ++k) virial[k] += vl[k];
            }
            free(gl);
        }
        if (order == 2) {
            *e2b = etot;
            counts[0] = n;
            counts[1] = nev;
        } else {
            *e3b = etot;
            counts[2] = n;
            counts[3] = nev;
        }
        free(list);
    }
    free(nat);
    free(fi);
    free(loc);
    free(tag);
    free(first);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* MB-nrg holders (see oracle.h). */

/* Variable and its dr derivative factor: value, and gg with d(value)/d(p1) = gg * (p1 - p2)
 * (tools/variable.cpp:49-129). */
static double nrg_var(int form, double k, double d, const double* p1, const double* p2, double* gvec) {
    gvec[0] = p1[0] - p2[0];
    gvec[1] = p1[1] - p2[1];
    gvec[2] = p1[2] - p2[2];
    const double rsq = gvec[0] * gvec[0] + gvec[1] * gvec[1] + gvec[2] * gvec[2];
    const double r = sqrt(rsq);
    double val, gg;
    if (form == 0) { /* v_exp0 */
        val = exp(k * (d - r));
        gg = -k * val / r;
    } else if (form == 1) { /* v_coul0 */
        const double e1 = exp(k * (d - r));
        const double rinv = 1.0 / r;
        val = e1 * rinv;
        gg = -(k + rinv) * val * rinv;
    } else { /* v_exp */
        val = exp(k * (-r));
        gg = -k * val / r;
    }
    gvec[0] *= gg;
    gvec[1] *= gg;
    gvec[2] *= gg;
    return val;
}

/* x1b_A1B4_v1x::eval with gradients (x1b_A1B4_deg5_exp0_v1x.cpp:371-470): the variables in
 * order, the polynomial, variable::grads per variable (:56-63), the virial rows (:442-462). */
double orc_1b_nrg(const orc_holder* h, const orc_poly* p, const double* a, const double* w, double* g, double* vir6) {
    double v[64], gv[64][3], gp[64];
    for (int j = 0; j < h->nvars; ++j)
        v[j] = nrg_var(h->form[j], h->k[j], h->d[j], w + 3 * h->s1[j], w + 3 * h->s2[j], gv[j]);
    const double e = orc_poly_eval(p, a, v, g || vir6 ? gp : NULL);
    if (!g && !vir6) return e;
    double gw[3 * 16];
    for (int i = 0; i < 3 * h->nreal; ++i) gw[i] = 0.0;
    for (int j = 0; j < h->nvars; ++j)
        for (int i = 0; i < 3; ++i) {
            const double dd = gp[j] * gv[j][i];
            gw[3 * h->s1[j] + i] += dd;
            gw[3 * h->s2[j] + i] -= dd;
        }
    if (g)
        for (int i = 0; i < 3 * h->nreal; ++i) g[i] = gw[i];
    if (vir6) {
        const int va[6] = {0, 0, 0, 1, 1, 2}, vb[6] = {0, 1, 2, 1, 2, 2};
        for (int q = 0; q < 6; ++q) {
            double t = -w[va[q]] * gw[vb[q]];
            for (int at = 1; at < h->nreal; ++at) t = t - w[3 * at + va[q]] * gw[3 * at + vb[q]];
            vir6[q] += t;
        }
    }
    return e;
}

/* f_switch of the MB-nrg holders (x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp:804-817,
 * mbnrg_3b_A1B4_C1D2_C1D2_deg3_v1.cpp:1319-1333). */
static double nrg_switch(double ri, double ro, double r, double* g) {
    if (r > ro) {
        *g = 0.0;
        return 0.0;
    } else if (r > ri) {
        const double t1 = M_PI / (ro - ri);
        const double x = (r - ri) * t1;
        *g = -sin(x) * t1 / 2.0;
        return (1.0 + cos(x)) / 2.0;
    }
    *g = 0.0;
    return 1.0;
}

double orc_nrg_holder(const orc_holder* h, const orc_poly* p, const double* a, const double* const* mons,
                      double* const* grads, double* vir, int* evaluated) {
    double xs[3 * 16], xg[3 * 16];
    int off[4] = {0, 0, 0, 0};
    for (int m = 0; m < h->nmon; ++m) off[m + 1] = off[m] + h->nat[m];
    for (int m = 0; m < h->nmon; ++m) memcpy(xs + 3 * off[m], mons[m], sizeof(double) * 3 * h->nat[m]);
    *evaluated = 0;
    /* first-atom distances */
    double dd[3][3], r[3];
    const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
    const int npair = h->nmon == 2 ? 1 : 3;
    for (int q = 0; q < npair; ++q) {
        const double* u = xs + 3 * off[pa[q]];
        const double* v = xs + 3 * off[pb[q]];
        dd[q][0] = u[0] - v[0];
        dd[q][1] = u[1] - v[1];
        dd[q][2] = u[2] - v[2];
        r[q] = sqrt(dd[q][0] * dd[q][0] + dd[q][1] * dd[q][1] + dd[q][2] * dd[q][2]);
    }
    if (h->nmon == 2 ? r[0] > h->r_out : (r[0] > h->r_out && r[1] > h->r_out && r[2] > h->r_out)) return 0.0;
    *evaluated = 1;
    double oh1[3], oh2[3];
    if (h->vsite) msite_setup_w(xs + 3 * h->osite, h->in_plane, h->out_of_plane, oh1, oh2, xs + 3 * h->nreal,
                                xs + 3 * h->nreal + 3);
    double v[64], gvv[64][3];
    for (int j = 0; j < h->nvars; ++j)
        v[j] = nrg_var(h->form[j], h->k[j], h->d[j], xs + 3 * h->s1[j], xs + 3 * h->s2[j], gvv[j]);
    double g[64];
    const double P = orc_poly_eval(p, a, v, g);
    double sw, s[3], gs[3];
    if (h->nmon == 2) {
        s[0] = nrg_switch(h->r_in, h->r_out, r[0], &gs[0]);
        sw = s[0];
    } else {
        for (int q = 0; q < 3; ++q) s[q] = nrg_switch(h->r_in, h->r_out, r[q], &gs[q]);
        sw = s[0] * s[1] + s[0] * s[2] + s[1] * s[2];
        for (int j = 0; j < h->nvars; ++j) g[j] *= sw; /* 3B scales the variable gradients first */
    }
    memset(xg, 0, sizeof(xg));
    for (int j = 0; j < h->nvars; ++j)
        for (int c = 0; c < 3; ++c) {
            const double t = g[j] * gvv[j][c];
            xg[3 * h->s1[j] + c] += t;
            xg[3 * h->s2[j] + c] -= t;
        }
    if (h->vsite) msite_grads_w(oh1, oh2, h->in_plane, h->out_of_plane, xg + 3 * h->nreal, xg + 3 * h->nreal + 3,
                                xg + 3 * h->osite);
    double E;
    if (h->nmon == 2) {
        for (int i = 0; i < 3 * h->nreal; ++i) xg[i] *= sw;
        const double gsw = gs[0] * (P / r[0]);
        E = P * sw;
        for (int c = 0; c < 3; ++c) {
            const double t = gsw * dd[0][c];
            xg[c] += t;
            xg[3 * off[1] + c] -= t;
        }
    } else {
        const double g12 = gs[0] * ((1.0 * s[1] + 1.0 * s[2]) * P / r[0]);
        const double g13 = gs[1] * ((s[0] * 1.0 + 1.0 * s[2]) * P / r[1]);
        const double g23 = gs[2] * ((s[0] * 1.0 + s[1] * 1.0) * P / r[2]);
        E = P * sw;
        for (int c = 0; c < 3; ++c) {
            xg[c] += 0.0 + (g12 * dd[0][c]) + (g13 * dd[1][c]);
            xg[3 * off[1] + c] += 0.0 - (g12 * dd[0][c]) + (g23 * dd[2][c]);
            xg[3 * off[2] + c] += 0.0 - (g13 * dd[1][c]) - (g23 * dd[2][c]);
        }
    }
    for (int m = 0; m < h->nmon; ++m)
        if (grads && grads[m])
            for (int i = 0; i < 3 * h->nat[m]; ++i) grads[m][i] += xg[3 * off[m] + i];
    if (vir) {
        const int va[6] = {0, 0, 0, 1, 1, 2}, vb[6] = {0, 1, 2, 1, 2, 2}, slot[6] = {0, 1, 2, 4, 5, 8};
        for (int q = 0; q < 6; ++q) {
            double acc = 0.0;
            for (int at = 0; at < h->nreal; ++at) acc += -xs[3 * at + va[q]] * xg[3 * at + vb[q]];
            vir[slot[q]] += acc;
        }
        vir[3] = vir[1];
        vir[6] = vir[2];
        vir[7] = vir[5];
    }
    return E;
}

/* Order of an n-mer's monomers for its holder (energy2b.cpp:243-254, 323-325;
 * energy3b.cpp:185-216).  Returns the family (0 mbpol 2B, 1 ch4h2o 2B, 2 ch4ch4 2B, 3 mbpol 3B,
 * 4 ch4h2o2 3B) or -1 when the n-mer has no term; perm[m] = list position of holder slot m. */
static int nmer_family(int order, const int* kinds, int* perm) {
    if (order == 2) {
        perm[0] = 0;
        perm[1] = 1;
        if (kinds[0] == 0 && kinds[1] == 0) return 0;
        if (kinds[0] == 1 && kinds[1] == 1) return 2;
        if (kinds[0] == 0 && kinds[1] == 1) return 1;
        if (kinds[0] == 1 && kinds[1] == 0) {
            perm[0] = 1;
            perm[1] = 0;
            return 1;
        }
        return -1;
    }
    /* the exchange network of e3b::get_3b_energy on the names: rank 0 = "ch4" < rank 1 = "h2o" */
    int k[3] = {kinds[0] == 1 ? 0 : 1, kinds[1] == 1 ? 0 : 1, kinds[2] == 1 ? 0 : 1}, ix[3] = {0, 1, 2}, t;
#define SWP(i, j) (t = k[i], k[i] = k[j], k[j] = t, t = ix[i], ix[i] = ix[j], ix[j] = t)
    if (k[2] < k[0] && k[2] < k[1]) SWP(0, 2);
    if (k[0] > k[1] && k[0] > k[2])
        SWP(0, 2);
    else if (k[1] > k[0] && k[1] > k[2])
        SWP(1, 2);
    if (k[0] > k[1]) SWP(0, 1);
#undef SWP
    perm[0] = ix[0];
    perm[1] = ix[1];
    perm[2] = ix[2];
    if (kinds[0] == 0 && kinds[1] == 0 && kinds[2] == 0) return 3;
    if (k[0] == 0 && k[1] == 1 && k[2] == 1 && kinds[perm[0]] == 1 && kinds[perm[1]] == 0 && kinds[perm[2]] == 0)
        return 4;
    return -1;
}

int orc_system(const orc_poly* p2, const orc_poly* p3, const orc_holder* const* hol, const orc_poly* const* hp,
               int fam_mask, int nmon, const int* kind, const int* nat, const int* tag, double* xyz,
               const double* box, double cutoff2b, double cutoff3b, double* e2b, double* e3b, double* grad,
               double* virial, long* counts) {
    double bi[9];
    const int pbc = box != NULL;
    int* fi = (int*)malloc(sizeof(int) * nmon);
    int* loc = (int*)malloc(sizeof(int) * nmon);
    double* first = (double*)malloc(sizeof(double) * 3 * nmon);
    int natoms = 0;
    for (int i = 0; i < nmon; ++i) {
        fi[i] = natoms;
        natoms += nat[i];
        loc[i] = 1;
    }
    if (pbc) {
        if (orc_invert_unit_cell(box, bi)) return -1;
        orc_fix_monomer_coordinates(xyz, box, bi, nmon, nat, fi);
    }
    for (int i = 0; i < nmon; ++i) memcpy(first + 3 * i, xyz + 3 * fi[i], 3 * sizeof(double));
    memset(grad, 0, sizeof(double) * 3 * natoms);
    memset(virial, 0, sizeof(double) * 9);
    *e2b = 0.0;
    *e3b = 0.0;
    for (int order = 2; order <= 3; ++order) {
        const double cut = order == 2 ? cutoff2b : cutoff3b;
        const long n = orc_clusters(order, cut, nmon, pbc, box, bi, first, loc, tag, 0, NULL, 0);
        int* list = (int*)malloc(sizeof(int) * order * (n > 0 ? n : 1));
        orc_clusters(order, cut, nmon, pbc, box, bi, first, loc, tag, 0, list, n);
        long nev = 0;
        double etot = 0.0;
#pragma omp parallel reduction(+ : nev, etot)
        {
            double* gl = (double*)calloc(3 * (size_t)natoms, sizeof(double));
            double vl[9] = {0};
#pragma omp for schedule(dynamic, 64)
            for (long t = 0; t < n; ++t) {
                const int* id = list + order * t;
                double w[3][15], g[3][15], v[9] = {0};
                int kd[3], perm[3];
                for (int m = 0; m < order; ++m) {
                    kd[m] = kind[id[m]];
                    memcpy(w[m], xyz + 3 * fi[id[m]], sizeof(double) * 3 * nat[id[m]]);
                    memset(g[m], 0, sizeof(g[m]));
                }
                if (pbc)
                    for (int m = 1; m < order; ++m) orc_close_image(box, bi, w[0], nat[id[m]], w[m]);
                const int fam = nmer_family(order, kd, perm);
                if (fam < 0 || !((fam_mask >> fam) & 1)) continue;
                int ev = 0;
                double e;
                if (fam == 0) {
                    e = orc_mbpol2b(p2, NULL, w[0], w[1], g[0], g[1], v, &ev);
                } else if (fam == 3) {
                    e = orc_mbpol3b(p3, NULL, w[perm[0]], w[perm[1]], w[perm[2]], g[perm[0]], g[perm[1]], g[perm[2]],
                                    v, &ev);
                } else {
                    const int h = fam == 4 ? 2 : fam - 1;
                    const double* ms[3] = {w[perm[0]], w[perm[1]], order == 3 ? w[perm[2]] : NULL};
                    double* gs[3] = {g[perm[0]], g[perm[1]], order == 3 ? g[perm[2]] : NULL};
                    e = orc_nrg_holder(hol[h], hp[h], NULL, ms, gs, v, &ev);
                }
                nev += ev;
                etot += e;
                for (int m = 0; m < order; ++m)
                    for (int j = 0; j < 3 * nat[id[m]]; ++j) gl[3 * (size_t)fi[id[m]] + j] += g[m][j];
                for (int q = 0; q < 9; ++q) vl[q] += v[q];
            }
#pragma omp critical(orc_reduce_sys)
            {
                for (long j = 0; j < 3L * natoms; ++j) grad[j] += gl[j];
                for (int q = 0; q < 9; ++q) virial[q] += vl[q];
            }
            free(gl);
        }
        if (order == 2) {
            *e2b = etot;
            counts[0] = n;
            counts[1] = nev;
        } else {
            *e3b = etot;
            counts[2] = n;
            counts[3] = nev;
        }
        free(list);
    }
    free(fi);
    free(loc);
    free(first);
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * MB-pol 1B, ps::pot_nasa (potential/1b/ps.cpp:346-633), one water at a time.  The operation
 * order of the reference is kept (no contraction): bond vectors and lengths (:393-433), cos of
 * the HOH angle (:437-440), the Morse and H-H terms (:456-466), the polynomial in x1, x2, x3
 * through the power tables fmat (:475-491) and the four sums (:505-526), energy (:533-539) and
 * the chain rule to H1, H2 and O (:542-582), virial (:584-603). */
double orc_ps_1b(const orc_ps* p, const double* w, double* g, double* vir6) {
    double roh1[3], roh2[3], rhh[3];
    for (int i = 0; i < 3; ++i) {
        roh1[i] = w[3 + i] - w[i];
        roh2[i] = w[6 + i] - w[i];
        rhh[i] = w[3 + i] - w[6 + i];
    }
    double d1 = roh1[0] * roh1[0] + roh1[1] * roh1[1];
    d1 = d1 + roh1[2] * roh1[2];
    double d2 = roh2[0] * roh2[0] + roh2[1] * roh2[1];
    d2 = d2 + roh2[2] * roh2[2];
    double d3 = rhh[0] * rhh[0] + rhh[1] * rhh[1];
    d3 = d3 + rhh[2] * rhh[2];
    d1 = sqrt(d1);
    d2 = sqrt(d2);
    d3 = sqrt(d3);
    const double costh = (roh1[0] * roh2[0] + roh1[1] * roh2[1] + roh1[2] * roh2[2]) / (d1 * d2);
    const double exp1 = exp(-p->alphaoh * (d1 - p->roh));
    const double exp2 = exp(-p->alphaoh * (d2 - p->roh));
    const double Va = p->deoh * (exp1 * (exp1 - 2.0) + exp2 * (exp2 - 2.0));
    const double Vb = p->phh1 * exp(-p->phh2 * d3);
    const double dVa1 = 2.0 * p->alphaoh * p->deoh * exp1 * (1.0 - exp1) / d1;
    const double dVa2 = 2.0 * p->alphaoh * p->deoh * exp2 * (1.0 - exp2) / d2;
    const double dVb = -p->phh2 * Vb / d3;
    const double x1 = (d1 - p->reoh) / p->reoh;
    const double x2 = (d2 - p->reoh) / p->reoh;
    const double x3 = costh - p->costhe;
    double f[3][16];
    for (int i = 0; i < 3; ++i) {
        f[i][0] = 0.0;
        f[i][1] = 1.0;
    }
    for (int j = 2; j < 16; ++j) {
        f[0][j] = f[0][j - 1] * x1;
        f[1][j] = f[1][j - 1] * x2;
        f[2][j] = f[2][j - 1] * x3;
    }
    const double efac = exp(-p->b1 * (pow(d1 - p->reoh, 2) + pow(d2 - p->reoh, 2)));
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int j = 1; j < 245; ++j) {
        const int I = p->idx1[j], J = p->idx2[j], K = p->idx3[j];
        const double c = p->c5z[j];
        s0 += c * (f[0][I] * f[1][J] + f[0][J] * f[1][I]) * f[2][K];
        s1 += c * ((double)(I - 1) * f[0][I - 1] * f[1][J] + (double)(J - 1) * f[0][J - 1] * f[1][I]) * f[2][K];
        s2 += c * ((double)(J - 1) * f[0][I] * f[1][J - 1] + (double)(I - 1) * f[0][J] * f[1][I - 1]) * f[2][K];
        s3 += c * (f[0][I] * f[1][J] + f[0][J] * f[1][I]) * (double)(K - 1) * f[2][K - 1];
    }
    const double Vc = 2 * p->c5z[0] + efac * s0;
    double e = Va + Vb + Vc;
    e += p->correction;
    e *= p->cm1;
    if (g) {
        const double dVcdr1 = (-2 * p->b1 * efac * (d1 - p->reoh) * s0 + efac * s1 / p->reoh) / d1;
        const double dVcdr2 = (-2 * p->b1 * efac * (d2 - p->reoh) * s0 + efac * s2 / p->reoh) / d2;
        const double dVcdcth = efac * s3;
        for (int i = 0; i < 3; ++i) {
            const double h1 = (dVa1 * roh1[i] + dVb * rhh[i] + dVcdr1 * roh1[i] +
                               dVcdcth * (roh2[i] / (d1 * d2) - costh * roh1[i] / (d1 * d1))) * p->cm1;
            const double h2 = (dVa2 * roh2[i] - dVb * rhh[i] + dVcdr2 * roh2[i] +
                               dVcdcth * (roh1[i] / (d1 * d2) - costh * roh2[i] / (d2 * d2))) * p->cm1;
            g[3 + i] = h1;
            g[6 + i] = h2;
            g[i] = -(h1 + h2);
        }
        if (vir6) {
            const int va[6] = {0, 0, 0, 1, 1, 2}, vb[6] = {0, 1, 2, 1, 2, 2};
            for (int q = 0; q < 6; ++q)
                vir6[q] += -w[va[q]] * g[vb[q]] - w[3 + va[q]] * g[3 + vb[q]] - w[6 + va[q]] * g[6 + vb[q]];
        }
    }
    return e;
}

/* ------------------------------------------------------------------------------------------ */
/* Further MB-nrg holders as data (see oracle.h, orc_xholder).                                 */

/* Variable forms 0-2 as nrg_var, and form 3: variable::v_coul, exp(-k r) / r
 * (tools/variable.cpp:109-129). */
static double xvar(int form, double k, double d, const double* p1, const double* p2, double* gvec) {
    if (form != 3) return nrg_var(form, k, d, p1, p2, gvec);
    gvec[0] = p1[0] - p2[0];
    gvec[1] = p1[1] - p2[1];
    gvec[2] = p1[2] - p2[2];
    const double rsq = gvec[0] * gvec[0] + gvec[1] * gvec[1] + gvec[2] * gvec[2];
    const double r = sqrt(rsq);
    const double exp1 = exp(k * (-r));
    const double rinv = 1.0 / r;
    const double val = exp1 * rinv;
    const double gg = -(k + rinv) * val * rinv;
    gvec[0] *= gg;
    gvec[1] *= gg;
    gvec[2] *= gg;
    return val;
}

/* The generated holders' eval with gradients, e.g. mbnrg_2b_A1_B1C2X2_deg6_v1.cpp:3789-3960,
 * mbnrg_3b_A1_B1C2X2_B1C2X2_deg4_v1.cpp:12955-13230, x2b_A1B2_A1B2_v1x.cpp:18485-18700,
 * mbnrg_3b_A1B2_A1B2_A1B2_deg4_v1.cpp:1223-1460, mbnrg_4b_A1B2_A1B2_A1B2_A1B2_deg4_v1.cpp:657-1100,
 * x2b-A1B2Z2_C-v1x.cpp:2546-2770 (alkali ion, water), x3b-h2o-ion-v1x-deg4-filtered.cpp:206-482:
 * first-atom separations of every monomer pair and the skip rule; the water lone-pair sites
 * (tools/water_monomer_lp.cpp:3-21); variables; P and its gradient; the gradient scaled by the
 * switch combination sw; variable gradients to the sites; lone-pair gradients back to O, H, H
 * (water_monomer_lp.cpp:25-48); the switch derivative along every pair; E = sw P; the virial
 * -sum_atoms x (x) g over the real atoms. */
double orc_xholder_eval(const orc_xholder* h, const orc_poly* p, const double* a, const double* const* mons,
                        double* const* grads, double* vir, int* evaluated) {
    double xs[3 * 24], xg[3 * 24];
    int off[5] = {0, 0, 0, 0, 0};
    for (int m = 0; m < h->nmon; ++m) off[m + 1] = off[m] + h->nat[m];
    for (int m = 0; m < h->nmon; ++m) memcpy(xs + 3 * off[m], mons[m], sizeof(double) * 3 * h->nat[m]);
    *evaluated = 0;
    double dd[6][3], r[6];
    int beyond = 0;
    for (int q = 0; q < h->npair; ++q) {
        const double* u = xs + 3 * off[h->pa[q]];
        const double* v = xs + 3 * off[h->pb[q]];
        dd[q][0] = u[0] - v[0];
        dd[q][1] = u[1] - v[1];
        dd[q][2] = u[2] - v[2];
        r[q] = sqrt(dd[q][0] * dd[q][0] + dd[q][1] * dd[q][1] + dd[q][2] * dd[q][2]);
        beyond += r[q] > h->r_out;
    }
    if (beyond >= h->skip_min) return 0.0;
    *evaluated = 1;
    double oh1[2][3], oh2[2][3];
    for (int w = 0; w < h->nvsite; ++w)
        msite_setup_w(xs + 3 * h->vo[w], h->in_plane, h->out_of_plane, oh1[w], oh2[w], xs + 3 * (h->nreal + 2 * w),
                      xs + 3 * (h->nreal + 2 * w + 1));
    double v[96], gvv[96][3], g[96];
    for (int j = 0; j < h->nvars; ++j)
        v[j] = xvar(h->form[j], h->k[j], h->d[j], xs + 3 * h->s1[j], xs + 3 * h->s2[j], gvv[j]);
    const double P = orc_poly_eval(p, a, v, g);
    double s[6], gs[6];
    for (int q = 0; q < h->npair; ++q) s[q] = nrg_switch(h->r_in, h->r_out, r[q], &gs[q]);
    double sw = 0.0;
    for (int t = 0; t < h->nterm; ++t) {
        double prod = h->tcoef[t];
        for (int q = 0; q < h->npair; ++q)
            if (h->tmask[t] >> q & 1) prod *= s[q];
        sw += prod;
    }
    for (int j = 0; j < h->nvars; ++j) g[j] *= sw;
    memset(xg, 0, sizeof(xg));
    for (int j = 0; j < h->nvars; ++j)
        for (int c = 0; c < 3; ++c) {
            const double t = g[j] * gvv[j][c];
            xg[3 * h->s1[j] + c] += t;
            xg[3 * h->s2[j] + c] -= t;
        }
    for (int w = 0; w < h->nvsite; ++w)
        msite_grads_w(oh1[w], oh2[w], h->in_plane, h->out_of_plane, xg + 3 * (h->nreal + 2 * w),
                      xg + 3 * (h->nreal + 2 * w + 1), xg + 3 * h->vo[w]);
    for (int q = 0; q < h->npair; ++q) {
        /* d sw / d s_q: the terms holding pair q, with s_q left out */
        double dsw = 0.0;
        for (int t = 0; t < h->nterm; ++t) {
            if (!(h->tmask[t] >> q & 1)) continue;
            double prod = h->tcoef[t];
            for (int u = 0; u < h->npair; ++u)
                if (u != q && (h->tmask[t] >> u & 1)) prod *= s[u];
            dsw += prod;
        }
        const double gq = gs[q] * dsw * P / r[q];
        for (int c = 0; c < 3; ++c) {
            xg[3 * off[h->pa[q]] + c] += gq * dd[q][c];
            xg[3 * off[h->pb[q]] + c] -= gq * dd[q][c];
        }
    }
    for (int m = 0; m < h->nmon; ++m)
        if (grads && grads[m])
            for (int i = 0; i < 3 * h->nat[m]; ++i) grads[m][i] += xg[3 * off[m] + i];
    if (vir) {
        const int va[6] = {0, 0, 0, 1, 1, 2}, vb[6] = {0, 1, 2, 1, 2, 2}, slot[6] = {0, 1, 2, 4, 5, 8};
        for (int q = 0; q < 6; ++q) {
            double acc = 0.0;
            for (int at = 0; at < h->nreal; ++at) acc += -xs[3 * at + va[q]] * xg[3 * at + vb[q]];
            vir[slot[q]] += acc;
        }
        vir[3] = vir[1];
        vir[6] = vir[2];
        vir[7] = vir[5];
    }
    return P * sw;
}
