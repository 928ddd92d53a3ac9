// MB-pol 2B device code shared by the stage kernels (eval_mbpol2b.hip) and the launch that
// computes the 2B and 3B variables together (eval_mbpol3b.hip): the variable table, M-sites,
// switch, dimer loading and the variables of one dimer.
//
// Restates x2o::x2b_v9x::eval (potential/2b/x2b-v9x.cpp:1514-1820): monomer 2 imaged next to
// monomer 1 (GetCloseDimerImage, sys_tools.cpp:322-365), the M-sites (monomer::setup,
// x2b-v9x.cpp:117-135), 31 variables v_exp / v_coul (x2b-v9x.cpp:54-91, 1640-1675) and the
// switch on r_OO (x2b-v9x.cpp:1349-1362).
#pragma once

#include "mbxg_device.hpp"
#include "mbxg_kernels.hpp"
#include "scatter.hpp"

namespace mbxg {

namespace {

// Sites: 0 Oa 1 Ha1 2 Ha2 3 Ob 4 Hb1 5 Hb2 6 Xa1 7 Xa2 8 Xb1 9 Xb2.
// kind: 0 = v_exp intra (d0 = 1), 1 = v_coul inter (d0 = 4), 2 = v_exp inter (d0 = 4).
struct V2 {
    unsigned char kind, k, s1, s2;
};
__constant__ const V2 kVar2[31] = {
    {0, 0, 1, 2}, {0, 0, 4, 5}, {0, 1, 0, 1}, {0, 1, 0, 2}, {0, 1, 3, 4}, {0, 1, 3, 5}, {1, 2, 1, 4}, {1, 2, 1, 5},
    {1, 2, 2, 4}, {1, 2, 2, 5}, {1, 3, 0, 4}, {1, 3, 0, 5}, {1, 3, 3, 1}, {1, 3, 3, 2}, {1, 4, 0, 3}, {2, 5, 6, 4},
    {2, 5, 6, 5}, {2, 5, 7, 4}, {2, 5, 7, 5}, {2, 5, 8, 1}, {2, 5, 8, 2}, {2, 5, 9, 1}, {2, 5, 9, 2}, {2, 6, 0, 8},
    {2, 6, 0, 9}, {2, 6, 3, 6}, {2, 6, 3, 7}, {2, 7, 6, 8}, {2, 7, 6, 9}, {2, 7, 7, 8}, {2, 7, 7, 9}};

__device__ __forceinline__ double kval(int k) {
    switch (k) {
        case 0: return p2::k_HH_intra;
        case 1: return p2::k_OH_intra;
        case 2: return p2::k_HH_coul;
        case 3: return p2::k_OH_coul;
        case 4: return p2::k_OO_coul;
        case 5: return p2::k_XH_main;
        case 6: return p2::k_XO_main;
        default: return p2::k_XX_main;
    }
}

// monomer::setup: M-sites of the water whose O, H1, H2 sit at s[3*o..].
__device__ __forceinline__ void msites(double* s, int o, int x1, int x2, double* oh1, double* oh2) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        oh1[i] = s[3 * (o + 1) + i] - s[3 * o + i];
        oh2[i] = s[3 * (o + 2) + i] - s[3 * o + i];
    }
    const double v[3] = {oh1[1] * oh2[2] - oh1[2] * oh2[1], oh1[2] * oh2[0] - oh1[0] * oh2[2],
                         oh1[0] * oh2[1] - oh1[1] * oh2[0]};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double in_plane = s[3 * o + i] + 0.5 * p2::in_plane_gamma * (oh1[i] + oh2[i]);
        const double out_of_plane = p2::out_of_plane_gamma * v[i];
        s[3 * x1 + i] = in_plane + out_of_plane;
        s[3 * x2 + i] = in_plane - out_of_plane;
    }
}

// monomer::grads: fold the X-site gradients back onto O, H1, H2.
__device__ __forceinline__ void msite_grads(const double* oh1, const double* oh2, const double* g1, const double* g2,
                                            double* grd) {
    const double gm[3] = {g1[0] - g2[0], g1[1] - g2[1], g1[2] - g2[2]};
    const double t1[3] = {oh2[1] * gm[2] - oh2[2] * gm[1], oh2[2] * gm[0] - oh2[0] * gm[2],
                          oh2[0] * gm[1] - oh2[1] * gm[0]};
    const double t2[3] = {oh1[1] * gm[2] - oh1[2] * gm[1], oh1[2] * gm[0] - oh1[0] * gm[2],
                          oh1[0] * gm[1] - oh1[1] * gm[0]};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double gsum = g1[i] + g2[i];
        const double in_plane = 0.5 * p2::in_plane_gamma * gsum;
        const double gh1 = in_plane + p2::out_of_plane_gamma * t1[i];
        const double gh2 = in_plane - p2::out_of_plane_gamma * t2[i];
        grd[i + 0] += gsum - (gh1 + gh2);
        grd[i + 3] += gh1;
        grd[i + 6] += gh2;
    }
}

// f_switch (x2b-v9x.cpp:1349-1362).
__device__ __forceinline__ double switch2(double r, double& g) {
    if (r > p2::r2f) {
        g = 0.0;
        return 0.0;
    } else if (r > p2::r2i) {
        const double t1 = M_PI / (p2::r2f - p2::r2i);
        const double x = (r - p2::r2i) * t1;
        g = -sin(x) * t1 / 2.0;
        return (1.0 + cos(x)) / 2.0;
    }
    g = 0.0;
    return 1.0;
}

// Variable value and its derivative factor gg (d value / d(p1-p2) = gg * (p1-p2)).
__device__ __forceinline__ double var2(const V2& d, const double* s, double& gg, double& d0, double& d1, double& d2) {
    d0 = s[3 * d.s1] - s[3 * d.s2];
    d1 = s[3 * d.s1 + 1] - s[3 * d.s2 + 1];
    d2 = s[3 * d.s1 + 2] - s[3 * d.s2 + 2];
    const double rsq = d0 * d0 + d1 * d1 + d2 * d2;
    const double r = sqrt(rsq);
    const double k = kval(d.k);
    if (d.kind == 1) {
        const double e1 = exp(k * (4.0 - r));
        const double rinv = 1.0 / r;
        const double val = e1 * rinv;
        gg = -(k + rinv) * val * rinv;
        return val;
    }
    const double e1 = exp(k * ((d.kind == 0 ? 1.0 : 4.0) - r));
    gg = -k * e1 / r;
    return e1;
}

// Sites of a dimer: monomer 2 imaged next to monomer 1, then the four M-sites.
__device__ __forceinline__ void load_dimer(const NmerCtx& N, int m1, int m2, double* s, double* oh1a, double* oh2a,
                                           double* oh1b, double* oh2b) {
    const double* a = N.xyz + 3 * (size_t)N.first_atom[m1];
    const double* b = N.xyz + 3 * (size_t)N.first_atom[m2];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        s[j] = a[j];
        s[9 + j] = b[j];
    }
    if (N.periodic) {
        const Box& B = *N.box;
        double dx, dy, dz;
        image_shift(B, s[0], s[1], s[2], s[9], s[10], s[11], dx, dy, dz);
#pragma unroll
        for (int j = 0; j < 3; ++j) image_apply(B, dx, dy, dz, s[9 + 3 * j], s[10 + 3 * j], s[11 + 3 * j]);
    }
    msites(s, 0, 6, 7, oh1a, oh2a);
    msites(s, 3, 8, 9, oh1b, oh2b);
}

// Stage 1 for the dimer of lane t: its 31 variables into xs (entry (j, t) at j*ld + t) and its
// switch into sw; lanes at or past min(*n_dev, cap) do nothing.
__device__ __forceinline__ void vars2_lane(int t, const int* __restrict__ list, const int* __restrict__ n_dev,
                                           int cap, const NmerCtx& N, double* __restrict__ xs,
                                           double* __restrict__ sw, int ld) {
    if (t >= min(*n_dev, cap)) return;
    double s[30], oh1a[3], oh2a[3], oh1b[3], oh2b[3];
    load_dimer(N, list[2 * t], list[2 * t + 1], s, oh1a, oh2a, oh1b, oh2b);
#pragma unroll
    for (int j = 0; j < 31; ++j) {
        double gg, d0, d1, d2;
        xs[(size_t)j * ld + t] = var2(kVar2[j], s, gg, d0, d1, d2);
    }
    const double dOO[3] = {s[0] - s[9], s[1] - s[10], s[2] - s[11]};
    const double rOO = sqrt(dOO[0] * dOO[0] + dOO[1] * dOO[1] + dOO[2] * dOO[2]);
    double gsw;
    sw[t] = switch2(rOO, gsw);
}

}  // namespace

}  // namespace mbxg
