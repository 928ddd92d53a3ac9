// MB-pol 2B energy + gradient for gfx950: one lane per evaluated dimer, one wavefront per block.
//
// Restates x2o::x2b_v9x::eval (potential/2b/x2b-v9x.cpp:1514-1820) for dimers with
// r_OO <= 6.5 A (the list builder compacts them), in three stages:
//   1. k_vars_mbpol2b: monomer 2 imaged next to monomer 1 (GetCloseDimerImage,
//      sys_tools.cpp:322-365), M-sites X1/X2 of each water (monomer::setup, x2b-v9x.cpp:117-135),
//      31 variables v_exp / v_coul (x2b-v9x.cpp:54-91, 1640-1675) and the switch on r_OO;
//   2. the polynomial and its gradient (poly_mbpol2b.hip);
//   3. k_grad_mbpol2b: E = sw*P, variable and M-site chain rule (x2b-v9x.cpp:95-103, 139-160),
//      switch derivative, per-dimer virial (x2b-v9x.cpp:1718-1800); gradients scattered with
//      hardware FP64 atomics, energy and virial reduced per wave.
// Stage buffers are variable-major (entry (j, t) at j*ld + t).  Compiled with -ffp-contract=off.
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

}  // namespace

// Stage 1: variables and switch.
__global__ void __launch_bounds__(64) k_vars_mbpol2b(const int* __restrict__ list, const int* __restrict__ n_dev,
                                                     int cap, NmerCtx N, double* __restrict__ xs,
                                                     double* __restrict__ sw, int ld) {
    const int t = blockIdx.x * 64 + threadIdx.x;
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

// Stage 3: energy, and with GRAD the chain rule, virial and gradient scatter.  Blocks of 128
// dimers (two wavefronts); one partial record per wavefront.
constexpr int kGradBlock2 = 128;

template <bool GRAD>
__global__ void __launch_bounds__(kGradBlock2) k_grad_mbpol2b(const int* __restrict__ list,
                                                              const int* __restrict__ n_dev, int cap, NmerCtx N,
                                                              const double* __restrict__ gs,
                                                              const double* __restrict__ ps,
                                                              const double* __restrict__ sws, int ld,
                                                              double* __restrict__ grad, double* __restrict__ e_out,
                                                              double* __restrict__ vir_out,
                                                              double* __restrict__ partials) {
    __shared__ GradTable<256, 9> T;
    const int t = blockIdx.x * kGradBlock2 + threadIdx.x;
    const int n = min(*n_dev, cap);
    if (GRAD) table_init(T);
    __syncthreads();
    double e_sum = 0.0;
    double vir[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (t < n) {
        const double sw = sws[t];
        const double P = ps[t];
        if (GRAD) {
            const int m1 = list[2 * t], m2 = list[2 * t + 1];
            double s[30], oh1a[3], oh2a[3], oh1b[3], oh2b[3];
            load_dimer(N, m1, m2, s, oh1a, oh2a, oh1b, oh2b);
            double xg[30];
#pragma unroll
            for (int j = 0; j < 30; ++j) xg[j] = 0.0;
#pragma unroll
            for (int j = 0; j < 31; ++j) {
                double gg, d0, d1, d2;
                var2(kVar2[j], s, gg, d0, d1, d2);
                const int s1 = kVar2[j].s1, s2 = kVar2[j].s2;
                const double gj = gs[(size_t)j * ld + t];
                const double q0 = gj * (d0 * gg), q1 = gj * (d1 * gg), q2 = gj * (d2 * gg);
                xg[3 * s1] += q0;
                xg[3 * s1 + 1] += q1;
                xg[3 * s1 + 2] += q2;
                xg[3 * s2] -= q0;
                xg[3 * s2 + 1] -= q1;
                xg[3 * s2 + 2] -= q2;
            }
            msite_grads(oh1a, oh2a, xg + 18, xg + 21, xg + 0);
            msite_grads(oh1b, oh2b, xg + 24, xg + 27, xg + 9);
            const double dOO[3] = {s[0] - s[9], s[1] - s[10], s[2] - s[11]};
            const double rOO = sqrt(dOO[0] * dOO[0] + dOO[1] * dOO[1] + dOO[2] * dOO[2]);
            double gsw;
            switch2(rOO, gsw);
            double gd[18];
#pragma unroll
            for (int j = 0; j < 18; ++j) gd[j] = sw * xg[j];
            const double gsc = gsw * (P / rOO);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double dd = gsc * dOO[c];
                gd[c] += dd;
                gd[9 + c] -= dd;
            }
            const int va[6] = {0, 0, 0, 1, 1, 2}, vb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                double acc = 0.0;
#pragma unroll
                for (int at = 0; at < 6; ++at) acc -= s[3 * at + va[q]] * gd[3 * at + vb[q]];
                vir[q] = acc;
            }
            table_add(T, m1, gd, 9, grad, 3 * (size_t)N.first_atom[m1]);
            table_add(T, m2, gd + 9, 9, grad, 3 * (size_t)N.first_atom[m2]);
        }
        const double e = sw * P;
        e_out[t] = e;
        e_sum = e;
        if (vir_out)
#pragma unroll
            for (int q = 0; q < 6; ++q) vir_out[(size_t)q * ld + t] = vir[q];
    }
    if (GRAD) {
        __syncthreads();
        table_flush(T, N.first_atom, N.nat, grad);
    }
    e_sum = wave_sum(e_sum);
    const int rec = blockIdx.x * (kGradBlock2 / 64) + threadIdx.x / 64;
    double* out = partials + (size_t)rec * kPartials;
    const bool lead = (threadIdx.x & 63) == 0;
    if (lead) out[0] = e_sum;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double vq = wave_sum(vir[q]);
        if (lead) out[1 + q] = vq;
    }
}

size_t workspace_mbpol2b(int cap) {
    const size_t ld = ((size_t)cap + 63) / 64 * 64;
    return (2 * 31 + 2) * ld;
}

// The stages of launch_mbpol2b on its own, for callers that run the polynomial stage
// elsewhere (launch_poly_mbpol_pair): variables and switch into ws, and the chain rule from ws.
hipError_t launch_mbpol2b_vars(const int* list, const int* n_dev, int cap, const NmerCtx& N, double* ws,
                               hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    const int nb = (cap + 63) / 64;
    const int ld = nb * 64;
    double* xs = ws;
    double* sw = xs + (size_t)(2 * 31 + 1) * ld;
    hipLaunchKernelGGL(k_vars_mbpol2b, dim3(nb), dim3(64), 0, st, list, n_dev, cap, N, xs, sw, ld);
    return hipGetLastError();
}

hipError_t launch_mbpol2b_chain(const int* list, const int* n_dev, int cap, const NmerCtx& N, int do_grads,
                                double* grad, double* e_out, double* vir_out, double* partials, double* ws,
                                hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    const int ld = (cap + 63) / 64 * 64;
    double* gs = ws + (size_t)31 * ld;
    double* ps = gs + (size_t)31 * ld;
    double* sw = ps + ld;
    const int ngb = (cap + kGradBlock2 - 1) / kGradBlock2;
    if (do_grads)
        hipLaunchKernelGGL(k_grad_mbpol2b<true>, dim3(ngb), dim3(kGradBlock2), 0, st, list, n_dev, cap, N, gs, ps, sw,
                           ld, grad, e_out, vir_out, partials);
    else
        hipLaunchKernelGGL(k_grad_mbpol2b<false>, dim3(ngb), dim3(kGradBlock2), 0, st, list, n_dev, cap, N, gs, ps,
                           sw, ld, grad, e_out, vir_out, partials);
    return hipGetLastError();
}

hipError_t launch_mbpol2b(const int* list, const int* n_dev, int cap, const NmerCtx& N, const double* coef,
                          int do_grads, double* grad, double* e_out, double* vir_out, double* partials, double* ws,
                          int split, hipStream_t st, hipEvent_t* ev) {
    if (cap <= 0) return hipSuccess;
    const int ld = (cap + 63) / 64 * 64;
    double* gs = ws + (size_t)31 * ld;
    double* ps = gs + (size_t)31 * ld;
    if (ev) (void)hipEventRecord(ev[0], st);
    hipError_t e = launch_mbpol2b_vars(list, n_dev, cap, N, ws, st);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[1], st);
    e = launch_poly_mbpol2b(ws, n_dev, cap, ld, coef, do_grads, gs, ps, split, st);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[2], st);
    e = launch_mbpol2b_chain(list, n_dev, cap, N, do_grads, grad, e_out, vir_out, partials, ws, st);
    if (ev) (void)hipEventRecord(ev[3], st);
    return e;
}

}  // namespace mbxg
