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
#include "mbpol2b_dev.hpp"

namespace mbxg {

// Stage 1: variables and switch.
__global__ void __launch_bounds__(64) k_vars_mbpol2b(const int* __restrict__ list, const int* __restrict__ n_dev,
                                                     int cap, NmerCtx N, double* __restrict__ xs,
                                                     double* __restrict__ sw, int ld) {
    vars2_lane(blockIdx.x * 64 + threadIdx.x, list, n_dev, cap, N, xs, sw, ld);
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
