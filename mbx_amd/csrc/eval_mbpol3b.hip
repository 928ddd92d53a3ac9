// MB-pol 3B energy + gradient for gfx950: one lane per evaluated trimer, one wavefront per block.
//
// Restates x2o::x3b_v2x::eval (potential/3b/x3b-v2x.cpp:1403-1806) for trimers that survive
// the 4.5 A rule (the list builder compacts them), in three stages:
//   1. k_vars_mbpol3b: monomers 2 and 3 imaged next to monomer 1 (GetCloseTrimerImage,
//      sys_tools.cpp:438-442); 36 variables exp(-k (d - d0)) (x3b-v2x.cpp:1548-1584) and the
//      switch product s = sab*sac + sab*sbc + sac*sbc (x3b-v2x.cpp:1614-1620);
//   2. the polynomial and its gradient (poly_mbpol3b.hip);
//   3. k_grad_mbpol3b: E = s*P, chain rule to atoms, switch derivatives and the per-trimer
//      virial (x3b-v2x.cpp:1622-1726); gradients scattered with hardware FP64 atomics, energy
//      and virial reduced per wave.
// Stage buffers are variable-major (entry (j, t) at j*ld + t).  Compiled with -ffp-contract=off.
#include "mbpol2b_dev.hpp"
#include "mbpol3b_chain.hpp"


namespace mbxg {

// Stage 1: variables and switch product.
__global__ void __launch_bounds__(64) k_vars_mbpol3b(const int* __restrict__ list, const int* __restrict__ n_dev,
                                                     int cap, NmerCtx N, double* __restrict__ xs,
                                                     double* __restrict__ sw, int ld) {
    vars3_lane(blockIdx.x * 64 + threadIdx.x, list, n_dev, cap, N, xs, sw, ld);
}

// Stage 1 of both MB-pol orders in one launch: blocks [0, nb3) the trimers, the rest the dimers
// (the 2B blocks take the slots the last 3B blocks leave, as in the paired polynomial launch).
__global__ void __launch_bounds__(64) k_vars_mbpol_pair(const int* __restrict__ list3, const int* __restrict__ n3,
                                                        int cap3, double* __restrict__ xs3, double* __restrict__ sw3,
                                                        int ld3, int nb3, const int* __restrict__ list2,
                                                        const int* __restrict__ n2, int cap2,
                                                        double* __restrict__ xs2, double* __restrict__ sw2, int ld2,
                                                        NmerCtx N) {
    if ((int)blockIdx.x < nb3)
        vars3_lane(blockIdx.x * 64 + threadIdx.x, list3, n3, cap3, N, xs3, sw3, ld3);
    else
        vars2_lane((blockIdx.x - nb3) * 64 + threadIdx.x, list2, n2, cap2, N, xs2, sw2, ld2);
}

// Stage 3: energy, and with GRAD the chain rule, virial and gradient scatter.  Blocks of 128
// trimers (two wavefronts); one partial record per wavefront.
constexpr int kGradBlock3 = 128;

template <bool GRAD>
__global__ void __launch_bounds__(kGradBlock3) __attribute__((amdgpu_waves_per_eu(3))) k_grad_mbpol3b(const int* __restrict__ list,
                                                              const int* __restrict__ n_dev, int cap, NmerCtx N,
                                                              const double* __restrict__ gs,
                                                              const double* __restrict__ ps,
                                                              const double* __restrict__ sw, int ld,
                                                              double* __restrict__ grad, double* __restrict__ e_out,
                                                              double* __restrict__ vir_out,
                                                              double* __restrict__ partials) {
    __shared__ GradTable<256, 9> T;
    const int t = blockIdx.x * kGradBlock3 + threadIdx.x;
    chain3<GRAD, false, 256>(t, blockIdx.x * (kGradBlock3 / 64) + threadIdx.x / 64, min(*n_dev, cap), list, N,
                             nullptr, gs, ps, sw, ld, grad, e_out, vir_out, partials, T);
}

size_t workspace_mbpol3b(int cap) {
    const size_t ld = ((size_t)cap + 63) / 64 * 64;
    return (2 * 36 + 2) * ld;
}

// The stages of launch_mbpol3b on its own, for callers that run the polynomial stage
// elsewhere (launch_poly_mbpol_pair): variables and switch into ws, and the chain rule from ws.
hipError_t launch_mbpol3b_vars(const int* list, const int* n_dev, int cap, const NmerCtx& N, double* ws,
                               hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    const int nb = (cap + 63) / 64;
    const int ld = nb * 64;
    double* xs = ws;
    double* sw = xs + (size_t)(2 * 36 + 1) * ld;
    hipLaunchKernelGGL(k_vars_mbpol3b, dim3(nb), dim3(64), 0, st, list, n_dev, cap, N, xs, sw, ld);
    return hipGetLastError();
}

hipError_t launch_mbpol_vars_pair(const int* list3, const int* n3_dev, int cap3, double* ws3, const int* list2,
                                  const int* n2_dev, int cap2, double* ws2, const NmerCtx& N, hipStream_t st) {
    if (cap3 <= 0 || cap2 <= 0) return hipErrorInvalidValue;
    const int nb3 = (cap3 + 63) / 64, nb2 = (cap2 + 63) / 64;
    const int ld3 = nb3 * 64, ld2 = nb2 * 64;
    // workspace layouts of launch_mbpol3b_vars / launch_mbpol2b_vars: xs first, sw after xs, gs and ps
    hipLaunchKernelGGL(k_vars_mbpol_pair, dim3(nb3 + nb2), dim3(64), 0, st, list3, n3_dev, cap3, ws3,
                       ws3 + (size_t)(2 * 36 + 1) * ld3, ld3, nb3, list2, n2_dev, cap2, ws2,
                       ws2 + (size_t)(2 * 31 + 1) * ld2, ld2, N);
    return hipGetLastError();
}

hipError_t launch_mbpol3b_chain(const int* list, const int* n_dev, int cap, const NmerCtx& N, int do_grads,
                                double* grad, double* e_out, double* vir_out, double* partials, double* ws,
                                hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    const int ld = (cap + 63) / 64 * 64;
    double* xs = ws;
    double* gs = xs + (size_t)36 * ld;
    double* ps = gs + (size_t)36 * ld;
    double* sw = ps + ld;
    const int ngb = (cap + kGradBlock3 - 1) / kGradBlock3;
    if (do_grads)
        hipLaunchKernelGGL(k_grad_mbpol3b<true>, dim3(ngb), dim3(kGradBlock3), 0, st, list, n_dev, cap, N, gs, ps,
                           sw, ld, grad, e_out, vir_out, partials);
    else
        hipLaunchKernelGGL(k_grad_mbpol3b<false>, dim3(ngb), dim3(kGradBlock3), 0, st, list, n_dev, cap, N, gs,
                           ps, sw, ld, grad, e_out, vir_out, partials);
    return hipGetLastError();
}

hipError_t launch_mbpol3b(const int* list, const int* n_dev, int cap, const NmerCtx& N, const double* coef,
                          int do_grads, double* grad, double* e_out, double* vir_out, double* partials, double* ws,
                          int split, hipStream_t st, hipEvent_t* ev) {
    if (cap <= 0) return hipSuccess;
    const int ld = (cap + 63) / 64 * 64;
    double* gs = ws + (size_t)36 * ld;
    double* ps = gs + (size_t)36 * ld;
    if (ev) (void)hipEventRecord(ev[0], st);
    hipError_t e = launch_mbpol3b_vars(list, n_dev, cap, N, ws, st);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[1], st);
    e = launch_poly_mbpol3b(ws, n_dev, cap, ld, coef, do_grads, gs, ps, split, st);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[2], st);
    e = launch_mbpol3b_chain(list, n_dev, cap, N, do_grads, grad, e_out, vir_out, partials, ws, st);
    if (ev) (void)hipEventRecord(ev[3], st);
    return e;
}

}  // namespace mbxg
