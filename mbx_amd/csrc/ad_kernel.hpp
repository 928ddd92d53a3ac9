// Differentiation kernels around the generated polynomial (gen/ad_<name>.inc).
//
// k_ad_tan evaluates the tangent of the polynomial's gradient: with variables x and their
// tangents xd (the derivative of x along a direction of the atom coordinates) it returns H xd,
// one lane per n-mer (P and grad P come from the plain E+G kernel, dP = grad P . xd).  The
// schedule's parts run as separate wavefronts (grid.y = part); each writes its own slice and the
// consumer sums the slices in part order.
//
// k_ad_basis accumulates, for every coefficient k, sum_n [alpha_n M_k(x_n) + sw_n dM_k(x_n)]
// over the n-mers of a block (M_k = the basis function multiplying coefficient k, dM_k its
// tangent), coefficients in slices (grid.y = slice).  Each block writes one row of per-block
// partial sums; k_ad_colsum adds the rows in a fixed order, so the result is deterministic.
//
// Buffers are variable-major: entry (j, t) at j*ld + t.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gen/poly_support.hpp"

namespace mbxg {

// Pieces of the gradient-tangent polynomial (one wavefront each per 64 n-mers).
constexpr int kAdParts = 16;

// Column sums of the block's LDS tile red[64][65]: lane j adds row j (its coefficient's values
// from the 64 lanes) in lane order.
__device__ __forceinline__ void ad_basis_flush(double* red, double* out, int cnt, int lane) {
    __syncthreads();
    if (lane < cnt) {
        double s = 0.0;
        for (int l = 0; l < 64; ++l) s += red[lane * 65 + l];
        out[lane] = s;
    }
    __syncthreads();
}

// An empty asm that claims to modify every variable: values computed from the variables before it
// cannot be reused after it (no common subexpressions across it), and nothing moves across it.
template <int NV>
__device__ __forceinline__ void ad_opaque(double* x, double* xd) {
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < NV; ++i) asm volatile("" : "+v"(x[i]), "+v"(xd[i]));
}

typedef void (*ad_tan_fn)(int, mbxg_cptr, const double*, const double*, double*);
typedef void (*ad_basis_fn)(int, double*, double*, double, double, double*, double*, int);

template <int NV, int PARTS, ad_tan_fn FN>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    k_ad_tan(const double* __restrict__ xs, const double* __restrict__ xds, int n, int ld, uint64_t coef_addr,
             double* __restrict__ gds) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    const int part = blockIdx.y;
    if (t >= n) return;
    const mbxg_cptr C = (mbxg_cptr)coef_addr;
    double x[NV], xd[NV], gd[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        x[j] = xs[(size_t)j * ld + t];
        xd[j] = xds[(size_t)j * ld + t];
        gd[j] = 0.0;
    }
    FN(part, C, x, xd, gd);
    const size_t base = (size_t)part * NV * ld;
#pragma unroll
    for (int j = 0; j < NV; ++j) gds[base + (size_t)j * ld + t] = gd[j];
}

// alpha_n = g_energy_n * sw_n + dsw_n (the weight of M_k), sw_n (the weight of dM_k).
template <int NV, int NCOEF, ad_basis_fn FN>
__global__ void __launch_bounds__(64) k_ad_basis(const double* __restrict__ xs, const double* __restrict__ xds,
                                                 const double* __restrict__ sw, const double* __restrict__ swd,
                                                 const double* __restrict__ g_energy, int n, int ld,
                                                 double* __restrict__ partial) {
    __shared__ double red[64 * 65];
    const int lane = threadIdx.x;
    const int t = blockIdx.x * 64 + lane;
    const bool active = t < n;
    double x[NV], xd[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        x[j] = active ? xs[(size_t)j * ld + t] : 0.0;
        xd[j] = active ? xds[(size_t)j * ld + t] : 0.0;
    }
    const double s = active ? sw[t] : 0.0;
    const double alpha = active ? g_energy[t] * s + swd[t] : 0.0;
    FN(blockIdx.y, x, xd, alpha, s, red, partial + (size_t)blockIdx.x * NCOEF, lane);
}


}  // namespace mbxg
