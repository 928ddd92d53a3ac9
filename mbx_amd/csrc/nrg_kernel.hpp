// MB-nrg families (CH4-H2O 2B, CH4-CH4 2B, CH4-H2O-H2O 3B): shared pieces.
//
// A holder is described as data (gen/holder_<name>.inc, generated from the holder's text):
// its monomers in argument order, its sites (real atoms in that order, plus the water's two
// lone-pair sites when it builds them), and its variables, each a function of one site pair:
//   form 0  exp(k (d - r))         tools/variable.cpp:49-64   (v_exp0)
//   form 1  exp(k (d - r)) / r     tools/variable.cpp:87-106  (v_coul0)
//   form 2  exp(-k r)              tools/variable.cpp:68-83   (v_exp)
// The polynomial comes in pieces, each its own translation unit and kernel (one lane per
// n-mer, partial value and gradient per piece); the chain-rule stage sums the pieces.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gen/poly_support.hpp"

namespace mbxg {

struct NrgVar {
    int form;
    double k, d;
    int s1, s2;
};

typedef double (*nrg_eg_fn)(mbxg_cptr, const double*, double*, double);
typedef double (*nrg_e_fn)(mbxg_cptr, const double*, double);

// One piece of a polynomial: gs/ps are the piece's own slices ([NV][ld] and [ld]).
template <int NV, nrg_eg_fn FN>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    k_poly_piece_eg(const double* __restrict__ xs, const int* __restrict__ n_dev, int cap, int ld,
                    uint64_t coef_addr, double* __restrict__ gs, double* __restrict__ ps) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    const int n = min(*n_dev, cap);
    if (t >= n) return;
    const mbxg_cptr C = (mbxg_cptr)coef_addr;
    double x[NV], g[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        x[j] = xs[(size_t)j * ld + t];
        g[j] = 0.0;
    }
    const double E = FN(C, x, g, 0.0);
#pragma unroll
    for (int j = 0; j < NV; ++j) gs[(size_t)j * ld + t] = g[j];
    ps[t] = E;
}

template <int NV, nrg_e_fn FN>
__global__ void __launch_bounds__(64) k_poly_piece_e(const double* __restrict__ xs, const int* __restrict__ n_dev,
                                                     int cap, int ld, uint64_t coef_addr, double* __restrict__ ps) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    const int n = min(*n_dev, cap);
    if (t >= n) return;
    const mbxg_cptr C = (mbxg_cptr)coef_addr;
    double x[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) x[j] = xs[(size_t)j * ld + t];
    ps[t] = FN(C, x, 0.0);
}

template <int NV, nrg_eg_fn FEG, nrg_e_fn FE>
hipError_t launch_poly_piece(const double* xs, const int* n_dev, int cap, int ld, const double* coef, int do_grads,
                             double* gs, double* ps, hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    const int nb = (cap + 63) / 64;
    const uint64_t ca = (uint64_t)(uintptr_t)coef;
    if (do_grads)
        hipLaunchKernelGGL((k_poly_piece_eg<NV, FEG>), dim3(nb), dim3(64), 0, st, xs, n_dev, cap, ld, ca, gs, ps);
    else
        hipLaunchKernelGGL((k_poly_piece_e<NV, FE>), dim3(nb), dim3(64), 0, st, xs, n_dev, cap, ld, ca, ps);
    return hipGetLastError();
}

// A piece of a further family's polynomial (gen/xpoly_<name>_p<k>.hip): always value and
// gradient.  Up to 48 variables the kernel runs two wavefronts per SIMD; above that (the water
// 4B polynomial, 66 variables: x and g alone take 264 registers) one.
template <int NV, nrg_eg_fn FN>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
    k_xpoly_piece_wide(const double* __restrict__ xs, const int* __restrict__ n_dev, int cap, int ld,
                       uint64_t coef_addr, double* __restrict__ gs, double* __restrict__ ps) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    const int n = min(*n_dev, cap);
    if (t >= n) return;
    const mbxg_cptr C = (mbxg_cptr)coef_addr;
    double x[NV], g[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        x[j] = xs[(size_t)j * ld + t];
        g[j] = 0.0;
    }
    const double E = FN(C, x, g, 0.0);
#pragma unroll
    for (int j = 0; j < NV; ++j) gs[(size_t)j * ld + t] = g[j];
    ps[t] = E;
}

template <int NV, nrg_eg_fn FN>
hipError_t launch_xpoly_piece(const double* xs, const int* n_dev, int cap, int ld, const double* coef, double* gs,
                              double* ps, hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    const int nb = (cap + 63) / 64;
    const uint64_t ca = (uint64_t)(uintptr_t)coef;
    if constexpr (NV > 48)
        hipLaunchKernelGGL((k_xpoly_piece_wide<NV, FN>), dim3(nb), dim3(64), 0, st, xs, n_dev, cap, ld, ca, gs, ps);
    else
        hipLaunchKernelGGL((k_poly_piece_eg<NV, FN>), dim3(nb), dim3(64), 0, st, xs, n_dev, cap, ld, ca, gs, ps);
    return hipGetLastError();
}

}  // namespace mbxg
