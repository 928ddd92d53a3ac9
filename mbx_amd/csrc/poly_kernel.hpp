// The polynomial stage: one lane per n-mer, variables in, gradient and value out.
//
// It runs as its own kernel so the register allocator sees nothing but the generated
// straight-line polynomial (x and g arrays plus a few trie temporaries); the variable and
// chain-rule stages, whose exp/sqrt code the instruction selector freely interleaves, would
// otherwise push the polynomial past the register file.
//
// Layout: xs and gs are variable-major ("SoA"), entry (j, t) at j*ld + t, so a wavefront's 64
// lanes read and write 64 consecutive doubles per variable.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gen/poly_support.hpp"

namespace mbxg {

typedef double (*poly_eg_fn)(mbxg_cptr, const double*, double*);
typedef double (*poly_e_fn)(mbxg_cptr, const double*);

// Energy and gradient.  Two waves per SIMD: the polynomial needs ~220 VGPRs.
template <int NV, poly_eg_fn FN>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    k_poly_eg(const double* __restrict__ xs, const int* __restrict__ n_dev, int cap, int ld, uint64_t coef_addr,
              double* __restrict__ gs, double* __restrict__ ps) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    const int n = min(*n_dev, cap);
    if (t >= n) return;
    const mbxg_cptr C = (mbxg_cptr)coef_addr;
    double x[NV], g[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) x[j] = xs[(size_t)j * ld + t];
    const double P = FN(C, x, g);
#pragma unroll
    for (int j = 0; j < NV; ++j) gs[(size_t)j * ld + t] = g[j];
    ps[t] = P;
}

// Energy only.
template <int NV, poly_e_fn FN>
__global__ void __launch_bounds__(64) k_poly_e(const double* __restrict__ xs, const int* __restrict__ n_dev, int cap,
                                               int ld, uint64_t coef_addr, double* __restrict__ ps) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    const int n = min(*n_dev, cap);
    if (t >= n) return;
    const mbxg_cptr C = (mbxg_cptr)coef_addr;
    double x[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) x[j] = xs[(size_t)j * ld + t];
    ps[t] = FN(C, x);
}

}  // namespace mbxg
