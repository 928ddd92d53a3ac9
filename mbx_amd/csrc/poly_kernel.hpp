// The polynomial stage: one lane per n-mer, variables in, gradient and value out.
//
// It runs as its own kernel so the register allocator sees nothing but the generated
// straight-line polynomial (x and g arrays plus a few trie temporaries); the variable and
// chain-rule stages, whose exp/sqrt code the instruction selector freely interleaves, would
// otherwise push the polynomial past the register file.  The one exception,
// k_poly_eg_pair_chain3 (poly_mbpol_pair.hip), runs the 3B chain rule strictly after a block's
// polynomial, behind a fence, so the two phases' registers are allocated one after the other.
//
// Layout: xs and gs are variable-major ("SoA"), entry (j, t) at j*ld + t, so a wavefront's 64
// lanes read and write 64 consecutive doubles per variable.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gen/poly_support.hpp"

namespace mbxg {

typedef double (*poly_eg_part_fn)(int, int, mbxg_cptr, const double*, double*);
typedef double (*poly_e_part_fn)(int, int, mbxg_cptr, const double*);

// The generated polynomial comes in Q pieces (linear: value and gradient are sums over the
// pieces).  A block holds 64 n-mers and P = blockDim.x / 64 wavefronts (P divides Q); wavefront
// w evaluates pieces w, w + P, ... for the block's 64 n-mers, and the P partial results are
// summed through LDS in a fixed order by wavefront 0, kRedChunk gradient components at a time.
// Splitting one n-mer's work over P wavefronts multiplies the number of wavefronts by P at 1/P
// of the length each, so the last round of wavefronts on the machine is short (less idle
// tail).  Dynamic LDS: poly_lds_bytes(P).
constexpr int kRedChunk = 12;
constexpr size_t poly_lds_bytes(int P) { return (size_t)(P - 1) * (kRedChunk + 1) * 64 * sizeof(double); }

// One block's work: 64 n-mers of group `grp`, P = blockDim.x / 64 wavefronts.
template <int NV, int Q, poly_eg_part_fn PART>
__device__ __forceinline__ void poly_eg_group(int grp, const double* __restrict__ xs, const int* __restrict__ n_dev,
                                              int cap, int ld, uint64_t coef_addr, double* __restrict__ gs,
                                              double* __restrict__ ps, double* red) {
    const int P = blockDim.x >> 6;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t = grp * 64 + lane;
    const int n = min(*n_dev, cap);
    if (grp * 64 >= n) return;  // the whole block is past the end (uniform per block)
    const bool active = t < n;
    const mbxg_cptr C = (mbxg_cptr)coef_addr;
    double x[NV], g[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        x[j] = active ? xs[(size_t)j * ld + t] : 0.0;
        g[j] = 0.0;
    }
    double E = PART(wv, P, C, x, g);
    if (P > 1) {
#pragma unroll
        for (int c0 = 0; c0 < NV; c0 += kRedChunk) {
            if (wv > 0) {
                double* r = red + (size_t)(wv - 1) * (kRedChunk + 1) * 64 + lane;
#pragma unroll
                for (int j = c0; j < c0 + kRedChunk && j < NV; ++j) r[(j - c0) * 64] = g[j];
                if (c0 == 0) r[kRedChunk * 64] = E;
            }
            __syncthreads();
            if (wv == 0) {
                for (int v = 1; v < P; ++v) {
                    const double* r = red + (size_t)(v - 1) * (kRedChunk + 1) * 64 + lane;
#pragma unroll
                    for (int j = c0; j < c0 + kRedChunk && j < NV; ++j) g[j] += r[(j - c0) * 64];
                    if (c0 == 0) E += r[kRedChunk * 64];
                }
            }
            __syncthreads();
        }
    }
    if (wv == 0 && active) {
#pragma unroll
        for (int j = 0; j < NV; ++j) gs[(size_t)j * ld + t] = g[j];
        ps[t] = E;
    }
}

// Energy and gradient, WPE waves per SIMD: 3 for MB-pol (at most 168 VGPRs), 2 for the
// 55-variable CH4-H2O-H2O polynomial (x and g alone take 220 VGPRs).
template <int NV, int Q, poly_eg_part_fn PART, int WPE>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
    k_poly_eg(const double* __restrict__ xs, const int* __restrict__ n_dev, int cap, int ld, uint64_t coef_addr,
              double* __restrict__ gs, double* __restrict__ ps) {
    extern __shared__ double red[];  // [P-1][kRedChunk+1][64]
    poly_eg_group<NV, Q, PART>(blockIdx.x, xs, n_dev, cap, ld, coef_addr, gs, ps, red);
}

// Two polynomials in one launch: blocks [0, nb_a) evaluate polynomial A (groups of 64 n-mers in
// order), the rest polynomial B.  Blocks are dispatched in index order, so B's blocks take the
// slots that A's last blocks leave idle.  Both use blockDim.x / 64 wavefronts per group.
struct PolyJob {
    const double* xs;
    const int* n_dev;
    int cap, ld;
    uint64_t coef_addr;
    double* gs;
    double* ps;
};

template <int NVA, int QA, poly_eg_part_fn PA, int NVB, int QB, poly_eg_part_fn PB>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3, 3)))
    k_poly_eg_pair(PolyJob A, PolyJob B, int nb_a) {
    extern __shared__ double red[];
    if ((int)blockIdx.x < nb_a)
        poly_eg_group<NVA, QA, PA>(blockIdx.x, A.xs, A.n_dev, A.cap, A.ld, A.coef_addr, A.gs, A.ps, red);
    else
        poly_eg_group<NVB, QB, PB>(blockIdx.x - nb_a, B.xs, B.n_dev, B.cap, B.ld, B.coef_addr, B.gs, B.ps, red);
}

// Energy only.
template <int NV, int Q, poly_e_part_fn PART>
__global__ void __launch_bounds__(512) k_poly_e(const double* __restrict__ xs, const int* __restrict__ n_dev, int cap,
                                                int ld, uint64_t coef_addr, double* __restrict__ ps) {
    extern __shared__ double red[];  // [P-1][64]
    const int P = blockDim.x >> 6;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t = blockIdx.x * 64 + lane;
    const int n = min(*n_dev, cap);
    if (blockIdx.x * 64 >= n) return;
    const bool active = t < n;
    const mbxg_cptr C = (mbxg_cptr)coef_addr;
    double x[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) x[j] = active ? xs[(size_t)j * ld + t] : 0.0;
    double E = PART(wv, P, C, x);
    if (wv > 0) red[(wv - 1) * 64 + lane] = E;
    if (P > 1) __syncthreads();
    if (wv == 0 && active) {
        for (int v = 1; v < P; ++v) E += red[(v - 1) * 64 + lane];
        ps[t] = E;
    }
}

// Launch helper: P wavefronts per 64 n-mers (P in {1, 2, 4, 8}, dividing Q).
template <int NV, int Q, poly_eg_part_fn PEG, poly_e_part_fn PE, int WPE = 3>
hipError_t launch_poly(const double* xs, const int* n_dev, int cap, int ld, const double* coef, int do_grads,
                       double* gs, double* ps, int P, hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    if (P != 1 && P != 2 && P != 4 && P != 8) return hipErrorInvalidValue;
    const int nb = (cap + 63) / 64;
    const uint64_t ca = (uint64_t)(uintptr_t)coef;
    if (do_grads)
        hipLaunchKernelGGL((k_poly_eg<NV, Q, PEG, WPE>), dim3(nb), dim3(64 * P), poly_lds_bytes(P), st, xs, n_dev, cap,
                           ld, ca, gs, ps);
    else
        hipLaunchKernelGGL((k_poly_e<NV, Q, PE>), dim3(nb), dim3(64 * P), (size_t)(P - 1) * 64 * sizeof(double), st,
                           xs, n_dev, cap, ld, ca, ps);
    return hipGetLastError();
}

}  // namespace mbxg
