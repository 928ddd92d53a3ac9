// 1B terms, one lane per monomer.
//
// MB-pol water: the Partridge-Schwenke monomer energy and gradient, ps::pot_nasa
// (potential/1b/ps.cpp:346-633).  The reference's operation order is kept (the build disables
// contraction): bond vectors and lengths (:393-433), cos of the HOH angle (:437-440), Morse and
// H-H terms (:456-466), the power tables fmat (:475-491), the four coefficient sums
// (gen/ps_1b.inc, :505-526), energy (:533-539), chain rule to H1, H2, O (:542-582) and the
// virial terms -sum_atoms w_a g_b (:584-603).
//
// MB-nrg CH4: x1b_A1B4_v1x::eval (potential/1b/x1b_A1B4_deg5_exp0_v1x.cpp:371-470): ten
// variables exp(k (d - r)) over the C-H and H-H pairs (holder data gen/holder_ch4_1b.inc), the
// degree-5 polynomial poly_1b_A1B4_deg5_exp0_v1x (gen/poly_ch4_1b.inc, value and gradient in one
// reverse-mode pass), the chain rule per variable in the reference's order (variable::grads,
// :56-63) and the virial -sum_atoms x_a g_b (:442-462).
//
// MB-nrg CO2: x1b_A1B2_v1x::eval (potential/1b/x1b_A1B2_deg4_v1x.cpp:160-215): three variables
// exp(-k (r - d)) over C-O1, C-O2 and O1-O2 (v_intra, :41-47; gen/holder_co2_1b.inc), the degree-4
// polynomial poly_1b_A1B2_deg4_v1x (gen/poly_co2_1b.inc), the chain rule g_intra (:50-61) and the
// virial -sum_atoms x_a g_b (:188-203), in the same kernel as CH4.
#include <hip/hip_runtime.h>

#include "gen/coef_ch4_1b.inc"
#include "gen/coef_co2_1b.inc"
#include "gen/holder_ch4_1b.inc"
#include "gen/holder_co2_1b.inc"
#include "gen/poly_ch4_1b.inc"
#include "gen/poly_co2_1b.inc"
#include "gen/ps_1b.inc"
#include "mbxg_kernels.hpp"

namespace mbxg {

namespace {

// e_out[t], and with grads g_out (3 doubles per atom, the water's 9 written) and vir_out[q * ld + t]
// (xx xy xz yy yz zz) for water mons[t] (first atom first_atom[mons[t]], atoms O H1 H2).
__global__ void __launch_bounds__(64) k_1b_ps(const double* __restrict__ xyz, const int* __restrict__ first_atom,
                                              const int* __restrict__ mons, int n, double* __restrict__ e_out,
                                              double* __restrict__ g_out, double* __restrict__ vir_out, int ld) {
    using namespace ps1b;
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= n) return;
    const size_t base = 3 * (size_t)first_atom[mons[t]];
    double w[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) w[i] = xyz[base + i];
    double roh1[3], roh2[3], rhh[3];
#pragma unroll
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
    const double exp1 = exp(-kAlphaoh * (d1 - kRoh));
    const double exp2 = exp(-kAlphaoh * (d2 - kRoh));
    const double Va = kDeoh * (exp1 * (exp1 - 2.0) + exp2 * (exp2 - 2.0));
    const double Vb = kPhh1 * exp(-kPhh2 * d3);
    const double x1 = (d1 - kReoh) / kReoh;
    const double x2 = (d2 - kReoh) / kReoh;
    const double x3 = costh - kCosthe;
    double f0[16], f1[16], f2[16];
    f0[0] = f1[0] = f2[0] = 0.0;
    f0[1] = f1[1] = f2[1] = 1.0;
#pragma unroll
    for (int j = 2; j < 16; ++j) {
        f0[j] = f0[j - 1] * x1;
        f1[j] = f1[j - 1] * x2;
        f2[j] = f2[j - 1] * x3;
    }
    const double r1 = d1 - kReoh, r2 = d2 - kReoh;
    const double efac = exp(-kB1 * (r1 * r1 + r2 * r2));
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    sums(f0, f1, f2, s0, s1, s2, s3);
    const double Vc = 2 * kC5z0 + efac * s0;
    double e = Va + Vb + Vc;
    e += kCorrection;
    e *= kCm1;
    e_out[t] = e;
    if (!g_out) return;
    const double dVa1 = 2.0 * kAlphaoh * kDeoh * exp1 * (1.0 - exp1) / d1;
    const double dVa2 = 2.0 * kAlphaoh * kDeoh * exp2 * (1.0 - exp2) / d2;
    const double dVb = -kPhh2 * Vb / d3;
    const double dVcdr1 = (-2 * kB1 * efac * (d1 - kReoh) * s0 + efac * s1 / kReoh) / d1;
    const double dVcdr2 = (-2 * kB1 * efac * (d2 - kReoh) * s0 + efac * s2 / kReoh) / d2;
    const double dVcdcth = efac * s3;
    double g[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double h1 = (dVa1 * roh1[i] + dVb * rhh[i] + dVcdr1 * roh1[i] +
                           dVcdcth * (roh2[i] / (d1 * d2) - costh * roh1[i] / (d1 * d1))) *
                          kCm1;
        const double h2 = (dVa2 * roh2[i] - dVb * rhh[i] + dVcdr2 * roh2[i] +
                           dVcdcth * (roh1[i] / (d1 * d2) - costh * roh2[i] / (d2 * d2))) *
                          kCm1;
        g[3 + i] = h1;
        g[6 + i] = h2;
        g[i] = -(h1 + h2);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) g_out[base + i] = g[i];
    if (vir_out) {
        const int va[6] = {0, 0, 0, 1, 1, 2}, vb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int q = 0; q < 6; ++q)
            vir_out[(size_t)q * ld + t] =
                -w[va[q]] * g[vb[q]] - w[3 + va[q]] * g[3 + vb[q]] - w[6 + va[q]] * g[6 + vb[q]];
    }
}

// An MB-nrg 1B holder H with polynomial POLY: the same outputs for monomers mons[0..n) of H's kind.
typedef double (*poly1b_fn)(const double* __restrict__, const double*, double*);

template <class H, poly1b_fn POLY>
__global__ void __launch_bounds__(64) k_1b_nrg(const double* __restrict__ xyz, const int* __restrict__ first_atom,
                                               const int* __restrict__ mons, int n, const double* __restrict__ coef,
                                               double* __restrict__ e_out, double* __restrict__ g_out,
                                               double* __restrict__ vir_out, int ld) {
    constexpr int NA = H::kNReal, NV = H::kNVars;
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= n) return;
    const size_t base = 3 * (size_t)first_atom[mons[t]];
    double w[3 * NA];
#pragma unroll
    for (int i = 0; i < 3 * NA; ++i) w[i] = xyz[base + i];
    // variable::v_exp0 (x1b_A1B4_deg5_exp0_v1x.cpp:67-84): v = exp(k (d - r)), and the vector
    // (p1 - p2) * (-k v / r) that the chain rule scales
    double v[NV], dv[NV][3];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const NrgVar V = H::kVars[j];
        const double* p1 = w + 3 * V.s1;
        const double* p2 = w + 3 * V.s2;
        double d[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double e1 = exp(V.k * (V.d - r));
        const double gg = -V.k * e1 / r;
        v[j] = e1;
#pragma unroll
        for (int i = 0; i < 3; ++i) dv[j][i] = d[i] * gg;
    }
    double gp[NV];
    const double E = POLY(coef, v, gp);
    e_out[t] = E;
    if (!g_out) return;
    double g[3 * NA];
#pragma unroll
    for (int i = 0; i < 3 * NA; ++i) g[i] = 0.0;
#pragma unroll
    for (int j = 0; j < NV; ++j) {  // variable::grads (:56-63), variables in order
        const NrgVar V = H::kVars[j];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double dd = gp[j] * dv[j][i];
            g[3 * V.s1 + i] += dd;
            g[3 * V.s2 + i] -= dd;
        }
    }
#pragma unroll
    for (int i = 0; i < 3 * NA; ++i) g_out[base + i] = g[i];
    if (vir_out) {
        const int va[6] = {0, 0, 0, 1, 1, 2}, vb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            double s = -w[va[q]] * g[vb[q]];
#pragma unroll
            for (int a = 1; a < NA; ++a) s = s - w[3 * a + va[q]] * g[3 * a + vb[q]];
            vir_out[(size_t)q * ld + t] = s;
        }
    }
}

}  // namespace

hipError_t launch_1b_ps(const double* xyz, const int* first_atom, const int* mons, int n, double* e_out,
                        double* g_out, double* vir_out, int ld, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_1b_ps, dim3((n + 63) / 64), dim3(64), 0, st, xyz, first_atom, mons, n, e_out, g_out,
                       vir_out, ld);
    return hipGetLastError();
}

hipError_t launch_1b_ch4(const double* xyz, const int* first_atom, const int* mons, int n, const double* coef,
                         double* e_out, double* g_out, double* vir_out, int ld, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_1b_nrg<Holder_ch4_1b, poly_ch4_1b>), dim3((n + 63) / 64), dim3(64), 0, st, xyz,
                       first_atom, mons, n, coef, e_out, g_out, vir_out, ld);
    return hipGetLastError();
}

hipError_t launch_1b_co2(const double* xyz, const int* first_atom, const int* mons, int n, const double* coef,
                         double* e_out, double* g_out, double* vir_out, int ld, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_1b_nrg<Holder_co2_1b, poly_co2_1b>), dim3((n + 63) / 64), dim3(64), 0, st, xyz,
                       first_atom, mons, n, coef, e_out, g_out, vir_out, ld);
    return hipGetLastError();
}

const double* ch4_1b_fit(int* n) {
    *n = kNCoef_ch4_1b;
    return kFit_ch4_1b;
}

const double* co2_1b_fit(int* n) {
    *n = kNCoef_co2_1b;
    return kFit_co2_1b;
}

// One workgroup: out[0] = sum of e[0..n), out[1 + q] = sum of vir[q * ld + 0..n) (zeros when vir is
// null); each thread sums a fixed stride, then a fixed tree (deterministic totals).
__global__ void __launch_bounds__(256) k_reduce_1b(const double* __restrict__ e, const double* __restrict__ vir, int n,
                                                   int ld, double* __restrict__ out) {
    __shared__ double s[kPartials][256];
    const int t = threadIdx.x;
    for (int q = 0; q < kPartials; ++q) {
        double a = 0.0;
        if (q == 0 || vir)
            for (int i = t; i < n; i += 256) a += q == 0 ? e[i] : vir[(size_t)(q - 1) * ld + i];
        s[q][t] = a;
    }
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
            for (int q = 0; q < kPartials; ++q) s[q][t] += s[q][t + h];
        __syncthreads();
    }
    if (t < kPartials) out[t] = s[t][0];
}

}  // namespace mbxg
