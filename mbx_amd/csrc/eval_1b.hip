// MB-pol 1B: the Partridge-Schwenke water monomer energy and gradient, ps::pot_nasa
// (potential/1b/ps.cpp:346-633), one lane per water.  The reference's operation order is kept
// (the build disables contraction): bond vectors and lengths (:393-433), cos of the HOH angle
// (:437-440), Morse and H-H terms (:456-466), the power tables fmat (:475-491), the four
// coefficient sums (gen/ps_1b.inc, :505-526), energy (:533-539), chain rule to H1, H2, O
// (:542-582) and the virial terms -sum_atoms w_a g_b (:584-603).
#include <hip/hip_runtime.h>

#include "gen/ps_1b.inc"
#include "mbxg_kernels.hpp"

namespace mbxg {

namespace {

// e_out[t], and with grads g_out (3 doubles per atom, the water's 9 written) and vir_out[q * n + t]
// (xx xy xz yy yz zz) for water mons[t] (first atom first_atom[mons[t]], atoms O H1 H2).
__global__ void __launch_bounds__(64) k_1b_ps(const double* __restrict__ xyz, const int* __restrict__ first_atom,
                                              const int* __restrict__ mons, int n, double* __restrict__ e_out,
                                              double* __restrict__ g_out, double* __restrict__ vir_out) {
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
            vir_out[(size_t)q * n + t] =
                -w[va[q]] * g[vb[q]] - w[3 + va[q]] * g[3 + vb[q]] - w[6 + va[q]] * g[6 + vb[q]];
    }
}

}  // namespace

hipError_t launch_1b_ps(const double* xyz, const int* first_atom, const int* mons, int n, double* e_out,
                        double* g_out, double* vir_out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_1b_ps, dim3((n + 63) / 64), dim3(64), 0, st, xyz, first_atom, mons, n, e_out, g_out,
                       vir_out);
    return hipGetLastError();
}

}  // namespace mbxg
