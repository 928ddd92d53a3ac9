// Differentiable MB-pol 2B / 3B evaluation of independent dimers and trimers.
//
// The reference evaluates x2o::x2b_v9x::eval (potential/2b/x2b-v9x.cpp:1514-1820) and
// x2o::x3b_v2x::eval (potential/3b/x3b-v2x.cpp:1403-1806) with its fitted coefficients and
// nonlinear parameters fixed at compile time, and returns energies and coordinate gradients only.
// Here the same energy E(x; a, theta) = sw(x) * P(v(x; theta); a) takes the linear coefficients
// a and the ten nonlinear parameters theta as inputs, and the adjoint returns the gradient of
//
//     phi = sum_n [ gE_n * E_n + gG_n . dE_n/dx_n ]
//
// with respect to the coordinates, a and theta: the vector-Jacobian product of the map
// (x, a, theta) -> (E, dE/dx), which is what an energy + force loss needs.  gG . dE/dx is the
// derivative of E along the direction u = gG, so phi's gradient is forward-over-reverse: every
// quantity of the energy evaluation is carried with its tangent along u (sites, variables v and
// vd, switch sw and swd, polynomial P and dP with its gradient and gradient tangent), and the
// reverse pass runs over that dual computation.
//
//   stage 1  k_ad_vars     sites (the M-sites of x2b-v9x.cpp:117-135 for 2B), variables and
//                          switch, with tangents
//   stage 2  P and grad P from the plain E+G polynomial kernel; in the backward pass also the
//            gradient tangent H vd by k_ad_tan (gen/ad_<name>_tan.inc), dP = grad P . vd
//   stage 3  k_ad_basis    coefficient gradient sum_n [alpha_n M_k + sw_n dM_k] (gen/ad_<name>_basis.inc)
//   stage 4  k_ad_chain    reverse pass to coordinates and theta; in the forward pass, E and dE/dx
//
// Parameter vectors theta (the reference's names and values, x2b-v9x.cpp:1330-1345 and
// x3b-v2x.cpp:1256-1271):
//   2B: k_HH_intra k_OH_intra k_HH_coul k_OH_coul k_OO_coul k_XH_main k_XO_main k_XX_main
//       in_plane_gamma out_of_plane_gamma
//   3B: kHH_intra kOH_intra kHH kOH kOO dHH_intra dOH_intra dHH dOH dOO
// Each n-mer is evaluated as given (no periodic imaging): coordinates are n-mer-local.  The
// reference's skip rules (2B r_OO > 6.5 A, 3B two O-O separations > 4.5 A) give E = 0 and zero
// derivatives, which the switch functions reproduce exactly.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "ad_kernel.hpp"
#include "gen/coef_mbpol_2b.inc"
#include "gen/coef_mbpol_3b.inc"
#include "mbxg.h"
#include "mbxg_device.hpp"
#include "mbxg_kernels.hpp"

namespace mbxg {

hipError_t launch_ad_tan_mbpol_2b(const double*, const double*, int, int, const double*, double*, hipStream_t);
hipError_t launch_ad_tan_mbpol_3b(const double*, const double*, int, int, const double*, double*, hipStream_t);
hipError_t launch_ad_basis_mbpol_2b(const double*, const double*, const double*, const double*, const double*, int,
                                    int, double*, hipStream_t);
hipError_t launch_ad_basis_mbpol_3b(const double*, const double*, const double*, const double*, const double*, int,
                                    int, double*, hipStream_t);

namespace {

// A variable of site pair (s1, s2) at separation r:
//   form 0  exp(k (d - r))          x2b-v9x.cpp:54-68   (v_exp, d = 1 intra / 4 inter)
//   form 1  exp(k (d - r)) / r      x2b-v9x.cpp:72-91   (v_coul, d = 4)
//   form 2  exp(-k (r - d))         x3b-v2x.cpp:52-62   (var, d a parameter)
// k = theta[kp]; d = theta[dp] when dp >= 0, else the constant d.
struct AdVar {
    int form, kp, dp;
    double d;
    int s1, s2;
};

constexpr int kNTheta = 10;

// Sites 0..5 = Oa Ha1 Ha2 Ob Hb1 Hb2, 6..9 = Xa1 Xa2 Xb1 Xb2 (x2b-v9x.cpp:1640-1675).
struct Ad2b {
    static constexpr int kNV = 31, kNAt = 6, kNS = 10, kPairs = 1, kParts = kAdParts;
    static constexpr bool kMsites = true;
    static constexpr double kRIn = p2::r2i, kROut = p2::r2f;
    static constexpr AdVar kVars[31] = {
        {0, 0, -1, 1.0, 1, 2}, {0, 0, -1, 1.0, 4, 5}, {0, 1, -1, 1.0, 0, 1}, {0, 1, -1, 1.0, 0, 2},
        {0, 1, -1, 1.0, 3, 4}, {0, 1, -1, 1.0, 3, 5}, {1, 2, -1, 4.0, 1, 4}, {1, 2, -1, 4.0, 1, 5},
        {1, 2, -1, 4.0, 2, 4}, {1, 2, -1, 4.0, 2, 5}, {1, 3, -1, 4.0, 0, 4}, {1, 3, -1, 4.0, 0, 5},
        {1, 3, -1, 4.0, 3, 1}, {1, 3, -1, 4.0, 3, 2}, {1, 4, -1, 4.0, 0, 3}, {0, 5, -1, 4.0, 6, 4},
        {0, 5, -1, 4.0, 6, 5}, {0, 5, -1, 4.0, 7, 4}, {0, 5, -1, 4.0, 7, 5}, {0, 5, -1, 4.0, 8, 1},
        {0, 5, -1, 4.0, 8, 2}, {0, 5, -1, 4.0, 9, 1}, {0, 5, -1, 4.0, 9, 2}, {0, 6, -1, 4.0, 0, 8},
        {0, 6, -1, 4.0, 0, 9}, {0, 6, -1, 4.0, 3, 6}, {0, 6, -1, 4.0, 3, 7}, {0, 7, -1, 4.0, 6, 8},
        {0, 7, -1, 4.0, 6, 9}, {0, 7, -1, 4.0, 7, 8}, {0, 7, -1, 4.0, 7, 9}};
    static constexpr int kPairA[3] = {0, 0, 0}, kPairB[3] = {3, 0, 0};
};

// Sites 0..8 = Oa Ha1 Ha2 Ob Hb1 Hb2 Oc Hc1 Hc2 (x3b-v2x.cpp:1548-1584); parameter set s uses
// k = theta[s], d = theta[5 + s].
struct Ad3b {
    static constexpr int kNV = 36, kNAt = 9, kNS = 9, kPairs = 3, kParts = kAdParts;
    static constexpr bool kMsites = false;
    static constexpr double kRIn = p3::r3i, kROut = p3::r3f;
    static constexpr AdVar kVars[36] = {
        {2, 0, 5, 0, 1, 2}, {2, 0, 5, 0, 4, 5}, {2, 0, 5, 0, 7, 8}, {2, 1, 6, 0, 0, 1}, {2, 1, 6, 0, 0, 2},
        {2, 1, 6, 0, 3, 4}, {2, 1, 6, 0, 3, 5}, {2, 1, 6, 0, 6, 7}, {2, 1, 6, 0, 6, 8}, {2, 2, 7, 0, 1, 4},
        {2, 2, 7, 0, 1, 5}, {2, 2, 7, 0, 1, 7}, {2, 2, 7, 0, 1, 8}, {2, 2, 7, 0, 2, 4}, {2, 2, 7, 0, 2, 5},
        {2, 2, 7, 0, 2, 7}, {2, 2, 7, 0, 2, 8}, {2, 2, 7, 0, 4, 7}, {2, 2, 7, 0, 4, 8}, {2, 2, 7, 0, 5, 7},
        {2, 2, 7, 0, 5, 8}, {2, 3, 8, 0, 0, 4}, {2, 3, 8, 0, 0, 5}, {2, 3, 8, 0, 0, 7}, {2, 3, 8, 0, 0, 8},
        {2, 3, 8, 0, 3, 1}, {2, 3, 8, 0, 3, 2}, {2, 3, 8, 0, 3, 7}, {2, 3, 8, 0, 3, 8}, {2, 3, 8, 0, 6, 1},
        {2, 3, 8, 0, 6, 2}, {2, 3, 8, 0, 6, 4}, {2, 3, 8, 0, 6, 5}, {2, 4, 9, 0, 0, 3}, {2, 4, 9, 0, 0, 6},
        {2, 4, 9, 0, 3, 6}};
    // O-O pairs ab, ac, bc; the switch is s_ab s_ac + s_ab s_bc + s_ac s_bc (x3b-v2x.cpp:1614-1620)
    static constexpr int kPairA[3] = {0, 0, 3}, kPairB[3] = {3, 6, 6};
};

// f_switch (x2b-v9x.cpp:1349-1362, x3b-v2x.cpp:1275-1288) with its first and second derivatives.
template <class F>
__device__ __forceinline__ double ad_switch(double r, double& f1, double& f2) {
    if (r > F::kROut) {
        f1 = f2 = 0.0;
        return 0.0;
    } else if (r > F::kRIn) {
        const double t1 = M_PI / (F::kROut - F::kRIn);
        const double x = (r - F::kRIn) * t1;
        const double c = cos(x);
        f1 = -sin(x) * t1 / 2.0;
        f2 = -c * t1 * t1 / 2.0;
        return (1.0 + c) / 2.0;
    }
    f1 = f2 = 0.0;
    return 1.0;
}

// Value f(r) of variable V and f1 = df/dr; with DER also f2 = d2f/dr2 and the parameter
// derivatives fk = df/dk, f1k = df1/dk, fd = df/dd, f1d = df1/dd.
template <bool DER>
__device__ __forceinline__ double ad_value(const AdVar& V, double k, double d, double r, double& f1, double& f2,
                                           double& fk, double& f1k, double& fd, double& f1d) {
    double f;
    if (V.form == 1) {
        const double e = exp(k * (d - r));
        const double rinv = 1.0 / r;
        f = e * rinv;
        const double kr = k + rinv;
        f1 = -kr * f;
        if (DER) {
            f2 = f * (kr * kr + rinv * rinv);
            fk = (d - r) * f;
            f1k = -f - kr * fk;
            fd = 0.0;
            f1d = 0.0;
        }
        return f;
    }
    f = V.form == 0 ? exp(k * (d - r)) : exp(-k * (r - d));
    f1 = -k * f;
    if (DER) {
        f2 = k * k * f;
        fk = (d - r) * f;
        f1k = -f - k * fk;
        fd = k * f;
        f1d = -k * fd;
    }
    return f;
}

__device__ __forceinline__ void cross(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Load the n-mer's atoms (and tangents), then the M-sites of each water (monomer::setup,
// x2b-v9x.cpp:117-135) with their tangents; oh[w] = (OH1, OH2) and uh[w] their tangents.
template <class F, bool TAN>
__device__ __forceinline__ void ad_sites(const double* __restrict__ xyz, const double* __restrict__ u, int t,
                                         const double* th, double* s, double* sd, double (*oh)[6],
                                         double (*uh)[6]) {
    const double* a = xyz + (size_t)t * 3 * F::kNAt;
#pragma unroll
    for (int i = 0; i < 3 * F::kNAt; ++i) s[i] = a[i];
    if (TAN) {  // a null direction is the zero direction
        const double* b = u ? u + (size_t)t * 3 * F::kNAt : nullptr;
#pragma unroll
        for (int i = 0; i < 3 * F::kNAt; ++i) sd[i] = b ? b[i] : 0.0;
    }
    if (F::kMsites) {
        const double g1 = th[8], g2 = th[9];
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const int o = 3 * w, x1 = 6 + 2 * w, x2 = 7 + 2 * w;
            double* oh1 = oh[w];
            double* oh2 = oh[w] + 3;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                oh1[i] = s[3 * (o + 1) + i] - s[3 * o + i];
                oh2[i] = s[3 * (o + 2) + i] - s[3 * o + i];
            }
            double v[3];
            cross(oh1, oh2, v);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double in_plane = s[3 * o + i] + 0.5 * g1 * (oh1[i] + oh2[i]);
                const double out_of_plane = g2 * v[i];
                s[3 * x1 + i] = in_plane + out_of_plane;
                s[3 * x2 + i] = in_plane - out_of_plane;
            }
            if (TAN) {
                double* u1 = uh[w];
                double* u2 = uh[w] + 3;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    u1[i] = sd[3 * (o + 1) + i] - sd[3 * o + i];
                    u2[i] = sd[3 * (o + 2) + i] - sd[3 * o + i];
                }
                double c1[3], c2[3];
                cross(u1, oh2, c1);
                cross(oh1, u2, c2);
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const double in_plane = sd[3 * o + i] + 0.5 * g1 * (u1[i] + u2[i]);
                    const double out_of_plane = g2 * (c1[i] + c2[i]);
                    sd[3 * x1 + i] = in_plane + out_of_plane;
                    sd[3 * x2 + i] = in_plane - out_of_plane;
                }
            }
        }
    }
}

// Per O-O pair q: separation d, length r, rho = d . (tangent of d), switch value and derivatives.
template <class F, bool TAN>
__device__ __forceinline__ void ad_pairs(const double* s, const double* sd, double (*d)[3], double (*dd)[3],
                                         double* r, double* rho, double* sv, double* f1, double* f2) {
#pragma unroll
    for (int q = 0; q < F::kPairs; ++q) {
        const int a = F::kPairA[q], b = F::kPairB[q];
        rho[q] = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            d[q][i] = s[3 * a + i] - s[3 * b + i];
            if (TAN) dd[q][i] = sd[3 * a + i] - sd[3 * b + i];
        }
        r[q] = sqrt(d[q][0] * d[q][0] + d[q][1] * d[q][1] + d[q][2] * d[q][2]);
        if (TAN) rho[q] = d[q][0] * dd[q][0] + d[q][1] * dd[q][1] + d[q][2] * dd[q][2];
        sv[q] = ad_switch<F>(r[q], f1[q], f2[q]);
    }
}

// Total switch and its tangent from the pair switches.
template <class F>
__device__ __forceinline__ void ad_total_switch(const double* sv, const double* f1, const double* r, const double* rho,
                                                double& sw, double& swd, double* sdot) {
#pragma unroll
    for (int q = 0; q < F::kPairs; ++q) sdot[q] = f1[q] * rho[q] / r[q];
    if (F::kPairs == 1) {
        sw = sv[0];
        swd = sdot[0];
    } else {
        sw = sv[0] * sv[1] + sv[0] * sv[2] + sv[1] * sv[2];
        swd = sdot[0] * (sv[1] + sv[2]) + sdot[1] * (sv[0] + sv[2]) + sdot[2] * (sv[0] + sv[1]);
    }
}

}  // namespace

// Stage 1: variables (and tangents) and the switch (and tangent).  Without TAN the tangent
// outputs, when given, are written as zeros.
template <class F, bool TAN>
__global__ void __launch_bounds__(64) k_ad_vars(const double* __restrict__ xyz, const double* __restrict__ u,
                                                const double* __restrict__ theta, int n, int ld,
                                                double* __restrict__ xs, double* __restrict__ xds,
                                                double* __restrict__ sw_out, double* __restrict__ swd_out) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= n) return;
    double th[kNTheta];
#pragma unroll
    for (int i = 0; i < kNTheta; ++i) th[i] = theta[i];
    double s[3 * F::kNS], sd[3 * F::kNS], oh[2][6], uh[2][6];
    ad_sites<F, TAN>(xyz, u, t, th, s, sd, oh, uh);
    static_for<F::kNV>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr AdVar V = F::kVars[j];
        double dv[3], ddv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            dv[i] = s[3 * V.s1 + i] - s[3 * V.s2 + i];
            if (TAN) ddv[i] = sd[3 * V.s1 + i] - sd[3 * V.s2 + i];
        }
        const double r = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        double f1, f2, fk, f1k, fd, f1d;
        const double f = ad_value<false>(V, th[V.kp], V.dp >= 0 ? th[V.dp >= 0 ? V.dp : 0] : V.d, r, f1, f2, fk, f1k, fd, f1d);
        xs[(size_t)j * ld + t] = f;
        if (TAN) {
            const double rho = dv[0] * ddv[0] + dv[1] * ddv[1] + dv[2] * ddv[2];
            xds[(size_t)j * ld + t] = f1 * rho / r;
        } else if (xds) {
            xds[(size_t)j * ld + t] = 0.0;
        }
    });
    double d[3][3], dd[3][3], r[3], rho[3], sv[3], f1[3], f2[3], sdot[3], sw, swd;
    ad_pairs<F, TAN>(s, sd, d, dd, r, rho, sv, f1, f2);
    ad_total_switch<F>(sv, f1, r, rho, sw, swd, sdot);
    sw_out[t] = sw;
    if (swd_out) swd_out[t] = TAN ? swd : 0.0;
}

// Stage 4.  Forward (BWD = false): energy_out = sw * P and dx_out = dE/dx from the plain
// polynomial's value and gradient.  Backward: dx_out = dphi/dx and per-block partial sums of
// dphi/dtheta (theta_partial[block * 16 + i]) from P, grad P, the gradient tangent gds and the
// variable tangents xds (dP = grad P . xds).  P and grad P come in `parts` slices, added here.
template <class F, bool BWD>
__global__ void __launch_bounds__(64) k_ad_chain(const double* __restrict__ xyz, const double* __restrict__ u,
                                                 const double* __restrict__ theta,
                                                 const double* __restrict__ g_energy, int n, int ld, int parts,
                                                 const double* __restrict__ gs, const double* __restrict__ gds,
                                                 const double* __restrict__ ps, const double* __restrict__ xds,
                                                 double* __restrict__ energy_out, double* __restrict__ dx_out,
                                                 double* __restrict__ theta_partial) {
    constexpr int NV = F::kNV;
    const int t = blockIdx.x * 64 + threadIdx.x;
    const bool active = t < n;
    double thb[kNTheta];
#pragma unroll
    for (int i = 0; i < kNTheta; ++i) thb[i] = 0.0;
    if (active) {
        double th[kNTheta];
#pragma unroll
        for (int i = 0; i < kNTheta; ++i) th[i] = theta[i];
        double s[3 * F::kNS], sd[3 * F::kNS], oh[2][6], uh[2][6];
        ad_sites<F, BWD>(xyz, u, t, th, s, sd, oh, uh);
        double P = 0.0, Pd = 0.0;
        for (int q = 0; q < parts; ++q) P += ps[(size_t)q * ld + t];
        double d[3][3], dd[3][3], r[3], rho[3], sv[3], f1[3], f2[3], sdot[3], sw, swd;
        ad_pairs<F, BWD>(s, sd, d, dd, r, rho, sv, f1, f2);
        ad_total_switch<F>(sv, f1, r, rho, sw, swd, sdot);
        const double gE = BWD ? g_energy[t] : 1.0;
        // adjoints of P and dP (= sw); those of sw and swd follow the variable loop, which also
        // accumulates dP = grad P . vd
        const double alpha = BWD ? gE * sw + swd : sw;
        double sb[3 * F::kNS], sdb[12];
#pragma unroll
        for (int i = 0; i < 3 * F::kNS; ++i) sb[i] = 0.0;
#pragma unroll
        for (int i = 0; i < 12; ++i) sdb[i] = 0.0;
        static_for<NV>([&](auto J) {
            constexpr int j = decltype(J)::value;
            constexpr AdVar V = F::kVars[j];
            double g = 0.0, gd = 0.0;
            for (int q = 0; q < parts; ++q) g += gs[((size_t)q * NV + j) * ld + t];
            if (BWD) {
                gd = gds[(size_t)j * ld + t];
                Pd += g * xds[(size_t)j * ld + t];
            }
            const double vb = BWD ? alpha * g + sw * gd : sw * g;  // dphi/dv
            const double vdb = sw * g;                              // dphi/dvd
            double dv[3], ddv[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                dv[i] = s[3 * V.s1 + i] - s[3 * V.s2 + i];
                if (BWD) ddv[i] = sd[3 * V.s1 + i] - sd[3 * V.s2 + i];
            }
            const double rr = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
            const double rinv = 1.0 / rr;
            const double k = th[V.kp], dpar = V.dp >= 0 ? th[V.dp >= 0 ? V.dp : 0] : V.d;
            double vf1, vf2, fk, f1k, fd, f1d;
            ad_value<BWD>(V, k, dpar, rr, vf1, vf2, fk, f1k, fd, f1d);
            if (!BWD) {
                const double c = vb * vf1 * rinv;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    sb[3 * V.s1 + i] += c * dv[i];
                    sb[3 * V.s2 + i] -= c * dv[i];
                }
            } else {
                const double rh = dv[0] * ddv[0] + dv[1] * ddv[1] + dv[2] * ddv[2];
                const double rdot = rh * rinv;
                const double cd = vb * vf1 * rinv + vdb * (vf2 - vf1 * rinv) * rdot * rinv;
                const double cdd = vdb * vf1 * rinv;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const double gi = cd * dv[i] + cdd * ddv[i];
                    sb[3 * V.s1 + i] += gi;
                    sb[3 * V.s2 + i] -= gi;
                }
                if (F::kMsites) {
                    if (V.s1 >= 6) {
#pragma unroll
                        for (int i = 0; i < 3; ++i) sdb[3 * (V.s1 - 6) + i] += cdd * dv[i];
                    }
                    if (V.s2 >= 6) {
#pragma unroll
                        for (int i = 0; i < 3; ++i) sdb[3 * (V.s2 - 6) + i] -= cdd * dv[i];
                    }
                }
                thb[V.kp] += vb * fk + vdb * rdot * f1k;
                if (V.dp >= 0) thb[V.dp >= 0 ? V.dp : 0] += vb * fd + vdb * rdot * f1d;
            }
        });
        const double a_sw = BWD ? gE * P + Pd : P;
        const double b_sw = P;
        // switches: pair q's value enters sw (weight a_q) and swd (weight b_q)
#pragma unroll
        for (int q = 0; q < F::kPairs; ++q) {
            double aq, bq;
            if (F::kPairs == 1) {
                aq = a_sw;
                bq = b_sw;
            } else {
                const int o1 = q == 0 ? 1 : 0, o2 = q == 2 ? 1 : 2;
                aq = a_sw * (sv[o1] + sv[o2]) + (BWD ? b_sw * (sdot[o1] + sdot[o2]) : 0.0);
                bq = b_sw * (sv[o1] + sv[o2]);
            }
            const double rinv = 1.0 / r[q];
            double cd = aq * f1[q] * rinv, cdd = 0.0;
            if (BWD) {
                const double rdot = rho[q] * rinv;
                cd += bq * (f2[q] - f1[q] * rinv) * rdot * rinv;
                cdd = bq * f1[q] * rinv;
            }
            const int a = F::kPairA[q], b = F::kPairB[q];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double gi = cd * d[q][i] + (BWD ? cdd * dd[q][i] : 0.0);
                sb[3 * a + i] += gi;
                sb[3 * b + i] -= gi;
            }
        }
        // M-sites back onto O, H1, H2 (monomer::grads, x2b-v9x.cpp:139-160), and the gammas
        if (F::kMsites) {
            const double g1 = th[8], g2 = th[9];
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int o = 3 * w, x1 = 6 + 2 * w, x2 = 7 + 2 * w;
                const double* oh1 = oh[w];
                const double* oh2 = oh[w] + 3;
                double Sp[3], Sm[3], c1[3], c2[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    Sp[i] = sb[3 * x1 + i] + sb[3 * x2 + i];
                    Sm[i] = sb[3 * x1 + i] - sb[3 * x2 + i];
                }
                cross(oh2, Sm, c1);
                cross(Sm, oh1, c2);
                double gh1[3], gh2[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    gh1[i] = 0.5 * g1 * Sp[i] + g2 * c1[i];
                    gh2[i] = 0.5 * g1 * Sp[i] + g2 * c2[i];
                }
                if (BWD) {
                    const double* u1 = uh[w];
                    const double* u2 = uh[w] + 3;
                    double Tp[3], Tm[3], e1[3], e2[3], v[3], vd1[3], vd2[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        Tp[i] = sdb[6 * w + i] + sdb[6 * w + 3 + i];
                        Tm[i] = sdb[6 * w + i] - sdb[6 * w + 3 + i];
                    }
                    cross(u2, Tm, e1);
                    cross(Tm, u1, e2);
                    cross(oh1, oh2, v);
                    cross(u1, oh2, vd1);
                    cross(oh1, u2, vd2);
                    double t8 = 0.0, t9 = 0.0;
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        gh1[i] += g2 * e1[i];
                        gh2[i] += g2 * e2[i];
                        t8 += Sp[i] * (oh1[i] + oh2[i]) + Tp[i] * (u1[i] + u2[i]);
                        t9 += Sm[i] * v[i] + Tm[i] * (vd1[i] + vd2[i]);
                    }
                    thb[8] += 0.5 * t8;
                    thb[9] += t9;
                }
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    sb[3 * o + i] += Sp[i] - (gh1[i] + gh2[i]);
                    sb[3 * (o + 1) + i] += gh1[i];
                    sb[3 * (o + 2) + i] += gh2[i];
                }
            }
        }
        double* dst = dx_out + (size_t)t * 3 * F::kNAt;
#pragma unroll
        for (int i = 0; i < 3 * F::kNAt; ++i) dst[i] = sb[i];
        if (!BWD) energy_out[t] = sw * P;
    }
    if (BWD) {
#pragma unroll
        for (int i = 0; i < kNTheta; ++i) {
            const double v = wave_sum(thb[i]);
            if (threadIdx.x == 0) theta_partial[(size_t)blockIdx.x * 16 + i] = v;
        }
    }
}

// out[k] = sum over rows r < nrows of partial[r * stride + k] for k < ncol: one wavefront per
// column, 64 rows at a time, a fixed butterfly order (deterministic).
__global__ void __launch_bounds__(64) k_ad_colsum(const double* __restrict__ partial, int nrows, int ncol, int stride,
                                                  double* __restrict__ out) {
    const int k = blockIdx.x;
    if (k >= ncol) return;
    double s = 0.0;
    for (int r = threadIdx.x; r < nrows; r += 64) s += partial[(size_t)r * stride + k];
    s = wave_sum(s);
    if (threadIdx.x == 0) out[k] = s;
}

// Sum the parts' slices of the gradient tangent into slice 0 (entry e of part q at
// q * rows + e, rows = nv * ld).  One thread per entry, parts added in order.
__global__ void __launch_bounds__(256) k_ad_sum_parts(double* __restrict__ gds, size_t rows, int parts) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= rows) return;
    double s = gds[e];
    for (int q = 1; q < parts; ++q) s += gds[q * rows + e];
    gds[e] = s;
}

__global__ void k_ad_set_int(int* p, int v) { *p = v; }

}  // namespace mbxg

using namespace mbxg;

namespace {

struct AdLayout {
    size_t xs, xds, sw, swd, gs, gds, ps, ndev, basis, theta, total;
};

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

bool ad_shape(int family, int& nv, int& nat, int& ncoef) {
    if (family == MBXG_AD_MBPOL_2B) {
        nv = Ad2b::kNV;
        nat = Ad2b::kNAt;
        ncoef = kNCoef_mbpol_2b;
        return true;
    }
    if (family == MBXG_AD_MBPOL_3B) {
        nv = Ad3b::kNV;
        nat = Ad3b::kNAt;
        ncoef = kNCoef_mbpol_3b;
        return true;
    }
    return false;
}

// Workspace: stage buffers for ld = n rounded up to 64 lanes; kAdParts slices of the gradient
// tangent, one slice of P and grad P (the plain kernel sums its parts itself).
AdLayout ad_layout(int nv, int ncoef, int n) {
    const size_t ld = (size_t)((n + 63) / 64) * 64, nblk = ld / 64, parts = kAdParts;
    AdLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes);
        return at;
    };
    L.xs = take(sizeof(double) * nv * ld);
    L.xds = take(sizeof(double) * nv * ld);
    L.sw = take(sizeof(double) * ld);
    L.swd = take(sizeof(double) * ld);
    L.gs = take(sizeof(double) * nv * ld);
    L.gds = take(sizeof(double) * parts * nv * ld);
    L.ps = take(sizeof(double) * ld);
    L.ndev = take(sizeof(int));
    L.basis = take(sizeof(double) * nblk * ncoef);
    L.theta = take(sizeof(double) * nblk * 16);
    L.total = o;
    return L;
}

int ad_err(int code, const char* fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return report_error(code, buf);
}

#define ADCHK(expr)                                                                                      \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) return ad_err(MBXG_EHIP, "%s: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

int ad_check(const char* who, int family, int n, const void* ws, size_t ws_bytes, int& nv, int& nat, int& ncoef,
             AdLayout& L) {
    if (!ad_shape(family, nv, nat, ncoef)) return ad_err(MBXG_EINVAL, "%s: unknown family %d", who, family);
    if (n < 0) return ad_err(MBXG_EINVAL, "%s: negative n-mer count %d", who, n);
    L = ad_layout(nv, ncoef, n);
    if (n > 0 && (!ws || ws_bytes < L.total))
        return ad_err(MBXG_EINVAL, "%s: workspace of %zu bytes, %zu needed for %d n-mers", who, ws_bytes, L.total, n);
    return MBXG_OK;
}

}  // namespace

extern "C" int mbxg_ad_shape(int family, int* n_atoms, int* n_coef, int* n_theta) {
    int nv, nat, ncoef;
    if (!ad_shape(family, nv, nat, ncoef)) return ad_err(MBXG_EINVAL, "mbxg_ad_shape: unknown family %d", family);
    if (n_atoms) *n_atoms = nat;
    if (n_coef) *n_coef = ncoef;
    if (n_theta) *n_theta = kNTheta;
    return MBXG_OK;
}

extern "C" int mbxg_ad_defaults(int family, double* coef, double* theta) {
    if (family == MBXG_AD_MBPOL_2B) {
        const double th[kNTheta] = {p2::k_HH_intra, p2::k_OH_intra, p2::k_HH_coul, p2::k_OH_coul,
                                    p2::k_OO_coul,  p2::k_XH_main,  p2::k_XO_main, p2::k_XX_main,
                                    p2::in_plane_gamma, p2::out_of_plane_gamma};
        if (coef)
            for (int i = 0; i < kNCoef_mbpol_2b; ++i) coef[i] = kFit_mbpol_2b[i];
        if (theta)
            for (int i = 0; i < kNTheta; ++i) theta[i] = th[i];
        return MBXG_OK;
    }
    if (family == MBXG_AD_MBPOL_3B) {
        const double th[kNTheta] = {p3::kHH_intra, p3::kOH_intra, p3::kHH, p3::kOH, p3::kOO,
                                    p3::dHH_intra, p3::dOH_intra, p3::dHH, p3::dOH, p3::dOO};
        if (coef)
            for (int i = 0; i < kNCoef_mbpol_3b; ++i) coef[i] = kFit_mbpol_3b[i];
        if (theta)
            for (int i = 0; i < kNTheta; ++i) theta[i] = th[i];
        return MBXG_OK;
    }
    return ad_err(MBXG_EINVAL, "mbxg_ad_defaults: unknown family %d", family);
}

extern "C" size_t mbxg_ad_workspace_bytes(int family, int n) {
    int nv, nat, ncoef;
    if (!ad_shape(family, nv, nat, ncoef) || n < 0) return 0;
    return ad_layout(nv, ncoef, n).total;
}

namespace {

// The forward stages over stage buffers xs (nv x ld), sw (ld), gs (nv x ld), ps (ld) and the
// device n-mer count ndev, wherever they live (a caller's workspace or the library's scratch).
int ad_forward_run(int family, int n, const double* xyz, const double* coef, const double* theta, double* energy,
                   double* grad, double* xs, double* sw, double* gs, double* ps, int* ndev, hipStream_t st) {
    const int ld = (n + 63) / 64 * 64, nb = ld / 64;
    hipLaunchKernelGGL(k_ad_set_int, dim3(1), dim3(1), 0, st, ndev, n);
    if (family == MBXG_AD_MBPOL_2B) {
        hipLaunchKernelGGL((k_ad_vars<Ad2b, false>), dim3(nb), dim3(64), 0, st, xyz, nullptr, theta, n, ld, xs,
                           nullptr, sw, nullptr);
        ADCHK(launch_poly_mbpol2b(xs, ndev, n, ld, coef, 1, gs, ps, 8, st));
        hipLaunchKernelGGL((k_ad_chain<Ad2b, false>), dim3(nb), dim3(64), 0, st, xyz, nullptr, theta, nullptr, n, ld,
                           1, gs, nullptr, ps, nullptr, energy, grad, nullptr);
    } else {
        hipLaunchKernelGGL((k_ad_vars<Ad3b, false>), dim3(nb), dim3(64), 0, st, xyz, nullptr, theta, n, ld, xs,
                           nullptr, sw, nullptr);
        ADCHK(launch_poly_mbpol3b(xs, ndev, n, ld, coef, 1, gs, ps, 8, st));
        hipLaunchKernelGGL((k_ad_chain<Ad3b, false>), dim3(nb), dim3(64), 0, st, xyz, nullptr, theta, nullptr, n, ld,
                           1, gs, nullptr, ps, nullptr, energy, grad, nullptr);
    }
    ADCHK(hipGetLastError());
    return MBXG_OK;
}

// ---- Forward without allocation -------------------------------------------------------------
// mbxg_ad_forward_noalloc keeps its stage buffers in this file-scope device array instead of a
// caller's workspace: forward-only buffers for at most kNoAllocMaxNmers n-mers of either family
// (sized for the 36 variables of the 3B one).  Each device holds its own copy for the life of the
// process; host code reaches it only through the address hipGetSymbolAddress returns on that
// device (the array's name in host code is not a device address).
constexpr int kNoAllocMaxNmers = 8192;
constexpr size_t fwd_bytes(int nv, int n) {
    // xs, sw, gs, ps and the count, each rounded up to kAlign bytes
    return ((sizeof(double) * nv * n + kAlign - 1) / kAlign * kAlign) * 2 +
           ((sizeof(double) * n + kAlign - 1) / kAlign * kAlign) * 2 + kAlign;
}
constexpr size_t kScratchBytes = fwd_bytes(36, kNoAllocMaxNmers);  // 4,849,920 bytes
static_assert(fwd_bytes(31, kNoAllocMaxNmers) <= kScratchBytes, "the 2B forward fits the 3B scratch");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};  // scratch's address on each device, set on first use there

// The scratch address on the current device: looked up once per device (with that device
// current), then reused.
int scratch_on_current_device(void** out) {
    int dev = -1;
    ADCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return ad_err(MBXG_EINVAL, "mbxg_ad_forward_noalloc: device %d outside 0..%d", dev, kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        ADCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return ad_err(MBXG_EHIP, "mbxg_ad_forward_noalloc: no device address for the scratch array");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return MBXG_OK;
}

}  // namespace

extern "C" int mbxg_ad_forward(int family, int n, const double* xyz, const double* coef, const double* theta,
                               double* energy, double* grad, void* workspace, size_t workspace_bytes, void* stream) {
    int nv, nat, ncoef;
    AdLayout L;
    if (int e = ad_check("mbxg_ad_forward", family, n, workspace, workspace_bytes, nv, nat, ncoef, L)) return e;
    if (n == 0) return MBXG_OK;
    if (!xyz || !coef || !theta || !energy || !grad) return ad_err(MBXG_EINVAL, "mbxg_ad_forward: null pointer");
    char* w = (char*)workspace;
    return ad_forward_run(family, n, xyz, coef, theta, energy, grad, (double*)(w + L.xs), (double*)(w + L.sw),
                          (double*)(w + L.gs), (double*)(w + L.ps), (int*)(w + L.ndev), (hipStream_t)stream);
}

extern "C" int mbxg_ad_forward_noalloc(int family, int n, const double* xyz, const double* coef,
                                       const double* theta, double* energy, double* grad, void* stream) {
    int nv, nat, ncoef;
    if (!ad_shape(family, nv, nat, ncoef)) return ad_err(MBXG_EINVAL, "mbxg_ad_forward_noalloc: unknown family %d", family);
    if (n < 0) return ad_err(MBXG_EINVAL, "mbxg_ad_forward_noalloc: negative n-mer count %d", n);
    if (n > kNoAllocMaxNmers)
        return ad_err(MBXG_EINVAL, "mbxg_ad_forward_noalloc: %d n-mers exceed the scratch limit of %d", n,
                      kNoAllocMaxNmers);
    if (n == 0) return MBXG_OK;
    if (!xyz || !coef || !theta || !energy || !grad) return ad_err(MBXG_EINVAL, "mbxg_ad_forward_noalloc: null pointer");
    void* base = nullptr;
    if (int e = scratch_on_current_device(&base)) return e;
    // the forward layout for this n inside the scratch; every stage writes its buffer before
    // reading it (the count by k_ad_set_int), so nothing needs clearing between calls
    const size_t ld = (size_t)(n + 63) / 64 * 64;
    char* w = (char*)base;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char* at = w + o;
        o = align_up(o + bytes);
        return at;
    };
    double* xs = (double*)take(sizeof(double) * nv * ld);
    double* gs = (double*)take(sizeof(double) * nv * ld);
    double* sw = (double*)take(sizeof(double) * ld);
    double* ps = (double*)take(sizeof(double) * ld);
    int* ndev = (int*)take(sizeof(int));
    if (o > kScratchBytes) return ad_err(MBXG_EINVAL, "mbxg_ad_forward_noalloc: layout exceeds the scratch");
    return ad_forward_run(family, n, xyz, coef, theta, energy, grad, xs, sw, gs, ps, ndev, (hipStream_t)stream);
}

extern "C" int mbxg_ad_backward(int family, int n, const double* xyz, const double* coef, const double* theta,
                                const double* g_energy, const double* g_grad, double* d_xyz, double* d_coef,
                                double* d_theta, void* workspace, size_t workspace_bytes, int from_forward,
                                void* stream) {
    int nv, nat, ncoef;
    AdLayout L;
    if (int e = ad_check("mbxg_ad_backward", family, n, workspace, workspace_bytes, nv, nat, ncoef, L)) return e;
    if ((n > 0 && !d_xyz) || !d_coef || !d_theta) return ad_err(MBXG_EINVAL, "mbxg_ad_backward: null output pointer");
    hipStream_t st = (hipStream_t)stream;
    if (n == 0) {
        hipLaunchKernelGGL(k_zero, dim3((ncoef + 255) / 256), dim3(256), 0, st, d_coef, (size_t)ncoef);
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, st, d_theta, (size_t)kNTheta);
        ADCHK(hipGetLastError());
        return MBXG_OK;
    }
    if (!xyz || !coef || !theta || !g_energy) return ad_err(MBXG_EINVAL, "mbxg_ad_backward: null input pointer");
    char* w = (char*)workspace;
    double* xs = (double*)(w + L.xs);
    double* xds = (double*)(w + L.xds);
    double* sw = (double*)(w + L.sw);
    double* swd = (double*)(w + L.swd);
    double* gs = (double*)(w + L.gs);
    double* gds = (double*)(w + L.gds);
    double* ps = (double*)(w + L.ps);
    int* ndev = (int*)(w + L.ndev);
    double* basis = (double*)(w + L.basis);
    double* thp = (double*)(w + L.theta);
    const int ld = (n + 63) / 64 * 64, nb = ld / 64;
    // a null g_grad is the zero direction (an energy-only loss)
    const size_t rows = (size_t)nv * ld;
    hipLaunchKernelGGL(k_ad_set_int, dim3(1), dim3(1), 0, st, ndev, n);
    if (family == MBXG_AD_MBPOL_2B) {
        hipLaunchKernelGGL((k_ad_vars<Ad2b, true>), dim3(nb), dim3(64), 0, st, xyz, g_grad, theta, n, ld, xs, xds, sw,
                           swd);
        if (!from_forward) ADCHK(launch_poly_mbpol2b(xs, ndev, n, ld, coef, 1, gs, ps, 8, st));
        ADCHK(launch_ad_tan_mbpol_2b(xs, xds, n, ld, coef, gds, st));
        ADCHK(launch_ad_basis_mbpol_2b(xs, xds, sw, swd, g_energy, n, ld, basis, st));
        hipLaunchKernelGGL(k_ad_sum_parts, dim3((rows + 255) / 256), dim3(256), 0, st, gds, rows, Ad2b::kParts);
        hipLaunchKernelGGL((k_ad_chain<Ad2b, true>), dim3(nb), dim3(64), 0, st, xyz, g_grad, theta, g_energy, n, ld, 1,
                           gs, gds, ps, xds, nullptr, d_xyz, thp);
    } else {
        hipLaunchKernelGGL((k_ad_vars<Ad3b, true>), dim3(nb), dim3(64), 0, st, xyz, g_grad, theta, n, ld, xs, xds, sw,
                           swd);
        if (!from_forward) ADCHK(launch_poly_mbpol3b(xs, ndev, n, ld, coef, 1, gs, ps, 8, st));
        ADCHK(launch_ad_tan_mbpol_3b(xs, xds, n, ld, coef, gds, st));
        ADCHK(launch_ad_basis_mbpol_3b(xs, xds, sw, swd, g_energy, n, ld, basis, st));
        hipLaunchKernelGGL(k_ad_sum_parts, dim3((rows + 255) / 256), dim3(256), 0, st, gds, rows, Ad3b::kParts);
        hipLaunchKernelGGL((k_ad_chain<Ad3b, true>), dim3(nb), dim3(64), 0, st, xyz, g_grad, theta, g_energy, n, ld, 1,
                           gs, gds, ps, xds, nullptr, d_xyz, thp);
    }
    hipLaunchKernelGGL(k_ad_colsum, dim3(ncoef), dim3(64), 0, st, basis, nb, ncoef, ncoef, d_coef);
    hipLaunchKernelGGL(k_ad_colsum, dim3(kNTheta), dim3(64), 0, st, thp, nb, kNTheta, 16, d_theta);
    ADCHK(hipGetLastError());
    return MBXG_OK;
}
