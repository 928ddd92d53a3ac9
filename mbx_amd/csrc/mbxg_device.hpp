// Device-side definitions shared by the list builder and the n-mer evaluation kernels.
//
// Geometry helpers restate, operation for operation, the reference's periodic-box arithmetic
// (paths relative to the reference's src/):
//   FixMonomerCoordinates   bblock/sys_tools.cpp:257-320
//   GetCloseDimerImage      bblock/sys_tools.cpp:322-365
//   XyzToCloud re-wrap      kdtree/kdtree_utils.h:215-232
// They run with FP contraction off so every product and sum rounds exactly as the
// reference's ISO C++ build does; list membership depends on those bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace mbxg {

// Monomer kinds known to the GPU path (the reference keys them by name, energy2b.cpp:243-254).
enum MonKind : int {
    MON_H2O = 0,
    MON_CH4 = 1,
    MON_OTHER = 2,  // a monomer no holder here takes (n-mers holding one are listed, not evaluated)
    MON_CO2 = 3,
    MON_F = 4,  // halide ions f-, cl-, br-, i-
    MON_CL = 5,
    MON_BR = 6,
    MON_I = 7,
    MON_LI = 8,  // alkali-metal ions li+, na+, k+, rb+, cs+
    MON_NA = 9,
    MON_K = 10,
    MON_RB = 11,
    MON_CS = 12
};

// Box description passed by value to kernels.  frac[] holds the coefficients of
// GetFractionalCoordinates (kdtree/kdtree_utils.h:165-194), computed on the host with libm.
struct Box {
    double b[9];
    double bi[9];
    double f00, f01, f02, f11, f12, f22;  // coord0 = f00*x - f01*y + f02*z, ...
    int periodic;
};

#pragma clang fp contract(off)

__device__ __forceinline__ void to_frac(const Box& B, double x, double y, double z, double& u, double& v, double& w) {
    u = B.bi[0] * x + B.bi[3] * y + B.bi[6] * z;
    v = B.bi[1] * x + B.bi[4] * y + B.bi[7] * z;
    w = B.bi[2] * x + B.bi[5] * y + B.bi[8] * z;
}

__device__ __forceinline__ void to_cart(const Box& B, double u, double v, double w, double& x, double& y, double& z) {
    x = B.b[0] * u + B.b[3] * v + B.b[6] * w;
    y = B.b[1] * u + B.b[4] * v + B.b[7] * w;
    z = B.b[2] * u + B.b[5] * v + B.b[8] * w;
}

// Shift of the n-mer member whose first atom is (x2,y2,z2) so it is the image closest to
// (x1,y1,z1) in fractional terms (sys_tools.cpp:330-343).
__device__ __forceinline__ void image_shift(const Box& B, double x1, double y1, double z1, double x2, double y2,
                                            double z2, double& dx, double& dy, double& dz) {
    double u1, v1, w1, u2, v2, w2;
    to_frac(B, x1, y1, z1, u1, v1, w1);
    to_frac(B, x2, y2, z2, u2, v2, w2);
    dx = floor(u2 - u1 + 0.5);
    dy = floor(v2 - v1 + 0.5);
    dz = floor(w2 - w1 + 0.5);
}

// Apply an image shift to one atom (sys_tools.cpp:345-356).
__device__ __forceinline__ void image_apply(const Box& B, double dx, double dy, double dz, double& x, double& y,
                                            double& z) {
    double u, v, w;
    to_frac(B, x, y, z, u, v, w);
    u -= dx;
    v -= dy;
    w -= dz;
    to_cart(B, u, v, w, x, y, z);
}

// d0*d0 + d1*d1 + d2*d2 in the reference's order (kdtree_utils.h:63-69).
__device__ __forceinline__ double dist2(double ax, double ay, double az, double bx, double by, double bz) {
    const double d0 = ax - bx;
    const double d1 = ay - by;
    const double d2 = az - bz;
    return d0 * d0 + d1 * d1 + d2 * d2;
}

#pragma clang fp contract(on)

// MB-pol 2B nonlinear parameters (potential/2b/x2b-v9x.cpp:1330-1345).
namespace p2 {
constexpr double k_HH_intra = -6.480884773303821e-01;
constexpr double k_OH_intra = 1.674518993682975e+00;
constexpr double k_HH_coul = 1.148231864355956e+00;
constexpr double k_OH_coul = 1.205989761123099e+00;
constexpr double k_OO_coul = 1.395357065790959e+00;
constexpr double k_XH_main = 7.347036852042255e-01;
constexpr double k_XO_main = 7.998249864422826e-01;
constexpr double k_XX_main = 7.960663960630585e-01;
constexpr double in_plane_gamma = -9.721486914088159e-02;
constexpr double out_of_plane_gamma = 9.859272078406150e-02;
constexpr double r2i = 4.500000000000000e+00;
constexpr double r2f = 6.500000000000000e+00;
}  // namespace p2

// MB-pol 3B parameters (potential/3b/x3b-v2x.cpp:1256-1271).
namespace p3 {
constexpr double r3i = 0.000000000000000e+00;
constexpr double r3f = 4.500000000000000e+00;
constexpr double kHH_intra = -2.254303257872797e+00;
constexpr double kOH_intra = -2.564771901404151e-01;
constexpr double kHH = 4.247920544074333e-01;
constexpr double kOH = 8.128985941165371e-01;
constexpr double kOO = 3.749457984616480e-02;
constexpr double dHH_intra = 1.690594510379166e+00;
constexpr double dOH_intra = -2.999999868517452e+00;
constexpr double dHH = 3.499031358429095e+00;
constexpr double dOH = 4.854042963514281e+00;
constexpr double dOO = 4.816312044947604e-08;
}  // namespace p3

// Per-block partial sums written by the evaluation kernels: energy, then the six distinct
// virial components (xx, xy, xz, yy, yz, zz).
constexpr int kPartials = 7;

// Compile-time loop: f(std::integral_constant<int, J>) for J = 0..N-1, so table entries indexed
// by J are constants and per-lane arrays indexed through them stay in registers.
template <class F, int... J>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, J...>) {
    (f(std::integral_constant<int, J>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace mbxg
