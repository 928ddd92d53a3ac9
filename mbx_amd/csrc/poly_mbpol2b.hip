// MB-pol 2B polynomial stage (31 variables, 1153 coefficients, degree 4): the generated
// reverse-mode schedule of gen/poly_mbpol_2b.inc in the one-lane-per-dimer kernel of
// poly_kernel.hpp.  Restates poly_2b_v6x::eval (potential/2b/poly-2b-v6x.cpp:12936) and its
// energy-only form (potential/2b/poly-2b-v6x-nogrd.cpp:41).
#include "gen/poly_mbpol_2b.inc"
#include "mbxg_kernels.hpp"
#include "poly_kernel.hpp"

namespace mbxg {

hipError_t launch_poly_mbpol2b(const double* xs, const int* n_dev, int cap, int ld, const double* coef, int do_grads,
                               double* gs, double* ps, hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    const int nb = (cap + 63) / 64;
    const uint64_t ca = (uint64_t)(uintptr_t)coef;
    if (do_grads)
        hipLaunchKernelGGL((k_poly_eg<31, poly_mbpol_2b>), dim3(nb), dim3(64), 0, st, xs, n_dev, cap, ld, ca, gs, ps);
    else
        hipLaunchKernelGGL((k_poly_e<31, poly_mbpol_2b_e>), dim3(nb), dim3(64), 0, st, xs, n_dev, cap, ld, ca, ps);
    return hipGetLastError();
}

}  // namespace mbxg
