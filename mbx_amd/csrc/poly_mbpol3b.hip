// MB-pol 3B polynomial stage (36 variables, 1163 coefficients, degree 4): the generated
// reverse-mode schedule of gen/poly_mbpol_3b.inc in the one-lane-per-trimer kernel of
// poly_kernel.hpp.  Restates poly_3b_v2x::eval (potential/3b/poly-3b-v2x.cpp:26982) and
// its energy-only form (potential/3b/poly-3b-v2x-nogrd.cpp:39).
#include "gen/poly_mbpol_3b.inc"
#include "mbxg_kernels.hpp"
#include "poly_kernel.hpp"

namespace mbxg {

hipError_t launch_poly_mbpol3b(const double* xs, const int* n_dev, int cap, int ld, const double* coef, int do_grads,
                               double* gs, double* ps, int P, hipStream_t st) {
    return launch_poly<36, 8, poly_mbpol_3b_parts, poly_mbpol_3b_e_parts>(xs, n_dev, cap, ld, coef, do_grads, gs,
                                                                                ps, P, st);
}

}  // namespace mbxg
