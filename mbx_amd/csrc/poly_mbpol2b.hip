// MB-pol 2B polynomial stage (31 variables, 1153 coefficients, degree 4): the generated
// reverse-mode schedule of gen/poly_mbpol_2b.inc in the one-lane-per-dimer kernel of
// poly_kernel.hpp.  Restates poly_2b_v6x::eval (potential/2b/poly-2b-v6x.cpp:12936) and its
// energy-only form (potential/2b/poly-2b-v6x-nogrd.cpp:41).
#include "gen/poly_mbpol_2b.inc"
#include "mbxg_kernels.hpp"
#include "poly_kernel.hpp"

namespace mbxg {

hipError_t launch_poly_mbpol2b(const double* xs, const int* n_dev, int cap, int ld, const double* coef, int do_grads,
                               double* gs, double* ps, int P, hipStream_t st) {
    return launch_poly<31, 8, poly_mbpol_2b_parts, poly_mbpol_2b_e_parts>(xs, n_dev, cap, ld, coef, do_grads, gs,
                                                                                ps, P, st);
}

}  // namespace mbxg
