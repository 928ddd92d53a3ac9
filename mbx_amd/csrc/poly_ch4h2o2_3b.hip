// CH4-H2O-H2O 3B polynomial stage (55 variables, 1254 coefficients, degree 3): the generated
// reverse-mode schedule of gen/poly_ch4h2o2_3b.inc in the blocked one-lane-per-trimer kernel of
// poly_kernel.hpp, two waves per SIMD.  Restates poly_A1B4_C1D2_C1D2_deg3_v1::eval (E+G,
// potential/3b/poly_3b_A1B4_C1D2_C1D2_deg3_grad_v1.cpp:39; E only,
// potential/3b/poly_3b_A1B4_C1D2_C1D2_deg3_nograd_v1.cpp:39).
#include "gen/poly_ch4h2o2_3b.inc"
#include "mbxg_kernels.hpp"
#include "poly_kernel.hpp"

namespace mbxg {

hipError_t launch_poly_ch4h2o2_3b(const double* xs, const int* n_dev, int cap, int ld, const double* coef,
                                  int do_grads, double* gs, double* ps, int P, hipStream_t st) {
    return launch_poly<55, 8, poly_ch4h2o2_3b_parts, poly_ch4h2o2_3b_e_parts, 2>(xs, n_dev, cap, ld, coef, do_grads,
                                                                                   gs, ps, P, st);
}

}  // namespace mbxg
