// The MB-pol 3B and 2B polynomial stages (gen/poly_mbpol_3b.inc, gen/poly_mbpol_2b.inc: the same
// schedules as poly_mbpol3b.hip / poly_mbpol2b.hip) with gradients in one launch of
// k_poly_eg_pair: the 3B blocks first, then the 2B ones.  Restates poly_3b_v2x::eval
// (potential/3b/poly-3b-v2x.cpp:26982) and poly_2b_v6x::eval (potential/2b/poly-2b-v6x.cpp:12936).
#include "gen/poly_mbpol_2b.inc"
#include "gen/poly_mbpol_3b.inc"
#include "mbxg_kernels.hpp"
#include "poly_kernel.hpp"

namespace mbxg {

hipError_t launch_poly_mbpol_pair(double* ws3, const int* n3_dev, int cap3, const double* coef3, double* ws2,
                                  const int* n2_dev, int cap2, const double* coef2, int split, hipStream_t st) {
    if (split != 1 && split != 2 && split != 4 && split != 8) return hipErrorInvalidValue;
    if (cap3 <= 0 || cap2 <= 0) return hipErrorInvalidValue;
    const int ld3 = (cap3 + 63) / 64 * 64, ld2 = (cap2 + 63) / 64 * 64;
    // workspace layout of launch_mbpol3b / launch_mbpol2b: xs [NV][ld], gs [NV][ld], ps [ld], sw [ld]
    const PolyJob A{ws3, n3_dev, cap3, ld3, (uint64_t)(uintptr_t)coef3, ws3 + (size_t)36 * ld3,
                    ws3 + (size_t)72 * ld3};
    const PolyJob B{ws2, n2_dev, cap2, ld2, (uint64_t)(uintptr_t)coef2, ws2 + (size_t)31 * ld2,
                    ws2 + (size_t)62 * ld2};
    const int nb3 = ld3 / 64, nb2 = ld2 / 64;
    hipLaunchKernelGGL((k_poly_eg_pair<36, 8, poly_mbpol_3b_parts, 31, 8, poly_mbpol_2b_parts>), dim3(nb3 + nb2),
                       dim3(64 * split), poly_lds_bytes(split), st, A, B, nb3);
    return hipGetLastError();
}

}  // namespace mbxg
