// The MB-pol 3B and 2B polynomial stages (gen/poly_mbpol_3b.inc, gen/poly_mbpol_2b.inc: the same
// schedules as poly_mbpol3b.hip / poly_mbpol2b.hip) with gradients in one launch of
// k_poly_eg_pair: the 3B blocks first, then the 2B ones.  Restates poly_3b_v2x::eval
// (potential/3b/poly-3b-v2x.cpp:26982) and poly_2b_v6x::eval (potential/2b/poly-2b-v6x.cpp:12936).
// k_poly_eg_pair_chain3 also runs the 3B chain rule (x3b-v2x.cpp:1622-1726) in each 3B block's
// tail.
#include <algorithm>

#include "gen/poly_mbpol_2b.inc"
#include "gen/poly_mbpol_3b.inc"
#include "mbpol3b_chain.hpp"
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

namespace {

struct Chain3Job {
    const int* list;
    const double* sw;
    double* grad;
    double* e_out;
    double* vir_out;
    double* partials;
};

// entries of the chain rule's LDS table (the distinct monomers of 64 trimers: the query monomer
// and its neighbours; a full table falls back to device atomics)
constexpr int kChainTable3 = 128;

// The paired launch with the 3B chain rule in the 3B blocks' tail: once a block's gradient is
// summed and written (poly_eg_group), its wavefront 0 runs chain3 for the block's 64 trimers with
// an LDS table of its own in the block's reduction buffer (no longer read by then); the block's
// other wavefronts end there, and their slots take further polynomial blocks.  The chain rule's
// gathers and atomics so wait beside other wavefronts' polynomial arithmetic instead of in a
// kernel of their own.  The 2B blocks are unchanged.
template <int NVA, int QA, poly_eg_part_fn PA, int NVB, int QB, poly_eg_part_fn PB>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3, 3)))
    k_poly_eg_pair_chain3(PolyJob A, PolyJob B, int nb_a, NmerCtx N, Chain3Job C) {
    extern __shared__ double red[];
    if ((int)blockIdx.x >= nb_a) {
        poly_eg_group<NVB, QB, PB>(blockIdx.x - nb_a, B.xs, B.n_dev, B.cap, B.ld, B.coef_addr, B.gs, B.ps, red);
        return;
    }
    poly_eg_group<NVA, QA, PA>(blockIdx.x, A.xs, A.n_dev, A.cap, A.ld, A.coef_addr, A.gs, A.ps, red);
    const int n = min(*A.n_dev, A.cap);
    if ((threadIdx.x >> 6) != 0 || (int)blockIdx.x * 64 >= n) return;
    // this wavefront's gs / ps stores are complete before it loads them back
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // the variables are read back (their exponentials would be VALU work in a VALU-bound launch)
    chain3<true, true, kChainTable3, true>(blockIdx.x * 64 + (threadIdx.x & 63), blockIdx.x, n, C.list, N, A.xs, A.gs,
                                           A.ps, C.sw, A.ld, C.grad, C.e_out, C.vir_out, C.partials,
                                           *reinterpret_cast<GradTable<kChainTable3, 9>*>(red));
}

}  // namespace

hipError_t launch_poly_mbpol_pair_chain3(double* ws3, const int* list3, const int* n3_dev, int cap3,
                                         const double* coef3, double* ws2, const int* n2_dev, int cap2,
                                         const double* coef2, int split, const NmerCtx& N, double* grad,
                                         double* e3_out, double* vir3_out, double* partials3, hipStream_t st) {
    if (split != 1 && split != 2 && split != 4 && split != 8) return hipErrorInvalidValue;
    if (cap3 <= 0 || cap2 <= 0) return hipErrorInvalidValue;
    const int ld3 = (cap3 + 63) / 64 * 64, ld2 = (cap2 + 63) / 64 * 64;
    // workspace layout of launch_mbpol3b / launch_mbpol2b: xs [NV][ld], gs [NV][ld], ps [ld], sw [ld]
    const PolyJob A{ws3, n3_dev, cap3, ld3, (uint64_t)(uintptr_t)coef3, ws3 + (size_t)36 * ld3,
                    ws3 + (size_t)72 * ld3};
    const PolyJob B{ws2, n2_dev, cap2, ld2, (uint64_t)(uintptr_t)coef2, ws2 + (size_t)31 * ld2,
                    ws2 + (size_t)62 * ld2};
    const Chain3Job C{list3, ws3 + (size_t)73 * ld3, grad, e3_out, vir3_out, partials3};
    const int nb3 = ld3 / 64, nb2 = ld2 / 64;
    const size_t lds = std::max(poly_lds_bytes(split), sizeof(GradTable<kChainTable3, 9>));
    hipLaunchKernelGGL((k_poly_eg_pair_chain3<36, 8, poly_mbpol_3b_parts, 31, 8, poly_mbpol_2b_parts>),
                       dim3(nb3 + nb2), dim3(64 * split), lds, st, A, B, nb3, N, C);
    return hipGetLastError();
}

}  // namespace mbxg
