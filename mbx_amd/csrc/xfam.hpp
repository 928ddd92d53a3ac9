// Further MB-nrg families (xfam.hip): halide-water 2B / 3B, CO2 2B / 3B, water 4B,
// alkali-metal-ion-water 2B / 3B.
#pragma once

#include <hip/hip_runtime.h>

#include "mbxg_kernels.hpp"

namespace mbxg {

// Families, in the order of their bits in the evaluated-family mask (bit 5 + f of
// mbxg_set_families) and of their energy records.
enum XFam : int {
    XF_F_H2O_2B = 0,
    XF_CL_H2O_2B = 1,
    XF_BR_H2O_2B = 2,
    XF_I_H2O_2B = 3,
    XF_CO2CO2_2B = 4,
    XF_F_H2O2_3B = 5,
    XF_CL_H2O2_3B = 6,
    XF_BR_H2O2_3B = 7,
    XF_I_H2O2_3B = 8,
    XF_CO2X3_3B = 9,
    XF_H2OX4_4B = 10,
    XF_LI_H2O_2B = 11,
    XF_NA_H2O_2B = 12,
    XF_K_H2O_2B = 13,
    XF_RB_H2O_2B = 14,
    XF_CS_H2O_2B = 15,
    XF_LI_H2O2_3B = 16,
    XF_NA_H2O2_3B = 17,
    XF_K_H2O2_3B = 18,
    XF_RB_H2O2_3B = 19,
    XF_CS_H2O2_3B = 20
};
constexpr int kXFam = 21;

// Evaluate lists of the families: entries of order + 1 ints (monomer ids in the holder's order,
// then the holder slot of the listed first monomer); at most cap[f] entries each.
struct XLists {
    int* list[kXFam];
    int cap[kXFam];
};

__host__ __device__ int xfam_order(int f);
size_t xfam_workspace(int f, int cap);  // doubles
const double* xfam_default_coef(int f, int* n);
size_t xclass_blocks(int n);

// Classification of n listed n-mers of one order: code[n], per-block counts blk / offsets
// blk_off ([xclass_blocks(n)][kXFam]) and the family totals tot[kXFam], all on the device.
hipError_t launch_xclass(int order, const int* list, int n, const NmerCtx& N, int fam_mask, int* code, int* blk,
                         int* blk_off, int* tot, hipStream_t st);
hipError_t launch_xemit(int order, const int* list, int n, const int* code, const int* blk_off, const XLists& X,
                        hipStream_t st);
// One family over its evaluate list: energies e_out, one kPartials record per 64 n-mers
// (partials_records(cap)), gradient added to grad; ws holds xfam_workspace(f, cap) doubles.
hipError_t launch_xfam(int f, const int* list, const int* n_dev, int cap, const NmerCtx& N, const double* coef,
                       double* grad, double* e_out, double* partials, double* ws, hipStream_t st);
// The further families' energy records (kXFam x kPartials) added to the per-order totals
// (records 2B, 3B, 1B, 4B).
hipError_t launch_xcombine(const double* xfam_tot, double* totals, hipStream_t st);
// Tetramers of the cloud (pass 0: counts and offsets[np + 1]; pass 1: out, 4 ints each, at most
// cap); *ovf receives the largest match count of a query that overflowed the match storage.
hipError_t launch_tetra(const PtInfo* pinfo, const double* cx, const double* cy, const double* cz, int np, int ns,
                        const NmerCtx& N, double c4sq, int pass, int* counts, int* offsets, int* out, int cap,
                        int* ovf, hipStream_t st);

}  // namespace mbxg
