// Kernel interfaces shared between the kernel translation units and the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include "mbxg_device.hpp"

namespace mbxg {

// Candidate grid for the neighbour search (cells in fractional coordinates under PBC,
// Cartesian cells over the bounding box in the gas phase).
struct Grid {
    int n[3];
    int lo_off[3];
    int hi_off[3];
    int periodic;
    double lo[3];
    double inv_w[3];
};

// Everything an n-mer kernel needs to find a monomer's atoms.  Coordinates are the wrapped
// working copy (3 doubles per real atom, internal monomer order).  The box stays in device
// memory (read with scalar loads where needed) so it does not pin 48 SGPRs in the kernels.
struct NmerCtx {
    const double* xyz;
    const int* first_atom;
    const int* kind;
    const int* tag;
    const unsigned char* islocal;
    const int* members;  // cloud point -> monomer
    const Box* box;
    int periodic;
};

struct QueryCtx;

__global__ void k_fix_monomers(double* xyz, const int* first_atom, const int* nat, int nmon, Box B);
__global__ void k_cloud(const double* xyz, const int* first_atom, const int* members, int np, Box B, Grid G,
                        double* cx, double* cy, double* cz, int* cell);
__global__ void k_bin_count(const int* cell, int np, int* cell_count);
__global__ void k_scan_cells(int* cell_count, int ncell, int* cell_start);
__global__ void k_bin_fill(const int* cell, int np, const int* cell_start, int* cursor, int* cell_pts);
__global__ void k_bin_sort(const int* cell_start, int ncell, int* cell_pts);
__global__ void k_scan_counts(const int2* counts, int n, int2* offsets);

// Launch helpers (defined in lists.hip, which owns QueryCtx).
hipError_t launch_dimers(const double* cx, const double* cy, const double* cz, const int* cell,
                         const int* cell_start, const int* cell_pts, int np, int nslot, const Grid& G,
                         const NmerCtx& N, double c2, int pass, int2* counts, const int2* offsets, int* out_list,
                         int* out_eval, int cap, int* overflow, hipStream_t st);
hipError_t launch_trimers(const double* cx, const double* cy, const double* cz, const int* cell,
                          const int* cell_start, const int* cell_pts, int np, int nslot, const Grid& G,
                          const NmerCtx& N, double c2, int pass, int2* counts, const int2* offsets, int* out_list,
                          int* out_eval, int cap, int* overflow, hipStream_t st);

// Evaluation of the compacted "evaluate" lists (eval_mbpol2b.hip, eval_mbpol3b.hip): variable
// stage, polynomial stage (poly_mbpol2b.hip, poly_mbpol3b.hip), chain-rule stage.  One lane per
// n-mer, one wavefront per block.  list holds the evaluated n-mers (monomer indices), n of them;
// e_out[t] receives n-mer t's energy; partials[block*kPartials + ...] its block sums; ws is a
// workspace of workspace_mbpolXb(n) doubles.
size_t workspace_mbpol2b(int n);
size_t workspace_mbpol3b(int n);
// ev (nullable) receives 4 events: before the variable stage, before the polynomial, before the
// chain-rule stage, after it.
hipError_t launch_mbpol2b(const int* list, int n, const NmerCtx& N, const double* coef, int do_grads, double* grad,
                          double* e_out, double* partials, double* ws, hipStream_t st, hipEvent_t* ev);
hipError_t launch_mbpol3b(const int* list, int n, const NmerCtx& N, const double* coef, int do_grads, double* grad,
                          double* e_out, double* partials, double* ws, hipStream_t st, hipEvent_t* ev);
// Polynomial stage alone: xs/gs variable-major with leading dimension ld; ps values.
hipError_t launch_poly_mbpol2b(const double* xs, int n, int ld, const double* coef, int do_grads, double* gs,
                               double* ps, hipStream_t st);
hipError_t launch_poly_mbpol3b(const double* xs, int n, int ld, const double* coef, int do_grads, double* gs,
                               double* ps, hipStream_t st);

// Sum nblocks partial records into out[kPartials] in a fixed order (deterministic totals).
__global__ void k_reduce_partials(const double* partials, int nblocks, double* out);

}  // namespace mbxg
