// Kernel interfaces shared between the kernel translation units and the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include "mbxg_device.hpp"

namespace mbxg {

// Candidate grid of one configuration.  Cell of a position y (fractional coordinates under
// PBC, where the grid spans the slots' extended range [-1, 1]^3; Cartesian in the gas phase,
// over the bounding box): floor((y - lo) * inv_w) per dimension, clamped to [0, n).  Every cell
// is at least one cutoff wide, so a query's matches lie in the 27 cells around its own.
struct Grid {
    int n[3];
    int periodic;
    double lo[3];
    double inv_w[3];
};

// Per cloud point p (a monomer, or an image of one): its first atom in the working coordinates
// (not imaged) and its tag and kind, so the n-mer kernel reads a match's data with one load.
struct PtInfo {
    double x, y, z;
    int tag, kind;
};

// Everything an n-mer kernel needs to find a monomer's atoms.  Coordinates are the wrapped
// working copy (3 doubles per real atom, internal monomer order).  The box stays in device
// memory (read with scalar loads where needed) so it does not pin 48 SGPRs in the kernels.
struct NmerCtx {
    const double* xyz;
    const int* first_atom;
    const int* nat;
    const int* kind;
    const int* tag;
    const unsigned char* islocal;
    const int* members;  // cloud point -> monomer
    const Box* box;
    int periodic;
};

// Records the message returned by mbxg_last_error and returns code (api.hip).
int report_error(int code, const char* msg);

// Cell grids: at most kMaxCellsPerDim cells per dimension (the gas-phase grid is sized on the
// device; the cell buffers are allocated for the maximum).
constexpr int kMaxCellsPerDim = 64;
constexpr int kMaxCells = kMaxCellsPerDim * kMaxCellsPerDim * kMaxCellsPerDim;

__global__ void k_fix_monomers(const double* in, double* out, const int* first_atom, const int* nat, int nmon, Box B);
__global__ void k_grid_gas(const double* xyz, const int* first_atom, const int* members, const int* pt_off,
                           const int* cell_bound, double cut, Grid* grids);
__global__ void k_cloud(const double* xyz, const int* first_atom, const int* members, const int* tag, const int* kind,
                        int np, Box B, const Grid* grids, const int* pconf, const int* cell_base, double* cx,
                        double* cy, double* cz, int* cell, PtInfo* pinfo);
// cell_pos[s] receives slot s's place among its cell's slots (the count before its own).
__global__ void k_bin_count(const int* cell, int ns, int* cell_count, int* cell_pos);
// Up to four device ranges to zero in one launch (null pointers skipped).
struct ZeroJobs {
    double* p[4];
    size_t n[4];
};
// k_zero over every range of Z, grid-stride.
__global__ void k_zero_multi(ZeroJobs Z);
// k_fix_monomers, k_cloud and k_bin_count in one launch of max(nmon, 8 np) threads (eight per
// cloud point, one slot pattern each), which also zero the ranges of Z.
__global__ void k_cloud_fix(const double* in, double* xyz, const int* first_atom, const int* nat, int nmon,
                            const int* members, const int* tag, const int* kind, int np, Box B, const Grid* grids,
                            const int* pconf, const int* cell_base, double* cx, double* cy, double* cz, int* cell,
                            PtInfo* pinfo, int* cell_count, int* cell_pos, ZeroJobs Z);
__global__ void k_scan_cells(int* cell_count, int ncell, int* cell_start);
__global__ void k_bin_fill(const int* cell, int ns, const int* cell_start, const int* cell_pos, const double* cx,
                           const double* cy, const double* cz, int* cell_slot, double* bx, double* by, double* bz);
__global__ void k_zero(double* p, size_t n);
__global__ void k_zero_int(int* p, size_t n);

// The n-mer kernel (lists.hip): one wavefront per query point; pass 0 counts, pass 1 writes.
// Points are grouped by configuration (pconf[p]); a configuration's cells start at
// cell_base[conf] and its grid is grids[conf], so lists never cross configurations.  Slots in
// cell order: cell_slot (slot index), bx/by/bz (coordinates), cell c holding
// [cell_start[c], cell_start[c+1]).
// Evaluated n-mer families (one evaluate list each): MB-pol 2B, MB-nrg CH4-H2O 2B, CH4-CH4 2B,
// MB-pol 3B, MB-nrg CH4-H2O-H2O 3B.
enum Family : int { FAM_MBPOL_2B = 0, FAM_CH4H2O_2B = 1, FAM_CH4CH4_2B = 2, FAM_MBPOL_3B = 3, FAM_CH4H2O2_3B = 4 };
constexpr int kFamilies = 5;
constexpr int family_order(int f) { return f < FAM_MBPOL_3B ? 2 : 3; }

// Wavefronts per query in the n-mer kernel; wpre / wpre_x hold each wavefront's prefix within
// its query (entry q * kQueryWaves + w), counts / offsets one entry per query.
constexpr int kQueryWaves = 4;

struct NmersArgs {
    const PtInfo* pinfo;
    const double* cx;
    const double* cy;
    const double* cz;
    const int* cell;
    const int* cell_start;
    const int* cell_slot;
    const double* bx;
    const double* by;
    const double* bz;
    const Grid* grids;
    const int* pconf;
    const int* cell_base;
    int np;
    double c2sq, c3sq;
    int4* counts;
    int4* counts_x;
    const int4* offsets;
    const int4* offsets_x;
    int4* wpre;
    int4* wpre_x;
    int* list2;
    int* list3;
    int* evals[kFamilies];
    int fam_mask;   // bit f: family f is evaluated (listed n-mers of a cleared family are not)
    // one-pass build (pass 2): wavefront w of query q stages its family-f n-mers at
    // stage[f] + (q * kQueryWaves + w) * scap[f] * order, at most scap[f] of them.  Every count
    // pass stores query q's largest per-wavefront counts in wmax[q] / wmax_x[q] (the layouts of
    // counts / counts_x; nullable)
    int* stage[kFamilies];
    int scap[kFamilies];
    int4* wmax;
    int4* wmax_x;
};

// out[0] / out[1] = componentwise maxima of wmax[0..n) / wmax_x[0..n) (one workgroup).
// The per-(query, wavefront) maxima of wmax / wmax_x into maxw[0..1] (skipped with wmax null) and
// the exclusive scans of counts and counts_x (.w maximised), totals at n, in one launch.
__global__ void k_scan_all(const int4* wmax, const int4* wmax_x, int4* maxw, const int4* counts, int4* offsets,
                           const int4* counts_x, int4* offsets_x, int n);

// The staged n-mers of a one-pass build, moved to the evaluate lists: block q (striding),
// wavefront w copies its family-f n-mers to evals[f] at the family's offset of query q plus the
// wavefront's prefix, bounded by fcap[f] (the list capacity in n-mers) and scap[f].
// The per-family energy/virial records of one compute (null partials: family not evaluated).
struct FinishJobs {
    const double* partials[kFamilies];
    const int* n_dev[kFamilies];
    int cap[kFamilies];
};
// One workgroup of kFinishThreads: each family's partial records summed in a fixed order into fam
// (kPartials per family, zeros for the families not evaluated), then the per-order records of
// totals as k_combine_families does (2B, 3B) and the 4B record zeroed.
constexpr int kFinishThreads = 512;
__global__ void k_finish(FinishJobs J, double* fam, double* totals);

struct CompactArgs {
    const int* stage[kFamilies];
    int scap[kFamilies];
    int* evals[kFamilies];
    int fcap[kFamilies];
};
__global__ void k_compact_evals(CompactArgs C, int np, const int4* counts, const int4* counts_x,
                                const int4* offsets, const int4* offsets_x, const int4* wpre, const int4* wpre_x);
// match_cap (a power of two) bounds the matches one query may hold.  With gscratch == nullptr
// they live in LDS (nmers_lds_bytes(match_cap) per block, one block per query); otherwise
// nblocks blocks stride over the queries, each with nmers_lds_bytes(match_cap) bytes of
// gscratch.
size_t nmers_lds_bytes(int match_cap);
hipError_t launch_nmers(const NmersArgs& A, const NmerCtx& N, int pass, int4 cap, int4 capx, int match_cap,
                        double* gscratch, int nblocks, hipStream_t st);

// Evaluation of the compacted "evaluate" lists (eval_mbpol2b.hip, eval_mbpol3b.hip): variable
// stage, polynomial stage (poly_mbpol2b.hip, poly_mbpol3b.hip), chain-rule stage.  One lane per
// n-mer.  The evaluated count n is read on the device from n_dev (clamped to cap): kernels are
// launched for cap n-mers and lanes past n retire at once, so the host never waits for the
// list builder.  e_out[t] receives n-mer t's energy and, when vir_out is not null,
// vir_out[q * ld + t] (q = xx xy xz yy yz zz, ld = cap rounded up to 64) its virial; partials
// one kPartials record per 64 n-mers (partials_records(cap) records); ws is a workspace of
// workspace_mbpolXb(cap) doubles.  ev (nullable) receives 4 events: before the variable stage, before the
// polynomial, before the chain-rule stage, after it.
inline size_t partials_records(int cap) { return ((size_t)cap + 127) / 128 * 2; }
size_t workspace_mbpol2b(int cap);
size_t workspace_mbpol3b(int cap);
hipError_t launch_mbpol2b(const int* list, const int* n_dev, int cap, const NmerCtx& N, const double* coef,
                          int do_grads, double* grad, double* e_out, double* vir_out, double* partials, double* ws,
                          int split, hipStream_t st, hipEvent_t* ev);
hipError_t launch_mbpol3b(const int* list, const int* n_dev, int cap, const NmerCtx& N, const double* coef,
                          int do_grads, double* grad, double* e_out, double* vir_out, double* partials, double* ws,
                          int split, hipStream_t st, hipEvent_t* ev);
// The variable and chain-rule stages alone (same workspace layout), and both MB-pol polynomial
// stages with gradients in one launch (poly_mbpol_pair.hip): the 3B blocks first, then the 2B
// ones, `split` wavefronts per 64 n-mers for both, so the short 2B blocks fill the machine while
// the last 3B blocks drain.
hipError_t launch_mbpol2b_vars(const int* list, const int* n_dev, int cap, const NmerCtx& N, double* ws,
                               hipStream_t st);
hipError_t launch_mbpol3b_vars(const int* list, const int* n_dev, int cap, const NmerCtx& N, double* ws,
                               hipStream_t st);
hipError_t launch_mbpol2b_chain(const int* list, const int* n_dev, int cap, const NmerCtx& N, int do_grads,
                                double* grad, double* e_out, double* vir_out, double* partials, double* ws,
                                hipStream_t st);
hipError_t launch_mbpol3b_chain(const int* list, const int* n_dev, int cap, const NmerCtx& N, int do_grads,
                                double* grad, double* e_out, double* vir_out, double* partials, double* ws,
                                hipStream_t st);
// 1B terms (eval_1b.hip): energies e_out[t] of monomers mons[0..n), and with g_out the gradient
// components of each (3 doubles per atom, written at its atoms) and vir_out[q * ld + t] (q = xx xy
// xz yy yz zz).  Water: MB-pol (Partridge-Schwenke); CH4: MB-nrg A1B4 with coefficients coef
// (ch4_1b_fit: the fitted ones).  k_reduce_1b sums n energies and virial rows into out[kPartials].
hipError_t launch_1b_ps(const double* xyz, const int* first_atom, const int* mons, int n, double* e_out,
                        double* g_out, double* vir_out, int ld, hipStream_t st);
hipError_t launch_1b_ch4(const double* xyz, const int* first_atom, const int* mons, int n, const double* coef,
                         double* e_out, double* g_out, double* vir_out, int ld, hipStream_t st);
const double* ch4_1b_fit(int* n);
// CO2 monomers (x1b_A1B2_v1x): the same outputs; co2_1b_fit: the fitted coefficients.
hipError_t launch_1b_co2(const double* xyz, const int* first_atom, const int* mons, int n, const double* coef,
                         double* e_out, double* g_out, double* vir_out, int ld, hipStream_t st);
const double* co2_1b_fit(int* n);
__global__ void k_reduce_1b(const double* e, const double* vir, int n, int ld, double* out);
hipError_t launch_poly_mbpol_pair(double* ws3, const int* n3_dev, int cap3, const double* coef3, double* ws2,
                                  const int* n2_dev, int cap2, const double* coef2, int split, hipStream_t st);
// launch_mbpol3b_vars and launch_mbpol2b_vars as one launch (workspaces as theirs).
hipError_t launch_mbpol_vars_pair(const int* list3, const int* n3_dev, int cap3, double* ws3, const int* list2,
                                  const int* n2_dev, int cap2, double* ws2, const NmerCtx& N, hipStream_t st);
// The same launch with the 3B chain rule (launch_mbpol3b_chain with gradients) in its tail:
// energies e3_out, gradient added to grad, virials vir3_out (nullable), partial records partials3.
hipError_t launch_poly_mbpol_pair_chain3(double* ws3, const int* list3, const int* n3_dev, int cap3,
                                         const double* coef3, double* ws2, const int* n2_dev, int cap2,
                                         const double* coef2, int split, const NmerCtx& N, double* grad,
                                         double* e3_out, double* vir3_out, double* partials3, hipStream_t st);
// MB-nrg families (eval_nrg.hip): the same three stages for FAM_CH4H2O_2B, FAM_CH4CH4_2B and
// FAM_CH4H2O2_3B; list entries in emission order (the kernels order each n-mer's monomers
// for its holder).  `split` (1, 2, 4 or 8) is the wavefronts per 64 n-mers of the blocked
// CH4-H2O-H2O polynomial; the pieced 2B polynomials ignore it.
size_t workspace_nrg_family(int family, int cap);
hipError_t launch_nrg_family(int family, const int* list, const int* n_dev, int cap, const NmerCtx& N,
                             const double* coef, int do_grads, double* grad, double* e_out, double* vir_out,
                             double* partials, double* ws, int split, hipStream_t st, hipEvent_t* ev);

// Polynomial stage alone: xs/gs variable-major with leading dimension ld (= cap rounded up
// to 64); ps values; n read from n_dev as above; `split` wavefronts (1, 2, 4 or 8) share each
// 64 n-mers (poly_kernel.hpp).
hipError_t launch_poly_mbpol2b(const double* xs, const int* n_dev, int cap, int ld, const double* coef, int do_grads,
                               double* gs, double* ps, int split, hipStream_t st);
hipError_t launch_poly_mbpol3b(const double* xs, const int* n_dev, int cap, int ld, const double* coef, int do_grads,
                               double* gs, double* ps, int split, hipStream_t st);
hipError_t launch_poly_ch4h2o2_3b(const double* xs, const int* n_dev, int cap, int ld, const double* coef,
                                  int do_grads, double* gs, double* ps, int split, hipStream_t st);

// Sum the partial records of min(*n_dev, n_cap) n-mers into out[kPartials] in a fixed order
// (deterministic totals).
__global__ void k_reduce_partials(const double* partials, const int* n_dev, int n_cap, double* out);

// Per-configuration totals in two fixed-order levels: block (c, s) of a (nconf, S) grid sums
// slice s of configuration c's evaluated n-mers, [offsets[pt_off[c]].comp,
// offsets[pt_off[c+1]].comp) clamped to cap, from e_out / vir_out (leading dimension ld) into
// part[c][s][kPartials] (energy, 6 virial components; zeros when vir_out is null); comp selects
// the int4 component (0..3) holding the family's count.  k_reduce_slices then adds the S
// slices of each configuration in order.
__global__ void k_reduce_configs(const double* e_out, const double* vir_out, int ld, const int4* offsets,
                                 const int* pt_off, int comp, int cap, double* part);
__global__ void k_reduce_slices(const double* part, int nconf, int nsl, double* out);
__global__ void k_combine_families(const double* fam, int nconf, double* out);

}  // namespace mbxg
