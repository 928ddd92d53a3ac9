// Cell-list dimer / trimer builder for gfx950.
//
// Produces the same n-mer sets as the reference's KD-tree enumeration
// (systools::AddClusters, bblock/sys_tools.cpp:472-641, over kdtutils::XyzToCloud,
// kdtree/kdtree_utils.h:196-279, with nanoflann's strict radius test
// kdtree/nanoflann.hpp:166-171):
//   * every local monomer contributes its first atom; under PBC the point is re-wrapped and
//     receives 7 image slots (slot s of point p lives at index p + s*np);
//   * a dimer (i, j) is listed from query i when some slot of j lies strictly inside the 2B
//     cutoff of i's point and tag[i] < tag[j];
//   * a trimer (i, j, k) is listed from query i when slots of j and k both lie inside the 3B
//     cutoff, j != k != i, and either |slot_j - slot_k|^2 > c^2 or tag[i] is the smallest.
// Every distance is computed from the same cloud coordinates with the same operation order,
// so membership decisions are bit-identical.  The KD-tree is replaced by a grid of cells
// (fractional coordinates under PBC, Cartesian over the bounding box in the gas phase), at
// least one cutoff wide; the grid only proposes candidates, every candidate slot is tested.
//
// One wavefront per query monomer gathers its matches at the larger of the two cutoffs once,
// sorts them by slot index in LDS (so the emission order does not depend on how points landed
// in cells), and precomputes per match the slot coordinates, the partner's first atom imaged
// next to the query's (GetCloseDimerImage, sys_tools.cpp:322-365) and its tag / kind.  Dimers
// and trimers are then decided from LDS only.  While emitting, each n-mer is also tested
// against its polynomial's own evaluation cutoff (x2b-v9x.cpp:1589-1600, x3b-v2x.cpp:1495-1515);
// those that pass go to a compacted "evaluate" list as well, so no lane of the polynomial
// kernels idles on a skipped n-mer.
//
// Lists are emitted in two passes (count, exclusive scan, write at offsets); in steady state
// the host does not wait between them: writes are bounded by the buffer capacities and the
// host checks the totals when it next synchronises (api.hip).
//
// This file is compiled with -ffp-contract=off.
#include "mbxg_device.hpp"
#include "mbxg_kernels.hpp"
#include "nrg_kernel.hpp"
#include "gen/holder_ch4ch4_2b.inc"
#include "gen/holder_ch4h2o2_3b.inc"
#include "gen/holder_ch4h2o_2b.inc"

namespace mbxg {

// ---------------------------------------------------------------------------------------------
// FixMonomerCoordinates (sys_tools.cpp:287-319): wrap each monomer's first atom into the cell
// and bring its other atoms next to it.  Reads the caller's coordinates, writes the working
// copy (a plain copy in the gas phase).  One thread per monomer.
// The first atom's working position (periodic): its fractional coordinates u0 v0 w0 wrapped
// into [-1/2, 1/2) and y0 = their Cartesian position.
__device__ __forceinline__ void fixed_first(const Box& B, const double* x, double& u0, double& v0, double& w0,
                                            double* y0) {
    to_frac(B, x[0], x[1], x[2], u0, v0, w0);
    u0 -= floor(u0 + 0.5);
    v0 -= floor(v0 + 0.5);
    w0 -= floor(w0 + 0.5);
    to_cart(B, u0, v0, w0, y0[0], y0[1], y0[2]);
}

__device__ __forceinline__ void fix_monomer(int i, const double* __restrict__ in, double* __restrict__ out,
                                            const int* __restrict__ first_atom, const int* __restrict__ nat,
                                            const Box& B) {
    const double* x = in + 3 * (size_t)first_atom[i];
    double* y = out + 3 * (size_t)first_atom[i];
    const int na = nat[i];
    if (!B.periodic) {
        for (int k = 0; k < 3 * na; ++k) y[k] = x[k];
        return;
    }
    double u0, v0, w0;
    fixed_first(B, x, u0, v0, w0, y);
    for (int a = 1; a < na; ++a) {
        double u, v, w;
        to_frac(B, x[3 * a], x[3 * a + 1], x[3 * a + 2], u, v, w);
        u -= floor(u - u0 + 0.5);
        v -= floor(v - v0 + 0.5);
        w -= floor(w - w0 + 0.5);
        to_cart(B, u, v, w, y[3 * a], y[3 * a + 1], y[3 * a + 2]);
    }
}

__global__ void k_fix_monomers(const double* __restrict__ in, double* __restrict__ out,
                               const int* __restrict__ first_atom, const int* __restrict__ nat, int nmon, Box B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nmon) fix_monomer(i, in, out, first_atom, nat, B);
}

// ---------------------------------------------------------------------------------------------
// Gas-phase grids: one workgroup per configuration computes the bounding box of its points
// [pt_off[c], pt_off[c+1]) and a grid of cells at least `cut` wide, with at most
// min(kMaxCellsPerDim^3, cell_bound[c]) cells (coarser cells stay correct: every cell is still
// at least one cutoff wide).
__global__ void __launch_bounds__(256) k_grid_gas(const double* __restrict__ xyz, const int* __restrict__ first_atom,
                                                  const int* __restrict__ members, const int* __restrict__ pt_off,
                                                  const int* __restrict__ cell_bound, double cut,
                                                  Grid* __restrict__ grids) {
    __shared__ double lo[3][256], hi[3][256];
    const int t = threadIdx.x;
    const int c0 = pt_off[blockIdx.x], c1 = pt_off[blockIdx.x + 1];
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    for (int p = c0 + t; p < c1; p += 256) {
        const double* x = xyz + 3 * (size_t)first_atom[members[p]];
        for (int c = 0; c < 3; ++c) {
            l[c] = fmin(l[c], x[c]);
            h[c] = fmax(h[c], x[c]);
        }
    }
    for (int c = 0; c < 3; ++c) {
        lo[c][t] = l[c];
        hi[c][t] = h[c];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int c = 0; c < 3; ++c) {
                lo[c][t] = fmin(lo[c][t], lo[c][t + s]);
                hi[c][t] = fmax(hi[c][t], hi[c][t + s]);
            }
        __syncthreads();
    }
    if (t == 0) {
        const bool any = c1 > c0;
        Grid g{};
        double ext[3];
        for (int c = 0; c < 3; ++c) {
            ext[c] = any ? hi[c][0] - lo[c][0] : 0.0;
            int n = (int)floor(ext[c] / cut);
            g.n[c] = n < 1 ? 1 : (n > kMaxCellsPerDim ? kMaxCellsPerDim : n);
        }
        const long bound = cell_bound[blockIdx.x];
        while ((long)g.n[0] * g.n[1] * g.n[2] > bound) {
            const int d = g.n[0] >= g.n[1] && g.n[0] >= g.n[2] ? 0 : (g.n[1] >= g.n[2] ? 1 : 2);
            g.n[d] -= 1;
        }
        for (int c = 0; c < 3; ++c) {
            g.lo[c] = any ? lo[c][0] : 0.0;
            g.inv_w[c] = ext[c] > 0 ? g.n[c] / ext[c] : 0.0;
        }
        g.periodic = 0;
        grids[blockIdx.x] = g;
    }
}

// ---------------------------------------------------------------------------------------------
// XyzToCloud (kdtree_utils.h:196-279): point p = first atom of cloud member p.  Slot 0 is the
// re-wrapped position; slots 1..7 are the images selected from its fractional coordinates.
// Every one of the 7 slots is written: each fractional coordinate admits at least one of the
// two image directions.
//
// Every slot is also given a cell of its configuration's grid.  Under PBC the grid spans the
// extended fractional range [-1, 1]^3 that the slots occupy (a slot's fractional position is
// its point's plus the image shift), so a query only visits slots near it instead of all 8
// slots of the points near it.  Cell index = cell_base[config] + local cell.  k_cloud runs one
// thread per point (cloud_point), k_cloud_fix eight, one slot pattern each (cloud_pattern).
// Slot pattern pat of point p (pat 0: the point's own slot; pat 1..7: the image whose non-zero
// directions are pat's bits m, n, l): emit(slot, x, y, z, cell) when the slot exists (always for
// pat 0; under PBC for the admissible images; never for images in the gas phase).  The patterns
// map to distinct slots (pat -> shift is one to one), so they may run in any order or in parallel.
// pinfo (nullable) receives the point's data from pattern 0.
template <class Emit>
__device__ __forceinline__ void cloud_pattern(int p, int pat, const double* q, int mp, const int* __restrict__ tag,
                                              const int* __restrict__ kind, int np, const Box& B,
                                              const Grid* __restrict__ grids, const int* __restrict__ pconf,
                                              const int* __restrict__ cell_base, PtInfo* __restrict__ pinfo,
                                              Emit emit) {
    const int conf = pconf[p];
    const Grid G = grids[conf];
    const int base = cell_base[conf];
    auto cell_of = [&](double a, double b, double c) {
        int ca = (int)floor((a - G.lo[0]) * G.inv_w[0]);
        int cb = (int)floor((b - G.lo[1]) * G.inv_w[1]);
        int cc = (int)floor((c - G.lo[2]) * G.inv_w[2]);
        ca = ca < 0 ? 0 : (ca >= G.n[0] ? G.n[0] - 1 : ca);
        cb = cb < 0 ? 0 : (cb >= G.n[1] ? G.n[1] - 1 : cb);
        cc = cc < 0 ? 0 : (cc >= G.n[2] ? G.n[2] - 1 : cc);
        return base + (cc * G.n[1] + cb) * G.n[0] + ca;
    };
    double qx = q[0], qy = q[1], qz = q[2];
    if (pat == 0 && pinfo) pinfo[p] = PtInfo{qx, qy, qz, tag[mp], kind[mp]};
    if (!B.periodic) {
        if (pat == 0) emit(p, qx, qy, qz, cell_of(qx, qy, qz));
        return;
    }
    double u, v, w;
    to_frac(B, qx, qy, qz, u, v, w);
    u -= floor(u + 0.5);
    v -= floor(v + 0.5);
    w -= floor(w + 0.5);
    to_cart(B, u, v, w, qx, qy, qz);
    to_frac(B, qx, qy, qz, u, v, w);  // cell placement only
    if (pat == 0) {
        emit(p, qx, qy, qz, cell_of(u, v, w));
        return;
    }
    // GetFractionalCoordinates (kdtree_utils.h:186-189)
    const double uvw0 = B.f00 * qx - B.f01 * qy + B.f02 * qz;
    const double uvw1 = B.f11 * qy + B.f12 * qz;
    const double uvw2 = B.f22 * qz;
    // XyzToCloud's loops over m, n, l in {0, -1, 1} (kdtree_utils.h:190-225) write image
    // slot `shift` once per admissible combination of its non-zero directions, the last one
    // kept.  Per dimension that is +1 when admissible, else -1: a coordinate of exactly 0
    // admits both, and the slot must still be emitted (and counted) once.
    auto dir = [](double f) { return !((f + 1) > 1.0 || (f + 1) < -1.0) ? 1 : (!((f - 1) > 1.0 || (f - 1) < -1.0) ? -1 : 0); };
    auto zero_ok = [](double f) { return !(f > 1.0 || f < -1.0); };
    const int m = (pat & 1) ? dir(uvw0) : 0, n = (pat & 2) ? dir(uvw1) : 0, l = (pat & 4) ? dir(uvw2) : 0;
    if ((pat & 1) ? !m : !zero_ok(uvw0)) return;
    if ((pat & 2) ? !n : !zero_ok(uvw1)) return;
    if ((pat & 4) ? !l : !zero_ok(uvw2)) return;
    int shift;
    if (m == 0 && n == 0)
        shift = 1;
    else if (m == 0 && l == 0)
        shift = 2;
    else if (n == 0 && l == 0)
        shift = 3;
    else if (m != 0 && n != 0 && l != 0)
        shift = 4;
    else if (m == 0)
        shift = 5;
    else if (n == 0)
        shift = 6;
    else
        shift = 7;
    const double sx = (double)m * B.b[0] + (double)n * B.b[3] + (double)l * B.b[6];
    const double sy = (double)m * B.b[1] + (double)n * B.b[4] + (double)l * B.b[7];
    const double sz = (double)m * B.b[2] + (double)n * B.b[5] + (double)l * B.b[8];
    emit(p + shift * np, qx + sx, qy + sy, qz + sz, cell_of(u + m, v + n, w + l));
}

// Every slot of point p (the image patterns, then the point's own slot), one thread.
template <class Emit>
__device__ __forceinline__ void cloud_point(int p, const double* q, int mp, const int* __restrict__ tag,
                                            const int* __restrict__ kind, int np, const Box& B,
                                            const Grid* __restrict__ grids, const int* __restrict__ pconf,
                                            const int* __restrict__ cell_base, PtInfo* __restrict__ pinfo,
                                            Emit emit) {
#pragma unroll 1
    for (int pat = 1; pat < 8; ++pat)
        cloud_pattern(p, pat, q, mp, tag, kind, np, B, grids, pconf, cell_base, nullptr, emit);
    cloud_pattern(p, 0, q, mp, tag, kind, np, B, grids, pconf, cell_base, pinfo, emit);
}

__global__ void k_cloud(const double* __restrict__ xyz, const int* __restrict__ first_atom,
                        const int* __restrict__ members, const int* __restrict__ tag, const int* __restrict__ kind,
                        int np, Box B, const Grid* __restrict__ grids, const int* __restrict__ pconf,
                        const int* __restrict__ cell_base, double* __restrict__ cx, double* __restrict__ cy,
                        double* __restrict__ cz, int* __restrict__ cell, PtInfo* __restrict__ pinfo) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int mp = members[p];
    const double* x0 = xyz + 3 * (size_t)first_atom[mp];
    const double q[3] = {x0[0], x0[1], x0[2]};
    cloud_point(p, q, mp, tag, kind, np, B, grids, pconf, cell_base, pinfo,
                [&](int s, double x, double y, double z, int c) {
                    cx[s] = x;
                    cy[s] = y;
                    cz[s] = z;
                    cell[s] = c;
                });
}

__device__ __forceinline__ void zero_jobs(const ZeroJobs& Z, size_t tid, size_t nthreads) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (Z.p[k])
            for (size_t j = tid; j < Z.n[k]; j += nthreads) Z.p[k][j] = 0.0;
}

__global__ void k_zero_multi(ZeroJobs Z) {
    zero_jobs(Z, (size_t)blockIdx.x * blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x);
}

// FixMonomerCoordinates, the cloud and the cell counts in one launch (one thread per monomer and
// cloud point): thread i fixes monomer i (into xyz) and emits point i's slots (as k_cloud) with
// its own copy of the point's fixed first atom, so no thread reads another's output; each slot
// adds one to its cell's count (zero on entry: k_scan_cells leaves it zero).
__global__ void k_cloud_fix(const double* __restrict__ in, double* __restrict__ xyz,
                            const int* __restrict__ first_atom, const int* __restrict__ nat, int nmon,
                            const int* __restrict__ members, const int* __restrict__ tag,
                            const int* __restrict__ kind, int np, Box B, const Grid* __restrict__ grids,
                            const int* __restrict__ pconf, const int* __restrict__ cell_base,
                            double* __restrict__ cx, double* __restrict__ cy, double* __restrict__ cz,
                            int* __restrict__ cell, PtInfo* __restrict__ pinfo, int* __restrict__ cell_count,
                            int* __restrict__ cell_pos, ZeroJobs Z) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    zero_jobs(Z, g, (size_t)gridDim.x * blockDim.x);
    if (g < nmon) fix_monomer(g, in, xyz, first_atom, nat, B);
    // eight lanes per cloud point, one slot pattern each (the grid covers max(nmon, 8 np))
    const int i = g >> 3, pat = g & 7;
    if (i >= np) return;
    const int mp = members[i];
    const double* x0 = in + 3 * (size_t)first_atom[mp];
    double q[3] = {x0[0], x0[1], x0[2]};
    if (B.periodic) {
        double u0, v0, w0;
        fixed_first(B, x0, u0, v0, w0, q);
    }
    cloud_pattern(i, pat, q, mp, tag, kind, np, B, grids, pconf, cell_base, pinfo,
                  [&](int s, double x, double y, double z, int c) {
                      cx[s] = x;
                      cy[s] = y;
                      cz[s] = z;
                      cell[s] = c;
                      cell_pos[s] = atomicAdd(&cell_count[c], 1);
                  });
}

// ---------------------------------------------------------------------------------------------
// Binning of slots into cells: count, scan, fill.  The fill also writes the slots' coordinates
// in cell order, so the gather reads them contiguously.  A slot's place inside its cell is the
// count its cell-count atomic returned (cell_pos), so the order inside a cell is left to the
// atomics; the n-mer kernel sorts each query's matches, which makes the output order independent
// of it.  cell_count is zero on entry: zeroed once when the cell layout is made, and left zero by
// every scan.
__global__ void k_bin_count(const int* __restrict__ cell, int ns, int* __restrict__ cell_count,
                            int* __restrict__ cell_pos) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns) cell_pos[s] = atomicAdd(&cell_count[cell[s]], 1);
}

// Single-workgroup exclusive scan of ncell cell counts; cell_start gets ncell+1 entries, and
// cell_count is left zero for the next build.
__global__ void __launch_bounds__(1024) k_scan_cells(int* __restrict__ cell_count, int ncell,
                                                     int* __restrict__ cell_start) {
    __shared__ int sh[1024];
    const int t = threadIdx.x;
    const int per = (ncell + 1023) / 1024;
    const int beg = t * per, end = min(ncell, beg + per);
    int acc = 0;
    for (int c = beg; c < end; ++c) acc += cell_count[c];
    sh[t] = acc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int o = sh[t] - acc;
    for (int c = beg; c < end; ++c) {
        cell_start[c] = o;
        o += cell_count[c];
        cell_count[c] = 0;
    }
    if (t == 1023) cell_start[ncell] = sh[t];
}

__global__ void k_bin_fill(const int* __restrict__ cell, int ns, const int* __restrict__ cell_start,
                           const int* __restrict__ cell_pos, const double* __restrict__ cx,
                           const double* __restrict__ cy, const double* __restrict__ cz, int* __restrict__ cell_slot,
                           double* __restrict__ bx, double* __restrict__ by, double* __restrict__ bz) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int at = cell_start[cell[s]] + cell_pos[s];
    cell_slot[at] = s;
    bx[at] = cx[s];
    by[at] = cy[s];
    bz[at] = cz[s];
}

// ---------------------------------------------------------------------------------------------
// The n-mer kernel: one wavefront (block) per query point.

namespace {

struct Lds {
    int* match;  // slot index, sorted
    int* mon;    // partner monomer
    int* tag;    // partner tag
    int* kind;   // partner kind
    int* flags;  // bit0: inside 2B cutoff, bit1: inside 3B cutoff
    int* m3;     // indices of the matches inside the 3B cutoff, in match order
    double* sx;  // slot coordinates
    double* sy;
    double* sz;
    double* ix;  // partner's first atom imaged next to the query's
    double* iy;
    double* iz;
    double* r;  // |query first atom - imaged partner first atom|
    double* p3;  // the matches inside the 3B cutoff packed in m3 order, kP3Doubles each: sx sy sz
                 // ix iy iz r, then int4 {mon, tag, kind, 0} at double 8 (five 16-byte reads)
};

// 7 doubles + 6 ints per match, and one packed 3B entry of 10 doubles.
constexpr int kP3Doubles = 10;
constexpr int kMatchDoubles = 10 + kP3Doubles;

__device__ __forceinline__ Lds carve(double* d, int cap) {
    Lds L;
    L.sx = d;
    L.sy = d + cap;
    L.sz = d + 2 * cap;
    L.ix = d + 3 * cap;
    L.iy = d + 4 * cap;
    L.iz = d + 5 * cap;
    L.r = d + 6 * cap;
    int* i = (int*)(d + 7 * cap);
    L.match = i;
    L.mon = i + cap;
    L.tag = i + 2 * cap;
    L.kind = i + 3 * cap;
    L.flags = i + 4 * cap;
    L.m3 = i + 5 * cap;
    L.p3 = d + 10 * cap;  // 16-byte aligned: cap is a multiple of 64
    return L;
}

}  // namespace

// One query, by the block's kQueryWaves wavefronts: gather, sort, per-match data, dimers,
// trimers.  L is the block's match storage (LDS, or a global scratch slice when the matches do
// not fit in LDS); sh_count and sh_w are block-shared.  The dimers and trimers of a query are
// enumerated in a fixed order and cut into kQueryWaves contiguous ranges, one per wavefront; the
// count pass stores each wavefront's prefix within the query (wpre, entry q * kQueryWaves + w)
// beside the query's counts, so the lists keep one deterministic order.
__device__ void nmers_query(int q, const PtInfo* __restrict__ pinfo, const double* __restrict__ cx,
                            const double* __restrict__ cy,
                            const double* __restrict__ cz, const int* __restrict__ cell,
                            const int* __restrict__ cell_start, const int* __restrict__ cell_slot,
                            const double* __restrict__ bx, const double* __restrict__ by,
                            const double* __restrict__ bz, int np, const Grid& G, int base, const NmerCtx& N,
                            double c2sq, double c3sq, int pass, int4* __restrict__ counts,
                            int4* __restrict__ counts_x, const int4* __restrict__ offsets,
                            const int4* __restrict__ offsets_x, int* __restrict__ list2, int* __restrict__ list3,
                            int* const* evals, int fam_mask, int4 cap, int4 capx, int match_cap,
                            const Lds& L, int* sh_count, int4* sh_w,
                            int4* __restrict__ wpre, int4* __restrict__ wpre_x, int* const* stage,
                            const int* scap, int4* __restrict__ wmax, int4* __restrict__ wmax_x) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int mi = N.members[q];
    if (!N.islocal[mi]) {
        if (pass != 1 && tid == 0) {
            counts[q] = make_int4(0, 0, 0, 0);
            counts_x[q] = make_int4(0, 0, 0, 0);
            if (wmax) wmax[q] = wmax_x[q] = make_int4(0, 0, 0, 0);
        }
        return;
    }
    const double cmsq = c2sq > c3sq ? c2sq : c3sq;

    // ---- gather: every slot strictly inside the larger cutoff, from the 27 cells around the
    // query's (clamped at the grid's faces).  The (up to) three cells along a of each of the 9
    // rows are contiguous in cell order; the rows' ranges are read first and the candidates of
    // all rows enumerated as one sequence, two chunks of 64 per step so each step's loads are
    // in flight together.
    const double qx = cx[q], qy = cy[q], qz = cz[q];
    const int ci = cell[q] - base;
    const int ia = ci % G.n[0], ib = (ci / G.n[0]) % G.n[1], ic = ci / (G.n[0] * G.n[1]);
    int rbeg[9], rlen[9];
    int nrow = 0;
    for (int cc = max(ic - 1, 0); cc <= min(ic + 1, G.n[2] - 1); ++cc) {
        for (int cb = max(ib - 1, 0); cb <= min(ib + 1, G.n[1] - 1); ++cb) {
            const int row = base + (cc * G.n[1] + cb) * G.n[0];
            rbeg[nrow] = row + max(ia - 1, 0);
            rlen[nrow] = row + min(ia + 1, G.n[0] - 1) + 1;
            ++nrow;
        }
    }
    int ncand = 0;
    for (int r = 0; r < 9; ++r) {
        if (r < nrow) {
            const int b0 = cell_start[rbeg[r]], e0 = cell_start[rlen[r]];
            rbeg[r] = b0;
            rlen[r] = e0 - b0;
            ncand += e0 - b0;
        }
    }
    // candidate c -> slot-array position (row by row, in order)
    auto position = [&](int c) {
        int at = -1;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            if (r < nrow && at < 0) {
                if (c < rlen[r])
                    at = rbeg[r] + c;
                else
                    c -= rlen[r];
            }
        }
        return at;
    };
    if (tid == 0) *sh_count = 0;
    __syncthreads();
    // each wavefront takes a contiguous share of the candidates, issues the loads of up to four
    // chunks of 64 before testing any, and appends its matches at a block-shared cursor (the
    // sort fixes the order)
    const int per_w = ((ncand + 63) / 64 + kQueryWaves - 1) / kQueryWaves;
    const int cbeg = wv * per_w * 64;
    for (int k0 = 0; k0 < per_w; k0 += 4) {
        double px[4], py[4], pz[4];
        int ps[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = cbeg + (k0 + k) * 64 + lane;
            const int pos = (k0 + k < per_w && c < ncand) ? position(c) : -1;
            ps[k] = -1;
            px[k] = py[k] = pz[k] = 0.0;
            if (pos >= 0) {
                px[k] = bx[pos];
                py[k] = by[pos];
                pz[k] = bz[pos];
                ps[k] = cell_slot[pos];
            }
        }
        unsigned long long m[4];
        int tot = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (ps[k] >= 0 && !(dist2(qx, qy, qz, px[k], py[k], pz[k]) < cmsq)) ps[k] = -1;
            m[k] = __ballot(ps[k] >= 0);
            tot += __popcll(m[k]);
        }
        int at0 = 0;
        if (lane == 0 && tot > 0) at0 = atomicAdd(sh_count, tot);
        at0 = __shfl(at0, 0, 64);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int pk = at0 + __popcll(m[k] & ((1ull << lane) - 1ull));
            if (ps[k] >= 0 && pk < match_cap) L.match[pk] = ps[k];
            at0 += __popcll(m[k]);
        }
    }
    __syncthreads();
    const int nm = *sh_count;
    if (nm > match_cap) {
        // The matches did not fit.  Which of them reached storage depends on the wavefronts'
        // timing, so the count and emit passes could keep different subsets and disagree on this
        // query's n-mers; it emits none in either pass.  The caller sees the overflow and
        // recomputes with more storage.
        if (pass != 1 && tid == 0) {
            counts[q] = make_int4(0, 0, 0, 0);
            counts_x[q] = make_int4(0, 0, 0, nm);
            if (wmax) wmax[q] = wmax_x[q] = make_int4(0, 0, 0, 0);
        }
        return;
    }
    // ---- sort the matches by slot index (distinct keys): each key's rank is the number of
    // smaller keys (every thread reads the same LDS word at a time); L.mon is free until the
    // per-match data are written
    for (int a = tid; a < nm; a += blockDim.x) {
        const int key = L.match[a];
        int rank = 0, b = 0;
        // 16 keys a step, four 16-byte reads in flight (L.match is 16-byte aligned: the match
        // storage holds a multiple of 64 matches)
        for (; b + 16 <= nm; b += 16) {
            const int4 k0 = *(const int4*)(L.match + b), k1 = *(const int4*)(L.match + b + 4);
            const int4 k2 = *(const int4*)(L.match + b + 8), k3 = *(const int4*)(L.match + b + 12);
            rank += (k0.x < key) + (k0.y < key) + (k0.z < key) + (k0.w < key) + (k1.x < key) + (k1.y < key) +
                    (k1.z < key) + (k1.w < key) + (k2.x < key) + (k2.y < key) + (k2.z < key) + (k2.w < key) +
                    (k3.x < key) + (k3.y < key) + (k3.z < key) + (k3.w < key);
        }
        for (; b < nm; ++b) rank += L.match[b] < key;
        L.mon[rank] = key;
    }
    __syncthreads();
    for (int a = tid; a < nm; a += blockDim.x) L.match[a] = L.mon[a];
    __syncthreads();

    // ---- per-match data (one payload load per match: PtInfo of its cloud point)
    const double* oi = N.xyz + 3 * (size_t)N.first_atom[mi];
    const double oix = oi[0], oiy = oi[1], oiz = oi[2];
    const int ti = N.tag[mi], ki = N.kind[mi];
    for (int a = tid; a < nm; a += blockDim.x) {
        const int s = L.match[a];
        const int p = s % np;
        const int mj = N.members[p];
        const PtInfo P = pinfo[p];
        const double x = cx[s], y = cy[s], z = cz[s];
        const double d = dist2(qx, qy, qz, x, y, z);
        L.sx[a] = x;
        L.sy[a] = y;
        L.sz[a] = z;
        L.mon[a] = mj;
        L.tag[a] = P.tag;
        L.kind[a] = P.kind;
        L.flags[a] = (d < c2sq ? 1 : 0) | (d < c3sq ? 2 : 0);
        double jx = P.x, jy = P.y, jz = P.z;
        if (N.periodic) {
            const Box& B = *N.box;
            double dx, dy, dz;
            image_shift(B, oix, oiy, oiz, jx, jy, jz, dx, dy, dz);
            image_apply(B, dx, dy, dz, jx, jy, jz);
        }
        L.ix[a] = jx;
        L.iy[a] = jy;
        L.iz[a] = jz;
        const double d0 = oix - jx, d1 = oiy - jy, d2 = oiz - jz;
        double rr = 0.0;
        rr += d0 * d0;
        rr += d1 * d1;
        rr += d2 * d2;
        L.r[a] = sqrt(rr);
    }
    __syncthreads();
    // compact the 3B matches (order preserved; wavefront 0)
    if (wv == 0) {
        int n3w = 0;
        for (int b = 0; b < nm; b += 64) {
            const int a = b + lane;
            const bool in3 = a < nm && (L.flags[a] & 2);
            const unsigned long long m = __ballot(in3);
            if (in3) L.m3[n3w + __popcll(m & ((1ull << lane) - 1ull))] = a;
            n3w += __popcll(m);
        }
        if (lane == 0) *sh_count = n3w;
    }
    __syncthreads();
    const int n3 = *sh_count;
    // the trimer loop reads each 3B match's data as one packed entry
    for (int u = tid; u < n3; u += blockDim.x) {
        const int a = L.m3[u];
        double* e = L.p3 + (size_t)kP3Doubles * u;
        *(double2*)e = make_double2(L.sx[a], L.sy[a]);
        *(double2*)(e + 2) = make_double2(L.sz[a], L.ix[a]);
        *(double2*)(e + 4) = make_double2(L.iy[a], L.iz[a]);
        *(double2*)(e + 6) = make_double2(L.r[a], 0.0);
        *(int4*)(e + 8) = make_int4(L.mon[a], L.tag[a], L.kind[a], 0);
    }
    __syncthreads();

    const unsigned long long below = (1ull << lane) - 1ull;
    int4 at = make_int4(0, 0, 0, 0), atx = make_int4(0, 0, 0, 0);
    if (pass == 1) {  // this wavefront's base: the query's offset plus the lower wavefronts' counts
        const int4 o = offsets[q], ox = offsets_x[q];
        const int4 w = wpre[q * kQueryWaves + wv], wx = wpre_x[q * kQueryWaves + wv];
        at = make_int4(o.x + w.x, o.y + w.y, o.z + w.z, o.w + w.w);
        atx = make_int4(ox.x + wx.x, ox.y + wx.y, ox.z + wx.z, ox.w + wx.w);
    }
    // this wavefront's range of a sequence of n items: whole chunks of 64, in order
    auto range = [&](int n, int& b0, int& b1) {
        const int chunks = (n + 63) / 64, per = (chunks + kQueryWaves - 1) / kQueryWaves;
        b0 = min(n, wv * per * 64);
        b1 = min(n, (wv + 1) * per * 64);
    };
    // evaluate-list bases and capacities per family (FAM_*)
    const int fbase[kFamilies] = {at.y, atx.x, atx.y, at.w, atx.z};
    const int fcap[kFamilies] = {cap.y, capx.x, capx.y, cap.w, capx.z};
    int nl2 = 0, nl3 = 0;
    int ne[kFamilies] = {0, 0, 0, 0, 0};

    // ---- dimers (AddClusters sys_tools.cpp:558-567).  Evaluation rules: MB-pol r_OO <= 6.5
    // (x2b-v9x.cpp:1589-1600); the MB-nrg holders skip r > r2f = 9 on the first atoms
    // (x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp:840, x2b_A1B4_A1B4_deg4_exp0_v1x.cpp:1489).
    if (c2sq > 0.0) {
        int b0, b1;
        range(nm, b0, b1);
        for (int b = b0; b < b1; b += 64) {
            const int a = b + lane;
            bool listed = false;
            int mj = 0, fam = -1;
            if (a < b1 && (L.flags[a] & 1)) {
                mj = L.mon[a];
                listed = ti < L.tag[a];
                if (listed) {
                    const int kj = L.kind[a];
                    const double r = L.r[a];
                    if (ki == MON_H2O && kj == MON_H2O) {
                        if (!(r > p2::r2f)) fam = FAM_MBPOL_2B;
                    } else if ((ki == MON_H2O && kj == MON_CH4) || (ki == MON_CH4 && kj == MON_H2O)) {
                        if (!(r > Holder_ch4h2o_2b::kROut)) fam = FAM_CH4H2O_2B;
                    } else if (ki == MON_CH4 && kj == MON_CH4) {
                        if (!(r > Holder_ch4ch4_2b::kROut)) fam = FAM_CH4CH4_2B;
                    }
                }
            }
            if (fam >= 0 && !((fam_mask >> fam) & 1)) fam = -1;
            const unsigned long long ml = __ballot(listed);
            if (pass == 1 && listed) {
                const int w = at.x + nl2 + __popcll(ml & below);
                if (w < cap.x) {
                    list2[2 * w] = mi;
                    list2[2 * w + 1] = mj;
                }
            }
            nl2 += __popcll(ml);
#pragma unroll
            for (int f = FAM_MBPOL_2B; f <= FAM_CH4CH4_2B; ++f) {
                const unsigned long long me = __ballot(fam == f);
                if (pass == 1 && fam == f) {
                    const int w = fbase[f] + ne[f] + __popcll(me & below);
                    if (w < fcap[f]) {
                        evals[f][2 * w] = mi;
                        evals[f][2 * w + 1] = mj;
                    }
                } else if (pass == 2 && fam == f) {
                    const int w = ne[f] + __popcll(me & below);
                    if (w < scap[f]) {
                        int* o = stage[f] + ((size_t)(q * kQueryWaves + wv) * scap[f] + w) * 2;
                        o[0] = mi;
                        o[1] = mj;
                    }
                }
                ne[f] += __popcll(me);
            }
        }
    }

    // ---- trimers (AddClusters sys_tools.cpp:570-592).  Evaluation rules: MB-pol skips a trimer
    // when any two O-O distances exceed 4.5 (x3b-v2x.cpp:1495-1515); the CH4-H2O-H2O holder
    // only when all three first-atom distances exceed 6 (mbnrg_3b_A1B4_C1D2_C1D2_deg3_v1.cpp:1364);
    // other mixtures have no 3B term (energy3b.cpp:258-260).
    if (c3sq > 0.0) {
        // pairs u < v of the 3B matches in (u, v) order: pair t lies in row u, which starts at
        // S(u) = u (2 n3 - u - 1) / 2 (a floating-point first guess, corrected exactly)
        const int total = n3 * (n3 - 1) / 2;
        int b0, b1;
        range(total, b0, b1);
        for (int b = b0; b < b1; b += 64) {
            const int t = b + lane;
            bool listed = false;
            int mj = 0, mk = 0, fam = -1;
            if (t < b1) {
                const double w = 2.0 * n3 - 1.0;
                int u = (int)((w - sqrt(w * w - 8.0 * t)) * 0.5);
                u = u < 0 ? 0 : (u > n3 - 2 ? n3 - 2 : u);
                while (u > 0 && u * (2 * n3 - u - 1) / 2 > t) --u;
                while (u < n3 - 2 && (u + 1) * (2 * n3 - u - 2) / 2 <= t) ++u;
                const int v = t - u * (2 * n3 - u - 1) / 2 + u + 1;
                {
                    const double* ej = L.p3 + (size_t)kP3Doubles * u;
                    const double* ek = L.p3 + (size_t)kP3Doubles * v;
                    const int4 ij = *(const int4*)(ej + 8), ik = *(const int4*)(ek + 8);
                    mj = ij.x;
                    mk = ik.x;
                    if (mj != mi && mk != mi && mj != mk) {
                        const double2 j01 = *(const double2*)ej, k01 = *(const double2*)ek;
                        const double2 j23 = *(const double2*)(ej + 2), k23 = *(const double2*)(ek + 2);
                        const double djk = dist2(j01.x, j01.y, j23.x, k01.x, k01.y, k23.x);
                        listed = djk > c3sq || (ti < ij.y && ti < ik.y);
                        const int nch4 = (ki == MON_CH4) + (ij.z == MON_CH4) + (ik.z == MON_CH4);
                        const int nh2o = (ki == MON_H2O) + (ij.z == MON_H2O) + (ik.z == MON_H2O);
                        if (listed && (nh2o == 3 || (nh2o == 2 && nch4 == 1))) {
                            const double2 j45 = *(const double2*)(ej + 4), k45 = *(const double2*)(ek + 4);
                            const double e0 = j23.y - k23.y, e1 = j45.x - k45.x, e2 = j45.y - k45.y;
                            double rr = 0.0;
                            rr += e0 * e0;
                            rr += e1 * e1;
                            rr += e2 * e2;
                            const double rab = ej[6], rac = ek[6], rbc = sqrt(rr);
                            if (nh2o == 3) {
                                const double f = p3::r3f;
                                if (!((rab > f && rac > f) || (rab > f && rbc > f) || (rac > f && rbc > f)))
                                    fam = FAM_MBPOL_3B;
                            } else {
                                const double f = Holder_ch4h2o2_3b::kROut;
                                if (!(rab > f && rac > f && rbc > f)) fam = FAM_CH4H2O2_3B;
                            }
                        }
                    }
                }
            }
            if (fam >= 0 && !((fam_mask >> fam) & 1)) fam = -1;
            const unsigned long long ml = __ballot(listed);
            if (pass == 1 && listed) {
                const int w = at.z + nl3 + __popcll(ml & below);
                if (w < cap.z) {
                    list3[3 * w] = mi;
                    list3[3 * w + 1] = mj;
                    list3[3 * w + 2] = mk;
                }
            }
            nl3 += __popcll(ml);
#pragma unroll
            for (int f = FAM_MBPOL_3B; f <= FAM_CH4H2O2_3B; ++f) {
                const unsigned long long me = __ballot(fam == f);
                if (pass == 1 && fam == f) {
                    const int w = fbase[f] + ne[f] + __popcll(me & below);
                    if (w < fcap[f]) {
                        evals[f][3 * w] = mi;
                        evals[f][3 * w + 1] = mj;
                        evals[f][3 * w + 2] = mk;
                    }
                } else if (pass == 2 && fam == f) {
                    const int w = ne[f] + __popcll(me & below);
                    if (w < scap[f]) {
                        int* o = stage[f] + ((size_t)(q * kQueryWaves + wv) * scap[f] + w) * 3;
                        o[0] = mi;
                        o[1] = mj;
                        o[2] = mk;
                    }
                }
                ne[f] += __popcll(me);
            }
        }
    }
    if (pass != 1) {  // the query's counts, and each wavefront's prefix within the query
        if (lane == 0) {
            sh_w[wv] = make_int4(nl2, ne[FAM_MBPOL_2B], nl3, ne[FAM_MBPOL_3B]);
            sh_w[kQueryWaves + wv] = make_int4(ne[FAM_CH4H2O_2B], ne[FAM_CH4CH4_2B], ne[FAM_CH4H2O2_3B], 0);
        }
        __syncthreads();
        if (tid == 0) {
            int4 a = make_int4(0, 0, 0, 0), ax = make_int4(0, 0, 0, 0);
            for (int w = 0; w < kQueryWaves; ++w) {
                wpre[q * kQueryWaves + w] = a;
                wpre_x[q * kQueryWaves + w] = ax;
                const int4 c = sh_w[w], cx = sh_w[kQueryWaves + w];
                a = make_int4(a.x + c.x, a.y + c.y, a.z + c.z, a.w + c.w);
                ax = make_int4(ax.x + cx.x, ax.y + cx.y, ax.z + cx.z, ax.w + cx.w);
            }
            counts[q] = a;
            counts_x[q] = make_int4(ax.x, ax.y, ax.z, nm);  // .w: the query's match count
            if (wmax) {
                int4 m = sh_w[0], mx = sh_w[kQueryWaves];
                for (int w = 1; w < kQueryWaves; ++w) {
                    const int4 c = sh_w[w], cx = sh_w[kQueryWaves + w];
                    m = make_int4(max(m.x, c.x), max(m.y, c.y), max(m.z, c.z), max(m.w, c.w));
                    mx = make_int4(max(mx.x, cx.x), max(mx.y, cx.y), max(mx.z, cx.z), max(mx.w, cx.w));
                }
                wmax[q] = m;
                wmax_x[q] = mx;
            }
        }
    }
}

// pass 0: counts[q] = {listed 2B, evaluated MB-pol 2B, listed 3B, evaluated MB-pol 3B},
// counts_x[q] = {evaluated CH4-H2O 2B, CH4-CH4 2B, CH4-H2O-H2O 3B, 0};
// pass 1: writes at offsets[q] / offsets_x[q], bounded by cap / capx (same orders).
// counts_x[q].w receives query q's match count (the scan maximises it: the host compares it with
// match_cap).  Blocks stride over the queries; with kLds the match storage is the dynamic LDS,
// otherwise block b uses gscratch + b * match_cap * kMatchDoubles doubles.  The two storages are
// separate instantiations so that the LDS one addresses its storage as LDS (ds_* instructions)
// rather than through generic pointers.
template <bool kLds>
__global__ void __launch_bounds__(64 * kQueryWaves) __attribute__((amdgpu_waves_per_eu(5)))
    k_nmers(NmersArgs A, NmerCtx N, int pass, int4 cap, int4 capx, int match_cap, double* __restrict__ gscratch) {
    extern __shared__ __attribute__((aligned(16))) double smem[];
    __shared__ int sh_count;
    __shared__ int4 sh_w[2 * kQueryWaves];
    double* store = kLds ? smem : gscratch + (size_t)blockIdx.x * match_cap * kMatchDoubles;
    const Lds L = carve(store, match_cap);
    for (int q = blockIdx.x; q < A.np; q += gridDim.x) {
        const int conf = A.pconf[q];
        const Grid G = A.grids[conf];
        nmers_query(q, A.pinfo, A.cx, A.cy, A.cz, A.cell, A.cell_start, A.cell_slot, A.bx, A.by, A.bz, A.np, G,
                    A.cell_base[conf], N, A.c2sq, A.c3sq, pass, A.counts, A.counts_x, A.offsets, A.offsets_x,
                    A.list2, A.list3, A.evals, A.fam_mask, cap, capx, match_cap, L, &sh_count, sh_w,
                    A.wpre, A.wpre_x, A.stage, A.scap, A.wmax, A.wmax_x);
        __syncthreads();
    }
}

// A family's component of the per-query count records: counts = {listed 2B, MB-pol 2B, listed
// 3B, MB-pol 3B}, counts_x = {CH4-H2O 2B, CH4-CH4 2B, CH4-H2O-H2O 3B, matches}.
__device__ __forceinline__ int fam_count(int f, const int4& c, const int4& cx) {
    return f == FAM_MBPOL_2B ? c.y : f == FAM_CH4H2O_2B ? cx.x : f == FAM_CH4CH4_2B ? cx.y : f == FAM_MBPOL_3B ? c.w : cx.z;
}

__global__ void __launch_bounds__(64 * kQueryWaves)
    k_compact_evals(CompactArgs C, int np, const int4* __restrict__ counts, const int4* __restrict__ counts_x,
                    const int4* __restrict__ offsets, const int4* __restrict__ offsets_x,
                    const int4* __restrict__ wpre, const int4* __restrict__ wpre_x) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = blockIdx.x; q < np; q += gridDim.x) {
        const int4 c = counts[q], cx = counts_x[q];
#pragma unroll
        for (int f = 0; f < kFamilies; ++f) {
            const int total = fam_count(f, c, cx);
            if (total <= 0) continue;  // wpre is only written for queries with n-mers
            const int order = family_order(f);
            const int k = q * kQueryWaves + wv;
            const int cur = fam_count(f, wpre[k], wpre_x[k]);
            const int next = wv + 1 < kQueryWaves ? fam_count(f, wpre[k + 1], wpre_x[k + 1]) : total;
            const int base = fam_count(f, offsets[q], offsets_x[q]) + cur;
            const int fit = min(next - cur, C.fcap[f] - base);  // n-mers of this wavefront the list holds
            const int staged = min(fit, C.scap[f]);
            const int* src = C.stage[f] + (size_t)k * C.scap[f] * order;
            int* dst = C.evals[f] + (size_t)base * order;
            // n-mers past the staging capacity were not kept: their entries name monomer 0, so
            // the evaluation reads valid coordinates (the caller sees maxw above the capacity and
            // recomputes)
            for (int i = lane; i < fit * order; i += 64) dst[i] = i < staged * order ? src[i] : 0;
        }
    }
}

// The three reductions after the n-mer count pass in one launch: the per-wavefront maxima
// (wmax may be null: skipped), the exclusive scan of counts and the scan of counts_x (.w
// maximised: running maxima, the total at n the largest).  Each thread takes a run of
// consecutive entries; runs are scanned within each wavefront by lane shuffles and across the
// block's 16 wavefronts through LDS.
namespace {

__device__ __forceinline__ int4 combine4(int4 a, int4 b, bool max_w) {
    return make_int4(a.x + b.x, a.y + b.y, a.z + b.z, max_w ? max(a.w, b.w) : a.w + b.w);
}

__device__ __forceinline__ int4 shfl_up4(int4 v, int o) {
    return make_int4(__shfl_up(v.x, o, 64), __shfl_up(v.y, o, 64), __shfl_up(v.z, o, 64), __shfl_up(v.w, o, 64));
}

}  // namespace

__global__ void __launch_bounds__(1024) k_scan_all(const int4* __restrict__ wmax, const int4* __restrict__ wmax_x,
                                                   int4* __restrict__ maxw, const int4* __restrict__ counts,
                                                   int4* __restrict__ offsets, const int4* __restrict__ counts_x,
                                                   int4* __restrict__ offsets_x, int n) {
    // Both scans and the two maxima in one pass: every thread's loads are issued together, and one
    // barrier separates the per-wavefront sums from their use.
    __shared__ int4 wsum[2][16];
    __shared__ int4 msum[2][16];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int per = (n + 1023) / 1024;
    const int beg = min(n, t * per), end = min(n, beg + per);
    const int4 zero = make_int4(0, 0, 0, 0);
    auto max4 = [](int4 a, int4 b) { return make_int4(max(a.x, b.x), max(a.y, b.y), max(a.z, b.z), max(a.w, b.w)); };
    int4 acc = zero, accx = zero, m = zero, mx = zero;
    for (int i = beg; i < end; ++i) {
        acc = combine4(acc, counts[i], false);
        accx = combine4(accx, counts_x[i], true);
    }
    if (wmax)
        for (int i = t; i < n; i += 1024) {
            m = max4(m, wmax[i]);
            mx = max4(mx, wmax_x[i]);
        }
    int4 incl = acc, inclx = accx;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int4 y = shfl_up4(incl, o), yx = shfl_up4(inclx, o);
        if (lane >= o) {
            incl = combine4(y, incl, false);
            inclx = combine4(yx, inclx, true);
        }
    }
    int4 excl = shfl_up4(incl, 1), exclx = shfl_up4(inclx, 1);
    if (lane == 0) excl = exclx = zero;
    if (wmax) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            m = make_int4(max(m.x, __shfl_xor(m.x, o, 64)), max(m.y, __shfl_xor(m.y, o, 64)),
                          max(m.z, __shfl_xor(m.z, o, 64)), max(m.w, __shfl_xor(m.w, o, 64)));
            mx = make_int4(max(mx.x, __shfl_xor(mx.x, o, 64)), max(mx.y, __shfl_xor(mx.y, o, 64)),
                           max(mx.z, __shfl_xor(mx.z, o, 64)), max(mx.w, __shfl_xor(mx.w, o, 64)));
        }
    }
    if (lane == 63) {
        wsum[0][wv] = incl;
        wsum[1][wv] = inclx;
    }
    if (lane == 0) {
        msum[0][wv] = m;
        msum[1][wv] = mx;
    }
    __syncthreads();
    int4 base = zero, basex = zero;
    for (int w = 0; w < wv; ++w) {
        base = combine4(base, wsum[0][w], false);
        basex = combine4(basex, wsum[1][w], true);
    }
    int4 o = combine4(base, excl, false), ox = combine4(basex, exclx, true);
    for (int i = beg; i < end; ++i) {
        offsets[i] = o;
        offsets_x[i] = ox;
        o = combine4(o, counts[i], false);
        ox = combine4(ox, counts_x[i], true);
    }
    if (t == 1023) {
        int4 tot = zero, totx = zero;
        for (int w = 0; w < 16; ++w) {
            tot = combine4(tot, wsum[0][w], false);
            totx = combine4(totx, wsum[1][w], true);
        }
        offsets[n] = tot;
        offsets_x[n] = totx;
    }
    if (wmax && t < 2) {
        int4 r = msum[t][0];
        for (int w = 1; w < 16; ++w) r = max4(r, msum[t][w]);
        maxw[t] = r;
    }
}

// Zero fills (graph-capturable replacements for memset nodes).
__global__ void k_zero(double* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}
__global__ void k_zero_int(int* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Deterministic sum of per-wave partial records (kPartials doubles each) into out[kPartials]:
// one workgroup, fixed strided accumulation and a fixed tree.  The record count comes from the
// evaluated n-mer count on the device (ceil(n / 64), n clamped to the capacity).
__global__ void __launch_bounds__(256) k_reduce_partials(const double* __restrict__ partials,
                                                         const int* __restrict__ n_dev, int n_cap,
                                                         double* __restrict__ out) {
    __shared__ double sh[kPartials][256];
    const int t = threadIdx.x;
    int n = *n_dev;
    n = n < n_cap ? n : n_cap;
    const int nb = (n + 63) / 64;
    double acc[kPartials];
    for (int q = 0; q < kPartials; ++q) acc[q] = 0.0;
    for (int b = t; b < nb; b += 256)
        for (int q = 0; q < kPartials; ++q) acc[q] += partials[(size_t)b * kPartials + q];
    for (int q = 0; q < kPartials; ++q) sh[q][t] = acc[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < kPartials; ++q) sh[q][t] += sh[q][t + s];
        __syncthreads();
    }
    if (t < kPartials) out[t] = sh[t][0];
}

__global__ void __launch_bounds__(256) k_reduce_configs(const double* __restrict__ e_out,
                                                        const double* __restrict__ vir_out, int ld,
                                                        const int4* __restrict__ offsets,
                                                        const int* __restrict__ pt_off, int comp, int cap,
                                                        double* __restrict__ part) {
    __shared__ double sh[kPartials][256];
    const int c = blockIdx.x, sl = blockIdx.y, nsl = gridDim.y, t = threadIdx.x;
    const int4 o0 = offsets[pt_off[c]], o1 = offsets[pt_off[c + 1]];
    auto pick = [comp](const int4& o) { return comp == 0 ? o.x : comp == 1 ? o.y : comp == 2 ? o.z : o.w; };
    const int cb = min(pick(o0), cap), ce = min(pick(o1), cap);
    const int chunk = (ce - cb + nsl - 1) / nsl;
    const int beg = cb + sl * chunk, end = min(ce, beg + chunk);
    double acc[kPartials];
    for (int q = 0; q < kPartials; ++q) acc[q] = 0.0;
    for (int i = beg + t; i < end; i += 256) {
        acc[0] += e_out[i];
        if (vir_out)
            for (int q = 0; q < 6; ++q) acc[1 + q] += vir_out[(size_t)q * ld + i];
    }
    for (int q = 0; q < kPartials; ++q) sh[q][t] = acc[q];
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (t < s2)
            for (int q = 0; q < kPartials; ++q) sh[q][t] += sh[q][t + s2];
        __syncthreads();
    }
    if (t < kPartials) part[((size_t)c * nsl + sl) * kPartials + t] = sh[t][0];
}

// Second level: out[c * kPartials + q] = sum over slices s (in order) of part[c][s][q].
__global__ void k_reduce_slices(const double* __restrict__ part, int nconf, int nsl, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nconf * kPartials) return;
    const int c = i / kPartials, q = i % kPartials;
    double acc = 0.0;
    for (int s = 0; s < nsl; ++s) acc += part[((size_t)c * nsl + s) * kPartials + q];
    out[i] = acc;
}

// Per-order records from per-family ones (kPartials doubles per record, nconf configurations):
// out[0][c] = fam[MB-pol 2B][c] + fam[CH4-H2O 2B][c] + fam[CH4-CH4 2B][c],
// out[1][c] = fam[MB-pol 3B][c] + fam[CH4-H2O-H2O 3B][c].
__global__ void __launch_bounds__(kFinishThreads) k_finish(FinishJobs J, double* __restrict__ fam, double* __restrict__ totals) {
    __shared__ double sh[kPartials][kFinishThreads];
    __shared__ double res[kFamilies][kPartials];
    const int t = threadIdx.x;
    for (int f = 0; f < kFamilies; ++f) {
        if (!J.partials[f]) {  // uniform: the family was not evaluated
            if (t < kPartials) res[f][t] = 0.0;
            continue;
        }
        int n = *J.n_dev[f];
        n = n < J.cap[f] ? n : J.cap[f];
        const int nb = (n + 63) / 64;
        double acc[kPartials];
        for (int q = 0; q < kPartials; ++q) acc[q] = 0.0;
        for (int b = t; b < nb; b += kFinishThreads)
            for (int q = 0; q < kPartials; ++q) acc[q] += J.partials[f][(size_t)b * kPartials + q];
        for (int q = 0; q < kPartials; ++q) sh[q][t] = acc[q];
        __syncthreads();
        for (int s = kFinishThreads / 2; s > 0; s >>= 1) {
            if (t < s)
                for (int q = 0; q < kPartials; ++q) sh[q][t] += sh[q][t + s];
            __syncthreads();
        }
        if (t < kPartials) res[f][t] = sh[t][0];
        __syncthreads();
    }
    __syncthreads();
    if (t < kPartials) {
        for (int f = 0; f < kFamilies; ++f) fam[f * kPartials + t] = res[f][t];
        totals[t] = res[FAM_MBPOL_2B][t] + res[FAM_CH4H2O_2B][t] + res[FAM_CH4CH4_2B][t];
        totals[kPartials + t] = res[FAM_MBPOL_3B][t] + res[FAM_CH4H2O2_3B][t];
        totals[3 * kPartials + t] = 0.0;
    }
}

__global__ void k_combine_families(const double* __restrict__ fam, int nconf, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nconf * kPartials) return;
    const size_t st = (size_t)nconf * kPartials;
    out[i] = fam[FAM_MBPOL_2B * st + i] + fam[FAM_CH4H2O_2B * st + i] + fam[FAM_CH4CH4_2B * st + i];
    out[st + i] = fam[FAM_MBPOL_3B * st + i] + fam[FAM_CH4H2O2_3B * st + i];
}

size_t nmers_lds_bytes(int match_cap) { return (size_t)match_cap * kMatchDoubles * sizeof(double); }

hipError_t launch_nmers(const NmersArgs& A, const NmerCtx& N, int pass, int4 cap, int4 capx, int match_cap,
                        double* gscratch, int nblocks, hipStream_t st) {
    if (A.np <= 0) return hipSuccess;
    const int nb = gscratch ? (nblocks < A.np ? nblocks : A.np) : A.np;
    const size_t lds = gscratch ? 0 : nmers_lds_bytes(match_cap);
    if (gscratch)
        hipLaunchKernelGGL(k_nmers<false>, dim3(nb), dim3(64 * kQueryWaves), 0, st, A, N, pass, cap, capx, match_cap,
                           gscratch);
    else
        hipLaunchKernelGGL(k_nmers<true>, dim3(nb), dim3(64 * kQueryWaves), lds, st, A, N, pass, cap, capx,
                           match_cap, gscratch);
    return hipGetLastError();
}

}  // namespace mbxg
