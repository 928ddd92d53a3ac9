// Cell-list dimer / trimer builder for gfx950.
//
// Produces the same n-mer sets as the reference's KD-tree enumeration
// (systools::AddClusters, bblock/sys_tools.cpp:472-641, over kdtutils::XyzToCloud,
// kdtree/kdtree_utils.h:196-279, with nanoflann's strict radius test
// kdtree/nanoflann.hpp:166-171):
//   * every local monomer contributes its first atom; under PBC the point is re-wrapped and
//     receives 7 image slots (slot s of point p lives at index p + s*np);
//   * a dimer (i, j) is listed from query i when some slot of j lies strictly inside the
//     cutoff of i's point and tag[i] < tag[j];
//   * a trimer (i, j, k) is listed from query i when slots of j and k are both inside the
//     cutoff, j != k != i, and either |slot_j - slot_k|^2 > c^2 or tag[i] is the smallest.
// Every distance is computed from the same cloud coordinates with the same operation order,
// so membership decisions are bit-identical.  The KD-tree is replaced by a grid of cells in
// fractional coordinates (one cell >= cutoff wide perpendicular to each face); the grid only
// proposes candidates, every candidate slot is tested exactly.
//
// The list is emitted in two passes (count, then write at scanned offsets) so its order is a
// deterministic function of the input.  While emitting, each n-mer is also imaged
// (GetCloseDimerImage, sys_tools.cpp:322-365) and tested against the polynomial's own
// evaluation cutoff (x2b-v9x.cpp:1589-1600 for MB-pol 2B; x3b-v2x.cpp:1495-1515 for MB-pol
// 3B); n-mers that pass are also written to a compacted "evaluate" list that feeds the
// polynomial kernels, so no wavefront idles on a skipped n-mer.
//
// This file is compiled with -ffp-contract=off.
#include "mbxg_device.hpp"
#include "mbxg_kernels.hpp"

namespace mbxg {

// ---------------------------------------------------------------------------------------------
// FixMonomerCoordinates (sys_tools.cpp:287-319): wrap each monomer's first atom into the cell
// and bring its other atoms next to it.  One thread per monomer.
__global__ void k_fix_monomers(double* __restrict__ xyz, const int* __restrict__ first_atom,
                               const int* __restrict__ nat, int nmon, Box B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nmon) return;
    double* x0 = xyz + 3 * (size_t)first_atom[i];
    double u0, v0, w0;
    to_frac(B, x0[0], x0[1], x0[2], u0, v0, w0);
    u0 -= floor(u0 + 0.5);
    v0 -= floor(v0 + 0.5);
    w0 -= floor(w0 + 0.5);
    to_cart(B, u0, v0, w0, x0[0], x0[1], x0[2]);
    for (int a = 1; a < nat[i]; ++a) {
        double* xa = x0 + 3 * a;
        double u, v, w;
        to_frac(B, xa[0], xa[1], xa[2], u, v, w);
        u -= floor(u - u0 + 0.5);
        v -= floor(v - v0 + 0.5);
        w -= floor(w - w0 + 0.5);
        to_cart(B, u, v, w, xa[0], xa[1], xa[2]);
    }
}

// ---------------------------------------------------------------------------------------------
// XyzToCloud (kdtree_utils.h:196-279): point p = first atom of cloud member p.  Slot 0 is the
// re-wrapped position; slots 1..7 are the images selected from its fractional coordinates.
// Also computes the point's grid cell.  One thread per point.
__global__ void k_cloud(const double* __restrict__ xyz, const int* __restrict__ first_atom,
                        const int* __restrict__ members, int np, Box B, Grid G, double* __restrict__ cx,
                        double* __restrict__ cy, double* __restrict__ cz, int* __restrict__ cell) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const double* x0 = xyz + 3 * (size_t)first_atom[members[p]];
    double qx = x0[0], qy = x0[1], qz = x0[2];
    if (B.periodic) {
        double u, v, w;
        to_frac(B, qx, qy, qz, u, v, w);
        u -= floor(u + 0.5);
        v -= floor(v + 0.5);
        w -= floor(w + 0.5);
        to_cart(B, u, v, w, qx, qy, qz);
        // GetFractionalCoordinates (kdtree_utils.h:186-189)
        const double uvw0 = B.f00 * qx - B.f01 * qy + B.f02 * qz;
        const double uvw1 = B.f11 * qy + B.f12 * qz;
        const double uvw2 = B.f22 * qz;
        const int ind[3] = {0, -1, 1};
        for (int m2 = 0; m2 < 3; ++m2) {
            const int m = ind[m2];
            if ((uvw0 + m) > 1.0 || (uvw0 + m) < -1.0) continue;
            for (int n2 = 0; n2 < 3; ++n2) {
                const int n = ind[n2];
                if ((uvw1 + n) > 1.0 || (uvw1 + n) < -1.0) continue;
                for (int l2 = 0; l2 < 3; ++l2) {
                    if (m2 == 0 && n2 == 0 && l2 == 0) continue;
                    const int l = ind[l2];
                    if ((uvw2 + l) > 1.0 || (uvw2 + l) < -1.0) continue;
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
                    const size_t s = (size_t)p + (size_t)shift * np;
                    cx[s] = qx + sx;
                    cy[s] = qy + sy;
                    cz[s] = qz + sz;
                }
            }
        }
        cx[p] = qx;
        cy[p] = qy;
        cz[p] = qz;
        // grid cell from the fractional position in [-0.5, 0.5)
        to_frac(B, qx, qy, qz, u, v, w);
        int ca = (int)floor((u + 0.5) * G.n[0]);
        int cb = (int)floor((v + 0.5) * G.n[1]);
        int cc = (int)floor((w + 0.5) * G.n[2]);
        ca = ca < 0 ? 0 : (ca >= G.n[0] ? G.n[0] - 1 : ca);
        cb = cb < 0 ? 0 : (cb >= G.n[1] ? G.n[1] - 1 : cb);
        cc = cc < 0 ? 0 : (cc >= G.n[2] ? G.n[2] - 1 : cc);
        cell[p] = (cc * G.n[1] + cb) * G.n[0] + ca;
    } else {
        cx[p] = qx;
        cy[p] = qy;
        cz[p] = qz;
        int ca = (int)floor((qx - G.lo[0]) * G.inv_w[0]);
        int cb = (int)floor((qy - G.lo[1]) * G.inv_w[1]);
        int cc = (int)floor((qz - G.lo[2]) * G.inv_w[2]);
        ca = ca < 0 ? 0 : (ca >= G.n[0] ? G.n[0] - 1 : ca);
        cb = cb < 0 ? 0 : (cb >= G.n[1] ? G.n[1] - 1 : cb);
        cc = cc < 0 ? 0 : (cc >= G.n[2] ? G.n[2] - 1 : cc);
        cell[p] = (cc * G.n[1] + cb) * G.n[0] + ca;
    }
}

// ---------------------------------------------------------------------------------------------
// Counting sort of points into cells: count, scan, fill, then sort each cell's members by point
// index so the cell contents (and hence the emission order) are deterministic.
__global__ void k_bin_count(const int* __restrict__ cell, int np, int* __restrict__ cell_count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) atomicAdd(&cell_count[cell[p]], 1);
}

// Single-workgroup exclusive scan of the cell counts; cell_start has ncell+1 entries.
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
        cell_count[c] = 0;  // reused as the fill cursor
    }
    if (t == 1023) cell_start[ncell] = sh[t];
}

__global__ void k_bin_fill(const int* __restrict__ cell, int np, const int* __restrict__ cell_start,
                           int* __restrict__ cursor, int* __restrict__ cell_pts) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int c = cell[p];
    cell_pts[cell_start[c] + atomicAdd(&cursor[c], 1)] = p;
}

__global__ void k_bin_sort(const int* __restrict__ cell_start, int ncell, int* __restrict__ cell_pts) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int beg = cell_start[c], end = cell_start[c + 1];
    for (int a = beg + 1; a < end; ++a) {
        const int v = cell_pts[a];
        int b = a - 1;
        while (b >= beg && cell_pts[b] > v) {
            cell_pts[b + 1] = cell_pts[b];
            --b;
        }
        cell_pts[b + 1] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// Neighbour search for one query per wavefront: gathers the matching cloud slots into LDS.
// Returns the number of matches (clamped at cap; *overflow set if exceeded).

struct QueryCtx {
    const double* cx;
    const double* cy;
    const double* cz;
    const int* cell;
    const int* cell_start;
    const int* cell_pts;
    int np;
    int nslot;  // 8 under PBC, else 1
};

__device__ int gather_matches(const QueryCtx& Q, const Grid& G, int i, double c2, int* s_match, int cap,
                              int* overflow) {
    const int lane = threadIdx.x & 63;
    const double qx = Q.cx[i], qy = Q.cy[i], qz = Q.cz[i];
    const int ci = Q.cell[i];
    const int ia = ci % G.n[0], ib = (ci / G.n[0]) % G.n[1], ic = ci / (G.n[0] * G.n[1]);
    int nmatch = 0;
    for (int oc = G.lo_off[2]; oc <= G.hi_off[2]; ++oc) {
        int cc = ic + oc;
        if (G.periodic) {
            cc = (cc + G.n[2]) % G.n[2];
        } else if (cc < 0 || cc >= G.n[2]) {
            continue;
        }
        for (int ob = G.lo_off[1]; ob <= G.hi_off[1]; ++ob) {
            int cb = ib + ob;
            if (G.periodic) {
                cb = (cb + G.n[1]) % G.n[1];
            } else if (cb < 0 || cb >= G.n[1]) {
                continue;
            }
            for (int oa = G.lo_off[0]; oa <= G.hi_off[0]; ++oa) {
                int ca = ia + oa;
                if (G.periodic) {
                    ca = (ca + G.n[0]) % G.n[0];
                } else if (ca < 0 || ca >= G.n[0]) {
                    continue;
                }
                const int c = (cc * G.n[1] + cb) * G.n[0] + ca;
                const int beg = Q.cell_start[c], end = Q.cell_start[c + 1];
                const int total = (end - beg) * Q.nslot;
                for (int base = 0; base < total; base += 64) {
                    const int t = base + lane;
                    int s = -1;
                    if (t < total) {
                        const int p = Q.cell_pts[beg + t / Q.nslot];
                        const int slot = t % Q.nslot;
                        const size_t si = (size_t)p + (size_t)slot * Q.np;
                        if (dist2(qx, qy, qz, Q.cx[si], Q.cy[si], Q.cz[si]) < c2) s = (int)si;
                    }
                    const unsigned long long mask = __ballot(s >= 0);
                    const int pos = __popcll(mask & ((1ull << lane) - 1ull));
                    if (s >= 0) {
                        const int w = nmatch + pos;
                        if (w < cap) s_match[w] = s;
                    }
                    nmatch += __popcll(mask);
                }
            }
        }
    }
    if (nmatch > cap) {
        if (lane == 0) atomicMax(overflow, nmatch);
        nmatch = cap;
    }
    __builtin_amdgcn_wave_barrier();
    return nmatch;
}

// ---------------------------------------------------------------------------------------------
// Evaluation rules (which listed n-mers reach a polynomial).

// 2B: MB-pol h2o-h2o evaluates iff r_OO <= r2f after imaging (x2b-v9x.cpp:1589-1600).
// Pairs without a GPU holder contribute nothing here (energy2b.cpp:372-374 returns 0 for an
// unknown pair); the host refuses to evaluate systems whose pairs would need one.
__device__ __forceinline__ bool eval2_rule(const NmerCtx& N, int m1, int m2) {
    const int k1 = N.kind[m1], k2 = N.kind[m2];
    if (!(k1 == MON_H2O && k2 == MON_H2O)) return false;
    const double* a = N.xyz + 3 * (size_t)N.first_atom[m1];
    const double* b = N.xyz + 3 * (size_t)N.first_atom[m2];
    double bx = b[0], by = b[1], bz = b[2];
    if (N.periodic) {
        const Box& B = *N.box;
        double dx, dy, dz;
        image_shift(B, a[0], a[1], a[2], bx, by, bz, dx, dy, dz);
        image_apply(B, dx, dy, dz, bx, by, bz);
    }
    const double d0 = a[0] - bx, d1 = a[1] - by, d2 = a[2] - bz;
    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    return !(r > p2::r2f);
}

// 3B: MB-pol h2o x3 is skipped when two O-O distances exceed r3f (x3b-v2x.cpp:1495-1515).
__device__ __forceinline__ bool eval3_rule(const NmerCtx& N, int m1, int m2, int m3) {
    const int k1 = N.kind[m1], k2 = N.kind[m2], k3 = N.kind[m3];
    if (!(k1 == MON_H2O && k2 == MON_H2O && k3 == MON_H2O)) return false;
    const double* a = N.xyz + 3 * (size_t)N.first_atom[m1];
    const double* b = N.xyz + 3 * (size_t)N.first_atom[m2];
    const double* c = N.xyz + 3 * (size_t)N.first_atom[m3];
    double bx = b[0], by = b[1], bz = b[2], cx = c[0], cy = c[1], cz = c[2];
    if (N.periodic) {
        const Box& B = *N.box;
        double dx, dy, dz;
        image_shift(B, a[0], a[1], a[2], bx, by, bz, dx, dy, dz);
        image_apply(B, dx, dy, dz, bx, by, bz);
        image_shift(B, a[0], a[1], a[2], cx, cy, cz, dx, dy, dz);
        image_apply(B, dx, dy, dz, cx, cy, cz);
    }
    const double ab0 = a[0] - bx, ab1 = a[1] - by, ab2 = a[2] - bz;
    const double ac0 = a[0] - cx, ac1 = a[1] - cy, ac2 = a[2] - cz;
    const double bc0 = bx - cx, bc1 = by - cy, bc2 = bz - cz;
    double sab = 0.0, sac = 0.0, sbc = 0.0;
    sab += ab0 * ab0;
    sab += ab1 * ab1;
    sab += ab2 * ab2;
    sac += ac0 * ac0;
    sac += ac1 * ac1;
    sac += ac2 * ac2;
    sbc += bc0 * bc0;
    sbc += bc1 * bc1;
    sbc += bc2 * bc2;
    const double rab = sqrt(sab), rac = sqrt(sac), rbc = sqrt(sbc);
    const double f = p3::r3f;
    const bool skip = (rab > f && rac > f) || (rab > f && rbc > f) || (rac > f && rbc > f);
    return !skip;
}

// ---------------------------------------------------------------------------------------------
// Dimers.  pass 0: counts[q] = {listed, evaluated}; pass 1: writes at offsets[q].
__global__ void __launch_bounds__(64) k_dimers(QueryCtx Q, Grid G, NmerCtx N, double c2, int pass,
                                               int2* __restrict__ counts, const int2* __restrict__ offsets,
                                               int* __restrict__ out_list, int* __restrict__ out_eval, int cap,
                                               int* overflow) {
    extern __shared__ int s_match[];
    const int q = blockIdx.x;  // cloud point index of the query
    if (q >= Q.np) return;
    const int lane = threadIdx.x;
    const int mi = N.members[q];
    if (!N.islocal[mi]) {
        if (pass == 0 && lane == 0) counts[q] = make_int2(0, 0);
        return;
    }
    const int nm = gather_matches(Q, G, q, c2, s_match, cap, overflow);
    const int tq = N.tag[mi];
    int base_l = pass ? offsets[q].x : 0;
    int base_e = pass ? offsets[q].y : 0;
    int nl = 0, ne = 0;
    for (int b = 0; b < nm; b += 64) {
        const int t = b + lane;
        bool listed = false, ev = false;
        int mj = 0;
        if (t < nm) {
            const int pj = s_match[t] % Q.np;
            mj = N.members[pj];
            listed = tq < N.tag[mj];
            if (listed) ev = eval2_rule(N, mi, mj);
        }
        const unsigned long long ml = __ballot(listed), me = __ballot(ev);
        const unsigned long long below = (1ull << lane) - 1ull;
        if (pass) {
            if (listed) {
                const int w = base_l + nl + __popcll(ml & below);
                out_list[2 * w] = mi;
                out_list[2 * w + 1] = mj;
            }
            if (ev) {
                const int w = base_e + ne + __popcll(me & below);
                out_eval[2 * w] = mi;
                out_eval[2 * w + 1] = mj;
            }
        }
        nl += __popcll(ml);
        ne += __popcll(me);
    }
    if (pass == 0 && lane == 0) counts[q] = make_int2(nl, ne);
}

// Trimers.  Pairs (a < b) of matches are spread over the lanes.
__global__ void __launch_bounds__(64) k_trimers(QueryCtx Q, Grid G, NmerCtx N, double c2, int pass,
                                                int2* __restrict__ counts, const int2* __restrict__ offsets,
                                                int* __restrict__ out_list, int* __restrict__ out_eval, int cap,
                                                int* overflow) {
    extern __shared__ int s_match[];
    const int q = blockIdx.x;
    if (q >= Q.np) return;
    const int lane = threadIdx.x;
    const int mi = N.members[q];
    if (!N.islocal[mi]) {
        if (pass == 0 && lane == 0) counts[q] = make_int2(0, 0);
        return;
    }
    const int nm = gather_matches(Q, G, q, c2, s_match, cap, overflow);
    const int tq = N.tag[mi];
    const long npair = (long)nm * (nm - 1) / 2;
    int base_l = pass ? offsets[q].x : 0;
    int base_e = pass ? offsets[q].y : 0;
    int nl = 0, ne = 0;
    for (long b = 0; b < npair; b += 64) {
        const long t = b + lane;
        bool listed = false, ev = false;
        int mj = 0, mk = 0;
        if (t < npair) {
            // invert t -> (a, c) with a < c, enumerating row-major over a
            // t = a*nm - a*(a+1)/2 + (c - a - 1)
            int a = (int)((2.0 * nm - 1.0 - sqrt((2.0 * nm - 1.0) * (2.0 * nm - 1.0) - 8.0 * (double)t)) * 0.5);
            if (a < 0) a = 0;
            while (a > 0 && (long)a * nm - (long)a * (a + 1) / 2 > t) --a;
            while ((long)(a + 1) * nm - (long)(a + 1) * (a + 2) / 2 <= t) ++a;
            const int c = (int)(t - ((long)a * nm - (long)a * (a + 1) / 2)) + a + 1;
            const int sa = s_match[a], sc = s_match[c];
            const int pj = sa % Q.np, pk = sc % Q.np;
            if (pj != q && pk != q && pj != pk) {
                mj = N.members[pj];
                mk = N.members[pk];
                const double djk = dist2(Q.cx[sa], Q.cy[sa], Q.cz[sa], Q.cx[sc], Q.cy[sc], Q.cz[sc]);
                listed = djk > c2 || (tq < N.tag[mj] && tq < N.tag[mk]);
                if (listed) ev = eval3_rule(N, mi, mj, mk);
            }
        }
        const unsigned long long ml = __ballot(listed), me = __ballot(ev);
        const unsigned long long below = (1ull << lane) - 1ull;
        if (pass) {
            if (listed) {
                const int w = base_l + nl + __popcll(ml & below);
                out_list[3 * w] = mi;
                out_list[3 * w + 1] = mj;
                out_list[3 * w + 2] = mk;
            }
            if (ev) {
                const int w = base_e + ne + __popcll(me & below);
                out_eval[3 * w] = mi;
                out_eval[3 * w + 1] = mj;
                out_eval[3 * w + 2] = mk;
            }
        }
        nl += __popcll(ml);
        ne += __popcll(me);
    }
    if (pass == 0 && lane == 0) counts[q] = make_int2(nl, ne);
}

// Exclusive scan of the per-query counts (single workgroup; totals written at offsets[n]).
__global__ void __launch_bounds__(1024) k_scan_counts(const int2* __restrict__ counts, int n,
                                                      int2* __restrict__ offsets) {
    __shared__ int sx[1024], sy[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int beg = t * per, end = min(n, beg + per);
    int ax = 0, ay = 0;
    for (int i = beg; i < end; ++i) {
        ax += counts[i].x;
        ay += counts[i].y;
    }
    sx[t] = ax;
    sy[t] = ay;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int vx = t >= off ? sx[t - off] : 0;
        int vy = t >= off ? sy[t - off] : 0;
        __syncthreads();
        sx[t] += vx;
        sy[t] += vy;
        __syncthreads();
    }
    int ox = sx[t] - ax, oy = sy[t] - ay;
    for (int i = beg; i < end; ++i) {
        offsets[i] = make_int2(ox, oy);
        ox += counts[i].x;
        oy += counts[i].y;
    }
    if (t == 1023) offsets[n] = make_int2(sx[t], sy[t]);
}

// Deterministic sum of per-block partial records (kPartials doubles each): one wavefront,
// lane-strided accumulation then a fixed butterfly.
__global__ void __launch_bounds__(64) k_reduce_partials(const double* __restrict__ partials, int nblocks,
                                                        double* __restrict__ out) {
    const int lane = threadIdx.x;
    for (int q = 0; q < kPartials; ++q) {
        double acc = 0.0;
        for (int b = lane; b < nblocks; b += 64) acc += partials[(size_t)b * kPartials + q];
        acc = wave_sum(acc);
        if (lane == 0) out[q] = acc;
    }
}

hipError_t launch_dimers(const double* cx, const double* cy, const double* cz, const int* cell,
                         const int* cell_start, const int* cell_pts, int np, int nslot, const Grid& G,
                         const NmerCtx& N, double c2, int pass, int2* counts, const int2* offsets, int* out_list,
                         int* out_eval, int cap, int* overflow, hipStream_t st) {
    QueryCtx Q{cx, cy, cz, cell, cell_start, cell_pts, np, nslot};
    hipLaunchKernelGGL(k_dimers, dim3(np), dim3(64), cap * sizeof(int), st, Q, G, N, c2, pass, counts, offsets,
                       out_list, out_eval, cap, overflow);
    return hipGetLastError();
}

hipError_t launch_trimers(const double* cx, const double* cy, const double* cz, const int* cell,
                          const int* cell_start, const int* cell_pts, int np, int nslot, const Grid& G,
                          const NmerCtx& N, double c2, int pass, int2* counts, const int2* offsets, int* out_list,
                          int* out_eval, int cap, int* overflow, hipStream_t st) {
    QueryCtx Q{cx, cy, cz, cell, cell_start, cell_pts, np, nslot};
    hipLaunchKernelGGL(k_trimers, dim3(np), dim3(64), cap * sizeof(int), st, Q, G, N, c2, pass, counts, offsets,
                       out_list, out_eval, cap, overflow);
    return hipGetLastError();
}

}  // namespace mbxg
