// Test infrastructure only: C entry points over the reference MBX leaves.
//
// This file is ours; it is compiled together with the reference's own leaf
// sources, read in place from the reference checkout by oracle/Makefile, into
// oracle/_ref/libmbxref.so.  Only tests/ and bench.py's CPU-baseline leg load
// that library.  Nothing here is part of the shipped HIP path.
//
// Wrapped reference entry points (paths relative to the reference src/):
//   poly_2b_v6x::eval           potential/2b/poly-2b-v6x.cpp:12936, -nogrd.cpp:41
//   poly_3b_v2x::eval           potential/3b/poly-3b-v2x.cpp:26982, -nogrd.cpp:39
//   x2b_v9x::eval               potential/2b/x2b-v9x.cpp:1366 (E), :1514 (E+G)
//   x3b_v2x::eval               potential/3b/x3b-v2x.cpp:1292 (E), :1403 (E+G)
//   systools::AddClusters       bblock/sys_tools.cpp:472
//   systools::FixMonomerCoordinates  bblock/sys_tools.cpp:257
//   systools::GetCloseDimerImage     bblock/sys_tools.cpp:322
//   InvertUnitCell              tools/math_tools.cpp:37
//   x2b_A1B2Z2_C1D4_v1x::eval   potential/2b/x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp:954 (E+G; water, methane)
//   mbnrg_A1B4_C1D2_C1D2_deg3_v1::eval  potential/3b/mbnrg_3b_A1B4_C1D2_C1D2_deg3_v1.cpp:1481 (E+G)
// The 2B+3B driver at the bottom restates the OpenMP orchestration of
// System::Get2B / System::Get3B (bblock/system.cpp:2425-2725, 2769-3073) over
// those leaves for a pure-water system; it is the reference-faithful CPU
// baseline.
#include "potential/1b/ps.h"
#include <omp.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bblock/sys_tools.h"
#include "potential/2b/poly-2b-v6x.h"
#include "potential/2b/x2b-v9x.h"
#include "potential/3b/poly-3b-v2x.h"
#include "potential/3b/x3b-v2x.h"
#include "tools/math_tools.h"
#include "potential/2b/x2b_A1B2Z2_C1D4_deg3_exp0_v1x.h"
#include "potential/3b/mbnrg_3b_A1B4_C1D2_C1D2_deg3_v1.h"
#include "potential/1b/x1b_A1B4_deg5_exp0_v1x.h"

extern "C" {

// Polynomial E+G, nd dimers, x and g variable-major [31][nd]; t workspace is internal.
void ref_poly2b(size_t nd, const double* a, const double* x, double* e, double* g) {
    std::vector<double> t(3200 * 8 + 64);
    double* ta = reinterpret_cast<double*>((reinterpret_cast<uintptr_t>(t.data()) + 127) & ~uintptr_t(127));
    size_t done = 0;
    // the E+G routine is written for blocks of <= 8 dimers (f0..f99 loop over 8 lanes)
    while (done < nd) {
        size_t n = std::min<size_t>(8, nd - done);
        std::vector<double> xb(31 * n), gb(31 * n, 0.0);
        for (int v = 0; v < 31; ++v)
            for (size_t i = 0; i < n; ++i) xb[v * n + i] = x[v * nd + done + i];
        std::vector<double> eb = x2o::poly_2b_v6x::eval(n, a, xb.data(), ta, gb.data());
        for (size_t i = 0; i < n; ++i) {
            e[done + i] = eb[i];
            for (int v = 0; v < 31; ++v) g[v * nd + done + i] = gb[v * n + i];
        }
        done += n;
    }
}

void ref_poly2b_nogrd(size_t nd, const double* a, const double* x, double* e) {
    std::vector<double> eb = x2o::poly_2b_v6x::eval(nd, a, x);
    std::copy(eb.begin(), eb.end(), e);
}

// Polynomial E+G for 8 trimers, x and g lane-major [36][8].
void ref_poly3b(const double* a, const double* x, double* e8, double* g) {
    std::vector<double> t(3460 * 8 + 64);
    double* ta = reinterpret_cast<double*>((reinterpret_cast<uintptr_t>(t.data()) + 127) & ~uintptr_t(127));
    std::vector<double> xx(x, x + 36 * 8);
    double* e = x2o::poly_3b_v2x::eval(a, xx.data(), ta, g);
    std::copy(e, e + 8, e8);
    delete[] e;
}

double ref_poly3b_nogrd(const double* a, const double* x) { return x2o::poly_3b_v2x::eval(a, x); }

double ref_x2b(const double* w1, const double* w2, size_t nd, double* g1, double* g2, double* virial) {
    if (!g1) return x2o::x2b_v9x::eval(w1, w2, nd);
    std::vector<double> v(9, 0.0);
    double e = x2o::x2b_v9x::eval(w1, w2, g1, g2, nd, &v);
    if (virial)
        for (int k = 0; k < 9; ++k) virial[k] += v[k];
    return e;
}

double ref_x3b(const double* w1, const double* w2, const double* w3, size_t nt, double* g1, double* g2, double* g3,
               double* virial) {
    if (!g1) return x2o::x3b_v2x::eval(w1, w2, w3, nt);
    std::vector<double> v(9, 0.0);
    double e = x2o::x3b_v2x::eval(w1, w2, w3, g1, g2, g3, nt, &v);
    if (virial)
        for (int k = 0; k < 9; ++k) virial[k] += v[k];
    return e;
}

int ref_invert_unit_cell(const double* box, double* box_inv) {
    std::vector<double> b(box, box + 9);
    std::vector<double> bi = InvertUnitCell(b);
    std::copy(bi.begin(), bi.end(), box_inv);
    return 0;
}

// xyz holds 3*nsites coordinates; monomer i starts at site first_index[i] and has nat[i] atoms.
void ref_fix_monomer_coordinates(double* xyz, size_t nsites, const double* box, const double* box_inv, size_t nmon,
                                 const size_t* nat, const size_t* first_index) {
    std::vector<double> v(xyz, xyz + 3 * nsites);
    systools::FixMonomerCoordinates(v, std::vector<double>(box, box + 9), std::vector<double>(box_inv, box_inv + 9),
                                    std::vector<size_t>(nat, nat + nmon),
                                    std::vector<size_t>(first_index, first_index + nmon));
    std::copy(v.begin(), v.end(), xyz);
}

void ref_close_dimer_image(const double* box, const double* box_inv, size_t nat1, size_t nat2, size_t nd, double* xyz1,
                           double* xyz2) {
    systools::GetCloseDimerImage(std::vector<double>(box, box + 9), std::vector<double>(box_inv, box_inv + 9), nat1,
                                 nat2, nd, xyz1, xyz2);
}

// Clusters of order nmax (2 or 3) over all monomers, queries in idxs order 0..nmon-1.
// Returns the number of n-mers found; writes min(found, cap) tuples to out.
size_t ref_add_clusters(int nmax, double cutoff, size_t nmon, int use_pbc, const double* box, const double* box_inv,
                        const double* xyz, size_t nsites, const size_t* first_index, const size_t* islocal,
                        const int* tag, int use_ghost, size_t* out, size_t cap) {
    std::vector<size_t> idxs(nmon);
    for (size_t i = 0; i < nmon; ++i) idxs[i] = i;
    std::vector<double> b, bi;
    if (use_pbc) {
        b.assign(box, box + 9);
        bi.assign(box_inv, box_inv + 9);
    }
    std::vector<size_t> nmers;
    systools::AddClusters(nmax, cutoff, idxs, nmon, use_pbc != 0, b, bi, std::vector<double>(xyz, xyz + 3 * nsites),
                          std::vector<size_t>(first_index, first_index + nmon),
                          std::vector<size_t>(islocal, islocal + nmon), std::vector<int>(tag, tag + nsites), nmers,
                          use_ghost != 0);
    size_t n = nmers.size() / nmax;
    std::copy(nmers.begin(), nmers.begin() + std::min(n, cap) * nmax, out);
    return n;
}

// ---------------------------------------------------------------------------------------------
// Reference-faithful 2B+3B E+G for nmon waters (3 atoms each, O H H, input = internal order),
// restating System::Get2B/Get3B's orchestration over the reference leaves:
//   per-thread KD-tree over rank-strided queries (system.cpp:2461-2474, 2808-2816),
//   pooled lists, <= batch n-mers per grab under a critical section (2495-2519, 2837-2864),
//   close image (2601-2604, 2934-2937), holder call, per-thread gradient pools (2630-2642),
//   striped condensation (2682-2722).  batch = 8 is the reference's value; a larger batch is the
//   "best-effort" CPU mode.  Tags are the input atom counters (system.cpp:866-869).
// xyz is wrapped in place (System::SetPBC -> FixMonomerCoordinates, system.cpp:2015).
// grad (9*nmon) is overwritten; virial (9) is overwritten.  counts[4] = listed 2B, listed 3B.
static double ref_nmer_term(int order, double cutoff, size_t nmon, const std::vector<double>& xyz, bool pbc,
                            const std::vector<double>& box, const std::vector<double>& box_inv, size_t batch,
                            int nthreads, double* grad, double* virial, size_t* nlisted) {
    std::vector<size_t> first_index(nmon), islocal(nmon, 1);
    std::vector<int> tag(3 * nmon);
    for (size_t i = 0; i < nmon; ++i) first_index[i] = 3 * i;
    for (size_t i = 0; i < 3 * nmon; ++i) tag[i] = int(i);
    std::vector<std::vector<size_t>> lists(nthreads);
#pragma omp parallel num_threads(nthreads)
    {
        int rank = omp_get_thread_num();
        std::vector<size_t> idxs;
        for (size_t i = rank; i < nmon; i += nthreads) idxs.push_back(i);
        systools::AddClusters(order, cutoff, idxs, nmon, pbc, box, box_inv, xyz, first_index, islocal, tag,
                              lists[rank], false);
    }
    std::vector<size_t> pool;
    for (auto& l : lists) pool.insert(pool.end(), l.begin(), l.end());
    const size_t n = pool.size() / order;
    *nlisted = n;
    std::vector<std::vector<double>> gpool(nthreads, std::vector<double>(9 * nmon, 0.0));
    std::vector<std::vector<double>> vpool(nthreads, std::vector<double>(9, 0.0));
    std::vector<double> epool(nthreads, 0.0);
    size_t next = 0;
#pragma omp parallel num_threads(nthreads)
    {
        int rank = omp_get_thread_num();
        std::vector<double> x1, x2, x3, g1, g2, g3, vir(9);
        while (true) {
            size_t start, cnt;
#pragma omp critical(ref_pool)
            {
                start = next;
                size_t remaining = n - start;
                size_t full_round = 8 * size_t(nthreads);
                size_t trunc = 8 * (remaining / full_round) + (remaining % full_round == 0 ? 0 : 8);
                cnt = std::min(trunc, std::min(batch, remaining));
                next += cnt;
            }
            if (cnt == 0) break;
            x1.assign(9 * cnt, 0.0);
            x2.assign(9 * cnt, 0.0);
            x3.assign(order == 3 ? 9 * cnt : 0, 0.0);
            for (size_t k = 0; k < cnt; ++k) {
                const size_t* t = &pool[(start + k) * order];
                std::copy(&xyz[9 * t[0]], &xyz[9 * t[0]] + 9, &x1[9 * k]);
                std::copy(&xyz[9 * t[1]], &xyz[9 * t[1]] + 9, &x2[9 * k]);
                if (order == 3) std::copy(&xyz[9 * t[2]], &xyz[9 * t[2]] + 9, &x3[9 * k]);
            }
            if (pbc) {
                systools::GetCloseDimerImage(box, box_inv, 3, 3, cnt, x1.data(), x2.data());
                if (order == 3) systools::GetCloseDimerImage(box, box_inv, 3, 3, cnt, x1.data(), x3.data());
            }
            g1.assign(9 * cnt, 0.0);
            g2.assign(9 * cnt, 0.0);
            g3.assign(order == 3 ? 9 * cnt : 0, 0.0);
            std::fill(vir.begin(), vir.end(), 0.0);
            if (order == 2)
                epool[rank] += x2o::x2b_v9x::eval(x1.data(), x2.data(), g1.data(), g2.data(), cnt, &vir);
            else
                epool[rank] += x2o::x3b_v2x::eval(x1.data(), x2.data(), x3.data(), g1.data(), g2.data(), g3.data(),
                                                   cnt, &vir);
            for (int k = 0; k < 9; ++k) vpool[rank][k] += vir[k];
            for (size_t k = 0; k < cnt; ++k) {
                const size_t* t = &pool[(start + k) * order];
                for (int j = 0; j < 9; ++j) {
                    gpool[rank][9 * t[0] + j] += g1[9 * k + j];
                    gpool[rank][9 * t[1] + j] += g2[9 * k + j];
                    if (order == 3) gpool[rank][9 * t[2] + j] += g3[9 * k + j];
                }
            }
        }
    }
    double e = 0.0;
    for (int r = 0; r < nthreads; ++r) {
        e += epool[r];
        for (size_t j = 0; j < 9 * nmon; ++j) grad[j] += gpool[r][j];
        for (int k = 0; k < 9; ++k) virial[k] += vpool[r][k];
    }
    return e;
}

int ref_water_2b3b(size_t nmon, double* xyz, const double* box9, double cutoff2b, double cutoff3b, size_t batch,
                   int nthreads, double* e2b, double* e3b, double* grad, double* virial, size_t* counts) {
    bool pbc = box9 != nullptr;
    std::vector<double> box, box_inv;
    std::vector<size_t> nat(nmon, 3), first_index(nmon);
    for (size_t i = 0; i < nmon; ++i) first_index[i] = 3 * i;
    std::vector<double> v(xyz, xyz + 9 * nmon);
    if (pbc) {
        box.assign(box9, box9 + 9);
        box_inv = InvertUnitCell(box);
        systools::FixMonomerCoordinates(v, box, box_inv, nat, first_index);
        std::copy(v.begin(), v.end(), xyz);
    }
    std::fill(grad, grad + 9 * nmon, 0.0);
    std::fill(virial, virial + 9, 0.0);
    *e2b = ref_nmer_term(2, cutoff2b, nmon, v, pbc, box, box_inv, batch, nthreads, grad, virial, &counts[0]);
    *e3b = ref_nmer_term(3, cutoff3b, nmon, v, pbc, box, box_inv, batch, nthreads, grad, virial, &counts[1]);
    return 0;
}

// MB-nrg CH4-H2O 2B holder, n dimers: w[9n] waters, c[15n] methanes (+= g, vir).
double ref_x2b_ch4h2o(const double* w, const double* c, size_t n, double* gw, double* gc, double* vir) {
    x2b_A1B2Z2_C1D4_deg3_exp0::x2b_A1B2Z2_C1D4_v1x pot("h2o", "ch4");
    std::vector<double> v(vir, vir + 9);
    const double e = pot.eval(w, c, gw, gc, n, &v);
    std::copy(v.begin(), v.end(), vir);
    return e;
}

// MB-nrg CH4-H2O-H2O 3B holder, n trimers: c[15n], w1[9n], w2[9n] (+= g, vir).
double ref_x3b_ch4h2o2(const double* c, const double* w1, const double* w2, size_t n, double* gc, double* g1,
                       double* g2, double* vir) {
    mbnrg_A1B4_C1D2_C1D2_deg3::mbnrg_A1B4_C1D2_C1D2_deg3_v1 pot("ch4", "h2o", "h2o");
    std::vector<double> v(vir, vir + 9);
    const double e = pot.eval(c, w1, w2, gc, g1, g2, n, &v);
    std::copy(v.begin(), v.end(), vir);
    return e;
}

// ps::pot_nasa (potential/1b/ps.cpp:346): energies of nw waters, gradient dr[9 nw] (nullable),
// virial[9] incremented (nullable).
void ref_ps_1b(const double* rr, double* dr, size_t nw, double* energies, double* virial) {
    std::vector<double> v(9, 0.0);
    if (virial) v.assign(virial, virial + 9);
    std::vector<double> e = ps::pot_nasa(rr, dr, nw, virial ? &v : nullptr);
    std::copy(e.begin(), e.end(), energies);
    if (virial) std::copy(v.begin(), v.end(), virial);
}

// x1b_A1B4_v1x("ch4")::eval with gradients over n CH4 monomers (15 doubles each).
void ref_ch4_1b(const double* xyz, double* grad, size_t n, double* energies, double* virial) {
    x1b_A1B4_deg5_exp0::x1b_A1B4_v1x pot("ch4");
    std::vector<double> v(virial, virial + 9);
    std::vector<double> e = pot.eval(xyz, grad, n, &v);
    std::copy(e.begin(), e.end(), energies);
    std::copy(v.begin(), v.end(), virial);
}

}  // extern "C"
