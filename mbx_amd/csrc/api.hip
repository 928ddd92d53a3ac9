// Host runtime of the mbxg engine: context, device buffers, the 2B/3B pipeline and the C ABI
// declared in include/mbxg.h.
//
// Pipeline of mbxg_compute (all on the context's stream, no host synchronisation in steady
// state):
//   k_grid_gas              (gas phase) candidate grid at the larger cutoff
//   k_cloud_fix             FixMonomerCoordinates into the working copy, cloud points with image
//                           slots and their cells, cell counts; zeroes the gradient and records
//   k_scan_cells, k_bin_fill  slots binned by cell
//   k_nmers                 per-query counts and (one pass) each family's evaluated n-mers staged
//                           per (query, wavefront); or count pass + emit pass when synchronised
//   k_scan_all              offsets; totals at offsets[np]
//   k_compact_evals         evaluate lists (one-pass builds)
//   1B, launch_mbpol2b/3b   variables, polynomials, chain rule over the evaluate lists; n-mer
//   / MB-nrg families       counts read on the device.  With both MB-pol orders and gradients:
//                           k_vars_mbpol_pair (both orders' variables), k_poly_eg_pair_chain3
//                           (both polynomials, the 3B chain rule in the 3B blocks' tail),
//                           k_grad_mbpol2b
//   k_finish                deterministic energy / virial totals per family and per order
//   run_xfam                (co2, ions, 4B) further families, synchronously
// Capacities (list buffers, match storage) come from a synchronised first call and carry
// headroom.  Every read of results synchronises and checks the totals against them; a call that
// exceeded them is recomputed synchronously with grown buffers, so results are always whole.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mbxg.h"
#include "gen/coef_ch4ch4_2b.inc"
#include "gen/coef_ch4h2o2_3b.inc"
#include "gen/coef_ch4h2o_2b.inc"
#include "gen/coef_mbpol_2b.inc"
#include "gen/coef_mbpol_3b.inc"
#include "mbxg_kernels.hpp"
#include "xfam.hpp"

using namespace mbxg;

namespace {

// Timing events: 0 start, 1 after fix, 2 after the list count pass, 3 after the emit pass,
// 4..7 2B stages, 8..11 3B stages, 12 end.
constexpr int kNumEvents = 15;

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return fail(MBXG_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                      \
    } while (0)

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t reserve(size_t want) {
        if (want <= n) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            n = 0;
        }
        size_t cap = want + want / 4 + 64;
        hipError_t e = hipMalloc((void**)&p, cap * sizeof(T));
        if (e == hipSuccess) n = cap;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Neighbour-search and list state (both orders share one gather).
struct ListState {
    // cell layout over all configurations (rebuilt when topology, batch, box or cutoff change)
    bool layout_ok = false;
    double layout_cut = -1.0;
    int ncell_total = 0;
    DBuf<Grid> d_grids;           // one per configuration
    DBuf<int> d_cell_base, d_cell_bound;
    DBuf<double> cx, cy, cz, bx, by, bz;
    DBuf<PtInfo> pinfo;           // per cloud point: first atom, tag, kind
    DBuf<int> cell, cell_count, cell_start, cell_pos, cell_slot;  // cell_pos: a slot's place in its cell
    DBuf<int4> counts, offsets, counts_x, offsets_x;
    DBuf<int4> wpre, wpre_x;      // per (query, wavefront): prefix within the query
    DBuf<int> list2, list3, evals[kFamilies];
    DBuf<double> gscratch;        // match storage when it does not fit in LDS
    // capacities in n-mers: cap = {list2, MB-pol 2B, list3, MB-pol 3B},
    // capx = {CH4-H2O 2B, CH4-CH4 2B, CH4-H2O-H2O 3B, unused}; tot / totx: totals, same layout
    int4 cap = {0, 0, 0, 0}, capx = {0, 0, 0, 0};
    int match_cap = 256;          // a multiple of 64
    int4 tot = {0, 0, 0, 0}, totx = {0, 0, 0, 0};
    bool built = false;           // the last compute built lists
    // one-pass builds: per-(query, wavefront) staging of each family's evaluated n-mers, sized
    // from the per-wavefront maxima (maxw) of the last synchronised build
    DBuf<int> stage[kFamilies];
    DBuf<int4> wmax, wmax_x, maxw;  // per-query largest per-wavefront counts, their maxima (2)
    int scap[kFamilies] = {0, 0, 0, 0, 0};
    bool staged = false;          // the last build was one-pass (listed lists not written)
    bool lists_written = false;   // list2 / list3 hold the last build's listed n-mers
    NmersArgs args{};             // the last build's arguments (to write the listed lists later)
    NmerCtx nctx{};
};

// Per-order evaluation buffers.
struct EvalState {
    DBuf<double> e_nmer, vir_nmer, partials, ws;
};

// Largest match storage kept in LDS (80 bytes per match): 512 matches = 40 KB per block.
constexpr int kMaxLdsMatches = 384;  // 160 bytes of LDS a match: at most 60 KB a block
constexpr int kScratchBlocks = 2048;
// slices per configuration in the per-configuration reductions (blocks = configurations x slices)
constexpr int kConfSlices = 16;

}  // namespace

struct mbxg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;               // the 2B evaluation runs here, beside the 3B one
    hipEvent_t fork = nullptr, join = nullptr;   // stream2 waits for `fork`, stream waits for `join`
    // topology
    int nmon = 0, natoms = 0, np = 0;
    std::vector<int> kind, nat, first_atom, tag, members;
    std::vector<unsigned char> islocal;
    DBuf<int> d_kind, d_nat, d_first, d_tag, d_members;
    DBuf<unsigned char> d_islocal;
    // state
    Box box{};
    double cut2 = 0.0, cut3 = 0.0;
    DBuf<double> d_in, d_xyz, d_grad, d_totals, d_coef2, d_coef3;
    DBuf<double> d_coefn[kFamilies];  // MB-nrg coefficient tables (families 1, 2, 4)
    DBuf<double> d_fam_tot;           // per family: energy + 6 virial components
    DBuf<Box> d_box;
    // configurations (mbxg_set_batch): monomers [mon_off[c], mon_off[c+1]) form configuration c;
    // its cloud points are [pt_off[c], pt_off[c+1])
    int nconf = 1;
    std::vector<int> mon_off, pt_off;
    DBuf<int> d_pconf, d_pt_off;
    DBuf<double> d_conf_tot;  // per configuration: 2B record, then 3B record (kPartials each)
    DBuf<double> d_conf_fam;  // per family and configuration (kPartials each)
    DBuf<double> d_conf_part; // per family, configuration and slice (kConfSlices per configuration)
    ListState L;
    EvalState fam[kFamilies];
    bool have_coords = false, have_results = false;
    bool need_sync = true;  // next compute sizes the buffers (first call, new topology/box/cutoffs)
    bool checked = false;   // the last compute's totals were checked against the capacities
    bool evaluated = false; // the last compute evaluated energies (not lists only)
    bool inputs_changed = false;  // coordinates, box, cutoffs or coefficients set since the last compute
    int last_terms = 0, last_grads = 0;
    // HIP graph replay of the steady-state pipeline (mbxg_set_graph)
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    int g_terms = -1, g_grads = -1;       // what gexec computes
    int eager_terms = -1, eager_grads = -1;  // last eager compute since the buffers were sized
    int fam_mask = (1 << (kFamilies + kXFam)) - 1;  // evaluated families (mbxg_set_families)
    int use_ghost = 0;                    // the cloud holds ghost monomers too (mbxg_set_use_ghost)
    // 1B term (MBXG_1B): the local waters, then the local CH4s; per monomer energy and virial rows
    std::vector<int> water_local, ch4_local, co2_local;
    DBuf<int> d_w1b;
    DBuf<double> d_e1b, d_v1b, d_coef1b, d_coef1b_co2;
    bool one_body_last = false;  // the last compute evaluated the 1B term
    DBuf<double> d_wrap_in, d_wrap_out;  // mbxg_fix_coordinates
    int n_h2o = 0, n_ch4 = 0;             // monomers of each kind (mbxg_set_topology)
    // wavefronts per 64 n-mers in the polynomial stages (mbxg_set_tuning / mbxg_autotune)
    int split2 = 4, split3 = 2, split3n = 4;
    int fuse_poly = 1;       // MB-pol 2B and 3B polynomials in one launch when both are evaluated
    int fuse_prep = 1;       // FixMonomerCoordinates, cloud and cell counts in one launch (k_cloud_fix)
    int fuse_chain = 1;      // the 3B chain rule in the paired polynomial launch's 3B blocks
    int fuse_vars = 1;       // the paired path's 2B and 3B variables in one launch
    bool fused_last = false;  // the last compute did so (its 2B polynomial time is in the 3B one)
    // stage timing (mbxg_set_timing): events bracketing each stage of the last compute
    bool timing = false;
    hipEvent_t ev[kNumEvents] = {};
    // further MB-nrg families and the 4B term (xfam.hip); they run synchronously
    double cut4 = 0.0;
    bool has_xkind = false;                  // the topology holds co2 or ion monomers
    DBuf<int> x_code, x_blk, x_off, x_tot;   // classification scratch; x_tot: [order 2, 3, 4][kXFam]
    DBuf<int> x_list[kXFam];
    int x_count[kXFam] = {};
    EvalState xf[kXFam];
    DBuf<double> d_xcoef[kXFam];
    DBuf<double> d_xfam_tot;                 // per further family: energy + 6 virial components
    DBuf<int> t_counts, t_off, t_list, t_ovf;  // tetramers
    int64_t n4_listed = 0;
    bool x_last = false;                     // the last compute ran the further families / 4B
};

int mbxg::report_error(int code, const char* msg) {
    g_err = msg;
    return code;
}

extern "C" const char* mbxg_last_error(void) { return g_err.c_str(); }

extern "C" int mbxg_create(mbxg_ctx** out, int device) {
    if (!out) return fail(MBXG_EINVAL, "mbxg_create: null output pointer");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(MBXG_EINVAL, "mbxg_create: device %d of %d", device, ndev);
    HIPCHK(hipSetDevice(device));
    mbxg_ctx* c = new mbxg_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
    if (e != hipSuccess) {
        delete c;
        return fail(MBXG_EHIP, "mbxg_create: stream/event creation failed: %s", hipGetErrorString(e));
    }
    int n1b = 0, n1b_co2 = 0;
    const double* fit1b = ch4_1b_fit(&n1b);
    const double* fit1b_co2 = co2_1b_fit(&n1b_co2);
    // totals: 2B record, 3B record, 1B record, 4B record (kPartials each)
    bool xbad = false;
    for (int f = 0; f < kXFam; ++f) {
        int nx = 0;
        xfam_default_coef(f, &nx);
        xbad = xbad || c->d_xcoef[f].reserve(nx);
    }
    if (xbad || c->d_xfam_tot.reserve(kXFam * kPartials) || c->d_coef2.reserve(kNCoef_mbpol_2b) ||
        c->d_coef3.reserve(kNCoef_mbpol_3b) || c->d_totals.reserve(4 * kPartials) || c->d_coef1b.reserve(n1b) ||
        c->d_coef1b_co2.reserve(n1b_co2) ||
        c->d_box.reserve(1) || c->d_fam_tot.reserve(kFamilies * kPartials) ||
        c->d_coefn[FAM_CH4H2O_2B].reserve(kNCoef_ch4h2o_2b) || c->d_coefn[FAM_CH4CH4_2B].reserve(kNCoef_ch4ch4_2b) ||
        c->d_coefn[FAM_CH4H2O2_3B].reserve(kNCoef_ch4h2o2_3b)) {
        delete c;
        return fail(MBXG_ENOMEM, "mbxg_create: device allocation failed");
    }
    HIPCHK(hipMemcpy(c->d_coef2.p, kFit_mbpol_2b, sizeof(double) * kNCoef_mbpol_2b, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_coef1b.p, fit1b, sizeof(double) * n1b, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_coef1b_co2.p, fit1b_co2, sizeof(double) * n1b_co2, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->d_totals.p, 0, sizeof(double) * 4 * kPartials));
    for (int f = 0; f < kXFam; ++f) {
        int nx = 0;
        const double* fit = xfam_default_coef(f, &nx);
        HIPCHK(hipMemcpy(c->d_xcoef[f].p, fit, sizeof(double) * nx, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->d_coef3.p, kFit_mbpol_3b, sizeof(double) * kNCoef_mbpol_3b, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_coefn[FAM_CH4H2O_2B].p, kFit_ch4h2o_2b, sizeof(double) * kNCoef_ch4h2o_2b,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_coefn[FAM_CH4CH4_2B].p, kFit_ch4ch4_2b, sizeof(double) * kNCoef_ch4ch4_2b,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_coefn[FAM_CH4H2O2_3B].p, kFit_ch4h2o2_3b, sizeof(double) * kNCoef_ch4h2o2_3b,
                     hipMemcpyHostToDevice));
    c->box.periodic = 0;
    *out = c;
    return MBXG_OK;
}

extern "C" int mbxg_destroy(mbxg_ctx* c) {
    if (!c) return MBXG_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    ListState& L = c->L;
    L.d_grids.release();
    L.d_cell_base.release();
    L.d_cell_bound.release();
    L.cx.release();
    L.pinfo.release();
    L.cy.release();
    L.cz.release();
    L.bx.release();
    L.by.release();
    L.bz.release();
    L.cell.release();
    L.cell_count.release();
    L.cell_pos.release();
    L.cell_start.release();
    L.cell_slot.release();
    L.counts.release();
    L.offsets.release();
    L.list2.release();
    L.list3.release();
    L.counts_x.release();
    L.offsets_x.release();
    L.wpre.release();
    L.wpre_x.release();
    for (DBuf<int>& b : L.evals) b.release();
    L.gscratch.release();
    c->d_pconf.release();
    c->d_pt_off.release();
    c->d_conf_tot.release();
    c->d_conf_fam.release();
    c->d_conf_part.release();
    c->d_fam_tot.release();
    for (DBuf<double>& b : c->d_coefn) b.release();
    for (DBuf<double>& b : c->d_xcoef) b.release();
    for (DBuf<int>& b : c->x_list) b.release();
    for (EvalState* o = c->xf; o != c->xf + kXFam; ++o) {
        o->e_nmer.release();
        o->partials.release();
        o->ws.release();
    }
    c->d_xfam_tot.release();
    c->x_code.release();
    c->x_blk.release();
    c->x_off.release();
    c->x_tot.release();
    c->t_counts.release();
    c->t_off.release();
    c->t_list.release();
    c->t_ovf.release();
    for (EvalState* o = c->fam; o != c->fam + kFamilies; ++o) {
        o->e_nmer.release();
        o->vir_nmer.release();
        o->partials.release();
        o->ws.release();
    }
    c->d_kind.release();
    c->d_nat.release();
    c->d_first.release();
    c->d_tag.release();
    c->d_members.release();
    c->d_islocal.release();
    c->d_in.release();
    c->d_xyz.release();
    c->d_grad.release();
    c->d_totals.release();
    c->d_w1b.release();
    c->d_e1b.release();
    c->d_v1b.release();
    c->d_coef1b.release();
    c->d_coef1b_co2.release();
    c->d_wrap_in.release();
    c->d_wrap_out.release();
    c->d_coef2.release();
    c->d_coef3.release();
    c->d_box.release();
    for (hipEvent_t& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    (void)hipEventDestroy(c->fork);
    (void)hipEventDestroy(c->join);
    (void)hipStreamDestroy(c->stream2);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MBXG_OK;
}

namespace {

// Split the monomers into configurations (CSR offsets over monomers; nullptr = one).  Cloud
// points keep monomer order, so each configuration's points are a contiguous range.
// Cloud members (AddClusters, sys_tools.cpp:502-515): every monomer with use_ghost, the local
// ones otherwise; queries are the local members either way (sys_tools.cpp:537).
int set_members(mbxg_ctx* c) {
    c->members.clear();
    for (int i = 0; i < c->nmon; ++i)
        if (c->use_ghost || c->islocal[i]) c->members.push_back(i);
    c->np = (int)c->members.size();
    if (c->d_members.reserve(c->nmon > 0 ? c->nmon : 1)) return fail(MBXG_ENOMEM, "members: device allocation failed");
    if (c->np > 0)
        HIPCHK(hipMemcpy(c->d_members.p, c->members.data(), sizeof(int) * c->np, hipMemcpyHostToDevice));
    return MBXG_OK;
}

int set_configs(mbxg_ctx* c, int nconf, const int* mon_off) {
    std::vector<int> off;
    if (!mon_off || nconf <= 1) {
        nconf = 1;
        off = {0, c->nmon};
    } else {
        off.assign(mon_off, mon_off + nconf + 1);
        if (off[0] != 0 || off[nconf] != c->nmon)
            return fail(MBXG_EINVAL, "mbxg_set_batch: offsets must run from 0 to the monomer count %d", c->nmon);
        for (int k = 0; k < nconf; ++k)
            if (off[k + 1] < off[k]) return fail(MBXG_EINVAL, "mbxg_set_batch: offsets must not decrease");
    }
    std::vector<int> pt_off(nconf + 1, 0), pconf(c->np > 0 ? c->np : 1, 0);
    int k = 0;
    for (int p = 0; p < c->np; ++p) {
        while (c->members[p] >= off[k + 1]) ++k;
        pconf[p] = k;
        ++pt_off[k + 1];
    }
    for (int q = 0; q < nconf; ++q) pt_off[q + 1] += pt_off[q];
    if (c->d_pconf.reserve(pconf.size()) || c->d_pt_off.reserve(nconf + 1) ||
        c->d_conf_tot.reserve((size_t)2 * kPartials * nconf) ||
        c->d_conf_fam.reserve((size_t)kFamilies * kPartials * nconf) ||
        c->d_conf_part.reserve((size_t)kFamilies * kConfSlices * kPartials * nconf))
        return fail(MBXG_ENOMEM, "configurations: device allocation failed");
    HIPCHK(hipMemcpy(c->d_pconf.p, pconf.data(), sizeof(int) * pconf.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_pt_off.p, pt_off.data(), sizeof(int) * (nconf + 1), hipMemcpyHostToDevice));
    c->nconf = nconf;
    c->mon_off = off;
    c->pt_off = pt_off;
    c->L.layout_ok = false;
    c->need_sync = true;
    c->have_results = false;
    return MBXG_OK;
}

}  // namespace

extern "C" int mbxg_set_topology(mbxg_ctx* c, int nmon, const int* kind, const int* nat, const int* tag,
                                 const unsigned char* islocal) {
    if (!c || nmon < 0 || (nmon > 0 && (!kind || !nat || !tag)))
        return fail(MBXG_EINVAL, "mbxg_set_topology: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    c->nmon = nmon;
    c->kind.assign(kind, kind + nmon);
    c->nat.assign(nat, nat + nmon);
    c->tag.assign(tag, tag + nmon);
    c->islocal.assign(nmon, 1);
    if (islocal)
        for (int i = 0; i < nmon; ++i) c->islocal[i] = islocal[i] ? 1 : 0;
    c->first_atom.resize(nmon);
    int acc = 0;
    for (int i = 0; i < nmon; ++i) {
        if (kind[i] == MBXG_H2O && nat[i] != 3) return fail(MBXG_EINVAL, "monomer %d: h2o needs 3 atoms", i);
        if (kind[i] == MBXG_CH4 && nat[i] != 5) return fail(MBXG_EINVAL, "monomer %d: ch4 needs 5 atoms", i);
        if (kind[i] == MBXG_CO2 && nat[i] != 3) return fail(MBXG_EINVAL, "monomer %d: co2 needs 3 atoms", i);
        if (kind[i] >= MBXG_F && kind[i] <= MBXG_CS && nat[i] != 1)
            return fail(MBXG_EINVAL, "monomer %d: an ion has 1 atom", i);
        if (kind[i] < MBXG_H2O || kind[i] > MBXG_CS) return fail(MBXG_EINVAL, "monomer %d: unknown kind %d", i, kind[i]);
        if (nat[i] < 1) return fail(MBXG_EINVAL, "monomer %d: no atoms", i);
        c->first_atom[i] = acc;
        acc += nat[i];
    }
    c->natoms = acc;
    c->n_h2o = (int)std::count(kind, kind + nmon, MBXG_H2O);
    c->n_ch4 = (int)std::count(kind, kind + nmon, MBXG_CH4);
    c->has_xkind = false;
    for (int i = 0; i < nmon; ++i) c->has_xkind = c->has_xkind || (kind[i] >= MBXG_CO2 && kind[i] <= MBXG_CS);
    const size_t nm = nmon > 0 ? nmon : 1;
    if (c->d_kind.reserve(nm) || c->d_nat.reserve(nm) || c->d_first.reserve(nm) || c->d_tag.reserve(nm) ||
        c->d_members.reserve(nm) || c->d_islocal.reserve(nm) || c->d_in.reserve(3 * (size_t)acc + 3) ||
        c->d_xyz.reserve(3 * (size_t)acc + 3) ||
        c->d_grad.reserve(3 * (size_t)acc + 3))
        return fail(MBXG_ENOMEM, "mbxg_set_topology: device allocation failed");
    if (nmon > 0) {
        HIPCHK(hipMemcpy(c->d_kind.p, c->kind.data(), sizeof(int) * nmon, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_nat.p, c->nat.data(), sizeof(int) * nmon, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_first.p, c->first_atom.data(), sizeof(int) * nmon, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_tag.p, c->tag.data(), sizeof(int) * nmon, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_islocal.p, c->islocal.data(), nmon, hipMemcpyHostToDevice));
    }
    if (int r = set_members(c)) return r;
    if (int r = set_configs(c, 1, nullptr)) return r;
    c->water_local.clear();
    c->ch4_local.clear();
    c->co2_local.clear();
    for (int i = 0; i < nmon; ++i) {
        if (c->kind[i] == MBXG_H2O && c->islocal[i]) c->water_local.push_back(i);
        if (c->kind[i] == MBXG_CH4 && c->islocal[i]) c->ch4_local.push_back(i);
        if (c->kind[i] == MBXG_CO2 && c->islocal[i]) c->co2_local.push_back(i);
    }
    std::vector<int> m1b(c->water_local);
    m1b.insert(m1b.end(), c->ch4_local.begin(), c->ch4_local.end());
    m1b.insert(m1b.end(), c->co2_local.begin(), c->co2_local.end());
    if (c->d_w1b.reserve(nm) || c->d_e1b.reserve(nm) || c->d_v1b.reserve(6 * nm))
        return fail(MBXG_ENOMEM, "mbxg_set_topology: device allocation failed");
    if (!m1b.empty())
        HIPCHK(hipMemcpy(c->d_w1b.p, m1b.data(), sizeof(int) * m1b.size(), hipMemcpyHostToDevice));
    c->have_coords = false;
    c->have_results = false;
    c->need_sync = true;
    return MBXG_OK;
}

extern "C" int mbxg_set_families(mbxg_ctx* c, int mask) {
    if (!c || (mask & ~((1 << (kFamilies + kXFam)) - 1)))
        return fail(MBXG_EINVAL, "mbxg_set_families: bad arguments");
    if (mask == c->fam_mask) return MBXG_OK;
    c->fam_mask = mask;
    c->need_sync = true;
    c->inputs_changed = true;
    return MBXG_OK;
}

extern "C" int mbxg_set_use_ghost(mbxg_ctx* c, int on) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_set_use_ghost: null context");
    on = on ? 1 : 0;
    if (on == c->use_ghost) return MBXG_OK;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->use_ghost = on;
    if (int r = set_members(c)) return r;
    // configuration ranges over the new cloud (needs a new list layout and a synchronised compute)
    if (int r = set_configs(c, c->nconf, c->nconf > 1 ? c->mon_off.data() : nullptr)) return r;
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->g_terms = c->g_grads = -1;
    c->eager_terms = c->eager_grads = -1;
    return MBXG_OK;
}


extern "C" int mbxg_set_batch(mbxg_ctx* c, int nconf, const int* mon_offsets) {
    if (!c || nconf < 1 || (nconf > 1 && !mon_offsets)) return fail(MBXG_EINVAL, "mbxg_set_batch: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    return set_configs(c, nconf, mon_offsets);
}

// The further family of a coefficient-table name, or -1.
static int xfam_index(const std::string& m) {
    static const char* names[kXFam] = {"f_h2o_2b",   "cl_h2o_2b",  "br_h2o_2b",  "i_h2o_2b",   "co2co2_2b",
                                       "f_h2o2_3b",  "cl_h2o2_3b", "br_h2o2_3b", "i_h2o2_3b",  "co2x3_3b",
                                       "h2ox4_4b",   "li_h2o_2b",  "na_h2o_2b",  "k_h2o_2b",   "rb_h2o_2b",
                                       "cs_h2o_2b",  "li_h2o2_3b", "na_h2o2_3b", "k_h2o2_3b",  "rb_h2o2_3b",
                                       "cs_h2o2_3b"};
    for (int f = 0; f < kXFam; ++f)
        if (m == names[f]) return f;
    return -1;
}

// InvertUnitCell (tools/math_tools.cpp:37-64), same operation order.
static int invert_unit_cell(const double* box, double* bi) {
    const double det = box[0] * (box[4] * box[8] - box[7] * box[5]) - box[1] * (box[3] * box[8] - box[5] * box[6]) +
                       box[2] * (box[3] * box[7] - box[4] * box[6]);
    if (std::fabs(det) < 1E-36) return -1;
    const double di = 1 / det;
    bi[0] = (box[4] * box[8] - box[7] * box[5]) * di;
    bi[1] = (box[2] * box[7] - box[1] * box[8]) * di;
    bi[2] = (box[1] * box[5] - box[2] * box[4]) * di;
    bi[3] = (box[5] * box[6] - box[3] * box[8]) * di;
    bi[4] = (box[0] * box[8] - box[2] * box[6]) * di;
    bi[5] = (box[3] * box[2] - box[0] * box[5]) * di;
    bi[6] = (box[3] * box[7] - box[6] * box[4]) * di;
    bi[7] = (box[6] * box[1] - box[0] * box[7]) * di;
    bi[8] = (box[0] * box[4] - box[3] * box[1]) * di;
    return 0;
}

extern "C" int mbxg_set_box(mbxg_ctx* c, const double* box9) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_set_box: null context");
    HIPCHK(hipSetDevice(c->device));
    if (!box9) {
        c->box = Box{};
        c->box.periodic = 0;
        HIPCHK(hipMemcpyAsync(c->d_box.p, &c->box, sizeof(Box), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->need_sync = true;
        c->inputs_changed = true;
        c->L.layout_ok = false;
        return MBXG_OK;
    }
    Box B{};
    std::memcpy(B.b, box9, sizeof(B.b));
    if (invert_unit_cell(B.b, B.bi)) return fail(MBXG_EINVAL, "mbxg_set_box: singular box");
    // GetFractionalCoordinates constants (kdtree/kdtree_utils.h:165-194), evaluated with libm.
    const double* box = B.b;
    const double A = box[0];
    const double Bl = std::sqrt(box[3] * box[3] + box[4] * box[4]);
    const double Cl = std::sqrt(box[6] * box[6] + box[7] * box[7] + box[8] * box[8]);
    const double AdotB = box[0] * box[3];
    const double AdotC = box[0] * box[6];
    const double BdotC = box[3] * box[6] + box[4] * box[7];
    const double gamma = std::acos(AdotB / A / Bl);
    const double beta = std::acos(AdotC / A / Cl);
    const double alpha = std::acos(BdotC / Bl / Cl);
    const double ca = std::cos(alpha), cb = std::cos(beta), cg = std::cos(gamma), sg = std::sin(gamma);
    const double omega = A * Bl * Cl * std::sqrt(1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg);
    B.f00 = 1 / A;
    B.f01 = cg / A / sg;
    B.f02 = Bl * Cl * (ca * cg - cb) / omega / sg;
    B.f11 = 1 / Bl / sg;
    B.f12 = A * Cl * (cb * cg - ca) / omega / sg;
    B.f22 = A * Bl * sg / omega;
    B.periodic = 1;
    c->box = B;
    HIPCHK(hipMemcpyAsync(c->d_box.p, &c->box, sizeof(Box), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->need_sync = true;
    c->inputs_changed = true;
    c->L.layout_ok = false;
    return MBXG_OK;
}

extern "C" int mbxg_set_cutoffs(mbxg_ctx* c, double c2, double c3) {
    if (!c || !(c2 >= 0.0) || !(c3 >= 0.0)) return fail(MBXG_EINVAL, "mbxg_set_cutoffs: bad arguments");
    if (c2 == c->cut2 && c3 == c->cut3) return MBXG_OK;
    c->cut2 = c2;
    c->cut3 = c3;
    c->need_sync = true;
    c->inputs_changed = true;
    c->L.layout_ok = false;
    return MBXG_OK;
}

extern "C" int mbxg_set_coefficients(mbxg_ctx* c, const char* model, const double* a, int n) {
    if (!c || !model) return fail(MBXG_EINVAL, "mbxg_set_coefficients: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    std::string m(model);
    double* dst;
    const double* fit;
    int need;
    if (m == "mbpol_2b") {
        dst = c->d_coef2.p;
        fit = kFit_mbpol_2b;
        need = kNCoef_mbpol_2b;
    } else if (m == "mbpol_3b") {
        dst = c->d_coef3.p;
        fit = kFit_mbpol_3b;
        need = kNCoef_mbpol_3b;
    } else if (m == "ch4h2o_2b") {
        dst = c->d_coefn[FAM_CH4H2O_2B].p;
        fit = kFit_ch4h2o_2b;
        need = kNCoef_ch4h2o_2b;
    } else if (m == "ch4ch4_2b") {
        dst = c->d_coefn[FAM_CH4CH4_2B].p;
        fit = kFit_ch4ch4_2b;
        need = kNCoef_ch4ch4_2b;
    } else if (m == "ch4h2o2_3b") {
        dst = c->d_coefn[FAM_CH4H2O2_3B].p;
        fit = kFit_ch4h2o2_3b;
        need = kNCoef_ch4h2o2_3b;
    } else if (m == "ch4_1b") {
        dst = c->d_coef1b.p;
        fit = ch4_1b_fit(&need);
    } else if (m == "co2_1b") {
        dst = c->d_coef1b_co2.p;
        fit = co2_1b_fit(&need);
    } else if (xfam_index(m) >= 0) {
        const int f = xfam_index(m);
        dst = c->d_xcoef[f].p;
        fit = xfam_default_coef(f, &need);
    } else {
        return fail(MBXG_EINVAL, "mbxg_set_coefficients: unknown model '%s'", model);
    }
    if (a && n != need) return fail(MBXG_EINVAL, "mbxg_set_coefficients: %s takes %d coefficients", model, need);
    HIPCHK(hipMemcpyAsync(dst, a ? a : fit, sizeof(double) * need, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->inputs_changed = true;
    return MBXG_OK;
}

extern "C" int mbxg_set_coords(mbxg_ctx* c, const double* xyz) {
    if (!c || (!xyz && c->natoms > 0)) return fail(MBXG_EINVAL, "mbxg_set_coords: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    if (c->natoms > 0)
        HIPCHK(hipMemcpyAsync(c->d_in.p, xyz, sizeof(double) * 3 * c->natoms, hipMemcpyHostToDevice, c->stream));
    c->have_coords = true;
    c->inputs_changed = true;
    return MBXG_OK;
}

extern "C" int mbxg_set_coords_device(mbxg_ctx* c, const double* d_xyz) {
    if (!c || (!d_xyz && c->natoms > 0)) return fail(MBXG_EINVAL, "mbxg_set_coords_device: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    if (c->natoms > 0)
        HIPCHK(hipMemcpyAsync(c->d_in.p, d_xyz, sizeof(double) * 3 * c->natoms, hipMemcpyDeviceToDevice,
                              c->stream));
    c->have_coords = true;
    c->inputs_changed = true;
    return MBXG_OK;
}

namespace {

// Candidate grid of a periodic cell over the slots' extended fractional range [-1, 1]^3: n
// cells per direction, each at least `cut` wide along the perpendicular (the cell's
// perpendicular width is 2 V / (n |b x c|) etc.).
Grid periodic_grid(const Box& box, double cut) {
    Grid G{};
    const double* b = box.b;
    auto cross = [](const double* u, const double* v, double* w) {
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
    };
    double bc[3], ca[3], ab[3];
    cross(b + 3, b + 6, bc);
    cross(b + 6, b + 0, ca);
    cross(b + 0, b + 3, ab);
    const double V = std::fabs(b[0] * bc[0] + b[1] * bc[1] + b[2] * bc[2]);
    const double w[3] = {V / std::sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]),
                         V / std::sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]),
                         V / std::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2])};
    for (int d = 0; d < 3; ++d) {
        int n = (int)std::floor(2.0 * w[d] / cut);
        if (n < 1) n = 1;
        if (n > kMaxCellsPerDim) n = kMaxCellsPerDim;
        G.n[d] = n;
        G.lo[d] = -1.0;
        G.inv_w[d] = 0.5 * n;
    }
    G.periodic = 1;
    return G;
}


int with_headroom(int n) { return n + n / 4 + 64; }

void launch_zero_int(int* p, size_t n, hipStream_t st) {
    const int nb = (int)std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_zero_int, dim3(nb > 0 ? nb : 1), dim3(256), 0, st, p, n);
}

void launch_zero(double* p, size_t n, hipStream_t st) {
    const int nb = (int)std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_zero, dim3(nb > 0 ? nb : 1), dim3(256), 0, st, p, n);
}

// Dimer and trimer lists at the larger cutoff (one gather for both orders).  With sync the
// totals and the largest match count are read back and the buffers sized from them.
// Capacity / total / device count of a family's evaluate list.
int fam_cap(const ListState& L, int f) {
    switch (f) {
        case FAM_MBPOL_2B: return L.cap.y;
        case FAM_CH4H2O_2B: return L.capx.x;
        case FAM_CH4CH4_2B: return L.capx.y;
        case FAM_MBPOL_3B: return L.cap.w;
        default: return L.capx.z;
    }
}

int fam_tot(const ListState& L, int f) {
    switch (f) {
        case FAM_MBPOL_2B: return L.tot.y;
        case FAM_CH4H2O_2B: return L.totx.x;
        case FAM_CH4CH4_2B: return L.totx.y;
        case FAM_MBPOL_3B: return L.tot.w;
        default: return L.totx.z;
    }
}

// Per-family largest per-wavefront counts from the two maxima records (counts / counts_x layout).
void wave_maxima(const int4* m, int* mw) {
    mw[FAM_MBPOL_2B] = m[0].y;
    mw[FAM_CH4H2O_2B] = m[1].x;
    mw[FAM_CH4CH4_2B] = m[1].y;
    mw[FAM_MBPOL_3B] = m[0].w;
    mw[FAM_CH4H2O2_3B] = m[1].z;
}

// Whether a family's stages are launched: it is evaluated (mbxg_set_families) and the topology
// holds the monomers it needs.  A family that is not launched has an empty evaluate list, and
// its totals keep the zeros they are cleared to before the families run.
bool family_launched(const mbxg_ctx* c, int f) {
    if (!((c->fam_mask >> f) & 1)) return false;
    switch (f) {
        case FAM_MBPOL_2B: return c->n_h2o >= 2;
        case FAM_CH4H2O_2B: return c->n_h2o >= 1 && c->n_ch4 >= 1;
        case FAM_CH4CH4_2B: return c->n_ch4 >= 2;
        case FAM_MBPOL_3B: return c->n_h2o >= 3;
        default: return c->n_h2o >= 2 && c->n_ch4 >= 1;
    }
}

// The int4 component of a family's count in offsets / offsets_x (k_reduce_configs' comp).
int fam_component(int f) {
    switch (f) {
        case FAM_MBPOL_2B: return 1;
        case FAM_CH4H2O_2B: return 0;
        case FAM_CH4CH4_2B: return 1;
        case FAM_MBPOL_3B: return 3;
        default: return 2;
    }
}

const int* fam_ndev(const ListState& L, int np, int f) {
    const int* a = (const int*)(L.offsets.p + np);
    const int* b = (const int*)(L.offsets_x.p + np);
    switch (f) {
        case FAM_MBPOL_2B: return a + 1;
        case FAM_CH4H2O_2B: return b + 0;
        case FAM_CH4CH4_2B: return b + 1;
        case FAM_MBPOL_3B: return a + 3;
        default: return b + 2;
    }
}

// Cell layout of all configurations: under PBC every configuration gets the same extended
// grid; in the gas phase each configuration's grid is sized on the device within a cell bound
// of 4 cells per point (at least 64, at most 64^3).
int prepare_layout(mbxg_ctx* c, double cut) {
    ListState& L = c->L;
    const int nconf = c->nconf;
    std::vector<int> bound(nconf), base(nconf);
    std::vector<Grid> grids;
    int total = 0;
    if (c->box.periodic) {
        const Grid G = periodic_grid(c->box, cut);
        grids.assign(nconf, G);
        for (int k = 0; k < nconf; ++k) bound[k] = G.n[0] * G.n[1] * G.n[2];
    } else {
        for (int k = 0; k < nconf; ++k) {
            const int npk = c->pt_off[k + 1] - c->pt_off[k];
            bound[k] = std::min(kMaxCells, std::max(64, 4 * npk));
        }
    }
    for (int k = 0; k < nconf; ++k) {
        base[k] = total;
        total += bound[k];
    }
    if (L.d_grids.reserve(nconf) || L.d_cell_base.reserve(nconf) || L.d_cell_bound.reserve(nconf) ||
        L.cell_count.reserve(total) || L.cell_start.reserve(total + 1))
        return fail(MBXG_ENOMEM, "cell layout: device allocation failed");
    hipStream_t st = c->stream;
    launch_zero_int(L.cell_count.p, total, st);  // every scan leaves the counts zero again
    HIPCHK(hipMemcpyAsync(L.d_cell_base.p, base.data(), sizeof(int) * nconf, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(L.d_cell_bound.p, bound.data(), sizeof(int) * nconf, hipMemcpyHostToDevice, st));
    if (!grids.empty())
        HIPCHK(hipMemcpyAsync(L.d_grids.p, grids.data(), sizeof(Grid) * nconf, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    L.ncell_total = total;
    L.layout_cut = cut;
    L.layout_ok = true;
    return MBXG_OK;
}

// Dimer and trimer lists at the larger cutoff (one gather for both orders).  With sync the
// totals and the largest match count are read back and the buffers sized from them.
//
// Two builds: the count pass, scan and emit pass (listed lists and evaluate lists), used when
// synchronising or when the listed lists are wanted (need_lists); otherwise one pass that counts
// and stages the evaluated n-mers per (query, wavefront), the scan, and a compaction into the
// evaluate lists (the listed lists are written later, if asked for, by mbxg_get_lists).
int build_lists(mbxg_ctx* c, const NmerCtx& N, double c2sq, double c3sq, double cmax, bool sync, bool need_lists,
                const ZeroJobs& Z, hipEvent_t* ev) {
    ListState& L = c->L;
    hipStream_t st = c->stream;
    const int np = c->np;
    const int nslot = c->box.periodic ? 8 : 1;
    const int ns = np * nslot;
    const double cut = cmax * (1.0 + 1e-9) + 1e-9;
    if (!L.layout_ok || L.layout_cut != cut)
        if (int r = prepare_layout(c, cut)) return r;
    if (L.cx.reserve(ns) || L.cy.reserve(ns) || L.cz.reserve(ns) || L.bx.reserve(ns) || L.by.reserve(ns) ||
        L.bz.reserve(ns) || L.cell.reserve(ns) || L.cell_pos.reserve(ns) || L.cell_slot.reserve(ns) || L.pinfo.reserve(np) ||
        L.counts.reserve(np) || L.wpre.reserve((size_t)np * kQueryWaves) || L.wpre_x.reserve((size_t)np * kQueryWaves) ||
        L.offsets.reserve(np + 1) || L.counts_x.reserve(np) || L.offsets_x.reserve(np + 1) ||
        L.wmax.reserve(np) || L.wmax_x.reserve(np) || L.maxw.reserve(2))
        return fail(MBXG_ENOMEM, "list buffers: device allocation failed");
    const int tb = 256, pb = (np + tb - 1) / tb, sb = (ns + tb - 1) / tb;
    // the gas-phase grids read the input coordinates (the working copy equals them there)
    if (!c->box.periodic)
        hipLaunchKernelGGL(k_grid_gas, dim3(c->nconf), dim3(256), 0, st, c->d_in.p, c->d_first.p, c->d_members.p,
                           c->d_pt_off.p, L.d_cell_bound.p, cut, L.d_grids.p);
    if (c->fuse_prep) {  // fix, cloud, cell counts and the zero fills Z in one launch (64-thread
                         // blocks: the cell-count atomics spread over many CUs; eight lanes a point)
        const int nt = std::max(c->nmon, 8 * np);
        hipLaunchKernelGGL(k_cloud_fix, dim3((nt + 63) / 64), dim3(64), 0, st, c->d_in.p, c->d_xyz.p, c->d_first.p,
                           c->d_nat.p, c->nmon, c->d_members.p, c->d_tag.p, c->d_kind.p, np, c->box, L.d_grids.p,
                           c->d_pconf.p, L.d_cell_base.p, L.cx.p, L.cy.p, L.cz.p, L.cell.p, L.pinfo.p,
                           L.cell_count.p, L.cell_pos.p, Z);
    } else {
        if (c->nmon > 0)
            hipLaunchKernelGGL(k_fix_monomers, dim3((c->nmon + 255) / 256), dim3(256), 0, st, c->d_in.p,
                               c->d_xyz.p, c->d_first.p, c->d_nat.p, c->nmon, c->box);
        hipLaunchKernelGGL(k_cloud, dim3(pb), dim3(tb), 0, st, c->d_xyz.p, c->d_first.p, c->d_members.p, c->d_tag.p,
                           c->d_kind.p, np, c->box, L.d_grids.p, c->d_pconf.p, L.d_cell_base.p, L.cx.p, L.cy.p, L.cz.p,
                           L.cell.p, L.pinfo.p);
        hipLaunchKernelGGL(k_bin_count, dim3(sb), dim3(tb), 0, st, L.cell.p, ns, L.cell_count.p, L.cell_pos.p);
    }
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, st, L.cell_count.p, L.ncell_total, L.cell_start.p);
    hipLaunchKernelGGL(k_bin_fill, dim3(sb), dim3(tb), 0, st, L.cell.p, ns, L.cell_start.p, L.cell_pos.p, L.cx.p,
                       L.cy.p, L.cz.p, L.cell_slot.p, L.bx.p, L.by.p, L.bz.p);
    HIPCHK(hipGetLastError());
    NmersArgs A{};
    A.pinfo = L.pinfo.p;
    A.cx = L.cx.p;
    A.cy = L.cy.p;
    A.cz = L.cz.p;
    A.cell = L.cell.p;
    A.cell_start = L.cell_start.p;
    A.cell_slot = L.cell_slot.p;
    A.bx = L.bx.p;
    A.by = L.by.p;
    A.bz = L.bz.p;
    A.grids = L.d_grids.p;
    A.pconf = c->d_pconf.p;
    A.cell_base = L.d_cell_base.p;
    A.np = np;
    A.c2sq = c2sq;
    A.c3sq = c3sq;
    A.counts = L.counts.p;
    A.counts_x = L.counts_x.p;
    A.offsets = L.offsets.p;
    A.offsets_x = L.offsets_x.p;
    A.wpre = L.wpre.p;
    A.wpre_x = L.wpre_x.p;
    A.fam_mask = c->fam_mask;
    A.wmax = L.wmax.p;
    A.wmax_x = L.wmax_x.p;
    const bool one_pass = !sync && !need_lists && L.scap[0] > 0;
    if (one_pass)
        for (int f = 0; f < kFamilies; ++f) {
            A.stage[f] = L.stage[f].p;
            A.scap[f] = L.scap[f];
        }

    auto scratch = [&]() -> double* {
        if (L.match_cap <= kMaxLdsMatches) return nullptr;
        const size_t need = (size_t)kScratchBlocks * L.match_cap * (nmers_lds_bytes(1) / sizeof(double));
        return L.gscratch.reserve(need) == hipSuccess ? L.gscratch.p : nullptr;
    };
    const int4 none = {0, 0, 0, 0};
    for (int attempt = 0;; ++attempt) {
        double* gs = scratch();
        if (L.match_cap > kMaxLdsMatches && !gs) return fail(MBXG_ENOMEM, "match scratch: device allocation failed");
        HIPCHK(launch_nmers(A, N, one_pass ? 2 : 0, none, none, L.match_cap, gs, kScratchBlocks, st));
        hipLaunchKernelGGL(k_scan_all, dim3(1), dim3(1024), 0, st, L.wmax.p, L.wmax_x.p, L.maxw.p, L.counts.p,
                           L.offsets.p, L.counts_x.p, L.offsets_x.p, np);
        HIPCHK(hipGetLastError());
        if (!sync) break;
        int4 tot, totx;
        int ovf = 0, mw[kFamilies];
        int4 m2[2];
        HIPCHK(hipMemcpyAsync(&tot, L.offsets.p + np, sizeof(int4), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&totx, L.offsets_x.p + np, sizeof(int4), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(m2, L.maxw.p, sizeof(m2), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        wave_maxima(m2, mw);
        ovf = totx.w;  // the largest match count of any query
        const int want = std::max(64, (ovf + ovf / 4 + 64) / 64 * 64);  // LDS per block: 80 bytes a match
        if (ovf > L.match_cap) {
            if (attempt > 0 || want > 2 * ns + 2)
                return fail(MBXG_EHIP, "neighbour search: match storage did not grow (%d matches)", ovf);
            L.match_cap = want;
            continue;  // counts were taken over clamped matches: count again
        }
        L.match_cap = want;  // smaller storage, more blocks per CU; still holds every match
        L.tot = tot;
        L.totx = totx;
        const int4 cap = {with_headroom(tot.x), with_headroom(tot.y), with_headroom(tot.z), with_headroom(tot.w)};
        const int4 capx = {with_headroom(totx.x), with_headroom(totx.y), with_headroom(totx.z), 0};
        L.cap = cap;
        L.capx = capx;
        bool bad = L.list2.reserve(2 * (size_t)cap.x) || L.list3.reserve(3 * (size_t)cap.z);
        for (int f = 0; f < kFamilies; ++f) {
            bad = bad || L.evals[f].reserve((size_t)family_order(f) * fam_cap(L, f) + 1);
            const int sc = mw[f] + mw[f] / 4 + 16;  // staging per (query, wavefront), with headroom
            bad = bad || L.stage[f].reserve((size_t)np * kQueryWaves * sc * family_order(f) + 1);
            if (!bad) L.scap[f] = sc;
        }
        if (bad) return fail(MBXG_ENOMEM, "n-mer lists: device allocation failed (%d + %d listed)", tot.x, tot.z);
        if (L.match_cap > kMaxLdsMatches && !scratch())
            return fail(MBXG_ENOMEM, "match scratch: device allocation failed");
        break;
    }
    if (ev) HIPCHK(hipEventRecord(ev[2], st));
    A.list2 = L.list2.p;
    A.list3 = L.list3.p;
    for (int f = 0; f < kFamilies; ++f) A.evals[f] = L.evals[f].p;
    L.args = A;
    L.nctx = N;
    L.staged = one_pass;
    L.lists_written = !one_pass;
    if (one_pass) {
        CompactArgs C;
        for (int f = 0; f < kFamilies; ++f) {
            C.stage[f] = L.stage[f].p;
            C.scap[f] = L.scap[f];
            C.evals[f] = L.evals[f].p;
            C.fcap[f] = fam_cap(L, f);
        }
        hipLaunchKernelGGL(k_compact_evals, dim3(std::min(np, 16384)), dim3(64 * kQueryWaves), 0, st, C, np,
                           L.counts.p, L.counts_x.p, L.offsets.p, L.offsets_x.p, L.wpre.p, L.wpre_x.p);
        HIPCHK(hipGetLastError());
        return 0;
    }
    HIPCHK(launch_nmers(A, N, 1, L.cap, L.capx, L.match_cap,
                        L.match_cap > kMaxLdsMatches ? L.gscratch.p : nullptr, kScratchBlocks, st));
    return 0;
}

void drop_graph(mbxg_ctx* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->g_terms = c->g_grads = -1;
    c->eager_terms = c->eager_grads = -1;
}

// One compute: lists, then the evaluation of the requested terms.
// The further MB-nrg families and the 4B term, after the others (synchronous: each order's
// family counts are read back to size its evaluate lists).  Tetramers are listed when the 4B term
// is wanted; listed dimers / trimers are classified when the topology holds co2 or halides.
int run_xfam(mbxg_ctx* c, const NmerCtx& N, bool want2, bool want3, bool want4, bool eval, int do_grads) {
    hipStream_t st = c->stream;
    ListState& L = c->L;
    const int np = c->np;
    c->n4_listed = 0;
    for (int f = 0; f < kXFam; ++f) c->x_count[f] = 0;
    launch_zero(c->d_xfam_tot.p, (size_t)kXFam * kPartials, st);
    if (c->x_tot.reserve(3 * kXFam)) return fail(MBXG_ENOMEM, "further families: device allocation failed");
    if (want4 && np >= 4) {
        const int ns = np * (c->box.periodic ? 8 : 1);
        if (c->t_counts.reserve(np) || c->t_off.reserve(np + 1) || c->t_ovf.reserve(1))
            return fail(MBXG_ENOMEM, "tetramers: device allocation failed");
        launch_zero_int(c->t_ovf.p, 1, st);
        const double c4sq = c->cut4 * c->cut4;
        HIPCHK(launch_tetra(L.pinfo.p, L.cx.p, L.cy.p, L.cz.p, np, ns, N, c4sq, 0, c->t_counts.p, c->t_off.p, nullptr,
                            0, c->t_ovf.p, st));
        int h[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(h, c->t_off.p + np, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(h + 1, c->t_ovf.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (h[1] > 0)
            return fail(MBXG_EINVAL, "tetramers: a monomer has %d neighbours within the 4B cutoff (at most 512)", h[1]);
        c->n4_listed = h[0];
        if (c->t_list.reserve(4 * (size_t)h[0] + 4)) return fail(MBXG_ENOMEM, "tetramers: device allocation failed");
        HIPCHK(launch_tetra(L.pinfo.p, L.cx.p, L.cy.p, L.cz.p, np, ns, N, c4sq, 1, c->t_counts.p, c->t_off.p,
                            c->t_list.p, h[0], c->t_ovf.p, st));
    }
    if (eval) {
        struct Ord {
            int order;
            const int* list;
            int n;
        };
        const Ord ords[3] = {{2, L.list2.p, want2 && c->has_xkind ? L.tot.x : 0},
                             {3, L.list3.p, want3 && c->has_xkind ? L.tot.z : 0},
                             {4, c->t_list.p, want4 ? (int)c->n4_listed : 0}};
        double* grad = do_grads ? c->d_grad.p : nullptr;
        for (const Ord& o : ords) {
            if (o.n <= 0) continue;
            const size_t nb = xclass_blocks(o.n);
            if (c->x_code.reserve(o.n) || c->x_blk.reserve(nb * kXFam) || c->x_off.reserve(nb * kXFam))
                return fail(MBXG_ENOMEM, "further families: device allocation failed");
            int* tot = c->x_tot.p + (o.order - 2) * kXFam;
            HIPCHK(launch_xclass(o.order, o.list, o.n, N, c->fam_mask >> kFamilies, c->x_code.p, c->x_blk.p,
                                 c->x_off.p, tot, st));
            int cnt[kXFam];
            HIPCHK(hipMemcpyAsync(cnt, tot, sizeof(cnt), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            XLists X;
            for (int f = 0; f < kXFam; ++f) {
                X.list[f] = nullptr;
                X.cap[f] = 0;
                if (xfam_order(f) != o.order || cnt[f] <= 0) continue;
                c->x_count[f] = cnt[f];
                EvalState& e = c->xf[f];
                if (c->x_list[f].reserve((size_t)(o.order + 1) * cnt[f]) || e.e_nmer.reserve(cnt[f]) ||
                    e.partials.reserve((size_t)kPartials * partials_records(cnt[f])) ||
                    e.ws.reserve(xfam_workspace(f, cnt[f])))
                    return fail(MBXG_ENOMEM, "further families: device allocation failed");
                X.list[f] = c->x_list[f].p;
                X.cap[f] = cnt[f];
            }
            HIPCHK(launch_xemit(o.order, o.list, o.n, c->x_code.p, c->x_off.p, X, st));
            for (int f = 0; f < kXFam; ++f) {
                if (!X.cap[f]) continue;
                EvalState& e = c->xf[f];
                HIPCHK(launch_xfam(f, X.list[f], tot + f, cnt[f], N, c->d_xcoef[f].p, grad, e.e_nmer.p, e.partials.p,
                                   e.ws.p, st));
                hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, st, e.partials.p, tot + f, cnt[f],
                                   c->d_xfam_tot.p + (size_t)kPartials * f);
                HIPCHK(hipGetLastError());
            }
        }
    }
    HIPCHK(launch_xcombine(c->d_xfam_tot.p, c->d_totals.p, st));
    return MBXG_OK;
}

int run(mbxg_ctx* c, int terms, int do_grads, bool sync) {
    hipStream_t st = c->stream;
    ListState& L = c->L;
    hipEvent_t* ev = c->timing ? c->ev : nullptr;
    if (ev)
        for (int k = 0; k < kNumEvents; ++k) HIPCHK(hipEventRecord(ev[k], st));  // stages skipped read 0
    if (ev) HIPCHK(hipEventRecord(ev[0], st));
    NmerCtx N{c->d_xyz.p, c->d_first.p, c->d_nat.p, c->d_kind.p, c->d_tag.p, c->d_islocal.p, c->d_members.p, c->d_box.p,
              c->box.periodic};
    const bool want2 = (terms & MBXG_2B) && c->cut2 > 0.0;
    const bool want3 = (terms & MBXG_3B) && c->cut3 > 0.0;
    const bool want4 = (terms & MBXG_4B) && c->cut4 > 0.0;
    const double c2sq = want2 ? c->cut2 * c->cut2 : 0.0;
    const double c3sq = want3 ? c->cut3 * c->cut3 : 0.0;
    // the further families read the listed lists, the tetramer search the cloud
    const bool xrun = (c->has_xkind && (want2 || want3)) || want4;
    L.built = (want2 || want3 || want4) && c->np >= 2;
    // FixMonomerCoordinates: here when no lists are built, else inside build_lists (fused with
    // the cloud for small periodic clouds)
    if (c->nmon > 0 && !L.built) {
        hipLaunchKernelGGL(k_fix_monomers, dim3((c->nmon + 255) / 256), dim3(256), 0, st, c->d_in.p, c->d_xyz.p,
                           c->d_first.p, c->d_nat.p, c->nmon, c->box);
        HIPCHK(hipGetLastError());
    }
    // the gradient, the 1B record when no 1B term runs and (batches) the family records start at
    // zero: zeroed by the list build's first kernel, or by one launch of their own
    c->one_body_last = (terms & MBXG_1B) && !(terms & MBXG_LISTS_ONLY);
    const bool batch = c->nconf > 1;
    ZeroJobs Z{};
    if (c->natoms > 0) {
        Z.p[0] = c->d_grad.p;
        Z.n[0] = 3 * (size_t)c->natoms;
    }
    if (!c->one_body_last) {
        Z.p[1] = c->d_totals.p + 2 * kPartials;
        Z.n[1] = kPartials;
    }
    if (batch) {
        Z.p[2] = c->d_fam_tot.p;
        Z.n[2] = (size_t)kFamilies * kPartials;
        Z.p[3] = c->d_conf_fam.p;
        Z.n[3] = (size_t)kFamilies * kPartials * c->nconf;
    }
    if (!(L.built && c->fuse_prep)) {
        hipLaunchKernelGGL(k_zero_multi, dim3(64), dim3(256), 0, st, Z);
        HIPCHK(hipGetLastError());
    }
    if (ev) HIPCHK(hipEventRecord(ev[1], st));
    if (L.built) {
        const double cmax = std::max(std::max(want2 ? c->cut2 : 0.0, want3 ? c->cut3 : 0.0), want4 ? c->cut4 : 0.0);
        if (int r = build_lists(c, N, c2sq, c3sq, cmax, sync, (terms & MBXG_LISTS_ONLY) != 0 || xrun, Z, ev))
            return r;
    } else {
        L.tot = make_int4(0, 0, 0, 0);
        L.totx = make_int4(0, 0, 0, 0);
    }
    if (ev) HIPCHK(hipEventRecord(ev[3], st));
    // 1B (System::Get1B, system.cpp:2269-2401): every local monomer's gradient is written at its own
    // atoms of the zeroed gradient before any 2B/3B kernel adds to it (they start after the fork)
    if (c->one_body_last) {
        const int nw = (int)c->water_local.size(), nc = (int)c->ch4_local.size();
        const int no = (int)c->co2_local.size(), n1 = nw + nc + no;
        double* g = do_grads ? c->d_grad.p : nullptr;
        double* v = do_grads ? c->d_v1b.p : nullptr;
        HIPCHK(launch_1b_ps(c->d_xyz.p, c->d_first.p, c->d_w1b.p, nw, c->d_e1b.p, g, v, n1, st));
        HIPCHK(launch_1b_ch4(c->d_xyz.p, c->d_first.p, c->d_w1b.p + nw, nc, c->d_coef1b.p, c->d_e1b.p + nw, g,
                             v ? v + nw : nullptr, n1, st));
        HIPCHK(launch_1b_co2(c->d_xyz.p, c->d_first.p, c->d_w1b.p + nw + nc, no, c->d_coef1b_co2.p,
                             c->d_e1b.p + nw + nc, g, v ? v + nw + nc : nullptr, n1, st));
        hipLaunchKernelGGL(k_reduce_1b, dim3(1), dim3(256), 0, st, c->d_e1b.p, v, n1, n1, c->d_totals.p + 2 * kPartials);
    }
    HIPCHK(hipGetLastError());
    c->evaluated = L.built && !(terms & MBXG_LISTS_ONLY);
    FinishJobs J{};  // single configuration: the families' records, reduced after the join
    if (c->evaluated) {
        // the 2B families run on stream2, concurrently with the 3B ones: their small blocks
        // fill the machine while the last 3B blocks drain
        HIPCHK(hipEventRecord(c->fork, st));
        HIPCHK(hipStreamWaitEvent(c->stream2, c->fork, 0));
        const int order_of[kFamilies] = {FAM_MBPOL_3B, FAM_CH4H2O2_3B, FAM_MBPOL_2B, FAM_CH4H2O_2B, FAM_CH4CH4_2B};
        auto active = [&](int f) {
            return (family_order(f) == 2 ? want2 : want3) && family_launched(c, f) && fam_cap(L, f) > 0;
        };
        // the MB-pol 2B and 3B polynomials share one launch (k_poly_eg_pair) when both run with
        // gradients; everything of both families then runs on the main stream
        const bool fuse = c->fuse_poly && do_grads && active(FAM_MBPOL_2B) && active(FAM_MBPOL_3B);
        c->fused_last = fuse;
        auto buffers = [&](int f) -> bool {  // false: allocation failed
            const int cap = fam_cap(L, f);
            EvalState& o = c->fam[f];
            const int ld = (cap + 63) / 64 * 64;
            const size_t wsz = f == FAM_MBPOL_2B   ? workspace_mbpol2b(cap)
                               : f == FAM_MBPOL_3B ? workspace_mbpol3b(cap)
                                                   : workspace_nrg_family(f, cap);
            return !(o.e_nmer.reserve(cap) || o.partials.reserve((size_t)kPartials * partials_records(cap)) ||
                     o.ws.reserve(wsz) || (batch && o.vir_nmer.reserve((size_t)6 * ld)));
        };
        auto reduce = [&](int f, hipStream_t sk) {
            const int cap = fam_cap(L, f);
            const int ld = (cap + 63) / 64 * 64;
            EvalState& o = c->fam[f];
            const int* nd = fam_ndev(L, c->np, f);
            if (!batch) {
                J.partials[f] = o.partials.p;
                J.n_dev[f] = nd;
                J.cap[f] = cap;
                return;
            }
            hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, sk, o.partials.p, nd, cap,
                               c->d_fam_tot.p + kPartials * f);
            if (batch) {
                double* part = c->d_conf_part.p + (size_t)f * kConfSlices * kPartials * c->nconf;
                hipLaunchKernelGGL(k_reduce_configs, dim3(c->nconf, kConfSlices), dim3(256), 0, sk, o.e_nmer.p,
                                   o.vir_nmer.p, ld, f == FAM_MBPOL_2B || f == FAM_MBPOL_3B ? L.offsets.p : L.offsets_x.p,
                                   c->d_pt_off.p, fam_component(f), cap, part);
                hipLaunchKernelGGL(k_reduce_slices, dim3((c->nconf * kPartials + 255) / 256), dim3(256), 0, sk, part,
                                   c->nconf, kConfSlices, c->d_conf_fam.p + (size_t)f * kPartials * c->nconf);
            }
        };
        for (int i = 0; i < kFamilies; ++i) {
            const int f = order_of[i];
            const bool two = family_order(f) == 2;
            if (!active(f) || (fuse && f == FAM_MBPOL_2B)) continue;
            const int cap = fam_cap(L, f);
            // with stage timing on, everything runs on the main stream so each stage is timed alone
            hipStream_t sk = (two && !ev) ? c->stream2 : st;
            EvalState& o = c->fam[f];
            if (!buffers(f) || (fuse && f == FAM_MBPOL_3B && !buffers(FAM_MBPOL_2B)))
                return fail(MBXG_ENOMEM, "evaluation buffers: device allocation failed");
            double* vir = batch ? o.vir_nmer.p : nullptr;
            const int* nd = fam_ndev(L, c->np, f);
            if (fuse && f == FAM_MBPOL_3B) {
                const int f2 = FAM_MBPOL_2B, cap2 = fam_cap(L, f2);
                EvalState& o2 = c->fam[f2];
                const int* nd2 = fam_ndev(L, c->np, f2);
                // all on the main stream (the 2B stages on stream2 beside the 3B ones measured
                // slower: 0.641-0.646 against 0.636-0.637 ms a step at 2048 waters)
                if (c->fuse_vars) {  // both orders' variables in one launch (timed as the 3B stage)
                    if (ev) HIPCHK(hipEventRecord(ev[4], st));
                    if (ev) HIPCHK(hipEventRecord(ev[5], st));
                    if (ev) HIPCHK(hipEventRecord(ev[8], st));
                    HIPCHK(launch_mbpol_vars_pair(L.evals[f].p, nd, cap, o.ws.p, L.evals[f2].p, nd2, cap2, o2.ws.p,
                                                  N, st));
                    if (ev) HIPCHK(hipEventRecord(ev[9], st));
                } else {
                    if (ev) HIPCHK(hipEventRecord(ev[4], st));
                    HIPCHK(launch_mbpol2b_vars(L.evals[f2].p, nd2, cap2, N, o2.ws.p, st));
                    if (ev) HIPCHK(hipEventRecord(ev[5], st));
                    if (ev) HIPCHK(hipEventRecord(ev[8], st));
                    HIPCHK(launch_mbpol3b_vars(L.evals[f].p, nd, cap, N, o.ws.p, st));
                    if (ev) HIPCHK(hipEventRecord(ev[9], st));
                }
                if (c->fuse_chain) {
                    HIPCHK(launch_poly_mbpol_pair_chain3(o.ws.p, L.evals[f].p, nd, cap, c->d_coef3.p, o2.ws.p, nd2,
                                                         cap2, c->d_coef2.p, c->split3, N, c->d_grad.p, o.e_nmer.p,
                                                         vir, o.partials.p, st));
                    if (ev) HIPCHK(hipEventRecord(ev[10], st));
                } else {
                    HIPCHK(launch_poly_mbpol_pair(o.ws.p, nd, cap, c->d_coef3.p, o2.ws.p, nd2, cap2, c->d_coef2.p,
                                                  c->split3, st));
                    if (ev) HIPCHK(hipEventRecord(ev[10], st));
                    HIPCHK(launch_mbpol3b_chain(L.evals[f].p, nd, cap, N, 1, c->d_grad.p, o.e_nmer.p, vir,
                                                o.partials.p, o.ws.p, st));
                }
                if (ev) HIPCHK(hipEventRecord(ev[11], st));
                if (ev) HIPCHK(hipEventRecord(ev[6], st));
                HIPCHK(launch_mbpol2b_chain(L.evals[f2].p, nd2, cap2, N, 1, c->d_grad.p, o2.e_nmer.p,
                                            batch ? o2.vir_nmer.p : nullptr, o2.partials.p, o2.ws.p, st));
                if (ev) HIPCHK(hipEventRecord(ev[7], st));
                reduce(f, st);
                reduce(f2, st);
                HIPCHK(hipGetLastError());
                continue;
            }
            if (f == FAM_MBPOL_2B)
                HIPCHK(launch_mbpol2b(L.evals[f].p, nd, cap, N, c->d_coef2.p, do_grads, c->d_grad.p, o.e_nmer.p, vir,
                                      o.partials.p, o.ws.p, c->split2, sk, ev ? ev + 4 : nullptr));
            else if (f == FAM_MBPOL_3B)
                HIPCHK(launch_mbpol3b(L.evals[f].p, nd, cap, N, c->d_coef3.p, do_grads, c->d_grad.p, o.e_nmer.p, vir,
                                      o.partials.p, o.ws.p, c->split3, sk, ev ? ev + 8 : nullptr));
            else {
                const bool timed = ev && f == FAM_CH4H2O2_3B;
                if (timed) HIPCHK(hipEventRecord(ev[13], st));
                HIPCHK(launch_nrg_family(f, L.evals[f].p, nd, cap, N, c->d_coefn[f].p, do_grads, c->d_grad.p,
                                         o.e_nmer.p, vir, o.partials.p, o.ws.p, c->split3n, sk, nullptr));
                if (timed) HIPCHK(hipEventRecord(ev[14], st));
            }
            reduce(f, sk);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(c->join, c->stream2));
        HIPCHK(hipStreamWaitEvent(st, c->join, 0));
    }
    // per-order totals: 2B = MB-pol + CH4-H2O + CH4-CH4, 3B = MB-pol + CH4-H2O-H2O (fixed order)
    if (!batch) {
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(kFinishThreads), 0, st, J, c->d_fam_tot.p, c->d_totals.p);
    } else {
        hipLaunchKernelGGL(k_combine_families, dim3(1), dim3(64), 0, st, c->d_fam_tot.p, 1, c->d_totals.p);
        launch_zero(c->d_totals.p + 3 * kPartials, kPartials, st);
    }
    c->x_last = xrun;
    if (xrun)  // then the further families' totals on top, and the 4B record
        if (int r = run_xfam(c, N, want2, want3, want4, !(terms & MBXG_LISTS_ONLY), do_grads)) return r;
    if (c->nconf > 1)
        hipLaunchKernelGGL(k_combine_families, dim3((c->nconf * kPartials + 63) / 64), dim3(64), 0, st,
                           c->d_conf_fam.p, c->nconf, c->d_conf_tot.p);
    HIPCHK(hipGetLastError());
    if (ev) HIPCHK(hipEventRecord(ev[12], st));
    c->have_results = true;
    c->checked = sync || !L.built;
    if (sync && L.built) {  // the list buffers are sized (a 1B-only compute sizes nothing)
        c->need_sync = false;
        drop_graph(c);  // buffers may have moved
    }
    return MBXG_OK;
}

// Wait for the last compute and make sure its lists fit their buffers; if not, compute again
// with buffers sized from a synchronised count.
int finish(mbxg_ctx* c) {
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->have_results || c->checked) return MBXG_OK;
    ListState& L = c->L;
    int4 tot, totx;
    int ovf = 0, mw[kFamilies] = {0, 0, 0, 0, 0};
    int4 m2[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    HIPCHK(hipMemcpyAsync(&tot, L.offsets.p + c->np, sizeof(int4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&totx, L.offsets_x.p + c->np, sizeof(int4), hipMemcpyDeviceToHost, c->stream));
    if (L.staged) HIPCHK(hipMemcpyAsync(m2, L.maxw.p, sizeof(m2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    wave_maxima(m2, mw);
    ovf = totx.w;  // the largest match count of any query
    bool staging_short = false;  // a one-pass build staged more n-mers per wavefront than it holds
    for (int f = 0; f < kFamilies; ++f) staging_short = staging_short || mw[f] > L.scap[f];
    if (tot.x > L.cap.x || tot.y > L.cap.y || tot.z > L.cap.z || tot.w > L.cap.w || totx.x > L.capx.x ||
        totx.y > L.capx.y || totx.z > L.capx.z || ovf > L.match_cap || staging_short) {
        c->need_sync = true;
        if (c->inputs_changed) {
            c->have_results = false;
            return fail(MBXG_ESTATE,
                        "the last compute outgrew its list buffers and its inputs have changed since; compute again");
        }
        if (int r = run(c, c->last_terms, c->last_grads, true)) return r;
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {
        L.tot = tot;
        L.totx = totx;
    }
    c->checked = true;
    return MBXG_OK;
}

}  // namespace

extern "C" int mbxg_compute(mbxg_ctx* c, int terms, int do_grads) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_compute: null context");
    if (!c->have_coords) return fail(MBXG_ESTATE, "mbxg_compute: coordinates not set");
    if (terms & ~(MBXG_1B | MBXG_2B | MBXG_3B | MBXG_LISTS_ONLY | MBXG_4B))
        return fail(MBXG_EINVAL, "mbxg_compute: unknown terms 0x%x", terms);
    if ((terms & MBXG_1B) && c->nconf > 1)
        return fail(MBXG_EINVAL, "mbxg_compute: the 1B term is not evaluated over batches");
    if (!(terms & MBXG_LISTS_ONLY))
        for (int i = 0; i < c->nmon; ++i)
            if (c->kind[i] == MBXG_OTHER)
                return fail(MBXG_EINVAL, "mbxg_compute: monomer %d has no GPU holder (lists only)", i);
    const bool xrun = (c->has_xkind && (terms & (MBXG_2B | MBXG_3B))) || ((terms & MBXG_4B) && c->cut4 > 0.0);
    if (xrun && c->nconf > 1)
        return fail(MBXG_EINVAL, "mbxg_compute: co2, halides and the 4B term are not evaluated over batches");
    HIPCHK(hipSetDevice(c->device));
    c->have_results = false;
    c->last_terms = terms;
    c->last_grads = do_grads;
    c->inputs_changed = false;
    const bool steady = !c->need_sync && !c->timing && !xrun;
    if (!(c->use_graph && steady && c->eager_terms == terms && c->eager_grads == do_grads)) {
        const bool sync = c->need_sync || xrun;
        if (int r = run(c, terms, do_grads, sync)) return r;
        if (!c->timing) {
            c->eager_terms = terms;
            c->eager_grads = do_grads;
        }
        return MBXG_OK;
    }
    // Steady state with the buffers already sized by an eager compute of the same terms:
    // capture the pipeline once, then replay it (no launches from the host but one).
    hipStream_t st = c->stream;
    if (!c->gexec || c->g_terms != terms || c->g_grads != do_grads) {
        if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int r = run(c, terms, do_grads, false);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (r) {
            if (graph) (void)hipGraphDestroy(graph);
            return r;
        }
        if (e != hipSuccess) return fail(MBXG_EHIP, "graph capture failed: %s", hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            c->gexec = nullptr;
            return fail(MBXG_EHIP, "graph instantiation failed: %s", hipGetErrorString(ei));
        }
        c->g_terms = terms;
        c->g_grads = do_grads;
    }
    HIPCHK(hipGraphLaunch(c->gexec, st));
    c->have_results = true;
    c->checked = !c->L.built;
    c->L.lists_written = !c->L.staged;  // the replayed build is the captured one
    return MBXG_OK;
}

extern "C" int mbxg_set_graph(mbxg_ctx* c, int on) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_set_graph: null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->use_graph = on != 0;
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->g_terms = c->g_grads = -1;
    return MBXG_OK;
}

extern "C" int mbxg_set_timing(mbxg_ctx* c, int on) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_set_timing: null context");
    HIPCHK(hipSetDevice(c->device));
    if (on && !c->ev[0])
        for (hipEvent_t& e : c->ev) HIPCHK(hipEventCreate(&e));
    c->timing = on != 0;
    return MBXG_OK;
}

extern "C" int mbxg_get_timings(mbxg_ctx* c, double* ms, int n) {
    if (!c || !ms) return fail(MBXG_EINVAL, "mbxg_get_timings: bad arguments");
    if (!c->timing || !c->have_results) return fail(MBXG_ESTATE, "mbxg_get_timings: no timed compute");
    if (int r = finish(c)) return r;
    // fix, list count pass, list emit pass, vars 2B, poly 2B, chain 2B, vars 3B, poly 3B, chain 3B, total,
    // CH4-H2O-H2O family (variables, polynomial, chain rule)
    const int pairs[11][2] = {{0, 1}, {1, 2}, {2, 3}, {4, 5}, {5, 6}, {6, 7}, {8, 9}, {9, 10}, {10, 11}, {0, 12}, {13, 14}};
    for (int k = 0; k < n && k < 11; ++k) {
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, c->ev[pairs[k][0]], c->ev[pairs[k][1]]));
        ms[k] = t;
    }
    if (c->fused_last && n > 4) ms[4] = 0.0;  // the 2B polynomial ran inside the 3B launch
    return MBXG_OK;
}

namespace {

bool valid_split(int v) { return v == 1 || v == 2 || v == 4 || v == 8; }

}  // namespace

extern "C" int mbxg_set_tuning(mbxg_ctx* c, const char* key, int value) {
    if (!c || !key) return fail(MBXG_EINVAL, "mbxg_set_tuning: bad arguments");
    const std::string k(key);
    if ((k == "split_2b" || k == "split_3b" || k == "split_ch4h2o2_3b") && !valid_split(value))
        return fail(MBXG_EINVAL, "mbxg_set_tuning: %s must be 1, 2, 4 or 8", key);
    if ((k == "fuse_poly" || k == "fuse_prep" || k == "fuse_chain" || k == "fuse_vars") && value != 0 && value != 1)
        return fail(MBXG_EINVAL, "mbxg_set_tuning: %s is 0 or 1", key);
    if (k == "split_2b")
        c->split2 = value;
    else if (k == "split_3b")
        c->split3 = value;
    else if (k == "split_ch4h2o2_3b")
        c->split3n = value;
    else if (k == "fuse_poly")
        c->fuse_poly = value;
    else if (k == "fuse_prep")
        c->fuse_prep = value;
    else if (k == "fuse_chain")
        c->fuse_chain = value;
    else if (k == "fuse_vars")
        c->fuse_vars = value;
    else
        return fail(MBXG_EINVAL, "mbxg_set_tuning: unknown key '%s'", key);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->g_terms = c->g_grads = -1;
    return MBXG_OK;
}

extern "C" int mbxg_get_tuning(mbxg_ctx* c, const char* key, int* value) {
    if (!c || !key || !value) return fail(MBXG_EINVAL, "mbxg_get_tuning: bad arguments");
    const std::string k(key);
    if (k == "split_2b")
        *value = c->split2;
    else if (k == "split_3b")
        *value = c->split3;
    else if (k == "split_ch4h2o2_3b")
        *value = c->split3n;
    else if (k == "fuse_poly")
        *value = c->fuse_poly;
    else if (k == "fuse_prep")
        *value = c->fuse_prep;
    else if (k == "fuse_chain")
        *value = c->fuse_chain;
    else if (k == "fuse_vars")
        *value = c->fuse_vars;
    else
        return fail(MBXG_EINVAL, "mbxg_get_tuning: unknown key '%s'", key);
    return MBXG_OK;
}

// Autotuner: time the polynomial stage of each order for every split on the current inputs
// (HIP events, stages run one at a time), keep the fastest, then recompute once so the results
// are those of the chosen schedule.
extern "C" int mbxg_autotune(mbxg_ctx* c, int terms, int do_grads, int reps) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_autotune: null context");
    if (!c->have_coords) return fail(MBXG_ESTATE, "mbxg_autotune: coordinates not set");
    if (reps < 1) reps = 1;
    const bool was_timing = c->timing;
    if (int r = mbxg_set_timing(c, 1)) return r;
    const int cand[4] = {1, 2, 4, 8};
    double best[2] = {1e300, 1e300};
    int pick[2] = {c->split2, c->split3};
    for (int i = 0; i < 4; ++i) {
        c->split2 = c->split3 = cand[i];
        double t[2] = {1e300, 1e300};
        for (int r = 0; r < reps + 1; ++r) {  // first repetition warms up
            if (int e = mbxg_compute(c, terms, do_grads)) return e;
            double ms[10];
            if (int e = mbxg_get_timings(c, ms, 10)) return e;
            if (r > 0) {
                t[0] = std::min(t[0], ms[4]);
                t[1] = std::min(t[1], ms[7]);
            }
        }
        for (int k = 0; k < 2; ++k)
            if (t[k] < best[k]) {
                best[k] = t[k];
                pick[k] = cand[i];
            }
    }
    c->split2 = pick[0];
    c->split3 = pick[1];
    if ((terms & MBXG_3B) && c->cut3 > 0.0 && family_launched(c, FAM_CH4H2O2_3B)) {
        double best3n = 1e300;
        int pick3n = c->split3n;
        for (int i = 0; i < 4; ++i) {
            c->split3n = cand[i];
            double t = 1e300;
            for (int r = 0; r < reps + 1; ++r) {
                if (int e = mbxg_compute(c, terms, do_grads)) return e;
                double ms[11];
                if (int e = mbxg_get_timings(c, ms, 11)) return e;
                if (r > 0) t = std::min(t, ms[10]);
            }
            if (t < best3n) {
                best3n = t;
                pick3n = cand[i];
            }
        }
        c->split3n = pick3n;
    }
    // The paired path's fusions (the 3B chain rule in the polynomial launch, both orders'
    // variables in one launch): every combination, timed over the whole compute.
    if (c->fuse_poly && do_grads && (terms & MBXG_2B) && (terms & MBXG_3B) && family_launched(c, FAM_MBPOL_2B) &&
        family_launched(c, FAM_MBPOL_3B)) {
        double bestf = 1e300;
        int pickc = c->fuse_chain, pickv = c->fuse_vars;
        for (int fc = 0; fc < 2; ++fc)
            for (int fv = 0; fv < 2; ++fv) {
                c->fuse_chain = fc;
                c->fuse_vars = fv;
                double t = 1e300;
                for (int r = 0; r < reps + 1; ++r) {
                    if (int e = mbxg_compute(c, terms, do_grads)) return e;
                    double ms[10];
                    if (int e = mbxg_get_timings(c, ms, 10)) return e;
                    if (r > 0) t = std::min(t, ms[9]);
                }
                if (t < bestf) {
                    bestf = t;
                    pickc = fc;
                    pickv = fv;
                }
            }
        c->fuse_chain = pickc;
        c->fuse_vars = pickv;
    }
    if (int r = mbxg_set_timing(c, was_timing ? 1 : 0)) return r;
    return mbxg_compute(c, terms, do_grads);
}

extern "C" int mbxg_get_results(mbxg_ctx* c, double* e2b, double* e3b, double* grad, double* virial) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_get_results: null context");
    if (!c->have_results) return fail(MBXG_ESTATE, "mbxg_get_results: nothing computed");
    if (int r = finish(c)) return r;
    double tot[4 * kPartials];
    HIPCHK(hipMemcpyAsync(tot, c->d_totals.p, sizeof(tot), hipMemcpyDeviceToHost, c->stream));
    if (grad && c->natoms > 0)
        HIPCHK(hipMemcpyAsync(grad, c->d_grad.p, sizeof(double) * 3 * c->natoms, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (e2b) *e2b = tot[0];
    if (e3b) *e3b = tot[kPartials];
    if (virial) {
        // xx xy xz yy yz zz -> 3x3 with the reference's mirroring
        const int slot[6] = {0, 1, 2, 4, 5, 8};
        for (int q = 0; q < 9; ++q) virial[q] = 0.0;
        for (int q = 0; q < 6; ++q)
            virial[slot[q]] = tot[2 * kPartials + 1 + q] + tot[1 + q] + tot[kPartials + 1 + q] + tot[3 * kPartials + 1 + q];
        virial[3] = virial[1];
        virial[6] = virial[2];
        virial[7] = virial[5];
    }
    return MBXG_OK;
}

extern "C" int mbxg_get_batch_results(mbxg_ctx* c, double* e2b, double* e3b, double* virial) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_get_batch_results: null context");
    if (!c->have_results) return fail(MBXG_ESTATE, "mbxg_get_batch_results: nothing computed");
    if (int r = finish(c)) return r;
    const int nc = c->nconf;
    std::vector<double> tot((size_t)2 * kPartials * nc, 0.0);
    if (nc > 1) {
        HIPCHK(hipMemcpyAsync(tot.data(), c->d_conf_tot.p, sizeof(double) * tot.size(), hipMemcpyDeviceToHost,
                              c->stream));
    } else {
        HIPCHK(hipMemcpyAsync(tot.data(), c->d_totals.p, sizeof(double) * 2 * kPartials, hipMemcpyDeviceToHost,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    const double* t2 = tot.data();
    const double* t3 = tot.data() + (size_t)kPartials * nc;
    const int slot[6] = {0, 1, 2, 4, 5, 8};
    for (int k = 0; k < nc; ++k) {
        if (e2b) e2b[k] = t2[k * kPartials];
        if (e3b) e3b[k] = t3[k * kPartials];
        if (virial) {
            double* v = virial + 9 * (size_t)k;
            for (int q = 0; q < 9; ++q) v[q] = 0.0;
            for (int q = 0; q < 6; ++q) v[slot[q]] = t2[k * kPartials + 1 + q] + t3[k * kPartials + 1 + q];
            v[3] = v[1];
            v[6] = v[2];
            v[7] = v[5];
        }
    }
    return MBXG_OK;
}

extern "C" int mbxg_get_counts(mbxg_ctx* c, int64_t* counts) {
    if (!c || !counts) return fail(MBXG_EINVAL, "mbxg_get_counts: bad arguments");
    if (int r = finish(c)) return r;
    const int4 t = c->have_results ? c->L.tot : make_int4(0, 0, 0, 0);
    const int4 u = c->have_results ? c->L.totx : make_int4(0, 0, 0, 0);
    counts[0] = t.x;
    counts[1] = c->evaluated ? (int64_t)t.y + u.x + u.y : 0;
    counts[2] = t.z;
    counts[3] = c->evaluated ? (int64_t)t.w + u.z : 0;
    return MBXG_OK;
}

extern "C" int mbxg_get_lists(mbxg_ctx* c, int32_t* dimers, int64_t cap2, int32_t* trimers, int64_t cap3) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_get_lists: null context");
    if (int r = finish(c)) return r;
    if (!c->have_results) return MBXG_OK;
    ListState& L = c->L;
    if (L.built && !L.lists_written) {
        // a one-pass build left the listed lists unwritten: the emit pass writes them from the
        // same cloud, counts and offsets (its evaluate-list writes repeat the compaction's)
        HIPCHK(launch_nmers(L.args, L.nctx, 1, L.cap, L.capx, L.match_cap,
                            L.match_cap > kMaxLdsMatches ? L.gscratch.p : nullptr, kScratchBlocks, c->stream));
        L.lists_written = true;
    }
    const int4 t = c->L.tot;
    if (dimers && t.x > 0) {
        const int64_t n = t.x < cap2 ? t.x : cap2;
        HIPCHK(hipMemcpyAsync(dimers, c->L.list2.p, sizeof(int32_t) * 2 * n, hipMemcpyDeviceToHost, c->stream));
    }
    if (trimers && t.z > 0) {
        const int64_t n = t.z < cap3 ? t.z : cap3;
        HIPCHK(hipMemcpyAsync(trimers, c->L.list3.p, sizeof(int32_t) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return MBXG_OK;
}

extern "C" int mbxg_get_nmer_energies(mbxg_ctx* c, int order, int32_t* nmers, double* energies, int64_t cap) {
    if (!c || (order != 2 && order != 3)) return fail(MBXG_EINVAL, "mbxg_get_nmer_energies: bad arguments");
    if (int r = finish(c)) return r;
    if (!c->have_results || !c->evaluated) return MBXG_OK;
    // families of this order, in family order (MB-pol first)
    int64_t done = 0;
    for (int f = 0; f < kFamilies && done < cap; ++f) {
        if (family_order(f) != order) continue;
        const int64_t n = std::min<int64_t>(fam_tot(c->L, f), cap - done);
        if (n <= 0) continue;
        if (nmers)
            HIPCHK(hipMemcpyAsync(nmers + order * done, c->L.evals[f].p, sizeof(int32_t) * order * n,
                                  hipMemcpyDeviceToHost, c->stream));
        if (energies)
            HIPCHK(hipMemcpyAsync(energies + done, c->fam[f].e_nmer.p, sizeof(double) * n, hipMemcpyDeviceToHost,
                                  c->stream));
        done += n;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return MBXG_OK;
}

extern "C" int mbxg_get_coords(mbxg_ctx* c, double* xyz) {
    if (!c || !xyz) return fail(MBXG_EINVAL, "mbxg_get_coords: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    if (c->natoms > 0)
        HIPCHK(hipMemcpyAsync(xyz, c->d_xyz.p, sizeof(double) * 3 * c->natoms, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MBXG_OK;
}

extern "C" int mbxg_device_outputs(mbxg_ctx* c, double** d_grad, double** d_totals) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_device_outputs: null context");
    if (d_grad) *d_grad = c->d_grad.p;
    if (d_totals) *d_totals = c->d_totals.p;
    return MBXG_OK;
}

extern "C" int mbxg_stream(mbxg_ctx* c, void** stream) {
    if (!c || !stream) return fail(MBXG_EINVAL, "mbxg_stream: bad arguments");
    *stream = (void*)c->stream;
    return MBXG_OK;
}

extern "C" int mbxg_synchronize(mbxg_ctx* c) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_synchronize: null context");
    return finish(c);
}

extern "C" int mbxg_eval_1b(mbxg_ctx* c, int do_grads, double* e1b, double* grad, double* virial) {
    if (!c || !e1b) return fail(MBXG_EINVAL, "mbxg_eval_1b: bad arguments");
    if (int r = mbxg_compute(c, MBXG_1B, do_grads)) return r;
    if (int r = mbxg_get_energies(c, e1b, nullptr, nullptr)) return r;
    if (int r = mbxg_get_results(c, nullptr, nullptr, do_grads ? grad : nullptr, virial)) return r;
    if (!do_grads && grad && c->natoms > 0) std::fill(grad, grad + 3 * (size_t)c->natoms, 0.0);
    return MBXG_OK;
}

extern "C" int mbxg_get_energies(mbxg_ctx* c, double* e1b, double* e2b, double* e3b) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_get_energies: null context");
    if (!c->have_results) return fail(MBXG_ESTATE, "mbxg_get_energies: nothing computed");
    if (int r = finish(c)) return r;
    double tot[3 * kPartials];
    HIPCHK(hipMemcpyAsync(tot, c->d_totals.p, sizeof(tot), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (e1b) *e1b = tot[2 * kPartials];
    if (e2b) *e2b = tot[0];
    if (e3b) *e3b = tot[kPartials];
    return MBXG_OK;
}

extern "C" int mbxg_fix_coordinates(mbxg_ctx* c, const double* xyz_in, double* xyz_out) {
    if (!c || (c->natoms > 0 && (!xyz_in || !xyz_out))) return fail(MBXG_EINVAL, "mbxg_fix_coordinates: bad arguments");
    if (c->natoms == 0) return MBXG_OK;
    HIPCHK(hipSetDevice(c->device));
    const size_t n3 = 3 * (size_t)c->natoms;
    if (c->d_wrap_in.reserve(n3) || c->d_wrap_out.reserve(n3))
        return fail(MBXG_ENOMEM, "mbxg_fix_coordinates: device allocation failed");
    hipStream_t st = c->stream;
    HIPCHK(hipMemcpyAsync(c->d_wrap_in.p, xyz_in, sizeof(double) * n3, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fix_monomers, dim3((c->nmon + 255) / 256), dim3(256), 0, st, c->d_wrap_in.p, c->d_wrap_out.p,
                       c->d_first.p, c->d_nat.p, c->nmon, c->box);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(xyz_out, c->d_wrap_out.p, sizeof(double) * n3, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return MBXG_OK;
}

extern "C" int mbxg_set_cutoff4b(mbxg_ctx* c, double c4) {
    if (!c || !(c4 >= 0.0)) return fail(MBXG_EINVAL, "mbxg_set_cutoff4b: bad arguments");
    if (c4 == c->cut4) return MBXG_OK;
    c->cut4 = c4;
    c->need_sync = true;
    c->inputs_changed = true;
    c->L.layout_ok = false;
    return MBXG_OK;
}

extern "C" int mbxg_get_term_energies(mbxg_ctx* c, double* e) {
    if (!c || !e) return fail(MBXG_EINVAL, "mbxg_get_term_energies: bad arguments");
    if (!c->have_results) return fail(MBXG_ESTATE, "mbxg_get_term_energies: nothing computed");
    if (int r = finish(c)) return r;
    double tot[4 * kPartials];
    HIPCHK(hipMemcpyAsync(tot, c->d_totals.p, sizeof(tot), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    e[0] = tot[2 * kPartials];
    e[1] = tot[0];
    e[2] = tot[kPartials];
    e[3] = tot[3 * kPartials];
    return MBXG_OK;
}

extern "C" int mbxg_get_xcounts(mbxg_ctx* c, int64_t* counts) {
    if (!c || !counts) return fail(MBXG_EINVAL, "mbxg_get_xcounts: bad arguments");
    if (int r = finish(c)) return r;
    for (int f = 0; f < kXFam; ++f) counts[f] = c->have_results && c->x_last ? c->x_count[f] : 0;
    counts[kXFam] = c->have_results && c->x_last ? c->n4_listed : 0;
    return MBXG_OK;
}

extern "C" int mbxg_get_tetramers(mbxg_ctx* c, int32_t* out, int64_t cap, int64_t* n) {
    if (!c || !n || (cap > 0 && !out)) return fail(MBXG_EINVAL, "mbxg_get_tetramers: bad arguments");
    if (int r = finish(c)) return r;
    *n = c->have_results && c->x_last ? c->n4_listed : 0;
    const int64_t k = std::min<int64_t>(*n, cap);
    if (k > 0) {
        HIPCHK(hipMemcpyAsync(out, c->t_list.p, sizeof(int32_t) * 4 * k, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return MBXG_OK;
}
