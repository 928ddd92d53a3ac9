// Host runtime of the mbxg engine: context, device buffers, the 2B/3B pipeline and the C ABI
// declared in include/mbxg.h.
//
// Pipeline of mbxg_compute (all on the context's stream):
//   k_fix_monomers         FixMonomerCoordinates on the working copy (PBC only)
//   per order (2, 3):      k_cloud -> cell binning -> k_dimers|k_trimers count pass ->
//                          k_scan_counts -> [totals to host, grow buffers] -> emit pass
//   k_mbpol2b, k_mbpol3b   polynomial evaluation over the compacted evaluate lists
//   k_reduce_partials      deterministic energy / virial totals
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mbxg.h"
#include "gen/coef_mbpol_2b.inc"
#include "gen/coef_mbpol_3b.inc"
#include "mbxg_kernels.hpp"

using namespace mbxg;

namespace {

// Timing events: 0 start, 1 after fix, 2 after 2B lists, 3 after 3B lists, 4..7 2B stages,
// 8..11 3B stages, 12 end.
constexpr int kNumEvents = 13;

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

// Per-order neighbour-search state.
struct OrderState {
    Grid grid{};
    int ncell = 0;
    DBuf<double> cx, cy, cz;
    DBuf<int> cell, cell_count, cell_start, cell_pts;
    DBuf<int2> counts, offsets;
    DBuf<int> list, eval;
    int64_t n_list = 0, n_eval = 0;
    DBuf<double> e_nmer, partials, ws;
};

}  // namespace

struct mbxg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // topology
    int nmon = 0, natoms = 0, np = 0;
    std::vector<int> kind, nat, first_atom, tag, members;
    std::vector<unsigned char> islocal;
    DBuf<int> d_kind, d_nat, d_first, d_tag, d_members;
    DBuf<unsigned char> d_islocal;
    // state
    Box box{};
    double cut2 = 0.0, cut3 = 0.0;
    DBuf<double> d_xyz, d_grad, d_totals, d_coef2, d_coef3, d_bbox;
    DBuf<Box> d_box;
    DBuf<int> d_overflow;
    OrderState o2, o3;
    bool have_coords = false, have_results = false;
    int match_cap = 1024;
    // stage timing (mbxg_set_timing): events bracketing each stage of the last compute
    bool timing = false;
    hipEvent_t ev[kNumEvents] = {};
};

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
    if (e != hipSuccess) {
        delete c;
        return fail(MBXG_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    if (c->d_coef2.reserve(kNCoef_mbpol_2b) || c->d_coef3.reserve(kNCoef_mbpol_3b) || c->d_totals.reserve(16) ||
        c->d_overflow.reserve(1) || c->d_bbox.reserve(8) || c->d_box.reserve(1)) {
        delete c;
        return fail(MBXG_ENOMEM, "mbxg_create: device allocation failed");
    }
    HIPCHK(hipMemcpy(c->d_coef2.p, kFit_mbpol_2b, sizeof(double) * kNCoef_mbpol_2b, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_coef3.p, kFit_mbpol_3b, sizeof(double) * kNCoef_mbpol_3b, hipMemcpyHostToDevice));
    c->box.periodic = 0;
    *out = c;
    return MBXG_OK;
}

extern "C" int mbxg_destroy(mbxg_ctx* c) {
    if (!c) return MBXG_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (OrderState* o : {&c->o2, &c->o3}) {
        o->cx.release();
        o->cy.release();
        o->cz.release();
        o->cell.release();
        o->cell_count.release();
        o->cell_start.release();
        o->cell_pts.release();
        o->counts.release();
        o->offsets.release();
        o->list.release();
        o->eval.release();
        o->e_nmer.release();
        o->partials.release();
        o->ws.release();
    }
    c->d_kind.release();
    c->d_nat.release();
    c->d_first.release();
    c->d_tag.release();
    c->d_members.release();
    c->d_islocal.release();
    c->d_xyz.release();
    c->d_grad.release();
    c->d_totals.release();
    c->d_coef2.release();
    c->d_coef3.release();
    c->d_bbox.release();
    c->d_box.release();
    c->d_overflow.release();
    for (hipEvent_t& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MBXG_OK;
}

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
        if (nat[i] < 1) return fail(MBXG_EINVAL, "monomer %d: no atoms", i);
        c->first_atom[i] = acc;
        acc += nat[i];
    }
    c->natoms = acc;
    // cloud members: every local monomer (AddClusters with use_ghost = false, sys_tools.cpp:508)
    c->members.clear();
    for (int i = 0; i < nmon; ++i)
        if (c->islocal[i]) c->members.push_back(i);
    c->np = (int)c->members.size();
    const size_t nm = nmon > 0 ? nmon : 1;
    if (c->d_kind.reserve(nm) || c->d_nat.reserve(nm) || c->d_first.reserve(nm) || c->d_tag.reserve(nm) ||
        c->d_members.reserve(nm) || c->d_islocal.reserve(nm) || c->d_xyz.reserve(3 * (size_t)acc + 3) ||
        c->d_grad.reserve(3 * (size_t)acc + 3))
        return fail(MBXG_ENOMEM, "mbxg_set_topology: device allocation failed");
    if (nmon > 0) {
        HIPCHK(hipMemcpy(c->d_kind.p, c->kind.data(), sizeof(int) * nmon, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_nat.p, c->nat.data(), sizeof(int) * nmon, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_first.p, c->first_atom.data(), sizeof(int) * nmon, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_tag.p, c->tag.data(), sizeof(int) * nmon, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_islocal.p, c->islocal.data(), nmon, hipMemcpyHostToDevice));
    }
    if (c->np > 0)
        HIPCHK(hipMemcpy(c->d_members.p, c->members.data(), sizeof(int) * c->np, hipMemcpyHostToDevice));
    c->have_coords = false;
    c->have_results = false;
    return MBXG_OK;
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
    return MBXG_OK;
}

extern "C" int mbxg_set_cutoffs(mbxg_ctx* c, double c2, double c3) {
    if (!c || !(c2 >= 0.0) || !(c3 >= 0.0)) return fail(MBXG_EINVAL, "mbxg_set_cutoffs: bad arguments");
    c->cut2 = c2;
    c->cut3 = c3;
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
    } else {
        return fail(MBXG_EINVAL, "mbxg_set_coefficients: unknown model '%s'", model);
    }
    if (a && n != need) return fail(MBXG_EINVAL, "mbxg_set_coefficients: %s takes %d coefficients", model, need);
    HIPCHK(hipMemcpyAsync(dst, a ? a : fit, sizeof(double) * need, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MBXG_OK;
}

extern "C" int mbxg_set_coords(mbxg_ctx* c, const double* xyz) {
    if (!c || (!xyz && c->natoms > 0)) return fail(MBXG_EINVAL, "mbxg_set_coords: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    if (c->natoms > 0)
        HIPCHK(hipMemcpyAsync(c->d_xyz.p, xyz, sizeof(double) * 3 * c->natoms, hipMemcpyHostToDevice, c->stream));
    c->have_coords = true;
    return MBXG_OK;
}

extern "C" int mbxg_set_coords_device(mbxg_ctx* c, const double* d_xyz) {
    if (!c || (!d_xyz && c->natoms > 0)) return fail(MBXG_EINVAL, "mbxg_set_coords_device: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    if (c->natoms > 0)
        HIPCHK(hipMemcpyAsync(c->d_xyz.p, d_xyz, sizeof(double) * 3 * c->natoms, hipMemcpyDeviceToDevice,
                              c->stream));
    c->have_coords = true;
    return MBXG_OK;
}

namespace {

// Bounding box of the cloud points (gas phase): one workgroup.
__global__ void __launch_bounds__(256) k_bbox(const double* xyz, const int* first_atom, const int* members, int np,
                                              double* out) {
    __shared__ double lo[3][256], hi[3][256];
    const int t = threadIdx.x;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    for (int p = t; p < np; p += 256) {
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
    if (t == 0)
        for (int c = 0; c < 3; ++c) {
            out[c] = lo[c][0];
            out[3 + c] = hi[c][0];
        }
}

int setup_grid(mbxg_ctx* c, double cutoff, Grid& G) {
    G = Grid{};
    const double cut = cutoff * (1.0 + 1e-9) + 1e-9;
    if (c->box.periodic) {
        const double* b = c->box.b;
        // perpendicular widths: V / |b x c|, V / |c x a|, V / |a x b|
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
            int n = (int)std::floor(w[d] / cut);
            if (n < 1) n = 1;
            if (n > 256) n = 256;
            G.n[d] = n;
            G.lo_off[d] = n >= 3 ? -1 : 0;
            G.hi_off[d] = n >= 3 ? 1 : n - 1;
        }
        G.periodic = 1;
    } else {
        if (c->np == 0) {
            G.n[0] = G.n[1] = G.n[2] = 1;
            return 0;
        }
        hipLaunchKernelGGL(k_bbox, dim3(1), dim3(256), 0, c->stream, c->d_xyz.p, c->d_first.p, c->d_members.p,
                           c->np, c->d_bbox.p);
        double bb[6];
        if (hipMemcpyAsync(bb, c->d_bbox.p, sizeof(bb), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return fail(MBXG_EHIP, "bounding box readback failed");
        for (int d = 0; d < 3; ++d) {
            const double ext = bb[3 + d] - bb[d];
            int n = (int)std::floor(ext / cut);
            if (n < 1) n = 1;
            if (n > 256) n = 256;
            G.n[d] = n;
            G.lo[d] = bb[d];
            G.inv_w[d] = ext > 0 ? n / ext : 0.0;
            G.lo_off[d] = -1;
            G.hi_off[d] = 1;
        }
        G.periodic = 0;
    }
    return 0;
}

// Lists of one order: cloud, binning, count, scan, emit.
int build_order(mbxg_ctx* c, int order, double cutoff, OrderState& o, const NmerCtx& N) {
    o.n_list = o.n_eval = 0;
    if (c->np < order || cutoff <= 0.0) return 0;
    if (int r = setup_grid(c, cutoff, o.grid)) return r;
    const Grid& G = o.grid;
    o.ncell = G.n[0] * G.n[1] * G.n[2];
    const int nslot = c->box.periodic ? 8 : 1;
    const size_t ns = (size_t)c->np * nslot;
    if (o.cx.reserve(ns) || o.cy.reserve(ns) || o.cz.reserve(ns) || o.cell.reserve(c->np) ||
        o.cell_count.reserve(o.ncell) || o.cell_start.reserve(o.ncell + 1) || o.cell_pts.reserve(c->np) ||
        o.counts.reserve(c->np) || o.offsets.reserve(c->np + 1))
        return fail(MBXG_ENOMEM, "list buffers: device allocation failed");
    hipStream_t st = c->stream;
    HIPCHK(hipMemsetAsync(o.cx.p, 0, ns * sizeof(double), st));
    HIPCHK(hipMemsetAsync(o.cy.p, 0, ns * sizeof(double), st));
    HIPCHK(hipMemsetAsync(o.cz.p, 0, ns * sizeof(double), st));
    const int tb = 256;
    hipLaunchKernelGGL(k_cloud, dim3((c->np + tb - 1) / tb), dim3(tb), 0, st, c->d_xyz.p, c->d_first.p,
                       c->d_members.p, c->np, c->box, G, o.cx.p, o.cy.p, o.cz.p, o.cell.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemsetAsync(o.cell_count.p, 0, sizeof(int) * o.ncell, st));
    hipLaunchKernelGGL(k_bin_count, dim3((c->np + tb - 1) / tb), dim3(tb), 0, st, o.cell.p, c->np, o.cell_count.p);
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, st, o.cell_count.p, o.ncell, o.cell_start.p);
    hipLaunchKernelGGL(k_bin_fill, dim3((c->np + tb - 1) / tb), dim3(tb), 0, st, o.cell.p, c->np, o.cell_start.p,
                       o.cell_count.p, o.cell_pts.p);
    hipLaunchKernelGGL(k_bin_sort, dim3((o.ncell + tb - 1) / tb), dim3(tb), 0, st, o.cell_start.p, o.ncell,
                       o.cell_pts.p);
    HIPCHK(hipGetLastError());
    const double c2 = cutoff * cutoff;
    auto launch = order == 2 ? launch_dimers : launch_trimers;
    for (int attempt = 0; attempt < 2; ++attempt) {
        HIPCHK(hipMemsetAsync(c->d_overflow.p, 0, sizeof(int), st));
        HIPCHK(launch(o.cx.p, o.cy.p, o.cz.p, o.cell.p, o.cell_start.p, o.cell_pts.p, c->np, nslot, G, N, c2, 0,
                      o.counts.p, nullptr, nullptr, nullptr, c->match_cap, c->d_overflow.p, st));
        hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, o.counts.p, c->np, o.offsets.p);
        HIPCHK(hipGetLastError());
        int2 tot;
        int ovf = 0;
        HIPCHK(hipMemcpyAsync(&tot, o.offsets.p + c->np, sizeof(int2), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&ovf, c->d_overflow.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (ovf > c->match_cap) {
            if (attempt == 0 && ovf <= 16384) {
                c->match_cap = (ovf + 255) / 256 * 256;
                continue;
            }
            return fail(MBXG_EINVAL, "neighbour search: %d matches for one monomer exceed the supported 16384",
                        ovf);
        }
        o.n_list = tot.x;
        o.n_eval = tot.y;
        break;
    }
    if (o.list.reserve((size_t)order * (o.n_list + 1)) || o.eval.reserve((size_t)order * (o.n_eval + 1)))
        return fail(MBXG_ENOMEM, "n-mer lists: device allocation failed (%lld listed)", (long long)o.n_list);
    HIPCHK(launch(o.cx.p, o.cy.p, o.cz.p, o.cell.p, o.cell_start.p, o.cell_pts.p, c->np, nslot, G, N, c2, 1,
                  o.counts.p, o.offsets.p, o.list.p, o.eval.p, c->match_cap, c->d_overflow.p, st));
    return 0;
}

}  // namespace

extern "C" int mbxg_compute(mbxg_ctx* c, int terms, int do_grads) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_compute: null context");
    if (!c->have_coords) return fail(MBXG_ESTATE, "mbxg_compute: coordinates not set");
    const bool lists_only = (terms & MBXG_LISTS_ONLY) != 0;
    if (!lists_only)
        for (int i = 0; i < c->nmon; ++i)
            if (c->kind[i] == MBXG_CH4)
                return fail(MBXG_EINVAL, "mbxg_compute: no GPU holder for ch4-containing n-mers yet (lists only)");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    c->have_results = false;
    hipEvent_t* ev = c->timing ? c->ev : nullptr;
    if (ev)
        for (int k = 0; k < kNumEvents; ++k) HIPCHK(hipEventRecord(ev[k], st));  // stages skipped read 0
    if (ev) HIPCHK(hipEventRecord(ev[0], st));
    if (c->box.periodic && c->nmon > 0) {
        hipLaunchKernelGGL(k_fix_monomers, dim3((c->nmon + 255) / 256), dim3(256), 0, st, c->d_xyz.p,
                           c->d_first.p, c->d_nat.p, c->nmon, c->box);
        HIPCHK(hipGetLastError());
    }
    NmerCtx N{c->d_xyz.p, c->d_first.p, c->d_kind.p, c->d_tag.p, c->d_islocal.p, c->d_members.p, c->d_box.p,
              c->box.periodic};
    c->o2.n_list = c->o2.n_eval = c->o3.n_list = c->o3.n_eval = 0;
    if (ev) HIPCHK(hipEventRecord(ev[1], st));
    if (terms & MBXG_2B)
        if (int r = build_order(c, 2, c->cut2, c->o2, N)) return r;
    if (ev) HIPCHK(hipEventRecord(ev[2], st));
    if (terms & MBXG_3B)
        if (int r = build_order(c, 3, c->cut3, c->o3, N)) return r;
    if (ev) HIPCHK(hipEventRecord(ev[3], st));
    if (lists_only) {
        if (ev) HIPCHK(hipEventRecord(ev[12], st));
        HIPCHK(hipMemsetAsync(c->d_totals.p, 0, sizeof(double) * 16, st));
        if (c->natoms > 0) HIPCHK(hipMemsetAsync(c->d_grad.p, 0, sizeof(double) * 3 * c->natoms, st));
        c->o2.n_eval = c->o3.n_eval = 0;
        c->have_results = true;
        return MBXG_OK;
    }
    if (c->natoms > 0) HIPCHK(hipMemsetAsync(c->d_grad.p, 0, sizeof(double) * 3 * c->natoms, st));
    HIPCHK(hipMemsetAsync(c->d_totals.p, 0, sizeof(double) * 16, st));
    OrderState* os[2] = {&c->o2, &c->o3};
    for (int k = 0; k < 2; ++k) {
        OrderState& o = *os[k];
        const int nb = (int)((o.n_eval + 63) / 64);
        const size_t wsz = k == 0 ? workspace_mbpol2b((int)o.n_eval) : workspace_mbpol3b((int)o.n_eval);
        if (o.e_nmer.reserve(o.n_eval + 1) || o.partials.reserve((size_t)kPartials * (nb + 1)) ||
            o.ws.reserve(wsz + 1))
            return fail(MBXG_ENOMEM, "evaluation buffers: device allocation failed");
        if (o.n_eval == 0) continue;
        if (k == 0)
            HIPCHK(launch_mbpol2b(o.eval.p, (int)o.n_eval, N, c->d_coef2.p, do_grads, c->d_grad.p, o.e_nmer.p,
                                  o.partials.p, o.ws.p, st, ev ? ev + 4 : nullptr));
        else
            HIPCHK(launch_mbpol3b(o.eval.p, (int)o.n_eval, N, c->d_coef3.p, do_grads, c->d_grad.p, o.e_nmer.p,
                                  o.partials.p, o.ws.p, st, ev ? ev + 8 : nullptr));
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(64), 0, st, o.partials.p, nb,
                           c->d_totals.p + kPartials * k);
        HIPCHK(hipGetLastError());
    }
    if (ev) HIPCHK(hipEventRecord(ev[12], st));
    c->have_results = true;
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
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    // fix, lists 2B, lists 3B, vars 2B, poly 2B, chain 2B, vars 3B, poly 3B, chain 3B, total
    const int pairs[10][2] = {{0, 1}, {1, 2}, {2, 3}, {4, 5}, {5, 6}, {6, 7}, {8, 9}, {9, 10}, {10, 11}, {0, 12}};
    for (int k = 0; k < n && k < 10; ++k) {
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, c->ev[pairs[k][0]], c->ev[pairs[k][1]]));
        ms[k] = t;
    }
    return MBXG_OK;
}

extern "C" int mbxg_get_results(mbxg_ctx* c, double* e2b, double* e3b, double* grad, double* virial) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_get_results: null context");
    if (!c->have_results) return fail(MBXG_ESTATE, "mbxg_get_results: nothing computed");
    HIPCHK(hipSetDevice(c->device));
    double tot[16];
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
        for (int q = 0; q < 6; ++q) virial[slot[q]] = tot[1 + q] + tot[kPartials + 1 + q];
        virial[3] = virial[1];
        virial[6] = virial[2];
        virial[7] = virial[5];
    }
    return MBXG_OK;
}

extern "C" int mbxg_get_counts(mbxg_ctx* c, int64_t* counts) {
    if (!c || !counts) return fail(MBXG_EINVAL, "mbxg_get_counts: bad arguments");
    counts[0] = c->o2.n_list;
    counts[1] = c->o2.n_eval;
    counts[2] = c->o3.n_list;
    counts[3] = c->o3.n_eval;
    return MBXG_OK;
}

extern "C" int mbxg_get_lists(mbxg_ctx* c, int32_t* dimers, int64_t cap2, int32_t* trimers, int64_t cap3) {
    if (!c) return fail(MBXG_EINVAL, "mbxg_get_lists: null context");
    HIPCHK(hipSetDevice(c->device));
    if (dimers && c->o2.n_list > 0) {
        const int64_t n = c->o2.n_list < cap2 ? c->o2.n_list : cap2;
        HIPCHK(hipMemcpyAsync(dimers, c->o2.list.p, sizeof(int32_t) * 2 * n, hipMemcpyDeviceToHost, c->stream));
    }
    if (trimers && c->o3.n_list > 0) {
        const int64_t n = c->o3.n_list < cap3 ? c->o3.n_list : cap3;
        HIPCHK(hipMemcpyAsync(trimers, c->o3.list.p, sizeof(int32_t) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return MBXG_OK;
}

extern "C" int mbxg_get_nmer_energies(mbxg_ctx* c, int order, int32_t* nmers, double* energies, int64_t cap) {
    if (!c || (order != 2 && order != 3)) return fail(MBXG_EINVAL, "mbxg_get_nmer_energies: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    OrderState& o = order == 2 ? c->o2 : c->o3;
    const int64_t n = o.n_eval < cap ? o.n_eval : cap;
    if (n > 0) {
        if (nmers)
            HIPCHK(hipMemcpyAsync(nmers, o.eval.p, sizeof(int32_t) * order * n, hipMemcpyDeviceToHost, c->stream));
        if (energies)
            HIPCHK(hipMemcpyAsync(energies, o.e_nmer.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
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
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MBXG_OK;
}
