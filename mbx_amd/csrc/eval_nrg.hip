// MB-nrg CH4 families on the GPU: variable and chain-rule stages for the holders described in
// gen/holder_<name>.inc, with the polynomial pieces of gen/poly_<name>_p<k>.hip between them.
//
// Restates, per n-mer (one lane each):
//   x2b_A1B2Z2_C1D4_v1x::eval   potential/2b/x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp:954-1220 (h2o, ch4)
//   x2b_A1B4_A1B4_v1x::eval     potential/2b/x2b_A1B4_A1B4_deg4_exp0_v1x.cpp:1600-1790  (ch4, ch4)
//   mbnrg_A1B4_C1D2_C1D2_deg3_v1::eval
//                               potential/3b/mbnrg_3b_A1B4_C1D2_C1D2_deg3_v1.cpp:1481-1790 (ch4, h2o, h2o)
// reached through e2b::get_2b_energy / e3b::get_3b_energy, which order the n-mer's monomers by
// name before dispatching (potential/2b/energy2b.cpp:243-254, 323-325;
// potential/3b/energy3b.cpp:185-216).  The n-mer's other monomers are imaged next to its first
// listed monomer first (GetCloseDimerImage / GetCloseTrimerImage, bblock/sys_tools.cpp:322-365,
// 438-442), as Get2B / Get3B do before the dispatch.
#include "gen/holder_ch4ch4_2b.inc"
#include "gen/holder_ch4h2o2_3b.inc"
#include "gen/holder_ch4h2o_2b.inc"
#include "mbxg_device.hpp"
#include "mbxg_kernels.hpp"
#include "nrg_kernel.hpp"
#include "scatter.hpp"

namespace mbxg {

namespace {

// Holder slot s -> position of its monomer in the listed n-mer.
template <class H>
__device__ __forceinline__ void holder_order(const NmerCtx& N, const int* ids, int* perm) {
    if (H::kMon == 2) {
        perm[0] = 0;
        perm[1] = 1;
        // the holder takes its monomers in its own order (water first for CH4-H2O)
        if (N.kind[ids[0]] != H::kKind[0]) {
            perm[0] = 1;
            perm[1] = 0;
        }
        return;
    }
    // e3b::get_3b_energy's exchange network on the names: rank 0 = "ch4" < rank 1 = "h2o"
    auto rank = [](int kind) { return kind == MON_CH4 ? 0 : 1; };
    int k[3] = {rank(N.kind[ids[0]]), rank(N.kind[ids[1]]), rank(N.kind[ids[2]])}, ix[3] = {0, 1, 2};
    auto swp = [&](int i, int j) {
        int t = k[i];
        k[i] = k[j];
        k[j] = t;
        t = ix[i];
        ix[i] = ix[j];
        ix[j] = t;
    };
    if (k[2] < k[0] && k[2] < k[1]) swp(0, 2);
    if (k[0] > k[1] && k[0] > k[2])
        swp(0, 2);
    else if (k[1] > k[0] && k[1] > k[2])
        swp(1, 2);
    if (k[0] > k[1]) swp(0, 1);
    perm[0] = ix[0];
    perm[1] = ix[1];
    perm[2] = ix[2];
}

template <class H>
constexpr int slot_site(int s) {
    return s == 0 ? 0 : (s == 1 ? H::kNat[0] : H::kNat[0] + H::kNat[1]);
}

// The n-mer's sites in holder order (real atoms, then the lone-pair sites), imaged next to the
// first listed monomer.  mon[s] receives slot s's monomer.
template <class H>
__device__ __forceinline__ void holder_sites(const NmerCtx& N, const int* ids, double* x, int* mon, double* oh1,
                                             double* oh2) {
    int perm[3];
    holder_order<H>(N, ids, perm);
    const double* f0 = N.xyz + 3 * (size_t)N.first_atom[ids[0]];
    const double x0 = f0[0], y0 = f0[1], z0 = f0[2];
#pragma unroll
    for (int s = 0; s < H::kMon; ++s) {
        const int m = ids[perm[s]];
        mon[s] = m;
        const double* a = N.xyz + 3 * (size_t)N.first_atom[m];
        double* dst = x + 3 * slot_site<H>(s);
#pragma unroll
        for (int j = 0; j < 3 * H::kNat[s]; ++j) dst[j] = a[j];
        if (N.periodic && perm[s] != 0) {
            const Box& B = *N.box;
            double dx, dy, dz;
            image_shift(B, x0, y0, z0, dst[0], dst[1], dst[2], dx, dy, dz);
#pragma unroll
            for (int at = 0; at < H::kNat[s]; ++at) image_apply(B, dx, dy, dz, dst[3 * at], dst[3 * at + 1], dst[3 * at + 2]);
        }
    }
    if (H::kVsite) {
        // monomer::setup (x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp:119-137) with MB-pol's gammas
        const double* o = x + 3 * H::kOSite;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            oh1[i] = o[i + 3] - o[i];
            oh2[i] = o[i + 6] - o[i];
        }
        const double v[3] = {oh1[1] * oh2[2] - oh1[2] * oh2[1], oh1[2] * oh2[0] - oh1[0] * oh2[2],
                             oh1[0] * oh2[1] - oh1[1] * oh2[0]};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double in_plane = o[i] + 0.5 * H::kInPlane * (oh1[i] + oh2[i]);
            const double out_of_plane = H::kOutOfPlane * v[i];
            x[3 * H::kNReal + i] = in_plane + out_of_plane;
            x[3 * H::kNReal + 3 + i] = in_plane - out_of_plane;
        }
    }
}

// f_switch of the MB-nrg holders (x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp:804-817,
// mbnrg_3b_A1B4_C1D2_C1D2_deg3_v1.cpp:1319-1333).
template <class H>
__device__ __forceinline__ double nrg_switch(double r, double& g) {
    if (r > H::kROut) {
        g = 0.0;
        return 0.0;
    } else if (r > H::kRIn) {
        const double t1 = M_PI / (H::kROut - H::kRIn);
        const double xx = (r - H::kRIn) * t1;
        g = -sin(xx) * t1 / 2.0;
        return (1.0 + cos(xx)) / 2.0;
    }
    g = 0.0;
    return 1.0;
}

// First-atom separations of the slots: 2B d[0] = slot0 - slot1; 3B d[0] = 0-1, d[1] = 0-2,
// d[2] = 1-2, with their lengths r, switch values s and derivatives gsw.
template <class H>
__device__ __forceinline__ void holder_switches(const double* x, double d[3][3], double* r, double* s, double* gsw) {
    const int np = H::kMon == 2 ? 1 : 3;
    const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
#pragma unroll
    for (int q = 0; q < np; ++q) {
        const double* u = x + 3 * slot_site<H>(pa[q]);
        const double* v = x + 3 * slot_site<H>(pb[q]);
        d[q][0] = u[0] - v[0];
        d[q][1] = u[1] - v[1];
        d[q][2] = u[2] - v[2];
        r[q] = sqrt(d[q][0] * d[q][0] + d[q][1] * d[q][1] + d[q][2] * d[q][2]);
        s[q] = nrg_switch<H>(r[q], gsw[q]);
    }
}

// Variable J's value from its site pair (tools/variable.cpp:49-129); gg with
// d(value)/d(site s1) = gg * (x[s1] - x[s2]).  Everything about the variable is a constant.
template <class H, int J>
__device__ __forceinline__ double nrg_value(const double* x, double* dvec, double& gg) {
    constexpr NrgVar V = H::kVars[J];
    dvec[0] = x[3 * V.s1] - x[3 * V.s2];
    dvec[1] = x[3 * V.s1 + 1] - x[3 * V.s2 + 1];
    dvec[2] = x[3 * V.s1 + 2] - x[3 * V.s2 + 2];
    const double r = sqrt(dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2]);
    double val;
    if (V.form == 0) {
        val = exp(V.k * (V.d - r));
        gg = -V.k * val / r;
    } else if (V.form == 1) {
        const double e1 = exp(V.k * (V.d - r));
        const double rinv = 1.0 / r;
        val = e1 * rinv;
        gg = -(V.k + rinv) * val * rinv;
    } else {
        val = exp(V.k * (-r));
        gg = -V.k * val / r;
    }
    return val;
}

}  // namespace

// Stage 1: variables and the total switch.
template <class H>
__global__ void __launch_bounds__(64) k_vars_nrg(const int* __restrict__ list, const int* __restrict__ n_dev, int cap,
                                                 NmerCtx N, double* __restrict__ xs, double* __restrict__ sw, int ld) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= min(*n_dev, cap)) return;
    double x[3 * H::kNSites], oh1[3], oh2[3];
    int mon[3];
    holder_sites<H>(N, list + H::kMon * (size_t)t, x, mon, oh1, oh2);
    static_for<H::kNVars>([&](auto J) {
        constexpr int j = decltype(J)::value;
        double dv[3], gg;
        xs[(size_t)j * ld + t] = nrg_value<H, j>(x, dv, gg);
    });
    double d[3][3], r[3], s[3], gs[3];
    holder_switches<H>(x, d, r, s, gs);
    sw[t] = H::kMon == 2 ? s[0] : s[0] * s[1] + s[0] * s[2] + s[1] * s[2];
}

constexpr int kGradBlockNrg = 128;

// Stage 3: energy; with GRAD the chain rule, virial and gradient scatter.  gs / ps hold the
// polynomial's pieces ([piece][var][ld], [piece][ld]), summed here in piece order.
template <class H, bool GRAD>
__global__ void __launch_bounds__(kGradBlockNrg) k_grad_nrg(const int* __restrict__ list, const int* __restrict__ n_dev,
                                                            int cap, NmerCtx N, const double* __restrict__ xs,
                                                            const double* __restrict__ gs,
                                                            const double* __restrict__ ps,
                                                            const double* __restrict__ sws, int ld,
                                                            double* __restrict__ grad, double* __restrict__ e_out,
                                                            double* __restrict__ vir_out,
                                                            double* __restrict__ partials) {
    __shared__ GradTable<128, 15> T;
    const int t = blockIdx.x * kGradBlockNrg + threadIdx.x;
    const int n = min(*n_dev, cap);
    if (GRAD) table_init(T);
    __syncthreads();
    double e_sum = 0.0;
    double vir[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (t < n) {
        double P = 0.0;
#pragma unroll
        for (int q = 0; q < H::kPieces; ++q) P += ps[(size_t)q * ld + t];
        double e;
        if (GRAD) {
            double x[3 * H::kNSites], oh1[3], oh2[3];
            int mon[3];
            holder_sites<H>(N, list + H::kMon * (size_t)t, x, mon, oh1, oh2);
            double d[3][3], r[3], s[3], gsw[3];
            holder_switches<H>(x, d, r, s, gsw);
            const double sw = H::kMon == 2 ? s[0] : s[0] * s[1] + s[0] * s[2] + s[1] * s[2];
            double xg[3 * H::kNSites];
#pragma unroll
            for (int i = 0; i < 3 * H::kNSites; ++i) xg[i] = 0.0;
            static_for<H::kNVars>([&](auto J) {
                constexpr int j = decltype(J)::value;
                double gj = 0.0;
#pragma unroll
                for (int q = 0; q < H::kPieces; ++q) gj += gs[((size_t)q * H::kNVars + j) * ld + t];
                if (H::kMon == 3) gj *= sw;  // the 3B holder scales the polynomial gradient first
                double dv[3], gg;
                nrg_value<H, j>(x, dv, gg);
                constexpr int a1 = H::kVars[j].s1, a2 = H::kVars[j].s2;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double tc = gj * (dv[c] * gg);
                    xg[3 * a1 + c] += tc;
                    xg[3 * a2 + c] -= tc;
                }
            });
            if (H::kVsite) {
                // monomer::grads (x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp:139-160)
                const double* g1 = xg + 3 * H::kNReal;
                const double* g2 = xg + 3 * H::kNReal + 3;
                double* grd = xg + 3 * H::kOSite;
                const double gm[3] = {g1[0] - g2[0], g1[1] - g2[1], g1[2] - g2[2]};
                const double t1[3] = {oh2[1] * gm[2] - oh2[2] * gm[1], oh2[2] * gm[0] - oh2[0] * gm[2],
                                      oh2[0] * gm[1] - oh2[1] * gm[0]};
                const double t2[3] = {oh1[1] * gm[2] - oh1[2] * gm[1], oh1[2] * gm[0] - oh1[0] * gm[2],
                                      oh1[0] * gm[1] - oh1[1] * gm[0]};
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const double gsum = g1[i] + g2[i];
                    const double in_plane = 0.5 * H::kInPlane * gsum;
                    const double gh1 = in_plane + H::kOutOfPlane * t1[i];
                    const double gh2 = in_plane - H::kOutOfPlane * t2[i];
                    grd[i + 0] += gsum - (gh1 + gh2);
                    grd[i + 3] += gh1;
                    grd[i + 6] += gh2;
                }
            }
            const int f1 = 3 * slot_site<H>(1);
            if (H::kMon == 2) {
#pragma unroll
                for (int i = 0; i < 3 * H::kNReal; ++i) xg[i] *= sw;
                const double g12 = gsw[0] * (P / r[0]);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double tc = g12 * d[0][c];
                    xg[c] += tc;
                    xg[f1 + c] -= tc;
                }
            } else {
                const int f2 = 3 * slot_site<H>(2);
                const double g12 = gsw[0] * ((1.0 * s[1] + 1.0 * s[2]) * P / r[0]);
                const double g13 = gsw[1] * ((s[0] * 1.0 + 1.0 * s[2]) * P / r[1]);
                const double g23 = gsw[2] * ((s[0] * 1.0 + s[1] * 1.0) * P / r[2]);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    xg[c] += 0.0 + (g12 * d[0][c]) + (g13 * d[1][c]);
                    xg[f1 + c] += 0.0 - (g12 * d[0][c]) + (g23 * d[2][c]);
                    xg[f2 + c] += 0.0 - (g13 * d[1][c]) - (g23 * d[2][c]);
                }
            }
            e = P * sw;
            const int va[6] = {0, 0, 0, 1, 1, 2}, vb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                double acc = 0.0;
#pragma unroll
                for (int at = 0; at < H::kNReal; ++at) acc += -x[3 * at + va[q]] * xg[3 * at + vb[q]];
                vir[q] = acc;
            }
#pragma unroll
            for (int s2 = 0; s2 < H::kMon; ++s2)
                table_add(T, mon[s2], xg + 3 * slot_site<H>(s2), 3 * H::kNat[s2], grad,
                          3 * (size_t)N.first_atom[mon[s2]]);
        } else {
            e = P * sws[t];
        }
        e_out[t] = e;
        e_sum = e;
        if (vir_out)
#pragma unroll
            for (int q = 0; q < 6; ++q) vir_out[(size_t)q * ld + t] = vir[q];
    }
    if (GRAD) {
        __syncthreads();
        table_flush(T, N.first_atom, N.nat, grad);
    }
    e_sum = wave_sum(e_sum);
    const int rec = blockIdx.x * (kGradBlockNrg / 64) + threadIdx.x / 64;
    double* out = partials + (size_t)rec * kPartials;
    const bool lead = (threadIdx.x & 63) == 0;
    if (lead) out[0] = e_sum;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double v = wave_sum(vir[q]);
        if (lead) out[1 + q] = v;
    }
}

namespace {

template <class H>
size_t workspace_nrg(int cap) {
    const size_t ld = ((size_t)cap + 63) / 64 * 64;
    return ((size_t)H::kNVars * (1 + H::kPieces) + H::kPieces + 1) * ld;
}

typedef hipError_t (*piece_launch)(const double*, const int*, int, int, const double*, int, double*, double*,
                                   hipStream_t);

// The polynomial stage of a pieced family: one kernel per piece, each into its own slice.
struct Pieced {
    const piece_launch* pieces;
    int count;
    hipError_t operator()(const double* xs, const int* n_dev, int cap, int ld, const double* coef, int do_grads,
                          double* gs, double* ps, int nv, hipStream_t st) const {
        for (int q = 0; q < count; ++q) {
            hipError_t e = pieces[q](xs, n_dev, cap, ld, coef, do_grads, gs + (size_t)q * nv * ld, ps + (size_t)q * ld, st);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
};

// The polynomial stage of a blocked family: one kernel, `split` wavefronts per 64 n-mers.
struct Blocked {
    hipError_t (*launch)(const double*, const int*, int, int, const double*, int, double*, double*, int, hipStream_t);
    int split;
    hipError_t operator()(const double* xs, const int* n_dev, int cap, int ld, const double* coef, int do_grads,
                          double* gs, double* ps, int, hipStream_t st) const {
        return launch(xs, n_dev, cap, ld, coef, do_grads, gs, ps, split, st);
    }
};

template <class H, class Poly>
hipError_t launch_nrg(const Poly& poly, const int* list, const int* n_dev, int cap, const NmerCtx& N,
                      const double* coef, int do_grads, double* grad, double* e_out, double* vir_out,
                      double* partials, double* ws, hipStream_t st, hipEvent_t* ev) {
    if (cap <= 0) return hipSuccess;
    const int nb = (cap + 63) / 64;
    const int ld = nb * 64;
    double* xs = ws;
    double* gs = xs + (size_t)H::kNVars * ld;
    double* ps = gs + (size_t)H::kPieces * H::kNVars * ld;
    double* sw = ps + (size_t)H::kPieces * ld;
    if (ev) (void)hipEventRecord(ev[0], st);
    hipLaunchKernelGGL(k_vars_nrg<H>, dim3(nb), dim3(64), 0, st, list, n_dev, cap, N, xs, sw, ld);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[1], st);
    e = poly(xs, n_dev, cap, ld, coef, do_grads, gs, ps, H::kNVars, st);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[2], st);
    const int ngb = (cap + kGradBlockNrg - 1) / kGradBlockNrg;
    if (do_grads)
        hipLaunchKernelGGL((k_grad_nrg<H, true>), dim3(ngb), dim3(kGradBlockNrg), 0, st, list, n_dev, cap, N, xs, gs,
                           ps, sw, ld, grad, e_out, vir_out, partials);
    else
        hipLaunchKernelGGL((k_grad_nrg<H, false>), dim3(ngb), dim3(kGradBlockNrg), 0, st, list, n_dev, cap, N, xs,
                           gs, ps, sw, ld, grad, e_out, vir_out, partials);
    e = hipGetLastError();
    if (ev) (void)hipEventRecord(ev[3], st);
    return e;
}

}  // namespace

// The polynomial pieces (gen/poly_<name>_p<k>.hip).
hipError_t launch_poly_ch4h2o_2b_p0(const double*, const int*, int, int, const double*, int, double*, double*,
                                    hipStream_t);
#define MBXG_CH4CH4_PIECE(k)                                                                                      \
    hipError_t launch_poly_ch4ch4_2b_p##k(const double*, const int*, int, int, const double*, int, double*, double*, \
                                          hipStream_t);
MBXG_CH4CH4_PIECE(0)
MBXG_CH4CH4_PIECE(1)
MBXG_CH4CH4_PIECE(2)
MBXG_CH4CH4_PIECE(3)
MBXG_CH4CH4_PIECE(4)
MBXG_CH4CH4_PIECE(5)
MBXG_CH4CH4_PIECE(6)
MBXG_CH4CH4_PIECE(7)
#undef MBXG_CH4CH4_PIECE

size_t workspace_nrg_family(int family, int cap) {
    switch (family) {
        case FAM_CH4H2O_2B:
            return workspace_nrg<Holder_ch4h2o_2b>(cap);
        case FAM_CH4CH4_2B:
            return workspace_nrg<Holder_ch4ch4_2b>(cap);
        case FAM_CH4H2O2_3B:
            return workspace_nrg<Holder_ch4h2o2_3b>(cap);
        default:
            return 0;
    }
}

hipError_t launch_nrg_family(int family, const int* list, const int* n_dev, int cap, const NmerCtx& N,
                             const double* coef, int do_grads, double* grad, double* e_out, double* vir_out,
                             double* partials, double* ws, int split, hipStream_t st, hipEvent_t* ev) {
    static const piece_launch p_ch4h2o[1] = {launch_poly_ch4h2o_2b_p0};
    static const piece_launch p_ch4ch4[8] = {launch_poly_ch4ch4_2b_p0, launch_poly_ch4ch4_2b_p1,
                                             launch_poly_ch4ch4_2b_p2, launch_poly_ch4ch4_2b_p3,
                                             launch_poly_ch4ch4_2b_p4, launch_poly_ch4ch4_2b_p5,
                                             launch_poly_ch4ch4_2b_p6, launch_poly_ch4ch4_2b_p7};
    switch (family) {
        case FAM_CH4H2O_2B:
            return launch_nrg<Holder_ch4h2o_2b>(Pieced{p_ch4h2o, 1}, list, n_dev, cap, N, coef, do_grads, grad, e_out, vir_out,
                                                partials, ws, st, ev);
        case FAM_CH4CH4_2B:
            return launch_nrg<Holder_ch4ch4_2b>(Pieced{p_ch4ch4, 8}, list, n_dev, cap, N, coef, do_grads, grad, e_out, vir_out,
                                                partials, ws, st, ev);
        case FAM_CH4H2O2_3B:
            return launch_nrg<Holder_ch4h2o2_3b>(Blocked{launch_poly_ch4h2o2_3b, split}, list, n_dev, cap, N, coef, do_grads, grad, e_out,
                                                 vir_out, partials, ws, st, ev);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace mbxg
