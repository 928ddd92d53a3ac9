// Further MB-nrg families on the GPU: halide-water 2B / 3B, CO2 2B / 3B, water 4B and
// alkali-metal-ion-water 2B / 3B.
//
// Restates, per n-mer (one lane each), the generated holders' eval with gradients:
//   mbnrg_A1_B1C2X2_deg6_v1          potential/2b/mbnrg_2b_A1_B1C2X2_deg6_v1.cpp:3789-3960  (f-, cl-, br-, i- with h2o)
//   mbnrg_A1_B1C2X2_B1C2X2_deg4_v1   potential/3b/mbnrg_3b_A1_B1C2X2_B1C2X2_deg4_v1.cpp:12955-13230
//   x2b_A1B2_A1B2_v1x                potential/2b/x2b_A1B2_A1B2_v1x.cpp:18485-18700 (co2, co2)
//   mbnrg_A1B2_A1B2_A1B2_deg4_v1     potential/3b/mbnrg_3b_A1B2_A1B2_A1B2_deg4_v1.cpp:1223-1460
//   mbnrg_A1B2_A1B2_A1B2_A1B2_deg4_v1 potential/4b/mbnrg_4b_A1B2_A1B2_A1B2_A1B2_deg4_v1.cpp:657-1100 (h2o x4)
//   x2b_h2o_ion_v2x                  potential/2b/x2b-A1B2Z2_C-v1x.cpp:2546-2770 (h2o with li+, na+, k+, rb+, cs+)
//   x3b_h2o_ion_v1x_deg4_filtered    potential/3b/x3b-h2o-ion-v1x-deg4-filtered.cpp:206-482
// with the holders as data (gen/xholder_<name>.inc: sites, variables, switching radii, the pair
// switch combination), reached through e2b::get_2b_energy / e3b::get_3b_energy /
// e4b::get_4b_energy: the n-mer's monomer names are sorted by those routines' exchange networks
// (potential/2b/energy2b.cpp:243-254, potential/3b/energy3b.cpp:185-216,
// potential/4b/energy4b.cpp:178-214) and passed to the holder in its argument order.  Every
// monomer of an n-mer is first imaged next to its first listed monomer (GetCloseNmerImage,
// bblock/sys_tools.cpp:453-468), as Get2B / Get3B / Get4B do before the dispatch.
//
// Pipeline per order (the engine's listed lists in, one evaluate list per family out):
//   k_xclass    family, holder order and skip rule of every listed n-mer; per-block counts
//   k_xscan     block offsets and family totals (one workgroup, fixed order)
//   k_xemit     each evaluated n-mer at its family's list, in listed order (deterministic)
// then per family: k_xvars (sites, variables, switch) -> polynomial pieces (gen/xpoly_*.hip)
// -> k_xgrad (chain rule, switch derivatives, energy, virial, gradient scatter).
// Tetramers (System::Get4B, AddClusters with nmax = 4, sys_tools.cpp:594-634): k_tetra counts
// and writes, per query, every match triple the center-3-neighbour rule keeps.
#include "gen/xholder_br_h2o2_3b.inc"
#include "gen/xholder_br_h2o_2b.inc"
#include "gen/xholder_cl_h2o2_3b.inc"
#include "gen/xholder_cl_h2o_2b.inc"
#include "gen/xholder_co2co2_2b.inc"
#include "gen/xholder_co2x3_3b.inc"
#include "gen/xholder_cs_h2o2_3b.inc"
#include "gen/xholder_cs_h2o_2b.inc"
#include "gen/xholder_f_h2o2_3b.inc"
#include "gen/xholder_f_h2o_2b.inc"
#include "gen/xholder_h2ox4_4b.inc"
#include "gen/xholder_i_h2o2_3b.inc"
#include "gen/xholder_i_h2o_2b.inc"
#include "gen/xholder_k_h2o2_3b.inc"
#include "gen/xholder_k_h2o_2b.inc"
#include "gen/xholder_li_h2o2_3b.inc"
#include "gen/xholder_li_h2o_2b.inc"
#include "gen/xholder_na_h2o2_3b.inc"
#include "gen/xholder_na_h2o_2b.inc"
#include "gen/xholder_rb_h2o2_3b.inc"
#include "gen/xholder_rb_h2o_2b.inc"
#include "gen/xcoef_br_h2o2_3b.inc"
#include "gen/xcoef_br_h2o_2b.inc"
#include "gen/xcoef_cl_h2o2_3b.inc"
#include "gen/xcoef_cl_h2o_2b.inc"
#include "gen/xcoef_co2co2_2b.inc"
#include "gen/xcoef_co2x3_3b.inc"
#include "gen/xcoef_cs_h2o2_3b.inc"
#include "gen/xcoef_cs_h2o_2b.inc"
#include "gen/xcoef_f_h2o2_3b.inc"
#include "gen/xcoef_f_h2o_2b.inc"
#include "gen/xcoef_h2ox4_4b.inc"
#include "gen/xcoef_i_h2o2_3b.inc"
#include "gen/xcoef_i_h2o_2b.inc"
#include "gen/xcoef_k_h2o2_3b.inc"
#include "gen/xcoef_k_h2o_2b.inc"
#include "gen/xcoef_li_h2o2_3b.inc"
#include "gen/xcoef_li_h2o_2b.inc"
#include "gen/xcoef_na_h2o2_3b.inc"
#include "gen/xcoef_na_h2o_2b.inc"
#include "gen/xcoef_rb_h2o2_3b.inc"
#include "gen/xcoef_rb_h2o_2b.inc"
#include "mbxg_device.hpp"
#include "mbxg_kernels.hpp"
#include "nrg_kernel.hpp"
#include "scatter.hpp"
#include "xfam.hpp"

namespace mbxg {

namespace {

// Name rank of each engine kind: the order of the reference's monomer id strings
// ("br-" < "ch4" < "cl-" < "co2" < "cs+" < "f-" < "h2o" < "i-" < "k+" < "li+" < "na+" < "rb+");
// kinds without a term rank last.
__device__ __forceinline__ int name_rank(int kind) {
    switch (kind) {
        case MON_BR: return 0;
        case MON_CH4: return 1;
        case MON_CL: return 2;
        case MON_CO2: return 3;
        case MON_CS: return 4;
        case MON_F: return 5;
        case MON_H2O: return 6;
        case MON_I: return 7;
        case MON_K: return 8;
        case MON_LI: return 9;
        case MON_NA: return 10;
        case MON_RB: return 11;
        default: return 100 + kind;
    }
}

// The exchange networks of get_2b/3b/4b_energy on name ranks: ix receives the positions of the
// n-mer's monomers in sorted order (strict comparisons, equal names keep their order).
template <int K>
__device__ __forceinline__ void sort_network(const int* rank_in, int* ix) {
    int k[4];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        k[i] = rank_in[i];
        ix[i] = i;
    }
    auto swp = [&](int i, int j) {
        int t = k[i];
        k[i] = k[j];
        k[j] = t;
        t = ix[i];
        ix[i] = ix[j];
        ix[j] = t;
    };
    if (K == 2) {
        if (k[1] < k[0]) swp(0, 1);
    } else if (K == 3) {
        if (k[2] < k[0] && k[2] < k[1]) swp(0, 2);
        if (k[0] > k[1] && k[0] > k[2])
            swp(0, 2);
        else if (k[1] > k[0] && k[1] > k[2])
            swp(1, 2);
        if (k[0] > k[1]) swp(0, 1);
    } else {
        if (k[0] > k[1]) swp(0, 1);
        if (k[1] > k[2]) swp(1, 2);
        if (k[2] > k[3]) swp(2, 3);
        if (k[0] > k[1]) swp(0, 1);
        if (k[1] > k[2]) swp(1, 2);
        if (k[0] > k[1]) swp(0, 1);
    }
}

// The families in the order of XFam: enum, holder, polynomial pieces.
#define MBXG_XFAMILIES(X)                              \
    X(XF_F_H2O_2B, f_h2o_2b, kPHalide2)                \
    X(XF_CL_H2O_2B, cl_h2o_2b, kPHalide2)              \
    X(XF_BR_H2O_2B, br_h2o_2b, kPHalide2)              \
    X(XF_I_H2O_2B, i_h2o_2b, kPHalide2)                \
    X(XF_CO2CO2_2B, co2co2_2b, kPCo2_2)                \
    X(XF_F_H2O2_3B, f_h2o2_3b, kPHalide3)              \
    X(XF_CL_H2O2_3B, cl_h2o2_3b, kPHalide3)            \
    X(XF_BR_H2O2_3B, br_h2o2_3b, kPHalide3)            \
    X(XF_I_H2O2_3B, i_h2o2_3b, kPHalide3)              \
    X(XF_CO2X3_3B, co2x3_3b, kPCo2_3)                  \
    X(XF_H2OX4_4B, h2ox4_4b, kPH2o_4)                  \
    X(XF_LI_H2O_2B, li_h2o_2b, kPAlkali2)              \
    X(XF_NA_H2O_2B, na_h2o_2b, kPAlkali2)              \
    X(XF_K_H2O_2B, k_h2o_2b, kPAlkali2)                \
    X(XF_RB_H2O_2B, rb_h2o_2b, kPAlkali2)              \
    X(XF_CS_H2O_2B, cs_h2o_2b, kPAlkali2)              \
    X(XF_LI_H2O2_3B, li_h2o2_3b, kPAlkali3)            \
    X(XF_NA_H2O2_3B, na_h2o2_3b, kPAlkali3)            \
    X(XF_K_H2O2_3B, k_h2o2_3b, kPAlkali3)              \
    X(XF_RB_H2O2_3B, rb_h2o2_3b, kPAlkali3)            \
    X(XF_CS_H2O2_3B, cs_h2o2_3b, kPAlkali3)

// Family descriptions for the classification: order, the holder's monomer kinds in argument
// order, the skip radius and how many first-atom separations beyond it skip the n-mer (all of
// them for the generated holders; two of three for x3b_h2o_ion_v1x_deg4_filtered,
// x3b-h2o-ion-v1x-deg4-filtered.cpp:295).
struct XDesc {
    int order;
    int hkind[4];
    double rout;
    int skip_min;
};

template <class H>
constexpr XDesc xdesc() {
    return XDesc{H::kMon, {H::kKind[0], H::kKind[1], H::kKind[2], H::kKind[3]}, H::kROut, H::kSkipMin};
}

#define MBXG_XDESC(F, name, pieces) xdesc<XHolder_##name>(),
__constant__ const XDesc kXDesc[kXFam] = {MBXG_XFAMILIES(MBXG_XDESC)};
#undef MBXG_XDESC

// Family of a listed n-mer (ids in listed order) with its holder order, or -1.  code packs the
// family (bits 0-7), the anchor (the holder slot of the listed first monomer, bits 8-11) and the
// holder order (listed position of holder slot s at bits 12 + 4 s).
template <int K>
__device__ __forceinline__ int classify(const NmerCtx& N, const int* ids, int fam_mask) {
    int kinds[4], rk[4], ix[4];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        kinds[i] = N.kind[ids[i]];
        rk[i] = name_rank(kinds[i]);
    }
    sort_network<K>(rk, ix);
    int fam = -1;
    for (int f = 0; f < kXFam && fam < 0; ++f) {
        const XDesc& D = kXDesc[f];
        if (D.order != K) continue;
        // the same multiset of names: every holder slot takes the first unused sorted position
        // with its kind
        bool ok = true;
        int used = 0;
        for (int s = 0; s < K && ok; ++s) {
            int got = -1;
            for (int j = 0; j < K; ++j)
                if (got < 0 && !((used >> j) & 1) && kinds[ix[j]] == D.hkind[s]) got = j;
            if (got < 0)
                ok = false;
            else
                used |= 1 << got;
        }
        if (ok) fam = f;
    }
    if (fam < 0 || !((fam_mask >> fam) & 1)) return -1;
    const XDesc& D = kXDesc[fam];
    int perm[4], used = 0, anchor = 0;
    for (int s = 0; s < K; ++s) {
        int got = 0;
        bool found = false;
        for (int j = 0; j < K; ++j)
            if (!found && !((used >> j) & 1) && kinds[ix[j]] == D.hkind[s]) {
                got = j;
                found = true;
            }
        used |= 1 << got;
        perm[s] = ix[got];
        if (perm[s] == 0) anchor = s;
    }
    // skip rule: at least skip_min first-atom separations (monomers imaged next to the listed
    // first one) beyond the outer radius (e.g. mbnrg_3b_A1_B1C2X2_B1C2X2_deg4_v1.cpp:13013, all of
    // them)
    double p[4][3];
    const double* f0 = N.xyz + 3 * (size_t)N.first_atom[ids[0]];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const double* a = N.xyz + 3 * (size_t)N.first_atom[ids[i]];
        p[i][0] = a[0];
        p[i][1] = a[1];
        p[i][2] = a[2];
        if (N.periodic && i > 0) {
            const Box& B = *N.box;
            double dx, dy, dz;
            image_shift(B, f0[0], f0[1], f0[2], p[i][0], p[i][1], p[i][2], dx, dy, dz);
            image_apply(B, dx, dy, dz, p[i][0], p[i][1], p[i][2]);
        }
    }
    int beyond = 0;
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
        for (int b = a + 1; b < K; ++b) {
            const double* u = p[perm[a]];
            const double* v = p[perm[b]];
            const double d0 = u[0] - v[0], d1 = u[1] - v[1], d2 = u[2] - v[2];
            const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            beyond += r > D.rout;
        }
    if (beyond >= D.skip_min) return -1;
    int code = fam | (anchor << 8);
#pragma unroll
    for (int s = 0; s < K; ++s) code |= perm[s] << (12 + 4 * s);
    return code;
}

constexpr int kXBlock = 256;

}  // namespace

// Pass 1: the code of every listed n-mer and each block's per-family counts (blk[b * kXFam + f]).
template <int K>
__global__ void __launch_bounds__(kXBlock) k_xclass(const int* __restrict__ list, int n, NmerCtx N, int fam_mask,
                                                    int* __restrict__ code, int* __restrict__ blk) {
    __shared__ int sh[kXBlock / 64][kXFam];
    const int t = blockIdx.x * kXBlock + threadIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int c = -1;
    if (t < n) {
        int ids[4];
#pragma unroll
        for (int i = 0; i < K; ++i) ids[i] = list[(size_t)K * t + i];
        c = classify<K>(N, ids, fam_mask);
        code[t] = c;
    }
    const int f = c < 0 ? -1 : (c & 255);
    for (int g = 0; g < kXFam; ++g) {
        const unsigned long long m = __ballot(f == g);
        if (lane == 0) sh[wv][g] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < kXFam) {
        int s = 0;
        for (int w = 0; w < kXBlock / 64; ++w) s += sh[w][threadIdx.x];
        blk[(size_t)blockIdx.x * kXFam + threadIdx.x] = s;
    }
}

// Pass 2 (one workgroup): exclusive scan of the block counts per family; tot[f] = totals.
__global__ void __launch_bounds__(1024) k_xscan(const int* __restrict__ blk, int nblk, int* __restrict__ off,
                                                int* __restrict__ tot) {
    __shared__ int sh[1024];
    const int t = threadIdx.x;
    for (int f = 0; f < kXFam; ++f) {
        int base = 0;
        for (int b0 = 0; b0 < nblk; b0 += 1024) {
            const int b = b0 + t;
            const int v = b < nblk ? blk[(size_t)b * kXFam + f] : 0;
            sh[t] = v;
            __syncthreads();
            for (int s = 1; s < 1024; s <<= 1) {
                const int a = t >= s ? sh[t - s] : 0;
                __syncthreads();
                sh[t] += a;
                __syncthreads();
            }
            if (b < nblk) off[(size_t)b * kXFam + f] = base + sh[t] - v;
            base += sh[1023];
            __syncthreads();
        }
        if (t == 0) tot[f] = base;
    }
}

// Pass 3: every evaluated n-mer to its family's list at its offset: K monomer ids in the
// holder's order, then the anchor slot.
template <int K>
__global__ void __launch_bounds__(kXBlock) k_xemit(const int* __restrict__ list, int n, const int* __restrict__ code,
                                                   const int* __restrict__ off, XLists X) {
    __shared__ int sh[kXBlock / 64][kXFam];
    const int t = blockIdx.x * kXBlock + threadIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = t < n ? code[t] : -1;
    const int f = c < 0 ? -1 : (c & 255);
    int mine = 0;
    for (int g = 0; g < kXFam; ++g) {
        const unsigned long long m = __ballot(f == g);
        if (lane == 0) sh[wv][g] = __popcll(m);
        if (f == g) mine = __popcll(m & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if (f < 0) return;
    int at = off[(size_t)blockIdx.x * kXFam + f] + mine;
    for (int w = 0; w < wv; ++w) at += sh[w][f];
    if (at >= X.cap[f]) return;
    int* o = X.list[f] + (size_t)(K + 1) * at;
#pragma unroll
    for (int s = 0; s < K; ++s) o[s] = list[(size_t)K * t + ((c >> (12 + 4 * s)) & 15)];
    o[K] = (c >> 8) & 15;
}

// ---------------------------------------------------------------------------------------------
// Holder evaluation

namespace {

template <class H>
constexpr int xslot_site(int s) {
    return s == 0 ? 0 : H::kNat[0] + (s > 1 ? H::kNat[1] : 0) + (s > 2 ? H::kNat[2] : 0);
}

// Sites of an evaluate-list entry in holder order: real atoms, every monomer imaged next to the
// anchor (the listed first monomer), then each water's two lone-pair sites
// (tools/water_monomer_lp.cpp:3-21).
template <class H>
__device__ __forceinline__ void xsites(const NmerCtx& N, const int* ent, double* x, double (*oh1)[3],
                                       double (*oh2)[3]) {
    const int anchor = ent[H::kMon];
    const double* f0 = N.xyz + 3 * (size_t)N.first_atom[ent[anchor]];
    const double x0 = f0[0], y0 = f0[1], z0 = f0[2];
#pragma unroll
    for (int s = 0; s < H::kMon; ++s) {
        const double* a = N.xyz + 3 * (size_t)N.first_atom[ent[s]];
        double* dst = x + 3 * xslot_site<H>(s);
#pragma unroll
        for (int j = 0; j < 3 * H::kNat[s]; ++j) dst[j] = a[j];
        if (N.periodic && s != anchor) {
            const Box& B = *N.box;
            double dx, dy, dz;
            image_shift(B, x0, y0, z0, dst[0], dst[1], dst[2], dx, dy, dz);
#pragma unroll
            for (int at = 0; at < H::kNat[s]; ++at) image_apply(B, dx, dy, dz, dst[3 * at], dst[3 * at + 1], dst[3 * at + 2]);
        }
    }
#pragma unroll
    for (int w = 0; w < H::kNVsite; ++w) {
        const double* o = x + 3 * H::kVO[w];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            oh1[w][i] = o[i + 3] - o[i];
            oh2[w][i] = o[i + 6] - o[i];
        }
        const double v[3] = {oh1[w][1] * oh2[w][2] - oh1[w][2] * oh2[w][1], oh1[w][2] * oh2[w][0] - oh1[w][0] * oh2[w][2],
                             oh1[w][0] * oh2[w][1] - oh1[w][1] * oh2[w][0]};
        double* x1 = x + 3 * (H::kNReal + 2 * w);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double in_plane = o[i] + 0.5 * H::kInPlane * (oh1[w][i] + oh2[w][i]);
            const double out_of_plane = H::kOutOfPlane * v[i];
            x1[i] = in_plane + out_of_plane;
            x1[3 + i] = in_plane - out_of_plane;
        }
    }
}

// Variable J (tools/variable.cpp:49-129; form 3 is variable::v_coul, :109-129) and gg with
// d(value)/d(site s1) = gg * (x[s1] - x[s2]).
template <class H, int J>
__device__ __forceinline__ double xvalue(const double* x, double* dvec, double& gg) {
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
    } else if (V.form == 2) {
        val = exp(V.k * (-r));
        gg = -V.k * val / r;
    } else {
        const double e1 = exp(V.k * (-r));
        const double rinv = 1.0 / r;
        val = e1 * rinv;
        gg = -(V.k + rinv) * val * rinv;
    }
    return val;
}

// f_switch of the holders (e.g. mbnrg_2b_A1_B1C2X2_deg6_v1.cpp:3685-3698).
template <class H>
__device__ __forceinline__ double xswitch(double r, double& g) {
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

// First-atom separations of the holder's pairs, their switches and the combination
// sw = sum_t c_t prod_{q in t} s_q (2B: s12; 3B: s12 s13 + s12 s23 + s13 s23 - 2 s12 s13 s23;
// 4B: the center-3-neighbour combination, mbnrg_4b_A1B2_A1B2_A1B2_A1B2_deg4_v1.cpp:640).
template <class H>
__device__ __forceinline__ double xswitches(const double* x, double (*d)[3], double* r, double* s, double* g) {
#pragma unroll
    for (int q = 0; q < H::kNPair; ++q) {
        const double* u = x + 3 * xslot_site<H>(H::kPA[q]);
        const double* v = x + 3 * xslot_site<H>(H::kPB[q]);
        d[q][0] = u[0] - v[0];
        d[q][1] = u[1] - v[1];
        d[q][2] = u[2] - v[2];
        r[q] = sqrt(d[q][0] * d[q][0] + d[q][1] * d[q][1] + d[q][2] * d[q][2]);
        s[q] = xswitch<H>(r[q], g[q]);
    }
    double sw = 0.0;
#pragma unroll
    for (int t = 0; t < H::kNTerm; ++t) {
        double prod = H::kTCoef[t];
#pragma unroll
        for (int q = 0; q < H::kNPair; ++q)
            if ((H::kTMask[t] >> q) & 1) prod *= s[q];
        sw += prod;
    }
    return sw;
}

}  // namespace

// Stage 1: variables and the switch combination.
template <class H>
__global__ void __launch_bounds__(64) k_xvars(const int* __restrict__ list, const int* __restrict__ n_dev, int cap,
                                              NmerCtx N, double* __restrict__ xs, double* __restrict__ sw, int ld) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= min(*n_dev, cap)) return;
    double x[3 * H::kNSites], oh1[2][3], oh2[2][3];
    xsites<H>(N, list + (size_t)(H::kMon + 1) * t, x, oh1, oh2);
    static_for<H::kNVars>([&](auto J) {
        constexpr int j = decltype(J)::value;
        double dv[3], gg;
        xs[(size_t)j * ld + t] = xvalue<H, j>(x, dv, gg);
    });
    double d[6][3], r[6], s[6], g[6];
    sw[t] = xswitches<H>(x, d, r, s, g);
}

constexpr int kGradBlockX = 128;

// Stage 3: energy; the polynomial gradient scaled by the switch combination, to the sites, the
// lone-pair sites' gradients back to O, H, H (water_monomer_lp.cpp:25-48), the switch derivative
// along every pair; virial -sum_atoms x (x) g; gradient scatter.  gs / ps hold the polynomial's
// pieces ([piece][var][ld], [piece][ld]), summed in piece order.  grad == nullptr: energies only
// (the reference computes the virial with the gradients only).
template <class H>
__global__ void __launch_bounds__(kGradBlockX) k_xgrad(const int* __restrict__ list, const int* __restrict__ n_dev,
                                                       int cap, NmerCtx N, const double* __restrict__ gs,
                                                       const double* __restrict__ ps, int ld,
                                                       double* __restrict__ grad, double* __restrict__ e_out,
                                                       double* __restrict__ partials) {
    __shared__ GradTable<128, 9> T;
    const int t = blockIdx.x * kGradBlockX + threadIdx.x;
    const int n = min(*n_dev, cap);
    if (grad) table_init(T);
    __syncthreads();
    double e_sum = 0.0;
    double vir[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (t < n && !grad) {
        const int* ent = list + (size_t)(H::kMon + 1) * t;
        double P = 0.0;
#pragma unroll
        for (int q = 0; q < H::kPieces; ++q) P += ps[(size_t)q * ld + t];
        double x[3 * H::kNSites], oh1[2][3], oh2[2][3];
        xsites<H>(N, ent, x, oh1, oh2);
        double d[6][3], r[6], s[6], gsw[6];
        e_sum = P * xswitches<H>(x, d, r, s, gsw);
        e_out[t] = e_sum;
    } else if (t < n) {
        const int* ent = list + (size_t)(H::kMon + 1) * t;
        double P = 0.0;
#pragma unroll
        for (int q = 0; q < H::kPieces; ++q) P += ps[(size_t)q * ld + t];
        double x[3 * H::kNSites], oh1[2][3], oh2[2][3];
        xsites<H>(N, ent, x, oh1, oh2);
        double d[6][3], r[6], s[6], gsw[6];
        const double sw = xswitches<H>(x, d, r, s, gsw);
        double xg[3 * H::kNSites];
#pragma unroll
        for (int i = 0; i < 3 * H::kNSites; ++i) xg[i] = 0.0;
        static_for<H::kNVars>([&](auto J) {
            constexpr int j = decltype(J)::value;
            double gj = 0.0;
#pragma unroll
            for (int q = 0; q < H::kPieces; ++q) gj += gs[((size_t)q * H::kNVars + j) * ld + t];
            gj *= sw;
            double dv[3], gg;
            xvalue<H, j>(x, dv, gg);
            constexpr int a1 = H::kVars[j].s1, a2 = H::kVars[j].s2;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double tc = gj * (dv[c] * gg);
                xg[3 * a1 + c] += tc;
                xg[3 * a2 + c] -= tc;
            }
        });
#pragma unroll
        for (int w = 0; w < H::kNVsite; ++w) {
            const double* g1 = xg + 3 * (H::kNReal + 2 * w);
            const double* g2 = g1 + 3;
            double* grd = xg + 3 * H::kVO[w];
            const double gm[3] = {g1[0] - g2[0], g1[1] - g2[1], g1[2] - g2[2]};
            const double t1[3] = {oh2[w][1] * gm[2] - oh2[w][2] * gm[1], oh2[w][2] * gm[0] - oh2[w][0] * gm[2],
                                  oh2[w][0] * gm[1] - oh2[w][1] * gm[0]};
            const double t2[3] = {oh1[w][1] * gm[2] - oh1[w][2] * gm[1], oh1[w][2] * gm[0] - oh1[w][0] * gm[2],
                                  oh1[w][0] * gm[1] - oh1[w][1] * gm[0]};
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
#pragma unroll
        for (int q = 0; q < H::kNPair; ++q) {
            double dsw = 0.0;  // d sw / d s_q
#pragma unroll
            for (int tt = 0; tt < H::kNTerm; ++tt) {
                if (!((H::kTMask[tt] >> q) & 1)) continue;
                double prod = H::kTCoef[tt];
#pragma unroll
                for (int u = 0; u < H::kNPair; ++u)
                    if (u != q && ((H::kTMask[tt] >> u) & 1)) prod *= s[u];
                dsw += prod;
            }
            const double gq = gsw[q] * dsw * P / r[q];
            const int ia = 3 * xslot_site<H>(H::kPA[q]), ib = 3 * xslot_site<H>(H::kPB[q]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                xg[ia + c] += gq * d[q][c];
                xg[ib + c] -= gq * d[q][c];
            }
        }
        const double e = P * sw;
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
            table_add(T, ent[s2], xg + 3 * xslot_site<H>(s2), 3 * H::kNat[s2], grad,
                      3 * (size_t)N.first_atom[ent[s2]]);
        e_out[t] = e;
        e_sum = e;
    }
    __syncthreads();
    if (grad) table_flush(T, N.first_atom, N.nat, grad);
    e_sum = wave_sum(e_sum);
    const int rec = blockIdx.x * (kGradBlockX / 64) + threadIdx.x / 64;
    double* out = partials + (size_t)rec * kPartials;
    const bool lead = (threadIdx.x & 63) == 0;
    if (lead) out[0] = e_sum;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double v = wave_sum(vir[q]);
        if (lead) out[1 + q] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// Tetramers: one workgroup per query monomer (grid-stride).  Matches: every cloud slot strictly
// inside the cutoff (the query's own slot included, as the radius search returns it), ranked by
// slot index so the order is canonical.  Triples j < k < l of matches are kept unless one of
// j, k, l has a tag no larger than the query's and lies within the cutoff of the other two
// (sys_tools.cpp:594-634; no self-image test there).  pass 0 counts per query, pass 1 writes at
// the scanned offsets.
constexpr int kTetraBlock = 256;
constexpr int kTetraMatches = 512;

__global__ void __launch_bounds__(kTetraBlock) k_tetra(const PtInfo* __restrict__ pinfo, const double* __restrict__ cx,
                                                       const double* __restrict__ cy, const double* __restrict__ cz,
                                                       int np, int ns, NmerCtx N, double c4sq, int pass,
                                                       int* __restrict__ counts, const int* __restrict__ offsets,
                                                       int* __restrict__ out, int cap, int* __restrict__ ovf) {
    __shared__ int match[kTetraMatches], ranked[kTetraMatches];
    __shared__ double mx[kTetraMatches], my[kTetraMatches], mz[kTetraMatches];
    __shared__ int mtag[kTetraMatches], mmon[kTetraMatches];
    __shared__ int nmatch, wsum[kTetraBlock / 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int q = blockIdx.x; q < np; q += gridDim.x) {
        const int mi = N.members[q];
        if (!N.islocal[mi]) {
            if (pass == 0 && tid == 0) counts[q] = 0;
            continue;
        }
        if (tid == 0) nmatch = 0;
        __syncthreads();
        const double qx = cx[q], qy = cy[q], qz = cz[q];
        for (int s0 = 0; s0 < ns; s0 += kTetraBlock) {
            const int s = s0 + tid;
            const bool in = s < ns && dist2(qx, qy, qz, cx[s], cy[s], cz[s]) < c4sq;
            const unsigned long long m = __ballot(in);
            int at = 0;
            if (lane == 0 && m) at = atomicAdd(&nmatch, __popcll(m));
            at = __shfl(at, 0, 64) + __popcll(m & ((1ull << lane) - 1ull));
            if (in && at < kTetraMatches) match[at] = s;
        }
        __syncthreads();
        const int nm = nmatch;
        if (nm > kTetraMatches) {  // emits nothing; the host sees the overflow and fails loudly
            if (tid == 0) {
                atomicMax(ovf, nm);
                if (pass == 0) counts[q] = 0;
            }
            __syncthreads();
            continue;
        }
        for (int a = tid; a < nm; a += kTetraBlock) {
            const int key = match[a];
            int rank = 0;
            for (int b = 0; b < nm; ++b) rank += match[b] < key;
            ranked[rank] = key;
        }
        __syncthreads();
        for (int a = tid; a < nm; a += kTetraBlock) {
            const int s = ranked[a];
            mx[a] = cx[s];
            my[a] = cy[s];
            mz[a] = cz[s];
            const int p = s % np;
            mtag[a] = pinfo[p].tag;
            mmon[a] = N.members[p];
        }
        __syncthreads();
        const int ti = pinfo[q].tag;
        // pair (j, k), j < k, per work item; each item loops over l > k
        const int npair = nm * (nm - 1) / 2;
        auto emit_pair = [&](int t, bool write, int base) {
            const double w = 2.0 * nm - 1.0;
            int j = (int)((w - sqrt(w * w - 8.0 * t)) * 0.5);
            j = j < 0 ? 0 : (j > nm - 2 ? nm - 2 : j);
            while (j > 0 && j * (2 * nm - j - 1) / 2 > t) --j;
            while (j < nm - 2 && (j + 1) * (2 * nm - j - 2) / 2 <= t) ++j;
            const int k = t - j * (2 * nm - j - 1) / 2 + j + 1;
            const double d2jk = dist2(mx[j], my[j], mz[j], mx[k], my[k], mz[k]);
            int cnt = 0;
            for (int l = k + 1; l < nm; ++l) {
                const double d2jl = dist2(mx[j], my[j], mz[j], mx[l], my[l], mz[l]);
                const double d2kl = dist2(mx[k], my[k], mz[k], mx[l], my[l], mz[l]);
                const bool drop = (ti >= mtag[j] && d2jk < c4sq && d2jl < c4sq) ||
                                  (ti >= mtag[k] && d2jk < c4sq && d2kl < c4sq) ||
                                  (ti >= mtag[l] && d2kl < c4sq && d2jl < c4sq);
                if (!drop) {
                    if (write && base + cnt < cap) {
                        int* o = out + 4 * (size_t)(base + cnt);
                        o[0] = mi;
                        o[1] = mmon[j];
                        o[2] = mmon[k];
                        o[3] = mmon[l];
                    }
                    ++cnt;
                }
            }
            return cnt;
        };
        // the block's work items in order: thread tid takes items tid, tid + B, ...; counts per
        // thread, a block scan gives each thread's start
        int mine = 0;
        for (int t = tid; t < npair; t += kTetraBlock) mine += emit_pair(t, false, 0);
        // block-wide exclusive scan of `mine` over threads in item order is not the item order:
        // items are interleaved, so the write pass walks the items in rounds and scans per round
        if (pass == 0) {
            int v = mine;
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) wsum[wv] = v;
            __syncthreads();
            if (tid == 0) {
                int s = 0;
                for (int w = 0; w < kTetraBlock / 64; ++w) s += wsum[w];
                counts[q] = s;
            }
            __syncthreads();
        } else {
            int base = offsets[q];
            for (int t0 = 0; t0 < npair; t0 += kTetraBlock) {
                const int t = t0 + tid;
                const int c = t < npair ? emit_pair(t, false, 0) : 0;
                // exclusive prefix of c over the round's items (threads in order)
                int incl = c;
                for (int o = 1; o < 64; o <<= 1) {
                    const int y = __shfl_up(incl, o, 64);
                    if (lane >= o) incl += y;
                }
                if (lane == 63) wsum[wv] = incl;
                __syncthreads();
                int pre = incl - c;
                for (int w = 0; w < wv; ++w) pre += wsum[w];
                int round = 0;
                for (int w = 0; w < kTetraBlock / 64; ++w) round += wsum[w];
                if (t < npair && c > 0) emit_pair(t, true, base + pre);
                base += round;
                __syncthreads();
            }
        }
        __syncthreads();
    }
}

// Exclusive scan of n per-query counts (one workgroup): off[0..n], off[n] = total.
__global__ void __launch_bounds__(1024) k_scan_int(const int* __restrict__ cnt, int n, int* __restrict__ off) {
    __shared__ int sh[1024];
    const int t = threadIdx.x;
    int base = 0;
    for (int b0 = 0; b0 < n; b0 += 1024) {
        const int b = b0 + t;
        const int v = b < n ? cnt[b] : 0;
        sh[t] = v;
        __syncthreads();
        for (int s = 1; s < 1024; s <<= 1) {
            const int a = t >= s ? sh[t - s] : 0;
            __syncthreads();
            sh[t] += a;
            __syncthreads();
        }
        if (b < n) off[b] = base + sh[t] - v;
        base += sh[1023];
        __syncthreads();
    }
    if (t == 0) off[n] = base;
}

// ---------------------------------------------------------------------------------------------
// Host launchers

#define MBXG_XPIECE(poly, k)                                                                           \
    hipError_t launch_xpoly_##poly##_p##k(const double*, const int*, int, int, const double*, double*, \
                                          double*, hipStream_t);
MBXG_XPIECE(halide_h2o_2b, 0)
MBXG_XPIECE(co2co2_2b, 0)
MBXG_XPIECE(co2co2_2b, 1)
MBXG_XPIECE(halide_h2o2_3b, 0)
MBXG_XPIECE(halide_h2o2_3b, 1)
MBXG_XPIECE(halide_h2o2_3b, 2)
MBXG_XPIECE(halide_h2o2_3b, 3)
MBXG_XPIECE(co2x3_3b, 0)
MBXG_XPIECE(co2x3_3b, 1)
MBXG_XPIECE(co2x3_3b, 2)
MBXG_XPIECE(co2x3_3b, 3)
MBXG_XPIECE(h2ox4_4b, 0)
MBXG_XPIECE(h2ox4_4b, 1)
MBXG_XPIECE(h2ox4_4b, 2)
MBXG_XPIECE(h2ox4_4b, 3)
MBXG_XPIECE(h2ox4_4b, 4)
MBXG_XPIECE(h2ox4_4b, 5)
MBXG_XPIECE(h2ox4_4b, 6)
MBXG_XPIECE(h2ox4_4b, 7)
MBXG_XPIECE(alkali_h2o_2b, 0)
MBXG_XPIECE(alkali_h2o2_3b, 0)
MBXG_XPIECE(alkali_h2o2_3b, 1)
#undef MBXG_XPIECE

namespace {

typedef hipError_t (*xpiece_launch)(const double*, const int*, int, int, const double*, double*, double*, hipStream_t);

template <class H>
size_t xworkspace(int cap) {
    const size_t ld = ((size_t)cap + 63) / 64 * 64;
    return ((size_t)H::kNVars * (1 + H::kPieces) + H::kPieces + 1) * ld;
}

template <class H>
hipError_t xlaunch(const xpiece_launch* pieces, const int* list, const int* n_dev, int cap, const NmerCtx& N,
                   const double* coef, double* grad, double* e_out, double* partials, double* ws, hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    const int nb = (cap + 63) / 64;
    const int ld = nb * 64;
    double* xs = ws;
    double* gs = xs + (size_t)H::kNVars * ld;
    double* ps = gs + (size_t)H::kPieces * H::kNVars * ld;
    double* sw = ps + (size_t)H::kPieces * ld;
    hipLaunchKernelGGL(k_xvars<H>, dim3(nb), dim3(64), 0, st, list, n_dev, cap, N, xs, sw, ld);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int q = 0; q < H::kPieces; ++q) {
        e = pieces[q](xs, n_dev, cap, ld, coef, gs + (size_t)q * H::kNVars * ld, ps + (size_t)q * ld, st);
        if (e != hipSuccess) return e;
    }
    const int ngb = (cap + kGradBlockX - 1) / kGradBlockX;
    hipLaunchKernelGGL(k_xgrad<H>, dim3(ngb), dim3(kGradBlockX), 0, st, list, n_dev, cap, N, gs, ps, ld, grad, e_out,
                       partials);
    return hipGetLastError();
}

const xpiece_launch kPHalide2[1] = {launch_xpoly_halide_h2o_2b_p0};
const xpiece_launch kPCo2_2[2] = {launch_xpoly_co2co2_2b_p0, launch_xpoly_co2co2_2b_p1};
const xpiece_launch kPHalide3[4] = {launch_xpoly_halide_h2o2_3b_p0, launch_xpoly_halide_h2o2_3b_p1,
                                    launch_xpoly_halide_h2o2_3b_p2, launch_xpoly_halide_h2o2_3b_p3};
const xpiece_launch kPCo2_3[4] = {launch_xpoly_co2x3_3b_p0, launch_xpoly_co2x3_3b_p1, launch_xpoly_co2x3_3b_p2,
                                  launch_xpoly_co2x3_3b_p3};
const xpiece_launch kPH2o_4[8] = {launch_xpoly_h2ox4_4b_p0, launch_xpoly_h2ox4_4b_p1, launch_xpoly_h2ox4_4b_p2,
                                  launch_xpoly_h2ox4_4b_p3, launch_xpoly_h2ox4_4b_p4, launch_xpoly_h2ox4_4b_p5,
                                  launch_xpoly_h2ox4_4b_p6, launch_xpoly_h2ox4_4b_p7};
const xpiece_launch kPAlkali2[1] = {launch_xpoly_alkali_h2o_2b_p0};
const xpiece_launch kPAlkali3[2] = {launch_xpoly_alkali_h2o2_3b_p0, launch_xpoly_alkali_h2o2_3b_p1};

}  // namespace

__host__ __device__ int xfam_order(int f) {
    if (f <= XF_CO2CO2_2B || (f >= XF_LI_H2O_2B && f <= XF_CS_H2O_2B)) return 2;
    return f == XF_H2OX4_4B ? 4 : 3;
}

size_t xfam_workspace(int f, int cap) {
    switch (f) {
#define MBXG_XW(F, name, pieces) \
    case F: return xworkspace<XHolder_##name>(cap);
        MBXG_XFAMILIES(MBXG_XW)
#undef MBXG_XW
        default: return 0;
    }
}

hipError_t launch_xfam(int f, const int* list, const int* n_dev, int cap, const NmerCtx& N, const double* coef,
                       double* grad, double* e_out, double* partials, double* ws, hipStream_t st) {
    switch (f) {
#define MBXG_XL(F, name, pieces) \
    case F: return xlaunch<XHolder_##name>(pieces, list, n_dev, cap, N, coef, grad, e_out, partials, ws, st);
        MBXG_XFAMILIES(MBXG_XL)
#undef MBXG_XL
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_xclass(int order, const int* list, int n, const NmerCtx& N, int fam_mask, int* code, int* blk,
                         int* blk_off, int* tot, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int nb = (n + kXBlock - 1) / kXBlock;
    if (order == 2)
        hipLaunchKernelGGL(k_xclass<2>, dim3(nb), dim3(kXBlock), 0, st, list, n, N, fam_mask, code, blk);
    else if (order == 3)
        hipLaunchKernelGGL(k_xclass<3>, dim3(nb), dim3(kXBlock), 0, st, list, n, N, fam_mask, code, blk);
    else
        hipLaunchKernelGGL(k_xclass<4>, dim3(nb), dim3(kXBlock), 0, st, list, n, N, fam_mask, code, blk);
    hipLaunchKernelGGL(k_xscan, dim3(1), dim3(1024), 0, st, blk, nb, blk_off, tot);
    return hipGetLastError();
}

hipError_t launch_xemit(int order, const int* list, int n, const int* code, const int* blk_off, const XLists& X,
                        hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int nb = (n + kXBlock - 1) / kXBlock;
    if (order == 2)
        hipLaunchKernelGGL(k_xemit<2>, dim3(nb), dim3(kXBlock), 0, st, list, n, code, blk_off, X);
    else if (order == 3)
        hipLaunchKernelGGL(k_xemit<3>, dim3(nb), dim3(kXBlock), 0, st, list, n, code, blk_off, X);
    else
        hipLaunchKernelGGL(k_xemit<4>, dim3(nb), dim3(kXBlock), 0, st, list, n, code, blk_off, X);
    return hipGetLastError();
}

size_t xclass_blocks(int n) { return ((size_t)n + kXBlock - 1) / kXBlock; }

// The further families' records onto the per-order totals: 2B record += the 2B families, 3B
// record += the 3B families, 4B record = the 4B family (fixed order).
__global__ void k_xcombine(const double* __restrict__ xt, double* __restrict__ totals) {
    const int q = threadIdx.x;
    if (q >= kPartials) return;
    double a[3] = {totals[q], totals[kPartials + q], 0.0};
    for (int f = 0; f < kXFam; ++f) a[xfam_order(f) - 2] += xt[(size_t)f * kPartials + q];
    totals[q] = a[0];
    totals[kPartials + q] = a[1];
    totals[3 * kPartials + q] = a[2];
}

hipError_t launch_xcombine(const double* xfam_tot, double* totals, hipStream_t st) {
    hipLaunchKernelGGL(k_xcombine, dim3(1), dim3(64), 0, st, xfam_tot, totals);
    return hipGetLastError();
}

hipError_t launch_tetra(const PtInfo* pinfo, const double* cx, const double* cy, const double* cz, int np, int ns,
                        const NmerCtx& N, double c4sq, int pass, int* counts, int* offsets, int* out, int cap,
                        int* ovf, hipStream_t st) {
    if (np <= 0) return hipSuccess;
    const int nb = np < 4096 ? np : 4096;
    hipLaunchKernelGGL(k_tetra, dim3(nb), dim3(kTetraBlock), 0, st, pinfo, cx, cy, cz, np, ns, N, c4sq, pass, counts,
                       offsets, out, cap, ovf);
    if (pass == 0) hipLaunchKernelGGL(k_scan_int, dim3(1), dim3(1024), 0, st, counts, np, offsets);
    return hipGetLastError();
}

const double* xfam_default_coef(int f, int* n) {
    switch (f) {
#define MBXG_XC(F, name, pieces) \
    case F:                      \
        *n = kNCoefX_##name;     \
        return kFitX_##name;
        MBXG_XFAMILIES(MBXG_XC)
#undef MBXG_XC
        default:
            *n = 0;
            return nullptr;
    }
}

}  // namespace mbxg
