// MB-pol 3B device code shared by the stage kernels (eval_mbpol3b.hip) and the paired polynomial
// launch that runs the 3B chain rule in its tail (poly_mbpol_pair.hip): the variable table,
// switches, trimer loading and the chain rule of 64 trimers (one lane each).
//
// Restates x2o::x3b_v2x::eval (potential/3b/x3b-v2x.cpp:1403-1806): variables
// exp(-k (d - d0)) (:1548-1584), the switch product s = sab*sac + sab*sbc + sac*sbc
// (:1614-1620), and the chain rule g_var (:63-76, 1622-1673), switch derivatives (:1676-1686)
// and per-trimer virial (:1687-1726).
#pragma once

#include "mbxg_device.hpp"
#include "mbxg_kernels.hpp"
#include "scatter.hpp"

namespace mbxg {

namespace {

// Variable table: atom slots 0..8 = Oa Ha1 Ha2 Ob Hb1 Hb2 Oc Hc1 Hc2.
struct V3 {
    unsigned char set, a1, a2;
};
__constant__ const V3 kVar3[36] = {
    {0, 1, 2}, {0, 4, 5}, {0, 7, 8}, {1, 0, 1}, {1, 0, 2}, {1, 3, 4}, {1, 3, 5}, {1, 6, 7}, {1, 6, 8},
    {2, 1, 4}, {2, 1, 5}, {2, 1, 7}, {2, 1, 8}, {2, 2, 4}, {2, 2, 5}, {2, 2, 7}, {2, 2, 8}, {2, 4, 7},
    {2, 4, 8}, {2, 5, 7}, {2, 5, 8}, {3, 0, 4}, {3, 0, 5}, {3, 0, 7}, {3, 0, 8}, {3, 3, 1}, {3, 3, 2},
    {3, 3, 7}, {3, 3, 8}, {3, 6, 1}, {3, 6, 2}, {3, 6, 4}, {3, 6, 5}, {4, 0, 3}, {4, 0, 6}, {4, 3, 6}};

__device__ __forceinline__ double kset(int s) {
    return s == 0 ? p3::kHH_intra : s == 1 ? p3::kOH_intra : s == 2 ? p3::kHH : s == 3 ? p3::kOH : p3::kOO;
}
__device__ __forceinline__ double dset(int s) {
    return s == 0 ? p3::dHH_intra : s == 1 ? p3::dOH_intra : s == 2 ? p3::dHH : s == 3 ? p3::dOH : p3::dOO;
}

// f_switch (x3b-v2x.cpp:1275-1288) with r3i = 0.
__device__ __forceinline__ double switch3(double r, double& g) {
    if (r > p3::r3f) {
        g = 0.0;
        return 0.0;
    } else if (r > p3::r3i) {
        const double t1 = M_PI / (p3::r3f - p3::r3i);
        const double x = (r - p3::r3i) * t1;
        g = -sin(x) * t1 / 2.0;
        return (1.0 + cos(x)) / 2.0;
    }
    g = 0.0;
    return 1.0;
}

// Load the 9 atoms of a water trimer, imaging monomers 2 and 3 next to monomer 1.
__device__ __forceinline__ void load_trimer(const NmerCtx& N, int m1, int m2, int m3, double* w) {
    const double* a = N.xyz + 3 * (size_t)N.first_atom[m1];
    const double* b = N.xyz + 3 * (size_t)N.first_atom[m2];
    const double* c = N.xyz + 3 * (size_t)N.first_atom[m3];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        w[j] = a[j];
        w[9 + j] = b[j];
        w[18 + j] = c[j];
    }
    if (N.periodic) {
        const Box& B = *N.box;
        double dx, dy, dz;
        image_shift(B, w[0], w[1], w[2], w[9], w[10], w[11], dx, dy, dz);
#pragma unroll
        for (int j = 0; j < 3; ++j) image_apply(B, dx, dy, dz, w[9 + 3 * j], w[10 + 3 * j], w[11 + 3 * j]);
        image_shift(B, w[0], w[1], w[2], w[18], w[19], w[20], dx, dy, dz);
#pragma unroll
        for (int j = 0; j < 3; ++j) image_apply(B, dx, dy, dz, w[18 + 3 * j], w[19 + 3 * j], w[20 + 3 * j]);
    }
}

// O-O separations and the switch product (x3b-v2x.cpp:1495-1509, 1614-1620); g* receive the
// switch derivatives.
__device__ __forceinline__ double switches(const double* w, double* dab, double* dac, double* dbc, double& rab,
                                           double& rac, double& rbc, double& sab, double& sac, double& sbc,
                                           double& gab, double& gac, double& gbc) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        dab[j] = w[j] - w[9 + j];
        dac[j] = w[j] - w[18 + j];
        dbc[j] = w[9 + j] - w[18 + j];
    }
    rab = sqrt(dab[0] * dab[0] + dab[1] * dab[1] + dab[2] * dab[2]);
    rac = sqrt(dac[0] * dac[0] + dac[1] * dac[1] + dac[2] * dac[2]);
    rbc = sqrt(dbc[0] * dbc[0] + dbc[1] * dbc[1] + dbc[2] * dbc[2]);
    sab = switch3(rab, gab);
    sac = switch3(rac, gac);
    sbc = switch3(rbc, gbc);
    return sab * sac + sab * sbc + sac * sbc;
}

// An empty asm over 27 doubles and an offset: the values are "rewritten" in place, so whatever
// computes them stays before the fence and whatever reads them (or addresses through off) after.
__device__ __forceinline__ void group_fence(size_t& off, double* v) {
    asm volatile(""
                 : "+v"(off), "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]), "+v"(v[5]), "+v"(v[6]),
                   "+v"(v[7]), "+v"(v[8]), "+v"(v[9]), "+v"(v[10]), "+v"(v[11]), "+v"(v[12]), "+v"(v[13]));
    asm volatile(""
                 : "+v"(off), "+v"(v[14]), "+v"(v[15]), "+v"(v[16]), "+v"(v[17]), "+v"(v[18]), "+v"(v[19]),
                   "+v"(v[20]), "+v"(v[21]), "+v"(v[22]), "+v"(v[23]), "+v"(v[24]), "+v"(v[25]), "+v"(v[26]));
}

// Stage 1 for the trimer of lane t: its 36 variables into xs (entry (j, t) at j*ld + t) and its
// switch product into sw; lanes at or past min(*n_dev, cap) do nothing.
__device__ __forceinline__ void vars3_lane(int t, const int* __restrict__ list, const int* __restrict__ n_dev,
                                           int cap, const NmerCtx& N, double* __restrict__ xs,
                                           double* __restrict__ sw, int ld) {
    if (t >= min(*n_dev, cap)) return;
    double w[27];
    load_trimer(N, list[3 * t], list[3 * t + 1], list[3 * t + 2], w);
#pragma unroll
    for (int j = 0; j < 36; ++j) {
        const int st = kVar3[j].set, a1 = kVar3[j].a1, a2 = kVar3[j].a2;
        const double d0 = w[3 * a1] - w[3 * a2];
        const double d1 = w[3 * a1 + 1] - w[3 * a2 + 1];
        const double d2 = w[3 * a1 + 2] - w[3 * a2 + 2];
        const double d = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        xs[(size_t)j * ld + t] = exp(-kset(st) * (d - dset(st)));
    }
    double dab[3], dac[3], dbc[3], rab, rac, rbc, sab, sac, sbc, gab, gac, gbc;
    sw[t] = switches(w, dab, dac, dbc, rab, rac, rbc, sab, sac, sbc, gab, gac, gbc);
}

// variables per group of the chain rule's software pipeline (a divisor of 36)
constexpr int kChainGroup3 = 4;

// E = s*P for the trimer of lane t (t < n), and with GRAD the chain rule to its atoms, switch
// derivatives, virial and gradient scatter through the LDS table T; the 64 lanes' energy and
// virial sums go to partial record `rec`.  WAVE: T belongs to this wavefront alone (lane-strided
// init and flush, no block barrier); otherwise every thread of the block calls this and T is the
// block's.  READX: the variables are read back from xs instead of recomputed.
template <bool GRAD, bool WAVE, int TB, bool READX = false>
__device__ __forceinline__ void chain3(int t, int rec, int n, const int* __restrict__ list, const NmerCtx& N,
                                       const double* __restrict__ xs, const double* __restrict__ gs, const double* __restrict__ ps,
                                       const double* __restrict__ sw, int ld, double* __restrict__ grad,
                                       double* __restrict__ e_out, double* __restrict__ vir_out,
                                       double* __restrict__ partials, GradTable<TB, 9>& T) {
    if (GRAD) {
        if (WAVE) {
            table_init_wave(T);
            wave_lds_order();
        } else {
            table_init(T);
        }
    }
    if (!WAVE) __syncthreads();
    double e_sum = 0.0;
    double vir[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int m1 = -1, m2 = -1, m3 = -1;
    double gd[27];
#pragma unroll
    for (int j = 0; j < 27; ++j) gd[j] = 0.0;
    if (t < n) {
        const double s = sw[t];
        const double P = ps[t];
        if (GRAD) {
            m1 = list[3 * t];
            m2 = list[3 * t + 1];
            m3 = list[3 * t + 2];
            double w[27];
            load_trimer(N, m1, m2, m3, w);
            // g_var (x3b-v2x.cpp:63-76, 1638-1673), kChainGroup3 variables at a time.  At each group
            // boundary the coordinates, the gradient accumulators and the stage offset pass
            // through empty asm statements: group g's arithmetic starts after the boundary, the
            // accumulations of group g - 1 end before it, and the loads of group g + 1 (issued
            // there, in flight while group g computes) follow it.  Left to itself the compiler
            // hoists every load and square root and sinks every accumulation: ~500 registers.
            constexpr int G = kChainGroup3;
            size_t toff = (size_t)t;
            double cg[G], ng[G], cx[G], nx[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                cg[k] = gs[(size_t)k * ld + toff];
                if (READX) cx[k] = xs[(size_t)k * ld + toff];
            }
#pragma unroll
            for (int g = 0; g < 36 / G; ++g) {
                group_fence(toff, w);
                group_fence(toff, gd);
                if (g < 36 / G - 1) {
#pragma unroll
                    for (int k = 0; k < G; ++k) {
                        ng[k] = gs[(size_t)(G * g + G + k) * ld + toff];
                        if (READX) nx[k] = xs[(size_t)(G * g + G + k) * ld + toff];
                    }
                }
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const int j = G * g + k;
                    const int st = kVar3[j].set, a1 = kVar3[j].a1, a2 = kVar3[j].a2;
                    const double d0 = w[3 * a1] - w[3 * a2];
                    const double d1 = w[3 * a1 + 1] - w[3 * a2 + 1];
                    const double d2 = w[3 * a1 + 2] - w[3 * a2 + 2];
                    const double d = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                    const double gj = cg[k] * s;
                    // the variable itself: read back (READX) or recomputed as k_vars_mbpol3b computed
                    // it (the same bits either way; alone, one exp is cheaper than the load)
                    const double xv = READX ? cx[k] : exp(-kset(st) * (d - dset(st)));
                    const double gg = -kset(st) * gj * xv / d;
                    gd[3 * a1] += gg * d0;
                    gd[3 * a1 + 1] += gg * d1;
                    gd[3 * a1 + 2] += gg * d2;
                    gd[3 * a2] -= gg * d0;
                    gd[3 * a2 + 1] -= gg * d1;
                    gd[3 * a2 + 2] -= gg * d2;
                }
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    cg[k] = ng[k];
                    if (READX) cx[k] = nx[k];
                }
            }
            group_fence(toff, gd);
            // switch derivatives (x3b-v2x.cpp:1676-1686)
            double dab[3], dac[3], dbc[3], rab, rac, rbc, sab, sac, sbc, gab, gac, gbc;
            switches(w, dab, dac, dbc, rab, rac, rbc, sab, sac, sbc, gab, gac, gbc);
            gab *= (sac + sbc) * P / rab;
            gac *= (sab + sbc) * P / rac;
            gbc *= (sab + sac) * P / rbc;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gd[c] += gab * dab[c] + gac * dac[c];
                gd[9 + c] += gbc * dbc[c] - gab * dab[c];
                gd[18 + c] -= gac * dac[c] + gbc * dbc[c];
            }
            // virial of this trimer (x3b-v2x.cpp:1687-1726): -sum_atoms w_a g_b
            const int va[6] = {0, 0, 0, 1, 1, 2}, vb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                double acc = 0.0;
#pragma unroll
                for (int at = 0; at < 9; ++at) acc -= w[3 * at + va[q]] * gd[3 * at + vb[q]];
                vir[q] = acc;
            }
        }
        const double e = s * P;
        e_out[t] = e;
        e_sum = e;
        if (vir_out)
#pragma unroll
            for (int q = 0; q < 6; ++q) vir_out[(size_t)q * ld + t] = vir[q];
    }
    if (GRAD) {
        // The evaluate list is grouped by query monomer, so usually every active lane of the
        // wavefront has the same monomer 1: its nine gradient values are then summed across the
        // wavefront (inactive lanes hold zeros) and enter the table once instead of 64 times.
        const bool act = t < n;
        const unsigned long long am = __ballot(act);
        if (am) {
            const int lane = threadIdx.x & 63, lead = __ffsll((long long)am) - 1;
            const int mq = __shfl(m1, lead, 64);
            if (__ballot(act && m1 != mq) == 0) {
                double r[9];
#pragma unroll
                for (int c = 0; c < 9; ++c) r[c] = wave_sum(gd[c]);
                if (lane == lead) table_add(T, mq, r, 9, grad, 3 * (size_t)N.first_atom[mq]);
            } else if (act) {
                table_add(T, m1, gd, 9, grad, 3 * (size_t)N.first_atom[m1]);
            }
            if (act) {
                table_add(T, m2, gd + 9, 9, grad, 3 * (size_t)N.first_atom[m2]);
                table_add(T, m3, gd + 18, 9, grad, 3 * (size_t)N.first_atom[m3]);
            }
        }
        if (WAVE) {
            wave_lds_order();
            table_flush_wave(T, N.first_atom, N.nat, grad);
        } else {
            __syncthreads();
            table_flush(T, N.first_atom, N.nat, grad);
        }
    }
    // one partial record per wavefront
    e_sum = wave_sum(e_sum);
    double* out = partials + (size_t)rec * kPartials;
    const bool lead = (threadIdx.x & 63) == 0;
    if (lead) out[0] = e_sum;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double v = wave_sum(vir[q]);
        if (lead) out[1 + q] = v;
    }
}

}  // namespace

}  // namespace mbxg
