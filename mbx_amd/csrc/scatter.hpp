// Gradient scatter for the chain-rule stages: "reduce in LDS, then atomics".
//
// A block of n-mers touches far fewer distinct monomers than it has n-mer slots (the evaluated
// lists are grouped by query monomer, and partners are its near neighbours).  Each lane adds
// its per-monomer atom gradients into a small open-addressing table in LDS keyed by monomer
// (LDS FP64 atomics); after a barrier the table is flushed to the global gradient with one
// set of device atomics per distinct monomer.  If the table fills, the lane falls back to
// device atomics directly, so the result never depends on the table size.
#pragma once

#include <hip/hip_runtime.h>

namespace mbxg {

template <int TB, int NA>
struct GradTable {
    int key[TB];
    double val[TB * NA];
};

template <int TB, int NA>
__device__ __forceinline__ void table_init(GradTable<TB, NA>& T) {
    for (int i = threadIdx.x; i < TB; i += blockDim.x) T.key[i] = -1;
    for (int i = threadIdx.x; i < TB * NA; i += blockDim.x) T.val[i] = 0.0;
}

// Add cnt (<= NA) values for monomer `mon`; falls back to device atomics at grad + goff if full.
template <int TB, int NA>
__device__ __forceinline__ void table_add(GradTable<TB, NA>& T, int mon, const double* g, int cnt, double* grad,
                                          size_t goff) {
    unsigned h = ((unsigned)mon * 2654435761u) % TB;
    for (int probe = 0; probe < TB; ++probe) {
        int k = __atomic_load_n(&T.key[h], __ATOMIC_RELAXED);
        if (k == -1) k = atomicCAS(&T.key[h], -1, mon);
        if (k == -1 || k == mon) {
#pragma unroll
            for (int c = 0; c < NA; ++c)
                if (c < cnt) unsafeAtomicAdd(&T.val[h * NA + c], g[c]);
            return;
        }
        h = (h + 1) % TB;
    }
#pragma unroll
    for (int c = 0; c < NA; ++c)
        if (c < cnt) unsafeAtomicAdd(grad + goff + c, g[c]);
}

// After a barrier: one set of device atomics per occupied entry (3 values per atom of the
// entry's monomer).
template <int TB, int NA>
__device__ __forceinline__ void table_flush(const GradTable<TB, NA>& T, const int* first_atom, const int* nat,
                                            double* grad) {
    for (int i = threadIdx.x; i < TB; i += blockDim.x) {
        const int mon = T.key[i];
        if (mon < 0) continue;
        double* gp = grad + 3 * (size_t)first_atom[mon];
        const int cnt = 3 * nat[mon];
#pragma unroll
        for (int c = 0; c < NA; ++c)
            if (c < cnt) unsafeAtomicAdd(gp + c, T.val[i * NA + c]);
    }
}

// A table that one wavefront owns alone (the other wavefronts of its block never touch it):
// lane-strided init and flush, no block barrier.  The LDS unit performs one wavefront's
// operations in issue order, so between the phases only the compiler has to be kept from moving
// them and the outstanding LDS operations have to be waited for: wave_lds_order.
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

template <int TB, int NA>
__device__ __forceinline__ void table_init_wave(GradTable<TB, NA>& T) {
    const int lane = threadIdx.x & 63;
    for (int i = lane; i < TB; i += 64) T.key[i] = -1;
    for (int i = lane; i < TB * NA; i += 64) T.val[i] = 0.0;
}

template <int TB, int NA>
__device__ __forceinline__ void table_flush_wave(const GradTable<TB, NA>& T, const int* first_atom, const int* nat,
                                                 double* grad) {
    for (int i = threadIdx.x & 63; i < TB; i += 64) {
        const int mon = T.key[i];
        if (mon < 0) continue;
        double* gp = grad + 3 * (size_t)first_atom[mon];
        const int cnt = 3 * nat[mon];
#pragma unroll
        for (int c = 0; c < NA; ++c)
            if (c < cnt) unsafeAtomicAdd(gp + c, T.val[i * NA + c]);
    }
}

}  // namespace mbxg
