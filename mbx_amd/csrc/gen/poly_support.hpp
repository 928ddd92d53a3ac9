// Support definitions for the generated polynomial code (gen/poly_*.inc).
#pragma once

#include <hip/hip_runtime.h>

// Coefficients live in the constant address space: uniform loads become scalar (SMEM) loads
// and their values arrive in SGPRs, which VALU FMAs take as an operand directly.
typedef const __attribute__((address_space(4))) double* mbxg_cptr;

// Node boundary: nothing is scheduled across it, and the coefficient pointer is redefined by
// an empty asm so coefficient loads are not merged across nodes (which would keep hundreds of
// coefficients live in SGPRs).  The "s" constraint keeps the pointer uniform.
#define MBXG_NODE_FENCE(C)                      \
    do {                                        \
        __builtin_amdgcn_sched_barrier(0);      \
        asm volatile("" : "+s"(C));             \
    } while (0)
