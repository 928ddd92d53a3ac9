// Support definitions for the generated polynomial code (gen/poly_*.inc).
#pragma once

#include <hip/hip_runtime.h>

// Coefficients live in the constant address space: uniform loads become scalar (SMEM) loads
// and their values arrive in SGPRs, which VALU FMAs take as an operand directly.
typedef const __attribute__((address_space(4))) double* mbxg_cptr;

// Region boundary.  Nothing is scheduled across it, and the coefficient pointer is redefined
// by an empty asm that also consumes `dep` (the previous region's last result): coefficient
// loads made through the new pointer can neither be merged with earlier loads nor hoisted
// above the previous region's arithmetic, so only about two regions' coefficients are ever
// live in SGPRs.  The "s" constraint keeps the pointer uniform.
#define MBXG_NODE_FENCE(C, dep)                   \
    do {                                          \
        __builtin_amdgcn_sched_barrier(0);        \
        asm volatile("" : "+s"(C) : "v"(dep));    \
    } while (0)
