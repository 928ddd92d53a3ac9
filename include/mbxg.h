/*
 * mbxg -- C ABI of the MI355X (gfx950) MBX many-body short-range engine.
 *
 * One context owns one HIP stream and all device memory it allocates; contexts are
 * independent (thread-safe across contexts, not within one).  Every function returns 0 on
 * success or a negative status; mbxg_last_error() describes the last failure of the calling
 * thread.  No function takes ownership of caller memory.  Units: Angstrom, kcal/mol;
 * gradients are +dE/dx (force = -grad), as in the reference.
 *
 * The engine replaces the reference's 2B/3B orchestration inside bblock::System
 * (src/bblock/system.cpp:2425 Get2B, :2769 Get3B) together with what they call:
 * systools::AddClusters (src/bblock/sys_tools.cpp:472), GetCloseDimerImage (:322),
 * FixMonomerCoordinates (:257), e2b::get_2b_energy (src/potential/2b/energy2b.cpp:206),
 * e3b::get_3b_energy (src/potential/3b/energy3b.cpp:139) and the MB-pol holders
 * x2b_v9x::eval (src/potential/2b/x2b-v9x.cpp:1514) / x3b_v2x::eval
 * (src/potential/3b/x3b-v2x.cpp:1403).  The C++ bblock::System of this project
 * (include/mbx/system.h) drives it; the reference-compatible extern "C" entry points
 * (include/mbx/external_call.h) sit on top of that.
 */
#ifndef MBXG_H
#define MBXG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mbxg_ctx mbxg_ctx;

/* Monomer kinds (the reference keys monomers by name: "h2o", "ch4", "co2", "f-", "cl-", "br-", "i-",
 * "li+", "na+", "k+", "rb+", "cs+"). */
enum {
    MBXG_H2O = 0, MBXG_CH4 = 1, MBXG_OTHER = 2, MBXG_CO2 = 3, MBXG_F = 4, MBXG_CL = 5, MBXG_BR = 6, MBXG_I = 7,
    MBXG_LI = 8, MBXG_NA = 9, MBXG_K = 10, MBXG_RB = 11, MBXG_CS = 12
};
/* Terms for mbxg_compute; MBXG_LISTS_ONLY builds the requested lists without evaluating. */
enum { MBXG_2B = 1, MBXG_3B = 2, MBXG_LISTS_ONLY = 4, MBXG_1B = 8, MBXG_4B = 16 };
/* Status codes. */
enum { MBXG_OK = 0, MBXG_EINVAL = -1, MBXG_EHIP = -2, MBXG_ESTATE = -3, MBXG_ENOMEM = -4 };

/* Text of the last error raised on this thread ("" if none). */
const char* mbxg_last_error(void);

/* Create a context on HIP device `device`. */
int mbxg_create(mbxg_ctx** ctx, int device);
int mbxg_destroy(mbxg_ctx* ctx);

/* Topology, in the engine's (internal) monomer order; monomer i's atoms follow monomer i-1's.
 * kind[i]: MBXG_H2O (3 atoms O H H), MBXG_CH4 (5 atoms C H H H H), MBXG_CO2 (3 atoms C O O), the
 * halide ions MBXG_F / MBXG_CL / MBXG_BR / MBXG_I and the alkali-metal ions MBXG_LI / MBXG_NA /
 * MBXG_K / MBXG_RB / MBXG_CS (1 atom each) or MBXG_OTHER (nat[i] atoms, listed but
 * without a holder here: only MBXG_LISTS_ONLY computes accept it).  tag[i] orders monomers
 * for n-mer ownership (the reference uses the first atom's tag, sys_tools.cpp:558-592).
 * islocal may be NULL (all local).  Replaces System::AddMonomer/Initialize bookkeeping
 * (system.cpp:848-870, 1794-1969). */
int mbxg_set_topology(mbxg_ctx* ctx, int nmon, const int* kind, const int* nat, const int* tag,
                      const unsigned char* islocal);

/* Box as 9 numbers (rows a, b, c), or NULL for the gas phase (System::SetPBC, system.cpp:712). */
int mbxg_set_box(mbxg_ctx* ctx, const double* box9);

/* 2B / 3B cutoffs (System::Set2bCutoff / Set3bCutoff, system.h:826,834).  0 disables a term. */
int mbxg_set_cutoffs(mbxg_ctx* ctx, double cutoff2b, double cutoff3b);

/* Evaluated families: bit MBXG_FAM_* set = evaluate those n-mers (default: all).  A cleared bit
 * is the System's ignore_2b_poly / ignore_3b_poly entry for that family (bblock/system.cpp:
 * 2606-2619, 2940-2953): the n-mers stay in the lists but are not evaluated. */
enum {
    MBXG_FAM_MBPOL_2B = 1,      /* h2o-h2o, MB-pol */
    MBXG_FAM_CH4H2O_2B = 2,     /* ch4-h2o, MB-nrg A1B2Z2_C1D4 */
    MBXG_FAM_CH4CH4_2B = 4,     /* ch4-ch4, MB-nrg A1B4_A1B4 */
    MBXG_FAM_MBPOL_3B = 8,      /* h2o x3, MB-pol */
    MBXG_FAM_CH4H2O2_3B = 16,   /* ch4-h2o-h2o, MB-nrg A1B4_C1D2_C1D2 */
    /* halide-water 2B, MB-nrg A1_B1C2X2 deg6 (energy2b.cpp:261-272) */
    MBXG_FAM_F_H2O_2B = 1 << 5, MBXG_FAM_CL_H2O_2B = 1 << 6, MBXG_FAM_BR_H2O_2B = 1 << 7, MBXG_FAM_I_H2O_2B = 1 << 8,
    MBXG_FAM_CO2CO2_2B = 1 << 9, /* co2-co2, x2b_A1B2_A1B2_v1x (energy2b.cpp:293-295) */
    /* halide-water-water 3B, MB-nrg A1_B1C2X2_B1C2X2 deg4 (energy3b.cpp:229-240) */
    MBXG_FAM_F_H2O2_3B = 1 << 10, MBXG_FAM_CL_H2O2_3B = 1 << 11, MBXG_FAM_BR_H2O2_3B = 1 << 12,
    MBXG_FAM_I_H2O2_3B = 1 << 13,
    MBXG_FAM_CO2X3_3B = 1 << 14, /* co2 x3, MB-nrg A1B2_A1B2_A1B2 deg4 (energy3b.cpp:247-249) */
    MBXG_FAM_H2OX4_4B = 1 << 15, /* h2o x4, MB-nrg A1B2_A1B2_A1B2_A1B2 deg4 (energy4b.cpp:218-221) */
    /* alkali-metal-ion-water 2B, x2b_h2o_ion_v2x (energy2b.cpp:258-282) */
    MBXG_FAM_LI_H2O_2B = 1 << 16, MBXG_FAM_NA_H2O_2B = 1 << 17, MBXG_FAM_K_H2O_2B = 1 << 18,
    MBXG_FAM_RB_H2O_2B = 1 << 19, MBXG_FAM_CS_H2O_2B = 1 << 20,
    /* alkali-metal-ion-water-water 3B, x3b_h2o_ion_v1x_deg4_filtered (energy3b.cpp:223-228) */
    MBXG_FAM_LI_H2O2_3B = 1 << 21, MBXG_FAM_NA_H2O2_3B = 1 << 22, MBXG_FAM_K_H2O2_3B = 1 << 23,
    MBXG_FAM_RB_H2O2_3B = 1 << 24, MBXG_FAM_CS_H2O2_3B = 1 << 25
};
int mbxg_set_families(mbxg_ctx* ctx, int mask);

/* Local/ghost decomposition (the LAMMPS path): with on = 1 the neighbour search holds every
 * monomer, ghosts (islocal = 0 in mbxg_set_topology) included, and n-mers are still found only
 * from local monomers, so local-ghost n-mers are evaluated once and ghost-only ones never; with
 * on = 0 (default) the ghosts are left out of the search.  Replaces the use_ghost argument of
 * System::TwoBodyEnergy / ThreeBodyEnergy / GetPairList (bblock/system.h:1088,1100,246), which
 * reaches systools::AddClusters (bblock/sys_tools.cpp:472-537). */
int mbxg_set_use_ghost(mbxg_ctx* ctx, int on);

/* The 1B term alone (mbxg_compute(ctx, MBXG_1B, do_grads), then its results): the energy of every
 * local monomer at the current coordinates (after FixMonomerCoordinates), as reached through
 * System::OneBodyEnergy / Get1B (bblock/system.cpp:2140-2158, 2269-2401) and e1b::get_1b_energy
 * (potential/1b/energy1b.cpp:118-147): water by the MB-pol Partridge-Schwenke potential
 * ps::pot_nasa (potential/1b/ps.cpp:346), CH4 by the MB-nrg x1b_A1B4_v1x holder
 * (potential/1b/x1b_A1B4_deg5_exp0_v1x.cpp:371), CO2 by x1b_A1B2_v1x
 * (potential/1b/x1b_A1B2_deg4_v1x.cpp:160); halide ions have no 1B term (get_1b_energy returns 0
 * for them, energy1b.cpp:175-177).  e1b receives the total (kcal/mol); with do_grads, grad (3
 * doubles per atom, nullable) is overwritten with the 1B gradient and virial (9, nullable) with
 * the 1B virial.  Synchronous.  Monomers without a holder (MBXG_OTHER): MBXG_EINVAL. */
int mbxg_eval_1b(mbxg_ctx* ctx, int do_grads, double* e1b, double* grad, double* virial);

/* Energies of the last compute, per term (each pointer nullable; a term not computed reads 0).
 * The 1B total is a fixed-order sum over the local monomers. */
int mbxg_get_energies(mbxg_ctx* ctx, double* e1b, double* e2b, double* e3b);

/* The 4B term (System::FourBodyEnergy / Get4B, bblock/system.cpp:2747-2767, 3075-3198): the
 * cutoff of the tetramer search (System::Set4bCutoff, system.h:842; 0, the default, disables the
 * term).  Tetramers come from AddClusters with nmax = 4 (bblock/sys_tools.cpp:594-634, the
 * center-3-neighbour rule) and are evaluated by e4b::get_4b_energy (potential/4b/energy4b.cpp:120;
 * water x4 only).  A compute with MBXG_4B, or over a topology holding co2 or ion monomers,
 * runs synchronously (eager, no graph replay). */
int mbxg_set_cutoff4b(mbxg_ctx* ctx, double cutoff4b);
/* e[0..3] = the 1B, 2B, 3B and 4B energies of the last compute (2B / 3B include every family). */
int mbxg_get_term_energies(mbxg_ctx* ctx, double* e);
/* counts[0..20]: evaluated n-mers of the further families, in the order of their MBXG_FAM_* bits
 * from MBXG_FAM_F_H2O_2B to MBXG_FAM_CS_H2O2_3B; counts[21]: listed tetramers. */
int mbxg_get_xcounts(mbxg_ctx* ctx, int64_t* counts);
/* The last compute's listed tetramers (4 internal monomer indices each, query monomer first), at
 * most cap of them; *n receives the number listed. */
int mbxg_get_tetramers(mbxg_ctx* ctx, int32_t* tetramers, int64_t cap, int64_t* n);

/* FixMonomerCoordinates (bblock/sys_tools.cpp:257-320) of host coordinates (3*natoms doubles,
 * internal order) under the current box, into xyz_out (may equal xyz_in): each monomer's first
 * atom wrapped into the cell (fractional x - floor(x + 0.5)), the others to their images nearest
 * to it; a copy in the gas phase.  The same kernel and arithmetic as the wrap of every compute,
 * on its own buffers (the last compute's results are untouched).  Synchronous.  Used by
 * System::SetPBC / GetRealXyz, which wrap the stored coordinates as the reference does. */
int mbxg_fix_coordinates(mbxg_ctx* ctx, const double* xyz_in, double* xyz_out);

/* Batches: split the monomers (in topology order) into nconf independent configurations,
 * configuration c holding monomers [mon_offsets[c], mon_offsets[c+1]) (mon_offsets[0] = 0,
 * mon_offsets[nconf] = nmon).  Lists never cross configurations; one compute evaluates all of
 * them in the same launches.  Configurations share the box (or are all gas phase).  nconf = 1
 * (mon_offsets may be NULL) is the single-system default that mbxg_set_topology restores.
 * Replaces the reference's per-configuration System calls (system.cpp:1995-2138) for many
 * configurations or path-integral beads at once. */
int mbxg_set_batch(mbxg_ctx* ctx, int nconf, const int* mon_offsets);

/* Replace the linear coefficients of a polynomial ("mbpol_2b": 1153, "mbpol_3b": 1163,
 * "ch4h2o_2b": 603, "ch4ch4_2b": 1345, "ch4h2o2_3b": 1254, "ch4_1b": 207, "co2_1b": 21,
 * "f_h2o_2b" / "cl_h2o_2b" /
 * "br_h2o_2b" / "i_h2o_2b": 906, "co2co2_2b": 2269, "f_h2o2_3b" / "cl_h2o2_3b" / "br_h2o2_3b" /
 * "i_h2o2_3b": 3173, "co2x3_3b": 1067, "h2ox4_4b": 398, "li_h2o_2b" / "na_h2o_2b" / "k_h2o_2b" /
 * "rb_h2o_2b" / "cs_h2o_2b": 429, "li_h2o2_3b" / "na_h2o2_3b" / "k_h2o2_3b" / "rb_h2o2_3b" /
 * "cs_h2o2_3b": 1016).  a == NULL restores the fitted values. */
int mbxg_set_coefficients(mbxg_ctx* ctx, const char* model, const double* a, int n);

/* Coordinates of all atoms (3*natoms doubles, internal order) from host or device memory.
 * mbxg_compute wraps them into its working copy under PBC the way System::SetPBC wraps xyz_
 * (sys_tools.cpp:257-320); the input copy itself is kept unchanged. */
int mbxg_set_coords(mbxg_ctx* ctx, const double* xyz);
int mbxg_set_coords_device(mbxg_ctx* ctx, const double* d_xyz);

/* Run lists + evaluation for `terms` (MBXG_1B | MBXG_2B | MBXG_3B) on the context stream.  The 1B
 * term (see mbxg_eval_1b) runs first and its gradient and virial join the others'; it is not
 * evaluated over batches (mbxg_set_batch with more than one configuration: MBXG_EINVAL).  Results stay
 * on the device until mbxg_get_results / the device accessors.  The first compute after a new
 * topology, box or cutoffs synchronises once to size the list buffers (with headroom); later
 * computes enqueue everything without waiting.  A compute whose lists outgrow the buffers is
 * detected by the next getter or mbxg_synchronize and recomputed there; if the inputs were
 * changed in between, that call fails with MBXG_ESTATE and the caller computes again. */
int mbxg_compute(mbxg_ctx* ctx, int terms, int do_grads);

/* Wait for the context stream and copy results: 2B and 3B energies, gradient of every computed
 * term (3*natoms, internal order; may be NULL), virial of every computed term (9; may be NULL;
 * the reference's ordering with [3]=[1], [6]=[2], [7]=[5]).  mbxg_get_energies gives the 1B
 * energy. */
int mbxg_get_results(mbxg_ctx* ctx, double* e2b, double* e3b, double* grad, double* virial);

/* Per-configuration results of the last compute: e2b[nconf], e3b[nconf], virial[9*nconf]
 * (each may be NULL).  The gradient of every atom is in mbxg_get_results' array. */
int mbxg_get_batch_results(mbxg_ctx* ctx, double* e2b, double* e3b, double* virial);

/* counts[0..3] = listed dimers, evaluated dimers, listed trimers, evaluated trimers. */
int mbxg_get_counts(mbxg_ctx* ctx, int64_t* counts);

/* Listed n-mers of the last compute (monomer indices, emission order), up to cap n-mers each;
 * either pointer may be NULL. */
int mbxg_get_lists(mbxg_ctx* ctx, int32_t* dimers, int64_t cap2, int32_t* trimers, int64_t cap3);

/* Evaluated n-mers (order 2 or 3) and their energies, up to cap n-mers. */
int mbxg_get_nmer_energies(mbxg_ctx* ctx, int order, int32_t* nmers, double* energies, int64_t cap);

/* Working (wrapped) coordinates, 3*natoms doubles. */
int mbxg_get_coords(mbxg_ctx* ctx, double* xyz);

/* Device pointers into the context (valid until the next compute that grows buffers):
 * gradient (3*natoms) and totals (21 doubles: e2b and its 6 virial components xx xy xz yy yz zz,
 * then e3b and its 6, then e1b and its 6). */
int mbxg_device_outputs(mbxg_ctx* ctx, double** d_grad, double** d_totals);

/* The context's HIP stream (hipStream_t), for callers that enqueue around the engine. */
int mbxg_stream(mbxg_ctx* ctx, void** stream);

/* Stage timing with HIP events on the context stream.  After a compute with timing on,
 * mbxg_get_timings writes up to 11 durations in ms: FixMonomerCoordinates, list count pass (grid,
 * cloud, binning, counts, scan), list emit pass, 2B variables, 2B polynomial, 2B chain rule, 3B variables, 3B polynomial, 3B chain rule, total,
 * CH4-H2O-H2O family (variables, polynomial and chain rule together).  A stage that did not run
 * reads about 0 (two back-to-back events). */
int mbxg_set_timing(mbxg_ctx* ctx, int on);
int mbxg_get_timings(mbxg_ctx* ctx, double* ms, int n);

/* HIP graph mode (off by default).  When on, a compute in steady state (buffers already sized by
 * an earlier compute of the same terms, stage timing off) replays a captured HIP graph of the
 * whole pipeline instead of launching its ~25 kernels one by one; the graph is captured on first
 * use and recaptured after anything that changes sizes, box, cutoffs or schedule.  Results are
 * read and checked exactly as in eager mode.  No counterpart in the reference. */
int mbxg_set_graph(mbxg_ctx* ctx, int on);

/* Schedule parameters.  "split_2b" / "split_3b": wavefronts (1, 2, 4 or 8) that share each
 * block of 64 dimers / trimers in the MB-pol polynomial stage, each evaluating 1/split of the
 * polynomial (defaults 4 and 2); "split_ch4h2o2_3b": the same for the CH4-H2O-H2O polynomial
 * (default 4); "fuse_poly" (0 or 1, default 1): the MB-pol 2B and 3B polynomials in one launch
 * when both run with gradients; "fuse_prep" (0 or 1, default 1): FixMonomerCoordinates, the
 * cloud and the cell counts in one launch; "fuse_chain" (0 or 1, default 1): with "fuse_poly", the
 * MB-pol 3B chain rule runs in the paired polynomial launch (each 3B block's tail) instead of a
 * kernel of its own; "fuse_vars" (0 or 1, default 1): with "fuse_poly", the 2B and 3B variables in
 * one launch.  mbxg_autotune times every split of both MB-pol orders
 * on the current inputs
 * (the polynomial stage, `reps` timed repetitions after one warm-up) and, when CH4-H2O-H2O trimers
 * are evaluated, every split of that family (its whole evaluation), keeps the fastest of each,
 * then, on the paired path, every combination of "fuse_chain" and "fuse_vars" (the whole compute),
 * and recomputes once with them.  No counterpart in the reference (its kernels have no
 * launch shape). */
int mbxg_set_tuning(mbxg_ctx* ctx, const char* key, int value);
int mbxg_get_tuning(mbxg_ctx* ctx, const char* key, int* value);
int mbxg_autotune(mbxg_ctx* ctx, int terms, int do_grads, int reps);

/* Block until the context's stream is idle and check that the last compute's n-mer lists fitted
 * their buffers (recomputing it with larger buffers if not).  Call before reading the device
 * pointers of mbxg_device_outputs; the host-side getters do this themselves. */
int mbxg_synchronize(mbxg_ctx* ctx);

/* ---- Differentiable evaluation of independent n-mers ---------------------------------------
 * The reference has no differentiable interface: x2b_v9x::eval (potential/2b/x2b-v9x.cpp:1514)
 * and x3b_v2x::eval (potential/3b/x3b-v2x.cpp:1403) fix the coefficients and nonlinear
 * parameters at compile time and return E and dE/dx.  These entry points take both as inputs and
 * back the torch.autograd.Function of the Python plugin (mbx_amd/autograd.py).
 *
 * Family: MBXG_AD_MBPOL_2B (water dimers, 6 atoms = O H H O H H, 1153 coefficients) or
 * MBXG_AD_MBPOL_3B (water trimers, 9 atoms, 1163 coefficients); 10 nonlinear parameters each
 * (2B: k_HH_intra k_OH_intra k_HH_coul k_OH_coul k_OO_coul k_XH_main k_XO_main k_XX_main
 * in_plane_gamma out_of_plane_gamma; 3B: kHH_intra kOH_intra kHH kOH kOO dHH_intra dOH_intra
 * dHH dOH dOO).  Coordinates are n-mer-local (no periodic imaging), row-major [n][3*atoms].
 * All array arguments of forward/backward are device pointers; work is enqueued on `stream`
 * (a hipStream_t; NULL = the default stream) and nothing is allocated: the caller passes a
 * device workspace of at least mbxg_ad_workspace_bytes(family, n) bytes. */
enum { MBXG_AD_MBPOL_2B = 0, MBXG_AD_MBPOL_3B = 1 };
int mbxg_ad_shape(int family, int* n_atoms, int* n_coef, int* n_theta);
/* Host arrays: the reference's fitted coefficients and nonlinear parameters. */
int mbxg_ad_defaults(int family, double* coef, double* theta);
size_t mbxg_ad_workspace_bytes(int family, int n);
/* energy[n] = E_n and grad[n][3*atoms] = dE_n/dx_n. */
int mbxg_ad_forward(int family, int n, const double* xyz, const double* coef, const double* theta, double* energy,
                    double* grad, void* workspace, size_t workspace_bytes, void* stream);
/* The forward without allocation: mbxg_ad_forward with the stage buffers in the library's own
 * device scratch (about 4.6 MiB per device, held until the process ends) instead of a caller's
 * workspace, for at most 8192 n-mers per call (more is an MBXG_EINVAL error).  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc/hipFree.
 * The scratch is looked up on the device current at the first call there; pass a stream of the
 * current device.  Every call on one device shares that device's scratch: calls that overlap on one
 * device (other streams, other threads) race and must be serialised by the caller.  Its contents
 * do not feed mbxg_ad_backward (call that with from_forward = 0). */
int mbxg_ad_forward_noalloc(int family, int n, const double* xyz, const double* coef, const double* theta,
                            double* energy, double* grad, void* stream);
/* With phi = sum_n g_energy[n] E_n + g_grad[n] . dE_n/dx_n: d_xyz = dphi/dx ([n][3*atoms]),
 * d_coef = dphi/da, d_theta = dphi/dtheta.  g_grad may be NULL (phi = sum g_energy E).
 * from_forward = 1: `workspace` is the one mbxg_ad_forward used on these same inputs (and
 * nothing has written it since), so the polynomial's value and gradient are taken from it
 * instead of being evaluated again; 0: any workspace. */
int mbxg_ad_backward(int family, int n, const double* xyz, const double* coef, const double* theta,
                     const double* g_energy, const double* g_grad, double* d_xyz, double* d_coef, double* d_theta,
                     void* workspace, size_t workspace_bytes, int from_forward, void* stream);

#ifdef __cplusplus
}
#endif
#endif
