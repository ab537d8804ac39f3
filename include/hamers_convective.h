/*
 * hamers_convective.h -- C-ABI of the MI355X convective-flux RHS library
 * (libhamers_amd.so).
 *
 * Drop-in boundary for HAMeRS' convective-flux hot path.  Every entry point
 * takes plain pointers and sizes; device pointers are HIP device addresses.
 * All calls return an int status (HM_OK = 0) and never abort; the message of
 * the last failure on a context is available from hm_last_error().
 *
 * Array layouts are SAMRAI's (x fastest, component-major):
 *   conservative variables: one array per equation over the ghost box
 *       idx = (i+g0) + (j+g1)*(n0+2g0) + (k+g2)*(n0+2g0)*(n1+2g1)
 *   convective flux (SideData, 0 ghosts), one array per axis a holding num_eqn
 *       components of (n_a+1) * prod_{b!=a} n_b faces each
 *   source (CellData, 0 ghosts): num_eqn components of prod n_b cells each.
 */
#ifndef HAMERS_CONVECTIVE_H
#define HAMERS_CONVECTIVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HM_MAX_EQN 12
#define HM_MAX_SPECIES 8

/* Status codes. */
enum {
    HM_OK = 0,
    HM_ERR_INVALID_ARGUMENT = 1,
    HM_ERR_UNSUPPORTED = 2,
    HM_ERR_DEVICE = 3,
    HM_ERR_OUT_OF_MEMORY = 4
};

/* Flow models: include/flow/flow_models/FlowModels.hpp (FLOW_MODEL::TYPE). */
enum {
    HM_MODEL_SINGLE_SPECIES = 0,
    HM_MODEL_FOUR_EQN_CONSERVATIVE = 1,
    HM_MODEL_FIVE_EQN_ALLAIRE = 2
};

/* Convective flux schemes: the deck strings "WCNS5_JS_HLLC_HLL" and
 * "WCNS6_LD_HLLC_HLL" of ConvectiveFluxReconstructorManager.cpp:40,79. */
enum {
    HM_SCHEME_WCNS5_JS_HLLC_HLL = 0,
    HM_SCHEME_WCNS6_LD_HLLC_HLL = 1
};

/*
 * Construction parameters of a reconstructor context.  Replaces the
 * constructor arguments of ConvectiveFluxReconstructorWCNS6_LD_HLLC_HLL /
 * ..._WCNS5_JS_HLLC_HLL (ConvectiveFluxReconstructorManager.cpp:14-153,
 * ConvectiveFluxReconstructorWCNS6-LD-HLLC-HLL.cpp:324-361) and of the flow
 * model's equation-of-state mixing rules (species_gamma, species_R).
 */
typedef struct hm_config {
    int dim;                 /* 1, 2 or 3 */
    int model;               /* HM_MODEL_* */
    int num_species;         /* 1 for HM_MODEL_SINGLE_SPECIES */
    int scheme;              /* HM_SCHEME_* */
    int constant_p;          /* WCNS exponent p (deck default 2) */
    int constant_q;          /* WCNS6-LD exponent q (deck default 4) */
    double constant_C;       /* WCNS6-LD C (deck default 1e9) */
    double constant_alpha_tau; /* WCNS6-LD alpha_tau (deck default 35) */
    double species_gamma[HM_MAX_SPECIES];
    double species_R[HM_MAX_SPECIES];
    int device;              /* HIP device ordinal the context runs on */
} hm_config;

/*
 * One patch as the reconstructor sees it: the SAMRAI hier::Patch box, the
 * CartesianPatchGeometry dx and the patch-data pointers that
 * computeConvectiveFluxAndSourceOnPatch reads and writes.  The patch's box
 * grown by max(ghost, 6) cells per side must hold fewer than 2^29 cells
 * (the kernels address one component of a patch array at a 32-bit byte
 * offset); a larger patch is rejected with HM_ERR_INVALID_ARGUMENT.
 */
typedef struct hm_patch_view {
    int n[3];                     /* interior cells per axis (1 on unused axes) */
    int ghost[3];                 /* ghost width of Q per axis (>= 4 on used axes) */
    double dx[3];                 /* grid spacing */
    const double* Q[HM_MAX_EQN];  /* device: one array per conservative variable over the ghost box;
                                     FIVE_EQN_ALLAIRE also passes the last species' volume fraction
                                     (SAMRAI stores all ns, the system evolves ns-1), i.e. num_eqn+1 */
    double* F[3];                 /* device: flux SideData array per axis (overwritten) */
    double* S;                    /* device: source CellData (accumulated, ADVECTIVE eqs) */
} hm_patch_view;

typedef struct hm_ctx hm_ctx;

/* Create / destroy a reconstructor context (owns device scratch and the
 * launch plan cache).  Replaces ConvectiveFluxReconstructorManager
 * construction (ConvectiveFluxReconstructorManager.cpp:14-153). */
int hm_ctx_create(const hm_config* cfg, hm_ctx** out);
int hm_ctx_destroy(hm_ctx* ctx);

/* ConvectiveFluxReconstructor::getConvectiveFluxNumberOfGhostCells
 * (ConvectiveFluxReconstructor.hpp:50-54; 4 for WCNS56, WCNS56-HLLC-HLL.cpp:22). */
int hm_get_num_ghosts(const hm_ctx* ctx, int* num_ghosts);

/* Number of equations of the configured flow model (FlowModel::getNumberOfEquations). */
int hm_get_num_eqn(const hm_ctx* ctx, int* num_eqn);

/* Number of conservative arrays a patch view must provide: num_eqn, plus one
 * (the last volume fraction) for HM_MODEL_FIVE_EQN_ALLAIRE
 * (FlowModelFiveEqnAllaire::getCellDataOfConservativeVariables). */
int hm_get_num_conservative_arrays(const hm_ctx* ctx, int* num_arrays);

/* Message describing the last failure on this context ("" if none). */
const char* hm_last_error(const hm_ctx* ctx);

/*
 * ConvectiveFluxReconstructor::computeConvectiveFluxAndSourceOnPatch
 * (ConvectiveFluxReconstructor.hpp:73-81, implemented for WCNS56 at
 * ConvectiveFluxReconstructorWCNS56-HLLC-HLL.cpp:39-2656): per-patch
 * ("plain") path.  F is overwritten with dt-scaled face fluxes; S is
 * accumulated (+=) for ADVECTIVE equations only.  `time` and `rk_step` are
 * accepted for signature parity and unused, as in the reference.
 * Asynchronous on `stream` (a hipStream_t, NULL = default stream).
 */
int hm_compute_convective_flux_and_source_on_patch(hm_ctx* ctx, const hm_patch_view* patch, double time,
                                                    double dt, int rk_step, void* stream);

/*
 * Allocation-free variant of hm_compute_convective_flux_and_source_on_patch
 * (no counterpart in the reference): for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same arguments, semantics and results (bit for bit); the
 * pipeline's scratch is a static device array of HM_NOALLOC_SCRATCH_DOUBLES
 * doubles (64 MiB) per device, set up by hm_ctx_create and held until the
 * process ends, instead of the context's arena.  A patch whose scratch does
 * not fit (a 2D 128^2 single-species patch needs 0.2 Mi doubles, a 3D 64^3
 * five-equation patch 6.8 Mi) is rejected with HM_ERR_INVALID_ARGUMENT.
 * Every call on a device shares that device's array: calls that overlap on
 * one device (on different streams, or from several threads or contexts)
 * race; order them on one stream or synchronize between them.
 */
#define HM_NOALLOC_SCRATCH_DOUBLES (8u << 20)
int hm_compute_convective_flux_and_source_on_patch_noalloc(hm_ctx* ctx, const hm_patch_view* patch, double time,
                                                            double dt, int rk_step, void* stream);

/*
 * Batched form of hm_compute_convective_flux_and_source_on_patch for all
 * patches of a level (uneven sizes
 * allowed): one work-list driven launch per kernel stage instead of one
 * launch sequence per patch.  Results are identical to calling the per-patch
 * entry point on every patch.  Replaces the serial patch loop of
 * RungeKuttaLevelIntegrator::advanceLevel (RungeKuttaLevelIntegrator.cpp:1715-1742)
 * around the reconstructor call.
 */
int hm_compute_convective_flux_and_source_on_patches(hm_ctx* ctx, const hm_patch_view* patches, int num_patches,
                                                      double time, double dt, int rk_step, void* stream);

/*
 * Launch schedule of the batched RHS and its autotuner (no counterpart in the
 * reference; the reference runs one patch at a time).  A schedule fixes how
 * a patch list's work is laid out over the GPU, not what is computed: every
 * schedule gives the same results bit for bit.
 *   tile_rows  single-species model: midpoint-face rows per tile of the fused
 *              flux along axes 1 and 2 (4, 8 or 24); others: 4 (ignored)
 *   xcd_remap  1: each of the eight XCDs (own L2) takes a contiguous range of
 *              block numbers; 0: the hardware's round-robin order
 *   order      0: work segments in patch order; 1: largest segments first
 *   persistent four- and five-equation models: 1 runs the midpoint flux as
 *              a persistent kernel (as many blocks as the device holds at
 *              once, taking the stage's blocks of work from a queue until it
 *              is empty) followed by a separate face-flux launch; 0: the
 *              midpoint and face flux fused, one block per tile of the patch's
 *              columns.  Single-species: 0
 * Without a pinned schedule a list uses the built-in rule (tile rows by the
 * call's size, xcd_remap 1 for the single-species model, order 0).
 * hm_autotune_on_patches times every candidate schedule on these views
 * (`repetitions` timed RHS calls each after one warm-up, median, HIP events on
 * `stream`), pins the fastest for the list and returns it with its time and
 * the built-in schedule's in ms[0], ms[1] (ms may be NULL).  The trials write
 * the views' F and accumulate into their S: tune on scratch outputs or before
 * the real call.  hm_set_schedule_on_patches pins a schedule;
 * hm_get_schedule_on_patches returns the one a call on the list would use.
 */
typedef struct hm_schedule {
    int tile_rows;
    int xcd_remap;
    int order;
    int persistent;
} hm_schedule;
int hm_autotune_on_patches(hm_ctx* ctx, const hm_patch_view* patches, int num_patches, double dt, int repetitions,
                           void* stream, hm_schedule* chosen, double* ms);
int hm_set_schedule_on_patches(hm_ctx* ctx, const hm_patch_view* patches, int num_patches,
                               const hm_schedule* schedule);
int hm_get_schedule_on_patches(hm_ctx* ctx, const hm_patch_view* patches, int num_patches, hm_schedule* schedule);

/*
 * Stage timing (instrumentation; no counterpart in the reference).  While
 * enabled, every RHS call records HIP events on its stream around each of the
 * pipeline stages: 0 derived cell data, 1 dilatation/vorticity, 2 fused
 * midpoint flux, 3 midpoint-to-face flux, 4 advective source.  hm_profile_read
 * waits for the recorded events and returns the summed milliseconds and launch
 * counts per stage since the last hm_profile_enable.
 */
#define HM_NUM_STAGES 5
int hm_profile_enable(hm_ctx* ctx, int enable);
int hm_profile_read(hm_ctx* ctx, double* stage_ms /* [HM_NUM_STAGES] */,
                    int64_t* stage_launches /* [HM_NUM_STAGES] */);

/*
 * Euler::computeSpectralRadiusesAndStableDtOnPatch (Euler.cpp:489-888) over
 * the interior cells of every given patch (the reference's default,
 * use_ghosts_to_compute_dt = false, reads current data without ghost cells),
 * reduced over the list as RungeKuttaLevelIntegrator::getLevelDt
 * (RungeKuttaLevelIntegrator.cpp:1175-1346) does before its CFL factor.  With
 * lambda_a = |u_a| + c (the model's sound speed): out[a] (a < dim) = max over
 * cells of lambda_a / dx_a; out[dim] = min over patches of the patch's stable
 * dt, 1 / max over cells of sum_a lambda_a / dx_a (for one patch, the vector
 * the reference returns); out[dim + 1] = the level dt before the CFL factor,
 * min(out[dim], 1 / sum_a out[a]).  Only the views' conservative arrays are
 * read (any ghost width); F and S are not used.  Synchronizes `stream`.
 */
int hm_compute_spectral_radiuses_and_stable_dt_on_patches(hm_ctx* ctx, const hm_patch_view* patches,
                                                          int num_patches, double* out, void* stream);

/* Block the host until all work this context queued on `stream` is done. */
int hm_synchronize(hm_ctx* ctx, void* stream);

/* ------------------------------------------------------------------------ */
/* Diffusive flux (four-equation conservative model)                         */
/* ------------------------------------------------------------------------ */

/*
 * Constant transport properties per species: the "constant" mixing rules of
 * the equations of shear viscosity, bulk viscosity, thermal conductivity and
 * mass diffusivity (src/util/mixing_rules/equations_of_<property>/constant/), i.e. the
 * deck values species_mu, species_mu_v, species_kappa, species_D and the
 * molecular weights species_M of the mixing weights Y_i / sqrt(M_i).
 */
typedef struct hm_transport {
    double species_mu[HM_MAX_SPECIES];
    double species_mu_v[HM_MAX_SPECIES];
    double species_kappa[HM_MAX_SPECIES];
    double species_D[HM_MAX_SPECIES];
    double species_M[HM_MAX_SPECIES];
} hm_transport;

/*
 * Give a FOUR_EQN_CONSERVATIVE context its diffusive flux reconstructor:
 * DiffusiveFluxReconstructorManager with the deck key "SIXTH_ORDER"
 * (DiffusiveFluxReconstructorManager.cpp:13-73) and the transport mixing rules.
 */
int hm_set_diffusive_transport(hm_ctx* ctx, const hm_transport* transport);

/* DiffusiveFluxReconstructor::getDiffusiveFluxNumberOfGhostCells: 6
 * (DiffusiveFluxReconstructorNodeSixthOrder.cpp:22). */
int hm_get_diffusive_num_ghosts(const hm_ctx* ctx, int* num_ghosts);

/*
 * DiffusiveFluxReconstructor::computeDiffusiveFluxOnPatch
 * (DiffusiveFluxReconstructor.hpp:73-79, node-based SIXTH_ORDER,
 * DiffusiveFluxReconstructorNode.cpp:32-1738): patch->F[a] is the diffusive
 * flux SideData of axis a (num_eqn components, overwritten with dt-scaled face
 * fluxes), patch->Q needs a ghost width >= 6, patch->S is not used.  The stage
 * update subtracts its divergence beside the convective flux's
 * (NavierStokes.cpp:1704-1750).
 */
int hm_compute_diffusive_flux_on_patch(hm_ctx* ctx, const hm_patch_view* patch, double time, double dt,
                                       int rk_step, void* stream);
int hm_compute_diffusive_flux_on_patches(hm_ctx* ctx, const hm_patch_view* patches, int num_patches,
                                         double time, double dt, int rk_step, void* stream);

/* ------------------------------------------------------------------------ */
/* Discrete adjoint (no counterpart in the reference: a new capability)      */
/* ------------------------------------------------------------------------ */

/*
 * One patch of a vector-Jacobian product of the convective RHS: the forward
 * input Q (as in hm_patch_view), the adjoints F_bar of the face fluxes (the
 * layout of F) and S_bar of the source, and the output Q_bar = (dF/dQ)^T F_bar
 * (+ (dS/dQ)^T S_bar) over the ghost box of Q, overwritten (zero outside the
 * 4-cell stencil halo).
 */
typedef struct hm_vjp_patch_view {
    int n[3];
    int ghost[3];
    double dx[3];
    const double* Q[HM_MAX_EQN];
    const double* F_bar[3];
    const double* S_bar;
    double* Q_bar[HM_MAX_EQN];
} hm_vjp_patch_view;

/*
 * Discrete adjoint of hm_compute_convective_flux_and_source_on_patch(es):
 * the transposed Jacobian of the forward map (branches of the scheme -- WENO
 * weight selection, positivity fallback, sensor and Riemann wave selection --
 * fixed at their forward values), for the single-species model
 * (HM_ERR_UNSUPPORTED otherwise).  Deterministic (no atomics).
 */
int hm_convective_flux_vjp_on_patch(hm_ctx* ctx, const hm_vjp_patch_view* patch, double time, double dt,
                                    int rk_step, void* stream);
int hm_convective_flux_vjp_on_patches(hm_ctx* ctx, const hm_vjp_patch_view* patches, int num_patches,
                                      double time, double dt, int rk_step, void* stream);

/* ------------------------------------------------------------------------ */
/* Runge-Kutta stage update                                                  */
/* ------------------------------------------------------------------------ */

#define HM_MAX_RK_COEFFS 5

/*
 * One patch of a stage update, as Euler::advanceSingleStepOnPatch sees it:
 * the scratch-context conservative arrays (interior overwritten) and, for
 * each coefficient n, the intermediate-context arrays.  Flux and source
 * layouts are those of hm_patch_view; F_accumulated / S_accumulated are the
 * scratch-context flux and source receiving gamma-weighted sums (NULL: skip).
 */
typedef struct hm_rk_patch_view {
    int n[3];
    int ghost[3];                 /* ghost width of Q */
    int ghost_intermediate[3];    /* ghost width of Q_intermediate */
    double dx[3];
    double* Q[HM_MAX_EQN];        /* num_eqn (+1 for FIVE_EQN_ALLAIRE: the last volume fraction,
                                     written as 1 - sum of the others) */
    const double* Q_intermediate[HM_MAX_RK_COEFFS][HM_MAX_EQN];  /* num_eqn arrays each */
    const double* F_intermediate[HM_MAX_RK_COEFFS][3];
    const double* S_intermediate[HM_MAX_RK_COEFFS];
    double* F_accumulated[3];
    double* S_accumulated;
    /* Navier-Stokes stage (NavierStokes.cpp:1003-2400): the intermediate
     * diffusive fluxes of each coefficient (NULL: the Euler form) and the
     * scratch diffusive flux receiving the gamma-weighted sums (NULL: skip). */
    const double* F_diffusive_intermediate[HM_MAX_RK_COEFFS][3];
    double* F_diffusive_accumulated[3];
} hm_rk_patch_view;

/*
 * Euler::advanceSingleStepOnPatch (Euler.cpp:1003-1680) with the zero fill of
 * RungeKuttaLevelIntegrator::advanceLevel before it: per interior cell
 *   Q = sum_n alpha[n] Q_int[n] + beta[n] (-sum_a (F_a,R - F_a,L)/dx_a + S_int[n])
 * (terms in the reference's order), the five-equation model's last volume
 * fraction = 1 - sum of the others after every nonzero beta
 * (FlowModelFiveEqnAllaire.cpp:1739-1890), and F_acc += gamma[n] F_int[n],
 * S_acc += gamma[n] S_int[n].  num_coeffs <= HM_MAX_RK_COEFFS; gamma may be NULL.
 * With diffusive fluxes given for a coefficient, its term is the Navier-Stokes
 * one, -sum_a (F_a,R - F_a,L + Fd_a,R - Fd_a,L)/dx_a + S (NavierStokes.cpp:1743-1748,
 * 2085-2092), and Fd_acc += gamma[n] Fd_int[n].
 */
int hm_advance_single_step_on_patch(hm_ctx* ctx, const hm_rk_patch_view* patch, int num_coeffs,
                                    const double* alpha, const double* beta, const double* gamma, void* stream);
int hm_advance_single_step_on_patches(hm_ctx* ctx, const hm_rk_patch_view* patches, int num_patches,
                                      int num_coeffs, const double* alpha, const double* beta,
                                      const double* gamma, void* stream);

/* ------------------------------------------------------------------------ */
/* Level integrator                                                          */
/* ------------------------------------------------------------------------ */

/*
 * One AMR level held on the device: the scratch conservative arrays of every
 * patch (ghost width 4), the intermediate copies, fluxes and sources of each
 * Runge-Kutta stage.  hm_level_advance runs the stage loop of
 * RungeKuttaLevelIntegrator::advanceLevel (RungeKuttaLevelIntegrator.cpp:1457-1939,
 * stages 1672-1745): same-level ghost fill (periodic images included), batched
 * convective RHS (hm_compute_convective_flux_and_source_on_patches) and the
 * stage update, with SSP-RK3 weights (RungeKuttaLevelIntegrator.cpp:3907-3924)
 * unless set otherwise.  Ghost cells no same-level patch covers keep what the
 * caller wrote there, unless they lie outside the domain on a side with a
 * physical boundary condition (hm_level_set_boundary_conditions).  Patch boxes
 * are (lo, n) in level index space; all (x, y, z)-ordered arrays of 3 per
 * patch.
 */
typedef struct hm_level hm_level;
int hm_level_create(hm_ctx* ctx, int num_patches, const int* patch_lo, const int* patch_n,
                    const int* domain_lo, const int* domain_n, const int* periodic, const double* dx,
                    hm_level** out);
int hm_level_destroy(hm_level* level);

/* Butcher-like tables of the level's scheme, row-major [num_stages][num_stages]
 * (entries above the diagonal ignored); gamma may be NULL. */
int hm_level_set_rk_coefficients(hm_level* level, int num_stages, const double* alpha, const double* beta,
                                 const double* gamma);

/* Host copies of one patch's interior (num_conservative_arrays components). */
int hm_level_write_patch(hm_level* level, int patch, const double* Q);
int hm_level_read_patch(hm_level* level, int patch, double* Q);

/* Device address of a patch's scratch arrays (components over the ghost box,
 * component stride = ghost-box volume) and the ghost width. */
int hm_level_patch_array(hm_level* level, int patch, double** Q, int* ghost);

/*
 * Physical boundary conditions of the level's non-periodic sides, applied to
 * every conservative variable after each same-level ghost fill (every Runge-
 * Kutta stage), as EulerBoundaryConditions::setPhysicalBoundaryConditions
 * (EulerBoundaryConditions.cpp:749-1100) applies the basic conditions of
 * BasicCartesianBoundaryUtilities1/2/3 (BasicCartesianBoundaryUtilities2.cpp:215-440
 * edge fills, 590-900 node fills; BasicBoundaryConditions.hpp:14-34):
 *   HM_BC_FLOW      copy of the boundary cell (zero gradient);
 *   HM_BC_REFLECT   mirror image about the boundary face, the momentum
 *                   component normal to it negated (scalar variables mirror);
 *   HM_BC_SYMMETRY  mirror image;
 *   HM_BC_DIRICHLET fixed conservative values;
 *   HM_BC_NONE      ghost cells outside the domain keep the caller's values
 *                   (the default).
 * conditions: 2 dim entries (xlo, xhi, ylo, yhi, zlo, zhi); periodic axes take
 * HM_BC_NONE.  dirichlet: [2 dim][num_conservative_arrays] values (rows of
 * non-DIRICHLET sides ignored), may be NULL when no side is DIRICHLET.  Cells
 * outside the domain along several axes (edges, corners) compose the rules of
 * those axes, as the reference's x-named node and edge conditions (XFLOW,
 * XREFLECT, ...) do; with DIRICHLET among them the first such axis in x, y, z
 * order gives the values.  A level with boundary conditions has no adjoint
 * (hm_level_advance_recorded returns HM_ERR_UNSUPPORTED).
 */
#define HM_BC_NONE 0
#define HM_BC_FLOW 1
#define HM_BC_REFLECT 2
#define HM_BC_SYMMETRY 3
#define HM_BC_DIRICHLET 4
int hm_level_set_boundary_conditions(hm_level* level, const int* conditions, const double* dirichlet);

/* Ghost fill of the scratch arrays: same level (periodic images included),
 * then the physical boundary conditions. */
int hm_level_fill_ghosts(hm_level* level, void* stream);

/* Schedule of the level's batched RHS (hm_schedule above): hm_level_autotune
 * times every candidate on the first stage's RHS of the current state and
 * makes the fastest the schedule of every stage (ms[0], ms[1]: its time and
 * the built-in schedule's; ms may be NULL); hm_level_set_schedule pins one.
 * Both drop the level's captured steps. */
int hm_level_autotune(hm_level* level, double dt, int repetitions, hm_schedule* chosen, double* ms);
int hm_level_set_schedule(hm_level* level, const hm_schedule* schedule);
int hm_level_get_schedule(hm_level* level, hm_schedule* schedule);

/* num_steps steps of size dt.  use_graph != 0 replays one step captured in a
 * hipGraph (captured once per dt) instead of issuing its launches. */
int hm_level_advance(hm_level* level, double dt, int num_steps, int use_graph, void* stream);

/*
 * num_steps steps whose dt is cfl times the level's stable dt, computed on the
 * device from the level state at the start of every step (the reduction of
 * hm_compute_spectral_radiuses_and_stable_dt_on_patches: getLevelDt with
 * cfl = min(cfl_init, cfl), RungeKuttaLevelIntegrator.cpp:1175-1346).  The RHS
 * kernels read that dt from device memory, so the host never waits for it;
 * use_graph != 0 replays one captured step for every dt.
 */
int hm_level_advance_cfl(hm_level* level, double cfl, int num_steps, int use_graph, void* stream);

/* The last hm_level_advance_cfl step's reduction (waits for the device):
 * info[0..2] spectral radiuses per axis, info[3] the patches' dt, info[4] the
 * level dt, info[5] the step's dt (info[4] * cfl), info[6] the sum of the step
 * dts of that hm_level_advance_cfl call. */
#define HM_DT_INFO_SLOTS 8
int hm_level_dt_info(hm_level* level, double* info);

/*
 * Discrete adjoint of a level's Runge-Kutta steps (no counterpart in the
 * reference; single-species model).  hm_level_advance_recorded advances like
 * hm_level_advance and keeps a copy of the level state every checkpoint_every
 * steps.  The caller then writes the adjoint of the final state (interior
 * cells, num_eqn components; e.g. dJ/dQ of an objective J) with
 * hm_level_write_adjoint and calls hm_level_adjoint, which recomputes each
 * checkpoint segment forward and runs the reverse stage loop (the transposed
 * stage update, hm_convective_flux_vjp_on_patches and the transposed ghost
 * fill); hm_level_read_adjoint then returns dJ/dQ of the initial state.  The
 * level state is left as hm_level_advance_recorded left it.
 */
int hm_level_advance_recorded(hm_level* level, double dt, int num_steps, int checkpoint_every, int use_graph,
                              void* stream);
int hm_level_write_adjoint(hm_level* level, int patch, const double* lam);
int hm_level_adjoint(hm_level* level, void* stream);
int hm_level_read_adjoint(hm_level* level, int patch, double* lam);

/*
 * Two-level hierarchy (refinement ratio 2): SAMRAI's time-refined advance as
 * RungeKuttaLevelIntegrator drives it (advanceLevel on the coarse level, twice
 * on the fine level with half the step, then synchronizeLevelWithCoarser,
 * RungeKuttaLevelIntegrator.cpp:2131-2225).  Before each fine step the fine
 * level's coarse-fine ghost cells (inside the domain, under no fine patch) are
 * filled from the coarse level, linearly interpolated in time between its
 * states at the start and the end of the coarse step and refined by
 * conservative linear interpolation (monotonized-central limited slopes, the
 * CONSERVATIVE_LINEAR_REFINE operator of FlowModelSingleSpecies.cpp:578-597);
 * after the two fine steps the coarse faces on the coarse-fine boundary take
 * the area average of the fine flux sums (each fine step's gamma-weighted flux
 * added to a sum that starts at zero, algs_upfluxsum2d.m4,
 * RungeKuttaLevelIntegrator.cpp:3060-3180), the coarse update is repeated from
 * the coarse state at the start of the step with those fluxes
 * (Euler::synchronizeFluxes, Euler.cpp:1682-1943; five-equation model: the
 * last volume fraction then 1 minus the others) and the coarse cells under the
 * fine level take the average of their fine cells (CONSERVATIVE_COARSEN).
 * A context with diffusive fluxes (four-equation Navier-Stokes form) refluxes
 * the diffusive flux the same way, as a flux variable of its own
 * (NavierStokes.cpp:393-402), and repeats the coarse update in the
 * Navier-Stokes form (NavierStokes::synchronizeFluxes, NavierStokes.cpp:2424-2700).
 * Both levels must belong to one context; the fine
 * domain is the coarse domain refined by 2, fine patch boxes start and end on
 * coarse cells, the fine level lies inside the coarse level with room for the
 * refinement stencil, and both levels carry the same periodic axes and
 * physical boundary conditions.  The levels keep working on their own
 * (hm_level_advance); destroy the hierarchy before its levels.
 */
typedef struct hm_hierarchy hm_hierarchy;
int hm_hierarchy_create(hm_level* coarse, hm_level* fine, int ratio, hm_hierarchy** out);
int hm_hierarchy_destroy(hm_hierarchy* hierarchy);
/* num_steps coarse steps of dt_coarse (each with two fine steps of dt_coarse / 2). */
int hm_hierarchy_advance(hm_hierarchy* hierarchy, double dt_coarse, int num_steps, void* stream);

/*
 * Trajectory storage of hm_level_adjoint (no counterpart in the reference).
 * Its recompute pass can keep the reconstructed characteristic face values of
 * every stage of a checkpoint segment (2 num_eqn doubles per midpoint face and
 * stage, checkpoint_every * num_stages stages), so that the reverse pass need
 * not reconstruct them again; the gradient is the same bit for bit either way.
 * hm_level_set_face_state_budget bounds that store in bytes: < 0 half of the
 * device memory free when hm_level_advance_recorded sets it up (the default),
 * 0 never.  hm_level_face_states_kept returns 1 when the last
 * hm_level_advance_recorded set a store up, else 0.
 */
int hm_level_set_face_state_budget(hm_level* level, int64_t bytes);
int hm_level_face_states_kept(hm_level* level);

#ifdef __cplusplus
}
#endif

#endif /* HAMERS_CONVECTIVE_H */
