/*
 * trixi_hip.h -- C ABI of the MI355X DGSEM right-hand side for the compressible
 * Euler equations on a TreeMesh (2D and 3D): uniform or refined (L2 mortars), periodic or with
 * Dirichlet or slip-wall boundaries.
 *
 * The library replaces the CPU methods that Trixi.jl dispatches from
 *   rhs_hyperbolic!(du_ode, u_ode, semi, t)
 *     src/semidiscretization/semidiscretization_hyperbolic.jl:578-597
 *   -> rhs_hyperbolic!(backend, du, u, t, mesh::TreeMesh{2,3}, equations,
 *                      boundary_conditions, source_terms, dg, cache)
 *     src/solvers/dgsem_tree/dg_2d.jl:113-186
 * A Julia method for a marker backend ccalls these entry points; everything
 * above it (ODEProblem, the 2N Runge-Kutta integrators, Stepsize/Analysis
 * callbacks) is unchanged.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative TRIXI_HIP_E* code on
 *    failure; trixi_hip_last_error(ctx) describes the last failure.
 *  - State vectors use the reference layout of u_ode (dg.jl:1169-1212):
 *    column-major [nvars, nnodes^ndims, nelements], Float64 at the boundary.
 *  - Element ids, interface ids and orientations are 0-based here (the Julia
 *    shim subtracts 1).
 *  - One context per host thread; a context owns one HIP stream (or uses the
 *    stream set by trixi_hip_set_stream) and its device memory. Not
 *    re-entrant. The library never calls exit().
 */
#ifndef TRIXI_HIP_H
#define TRIXI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define TRIXI_HIP_OK 0
#define TRIXI_HIP_EINVAL (-1)   /* bad argument / unsupported configuration */
#define TRIXI_HIP_EHIP (-2)     /* a HIP runtime call failed */
#define TRIXI_HIP_ENOMEM (-3)   /* device allocation failed */
#define TRIXI_HIP_ESTATE (-4)   /* call not valid in the context's state */

/* precision of the bulk flux arithmetic */
#define TRIXI_HIP_FP64 0 /* plain FP64 everywhere */
#define TRIXI_HIP_FP32 1 /* FP32 fluxes, FP64 RK registers and reductions */

/* volume integral (src/solvers/dg.jl:135-221, 574-582) */
#define TRIXI_HIP_VI_WEAK 0               /* VolumeIntegralWeakForm */
#define TRIXI_HIP_VI_FLUX_DIFFERENCING 1  /* VolumeIntegralFluxDifferencing(volume_flux) */
#define TRIXI_HIP_VI_SHOCK_CAPTURING_HG 2 /* VolumeIntegralShockCapturingHG(indicator; volume_flux_dg, volume_flux_fv) */
#define TRIXI_HIP_VI_PURE_FV 3            /* VolumeIntegralPureLGLFiniteVolume(volume_flux_fv) */

/* numerical flux (src/equations/numerical_fluxes.jl, compressible_euler_{2,3}d.jl) */
#define TRIXI_HIP_FLUX_LLF 0       /* FluxLaxFriedrichs() = flux_lax_friedrichs (max_abs_speed) */
#define TRIXI_HIP_FLUX_LLF_NAIVE 1 /* FluxLaxFriedrichs(max_abs_speed_naive) */
#define TRIXI_HIP_FLUX_RANOCHA 2   /* flux_ranocha */
#define TRIXI_HIP_FLUX_CENTRAL 3   /* flux_central */
#define TRIXI_HIP_FLUX_HLLC 4      /* flux_hllc (compressible_euler_3d.jl:1417-1541, 2D :1714-1816); surface or FV flux */

/* boundary conditions of a non-periodic TreeMesh (trixi_hip_config.boundary_condition) */
#define TRIXI_HIP_BC_NONE 0      /* fully periodic mesh: no boundaries */
#define TRIXI_HIP_BC_DIRICHLET 1 /* BoundaryConditionDirichlet(initial_condition) on every boundary
                                    (equations.jl:159-182); the boundary value function is the
                                    one set by trixi_hip_set_exact_solution */
#define TRIXI_HIP_BC_SLIP_WALL 2 /* boundary_condition_slip_wall on every boundary (TreeMesh methods of
                                    compressible_euler_{2,3}d.jl:319-420): wall flux from the inner
                                    state (Toro's pressure Riemann solution), any surface flux */

/* source terms */
#define TRIXI_HIP_SRC_NONE 0
#define TRIXI_HIP_SRC_CONVERGENCE_TEST 1 /* source_terms_convergence_test */

/* initial conditions known to the device (for error norms, exact solutions) */
#define TRIXI_HIP_IC_CONSTANT 0
#define TRIXI_HIP_IC_CONVERGENCE_TEST 1
#define TRIXI_HIP_IC_WEAK_BLAST_WAVE 2
#define TRIXI_HIP_IC_ISENTROPIC_VORTEX 3
#define TRIXI_HIP_IC_KELVIN_HELMHOLTZ 4
#define TRIXI_HIP_IC_FOURIER_MODES 5
#define TRIXI_HIP_IC_DENSITY_PULSE 6 /* elixir_euler_amr.jl:15-26 */
#define TRIXI_HIP_IC_BLAST_WAVE 7    /* 2D, elixir_euler_blast_wave_amr.jl:14-33 */
#define TRIXI_HIP_IC_BLOB 8          /* 2D, elixir_euler_blob_amr.jl:13-58 */

/* indicator variables of IndicatorHennemannGassner (config hg_variable,
 * trixi_hip_indicator_hg) */
#define TRIXI_HIP_VAR_DENSITY_PRESSURE 0 /* density_pressure, euler_2d.jl:2205-2210 */
#define TRIXI_HIP_VAR_DENSITY 1          /* Trixi.density */
#define TRIXI_HIP_VAR_PRESSURE 2         /* Trixi.pressure (euler_2d.jl) */

/* indicator modes for trixi_hip_indicator / trixi_hip_set_indicator */
#define TRIXI_HIP_IND_HG 0                /* IndicatorHennemannGassner */
#define TRIXI_HIP_IND_MLP 1               /* learned indicator (2D p3), network output only */
#define TRIXI_HIP_IND_MLP_WITH_FALLBACK 2 /* learned indicator; HG where an input leaves the trained range */

/* learned indicator: 19 -> 64 -> 64 -> 1 perceptron, FP32 parameters in the
 * order W1[64][19], b1[64], W2[64][64], b2[64], w3[64], b3 (row-major, rows =
 * output units); features [E][32] FP32, the first 19 used */
#define TRIXI_HIP_MLP_NPARAMS 5505
#define TRIXI_HIP_MLP_NFEATURES 19
#define TRIXI_HIP_MLP_FEATURE_STRIDE 32

/* where the pointers passed to trixi_hip_rhs live */
#define TRIXI_HIP_HOST 0
#define TRIXI_HIP_DEVICE 1

typedef struct trixi_hip_config {
  int32_t ndims;           /* 2 or 3 */
  int32_t polydeg;         /* N, nnodes = N + 1 (supported: 2..4) */
  int32_t nelements;       /* E */
  int32_t ninterfaces;     /* I */
  double gamma;            /* CompressibleEulerEquations{2,3}D(gamma) */
  int32_t precision;       /* TRIXI_HIP_FP64 | TRIXI_HIP_FP32 */
  int32_t volume_integral; /* TRIXI_HIP_VI_* */
  int32_t volume_flux;     /* TRIXI_HIP_FLUX_RANOCHA | TRIXI_HIP_FLUX_CENTRAL (FD and HG high order part) */
  int32_t surface_flux;    /* TRIXI_HIP_FLUX_* used at interfaces */
  int32_t fv_flux;         /* TRIXI_HIP_FLUX_* used for the FV subcells (HG, pure FV) */
  double alpha_max;        /* IndicatorHennemannGassner(alpha_max, alpha_min, alpha_smooth) */
  double alpha_min;
  int32_t alpha_smooth;
  int32_t source_terms;    /* TRIXI_HIP_SRC_* */
  int32_t elements_per_block; /* 0 = default tile of each element-kernel variant (see
                                 trixi_hip_autotune); else the workgroup tile (elements) */
  int32_t threads_per_block;  /* 0 = default; else threads per workgroup of that tile */
  int32_t nboundaries;        /* B: boundary faces of a non-periodic mesh (0 if periodic) */
  int32_t boundary_condition; /* TRIXI_HIP_BC_*; must be set when nboundaries > 0 */
  int32_t nmortars;           /* M: L2 mortars of a refined (non-conforming) mesh, 0 if uniform */
  int32_t hg_variable;        /* TRIXI_HIP_VAR_*: variable of the shock-capturing indicator */
} trixi_hip_config;

/* Mesh data as held in the reference cache (containers_{2,3}d.jl). */
typedef struct trixi_hip_mesh {
  const int32_t *neighbor_ids;     /* [2, I] interfaces.neighbor_ids (left, right), 0-based */
  const int32_t *orientations;     /* [I] interfaces.orientations, 0-based (x=0, y=1, z=2) */
  const double *inverse_jacobian;  /* [E] elements.inverse_jacobian */
  const double *element_centers;   /* [ndims, E] tree.coordinates[:, elements.cell_ids] */
  /* boundaries container (containers_2d.jl:475-543, containers_3d.jl), B entries;
   * every element face is either in an interface or a boundary */
  const int32_t *boundary_neighbor_ids;   /* [B] boundaries.neighbor_ids, 0-based element */
  const int32_t *boundary_orientations;   /* [B] boundaries.orientations, 0-based */
  const int32_t *boundary_neighbor_sides; /* [B] boundaries.neighbor_sides: 1 = element left of
                                             the boundary (its + face), 2 = right (its - face) */
  /* L2 mortars container (containers_3d.jl:652-788, containers_2d.jl:700-800), M
   * entries; neighbor_ids[k + (S + 1) m], S = 2^(ndims-1) small elements (3D: lower
   * left, lower right, upper left, upper right = first free dimension -/+, second
   * free dimension -/+; 2D: lower, upper) then the large element, 0-based */
  const int32_t *mortar_neighbor_ids;  /* [S + 1, M] */
  const int32_t *mortar_large_sides;   /* [M] 1 = large element left of the mortar, 2 = right */
  const int32_t *mortar_orientations;  /* [M] 0-based orientation */
} trixi_hip_mesh;

/* LobattoLegendreBasis fields (basis_lobatto_legendre.jl:17-86), row-major
 * matrices M[r*n + c] = M[r, c] (r, c 0-based). The analysis basis is the
 * SolutionAnalyzer (basis_lobatto_legendre.jl:274-290). */
typedef struct trixi_hip_basis {
  const double *nodes;                        /* [n] */
  const double *weights;                      /* [n] */
  const double *derivative_split;             /* [n, n] Dsplit */
  const double *derivative_hat;               /* [n, n] Dhat */
  const double *inverse_vandermonde_legendre; /* [n, n] */
  int32_t analysis_nnodes;                    /* na (2N + 1) */
  const double *analysis_weights;             /* [na] */
  const double *analysis_vandermonde;         /* [na, n] interpolation matrix */
  /* LobattoLegendreMortarL2 (basis_lobatto_legendre.jl:176-205, l2projection.jl),
   * [n, n] each; may be NULL when the mesh has no mortars */
  const double *mortar_forward_upper;
  const double *mortar_forward_lower;
  const double *mortar_reverse_upper;
  const double *mortar_reverse_lower;
} trixi_hip_basis;

typedef struct trixi_hip_ctx trixi_hip_ctx;

/* Build a context: copies mesh and basis to the device and allocates the
 * resident state. Replaces create_cache(mesh::TreeMesh, ...) (dg_2d.jl:14-37)
 * for the device. E and I are fixed for the context's lifetime. */
int trixi_hip_create(const trixi_hip_config *cfg, const trixi_hip_mesh *mesh,
                     const trixi_hip_basis *basis, trixi_hip_ctx **out);
int trixi_hip_destroy(trixi_hip_ctx *ctx);
const char *trixi_hip_last_error(const trixi_hip_ctx *ctx);

/* Use an external HIP stream (e.g. torch.cuda.current_stream().cuda_stream);
 * NULL restores the context's own stream. */
int trixi_hip_set_stream(trixi_hip_ctx *ctx, void *hip_stream);
int trixi_hip_synchronize(trixi_hip_ctx *ctx);

/* Resident state u (FP64 RK register; in FP32 mode also its FP32 shadow). */
int trixi_hip_upload_u(trixi_hip_ctx *ctx, const double *u_host);
int trixi_hip_download_u(trixi_hip_ctx *ctx, double *u_host);
int trixi_hip_set_u_device(trixi_hip_ctx *ctx, const double *u_dev);
int trixi_hip_get_u_device(trixi_hip_ctx *ctx, double *u_dev);

/* rhs_hyperbolic!(du, u, t) -- semidiscretization_hyperbolic.jl:578-597 and
 * dg_2d.jl:113-186. FP64 arrays in the reference layout, host or device
 * (where = TRIXI_HIP_HOST | TRIXI_HIP_DEVICE). In FP32 mode u is rounded to
 * FP32, the fluxes run in FP32 and du is widened to FP64. */
int trixi_hip_rhs(trixi_hip_ctx *ctx, double *du, const double *u, double t, int where);

/* The same operator on device arrays in the context's native precision
 * (float in FP32 mode, double in FP64 mode); no conversion, no allocation. */
int trixi_hip_rhs_native(trixi_hip_ctx *ctx, void *du_dev, const void *u_dev, double t);

/* One 2N low-storage stage on the resident state (methods_2N.jl:146-159):
 *   du = rhs(u, t_stage); u_tmp = du - a_stage*u_tmp; u += b_stage_dt*u_tmp
 * u is FP64; u_tmp is stored in the RHS precision (FP32 in FP32 mode) and updated in FP64. first_stage != 0 zeroes u_tmp first (methods_2N.jl:143). */
int trixi_hip_rk_stage(trixi_hip_ctx *ctx, double a_stage, double b_stage_dt,
                       double t_stage, int first_stage);

/* One full step of a 2N scheme on the resident state (step!, methods_2N.jl:131-168),
 * replayed from a captured hipGraph when use_graph != 0. */
int trixi_hip_rk_step(trixi_hip_ctx *ctx, const double *a, const double *b, const double *c,
                      int nstages, double t, double dt, int use_graph);

/* A whole run of a 2N scheme with StepsizeCallback(cfl) semantics on the resident
 * state, without host synchronisation per step: solve!/step! (time_integration.jl:102-123,
 * methods_2N.jl:131-168), the step size callback (stepsize.jl:93-155) and
 * limit_dt! (time_integration.jl:46-55). After every step dt = cfl * max_dt(u);
 * the step that would pass t_end (or land within sqrt(eps) of it) is shortened to
 * end there. One captured graph (stages + step control) is replayed; the host
 * reads the device control block every poll_every steps, so up to poll_every - 1
 * replays run after the end and leave the state unchanged. t_out and steps_out
 * receive the final time and the number of steps taken. Returns TRIXI_HIP_ESTATE
 * with "time step size `dt` is NaN" when dt becomes NaN (methods_2N.jl:138-140). */
int trixi_hip_solve(trixi_hip_ctx *ctx, const double *a, const double *b, const double *c,
                    int nstages, double t0, double t_end, double cfl, int max_steps,
                    int poll_every, double *t_out, int *steps_out);

/* max_dt (stepsize_dg{2,3}d.jl:8-32) of the resident state times cfl
 * (calculate_dt, stepsize.jl:146-155). FP64 reduction. */
int trixi_hip_max_dt(trixi_hip_ctx *ctx, double cfl, double *dt);

/* No counterpart in the reference. trixi_hip_max_dt for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops):
 * no context, no hipMalloc/hipFree, no host synchronisation. Inputs are the
 * FP64 state u_dev [V, (polydeg+1)^ndims, nelements] and inverse_jacobian_dev
 * [nelements] in device memory; cfl * max_dt is written to dt_dev[0] on
 * hip_stream (NULL: the null stream), enqueued only. The partial maxima go to
 * a fixed device array of 8192 doubles (64 KiB) that each GPU holds until the
 * process ends; larger meshes are reduced with several element groups per
 * workgroup, so every mesh size fits. That array is shared by all calls on one
 * GPU: calls that overlap on one GPU (other streams, other host threads) race.
 * Same kernels and result as trixi_hip_max_dt. Errors: TRIXI_HIP_EINVAL for
 * null pointers, nelements outside [1, 2^31 - 257] or an unsupported
 * (ndims, polydeg); trixi_hip_last_error(NULL) describes the failure. */
int trixi_hip_max_dt_noalloc(int ndims, int polydeg, int nelements, double gamma, const double *u_dev,
                             const double *inverse_jacobian_dev, double cfl, double *dt_dev, void *hip_stream);

/* Blending factors alpha[E] of IndicatorHennemannGassner (indicators.jl:119-156,
 * indicators_{2,3}d.jl) for the resident state, after smoothing. */
int trixi_hip_indicator(trixi_hip_ctx *ctx, int mode, double *alpha_host);

/* IndicatorHennemannGassner(semi; alpha_max, alpha_min, alpha_smooth, variable)
 * (indicators.jl:50-156, indicators_{2,3}d.jl) of the resident FP64 state with
 * its own parameters, as the reference builds a separate indicator for an AMR
 * controller (e.g. elixir_euler_kelvin_helmholtz_instability_amr.jl); smoothing
 * also across mortars. alpha_host[E] after smoothing (if alpha_smooth). */
int trixi_hip_indicator_hg(trixi_hip_ctx *ctx, double alpha_max, double alpha_min, int alpha_smooth, int variable,
                           double *alpha_host);

/* IndicatorMax(semi; variable) (indicators_3d.jl:239-259, indicators_2d.jl):
 * alpha_host[e] = maximum over the nodes of element e of conserved variable
 * `variable` (0 = rho, as `variable = first`) of the resident state. The AMR
 * controller (ControllerThreeLevel, amr.jl:1116-1153) reads it on the host. */
int trixi_hip_indicator_max(trixi_hip_ctx *ctx, int variable, double *alpha_host);

/* Solution transfer of one AMR adaptation (refine!/coarsen! of the DG solver,
 * amr_dg.jl and amr_dg3d.jl:138-230, 337-430, with LobattoLegendreAdaptorL2,
 * basis_lobatto_legendre.jl:334-360). `to` is a context created on the adapted
 * mesh (same ndims and polydeg, same device); its resident state is set from
 * the resident state of `from`, element by element of `to` (host arrays of
 * to's nelements):
 *   kind[e] = 0       copy of element src[e] of `from`;
 *   kind[e] = 1 + q   child q of element src[e] (bit d of q = upper half in
 *                     dimension d: forward_upper, else forward_lower);
 *   kind[e] = -1      L2 projection of the 2^D elements src[e] .. src[e]+2^D-1
 *                     (children in child order: reverse_upper / reverse_lower).
 * The matrices are the mortar matrices `to` was created with. Synchronizes
 * both contexts' streams; `from` may be destroyed afterwards. */
int trixi_hip_transfer(trixi_hip_ctx *to, trixi_hip_ctx *from, const int32_t *kind, const int32_t *src);

/* Learned troubled-cell indicator (2D polydeg 3 contexts). No counterpart in
 * Trixi.jl's source tree (its neural-network indicators were moved to a
 * separate package, NEWS.md:495); it replaces the HG blending factor of
 * VolumeIntegralShockCapturingHG (indicators.jl:119-156) when selected with
 * trixi_hip_set_indicator. The network reads features computed from the HG
 * indicator's own quantities (indicators_2d.jl:30-102), alpha = alpha_max *
 * sigmoid(output), clipped as HG clips its factor (indicators_2d.jl:86-99: below
 * alpha_min 0, above 1 - alpha_min 1, at most alpha_max); smoothing
 * (indicators_2d.jl:104-135) is applied as for HG. */
int trixi_hip_mlp_set_params(trixi_hip_ctx *ctx, const float *params, const float *feature_lo,
                             const float *feature_hi); /* params [NPARAMS]; box [NFEATURES] each; resets Adam */
int trixi_hip_mlp_get_params(trixi_hip_ctx *ctx, float *params);
/* blending-factor source of the shock-capturing RHS (TRIXI_HIP_IND_*) */
int trixi_hip_set_indicator(trixi_hip_ctx *ctx, int mode);
/* features [E][32] and raw HG alpha [E] of the resident state (either may be NULL) */
int trixi_hip_mlp_features(trixi_hip_ctx *ctx, float *features_host, double *alpha_hg_host);
/* y[n] = sigmoid(network(x[n][32])), device arrays */
int trixi_hip_mlp_forward(trixi_hip_ctx *ctx, const float *x_dev, int n, float *y_dev);
/* loss = sum (y - target)^2 / n and its gradient [NPARAMS] (host; either may be NULL) */
int trixi_hip_mlp_loss_grad(trixi_hip_ctx *ctx, const float *x_dev, const float *target_dev, int n, float *loss,
                            float *grad_host);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the loss above; loss may be NULL (no sync) */
int trixi_hip_mlp_train_step(trixi_hip_ctx *ctx, const float *x_dev, const float *target_dev, int n,
                             float learning_rate, float *loss);

/* Exact solution used by calc_error_norms (analysis_dg{2,3}d.jl). params:
 * IC_ISENTROPIC_VORTEX {p_background}; IC_FOURIER_MODES {A_rho, A_v} and
 * modes[4*8*4] = (kx, ky, kz, phase) per field (rho, v1, v2, v3) and mode. */
int trixi_hip_set_exact_solution(trixi_hip_ctx *ctx, int ic_kind, const double *params,
                                 int nparams, const double *modes);

/* compute_coefficients(func, t, semi) into the resident state on the device
 * (semidiscretization.jl:224-242; compute_coefficients! dg.jl:1293-1324, the
 * reference's GPU kernel): u at every node = the exact solution set with
 * trixi_hip_set_exact_solution at node coordinates center + (1 /
 * inverse_jacobian) * node (init_elements!, containers_3d.jl:87-128), FP64,
 * plus the FP32 shadow in FP32 mode. Asynchronous on the context's stream.
 * The device's transcendental functions can differ from the host's libm in
 * the last bit, so the state can differ from a host evaluation at round-off. */
int trixi_hip_compute_coefficients(trixi_hip_ctx *ctx, double t);

/* calc_error_norms(cons2cons, ...) on the resident state (analysis_dg3d.jl:123-161). */
int trixi_hip_error_norms(trixi_hip_ctx *ctx, double t, double *l2, double *linf);

/* integrate_via_indices of the conserved variables and of entropy_math,
 * normalised by the domain volume (analysis_dg3d.jl:364-386). */
int trixi_hip_integrals(trixi_hip_ctx *ctx, double *cons_integrals, double *entropy_integral);

/* AnalysisCallback(semi; interval) inside trixi_hip_solve (analysis.jl:100-236):
 * the solve analyses the initial state, every interval-th step and the final
 * state on the device, without host synchronisation, into a ring of capacity
 * records. interval 0 switches it off. */
int trixi_hip_set_analysis(trixi_hip_ctx *ctx, int interval, int capacity);
/* The newest min(made, capacity, max_records) records of the last solve, oldest
 * first, each [3 + 3V] = {t, iter, l2 [V], linf [V], integrals of u [V],
 * entropy integral} as trixi_hip_error_norms and trixi_hip_integrals give them. */
int trixi_hip_get_analysis(trixi_hip_ctx *ctx, double *records, int max_records, int *count);

/* Device allocations and host synchronisations (stream synchronisations and
 * synchronous copies) the library has made for this context since it was
 * created, creation included. The counterpart of the reference's allocation
 * test of the RHS, @test_allocations(Trixi.rhs_hyperbolic!, semi, sol, 1000)
 * (test/test_tree_3d_euler.jl:23): after set-up, trixi_hip_rhs_native,
 * trixi_hip_rk_stage and trixi_hip_rk_step add to neither count, and
 * trixi_hip_solve synchronises once per poll of its control block. Either
 * pointer may be NULL. */
int trixi_hip_get_counters(const trixi_hip_ctx *ctx, int64_t *device_allocations, int64_t *host_synchronizations);

/* Workgroup tile of the element kernel (trixi_hip_get_tile: the RHS variant;
 * trixi_hip_get_stage_tile: the fused 2N-stage variant, 0 when the stage runs
 * unfused): elements per block and threads. */
int trixi_hip_get_tile(const trixi_hip_ctx *ctx, int *elements_per_block, int *threads_per_block);
int trixi_hip_get_stage_tile(const trixi_hip_ctx *ctx, int *elements_per_block, int *threads_per_block);

/* Tile autotuner. No counterpart in the reference (its GPU launch shapes are
 * fixed). Times every instantiated workgroup tile of the context's element
 * kernel, the RHS variant and the fused 2N-stage variant separately (3 rounds
 * of `repetitions` launches on the context's stream, best round), on the
 * resident state (upload it first), and makes the fastest tile of each variant
 * the context's and, for contexts created later in this process with a default
 * tile, their default. Tiles give bitwise-equal results in FP64; in FP32 the
 * log-mean series/exact choice is made per wavefront, so states with large
 * jumps can differ at round-off level between tiles. The resident state u is
 * unchanged; the 2N register u_tmp is overwritten (call between steps).
 * timings (up to max_timings entries, may be NULL with max_timings 0) receive
 * the measured candidates, ntimings their number. Default tiles without
 * autotuning come from the table of measured sweeps compiled into the
 * library (tile_table.inc), else one wavefront per direction. */
typedef struct trixi_hip_tile_timing {
  int32_t stage;              /* 0: RHS kernel, 1: fused-stage kernel */
  int32_t elements_per_block;
  int32_t threads_per_block;
  double microseconds;        /* time per launch, best of the 3 rounds */
} trixi_hip_tile_timing;
int trixi_hip_autotune(trixi_hip_ctx *ctx, int repetitions, trixi_hip_tile_timing *timings, int max_timings,
                       int *ntimings);

/* Host-side TreeMesh adaptation and connectivity (no device work; callable
 * without a GPU). A tree is held as its leaves in element (depth-first)
 * order: level, integer coordinates at that level ([n][ndims], 0 <= c <
 * 2^level, level <= 19) and centre ([n][ndims]).
 *   trixi_hip_tree_create   the leaves of a TreeMesh (tree_mesh.jl:128-185);
 *                           each leaf's origin starts as itself (kind 0, src i)
 *   trixi_hip_tree_refine   refine! of the marked leaves with rebalance!
 *                           (abstract_tree.jl:367-393, 453-490); *nrefined =
 *                           cells refined, rebalancing included
 *   trixi_hip_tree_coarsen  the AMR callback's coarsening (amr.jl:304-370)
 *                           with the balance guard of coarsen!
 *                           (abstract_tree.jl:516-600), highest level first:
 *                           a parent is coarsened when all its 2^ndims
 *                           children are among the candidate leaves (levels,
 *                           coordinates as of before this adaptation's
 *                           refinement) and no child has a non-sibling face
 *                           neighbour with children
 *   trixi_hip_tree_sizes    {leaves, interfaces, mortars, boundaries}
 *   trixi_hip_tree_leaves   the leaves and their origin in this tree's
 *                           adaptations: kind 0 = old leaf src, 1 + q = child
 *                           q of old leaf src, -1 = parent of old leaves src ..
 *                           src + 2^ndims - 1 (trixi_hip_transfer's map)
 *   trixi_hip_tree_connectivity  init_interfaces!, init_mortars!,
 *                           init_boundaries! (containers_2d.jl:228-290,
 *                           475-543, 707-800; containers_3d.jl:229-279,
 *                           391-480, 652-788), 0-based: element neighbours
 *                           [n][2 ndims] (-1: none or boundary), interfaces
 *                           [I][2] + orientations, mortars [M][2^(ndims-1) +
 *                           1] (small elements, large element) + large sides
 *                           + orientations, boundaries + orientations +
 *                           neighbour sides; NULL outputs are skipped.
 * create returns NULL for invalid arguments. */
typedef struct trixi_hip_tree trixi_hip_tree;
trixi_hip_tree *trixi_hip_tree_create(int ndims, const int *periodicity, double length_level_0, int nleaves,
                                      const int32_t *levels, const int64_t *coordinates, const double *centers);
void trixi_hip_tree_destroy(trixi_hip_tree *t);
const char *trixi_hip_tree_last_error(const trixi_hip_tree *t);
int trixi_hip_tree_refine(trixi_hip_tree *t, const uint8_t *marked, int64_t *nrefined);
int trixi_hip_tree_coarsen(trixi_hip_tree *t, int ncandidates, const int32_t *levels, const int64_t *coordinates,
                           int64_t *ncoarsened);
int trixi_hip_tree_sizes(trixi_hip_tree *t, int64_t *sizes);
int trixi_hip_tree_leaves(const trixi_hip_tree *t, int32_t *levels, int64_t *coordinates, double *centers,
                          int32_t *kind, int32_t *src);
int trixi_hip_tree_connectivity(trixi_hip_tree *t, int32_t *element_neighbors, int32_t *neighbor_ids,
                                int32_t *orientations, int32_t *mortar_neighbor_ids, int32_t *mortar_large_sides,
                                int32_t *mortar_orientations, int32_t *boundary_neighbor_ids,
                                int32_t *boundary_orientations, int32_t *boundary_neighbor_sides);

#ifdef __cplusplus
}
#endif

#endif /* TRIXI_HIP_H */
