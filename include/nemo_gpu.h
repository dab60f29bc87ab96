/*
 * nemo_gpu.h -- C-ABI of the MI355X (gfx950) spatial operator and RK stage
 * update for nemo.
 *
 * The reference has no device boundary: its Fortran game loop calls the
 * kernel/equations/timedisc modules directly (call chain in the reference:
 * driver_mod.f90:62-150 -> hook_evolve_500 -> rkstep_cb -> kernel_evolve).
 * This header is the thin layer a Fortran host binds with ISO_C_BINDING
 * (nemo_amd/fortran/nemo_gpu_mod.f90), in the style of the reference's own
 * C-ABI bindings (mesh/p4est/wrapper/p4wrap_interfaces_mod.f90:46-120):
 * plain pointers, sizes and an opaque handle, `int` status returns.
 *
 * Conventions
 *   - Every entry point returns NEMO_OK (0) or a positive error code; the
 *     message of the last error on the calling thread is available from
 *     nemo_gpu_error_message().  This replaces raise_error / rt%stat
 *     (utils/status_mod.f90:7-21, runtime_mod.f90:55-61).
 *   - Host state arrays use the reference's per-quad payload layout stacked
 *     over elements: Fortran state(N_NODES,N_NODES,N_NODES,N_VARS,nelems),
 *     i.e. C u[elem][var][node] with node = i + M*j + M*M*k, M = N+1
 *     (kernel types: kernel/mcfv/3d/kernel_types_mod.f90:6-21).
 *   - Face ids: 0..5 = NOR, SOU, WES, EAS, FRO, BAC of utils/constants_mod.f90:
 *     34-40 (x-, x+, y-, y+, z-, z+).
 *   - One host thread per handle; all device work is ordered on the
 *     handle's HIP stream.
 */
#ifndef NEMO_GPU_H
#define NEMO_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NEMO_OK 0
#define NEMO_ERR_ARG 1         /* invalid argument or unsupported combination */
#define NEMO_ERR_HIP 2         /* HIP runtime error */
#define NEMO_ERR_STATE 3       /* positivity repair failed (kernel_repair false) */
#define NEMO_ERR_NOT_READY 4   /* mesh or state not uploaded yet */

/* equations (preproc.py selection of source/equations/...) */
#define NEMO_EQ_EULER 0
#define NEMO_EQ_GLM_MHD 1
/* volume integral */
#define NEMO_VOL_WEAK 0        /* kernel/dgsem/weakform (D-hat f) */
#define NEMO_VOL_CENTRAL 1     /* flux differencing, two_point/standard */
#define NEMO_VOL_EC 2          /* flux differencing, two_point/chandrashekar */
/* surface Riemann solver (riemann_mod slots, scripts/nemo.py:398-415) */
#define NEMO_SURF_RUSANOV 0    /* equations/euler/riemann/rusanov */
#define NEMO_SURF_HLLC 1       /* equations/euler/riemann/hllc (HLLC/HLLE hybrid) */
#define NEMO_SURF_HLLD 2       /* equations/mhd/riemann/hlld, GLM-coupled (GLM-MHD) */
/* troubled-cell indicator for DG/FV blending */
#define NEMO_IND_NONE 0        /* pure DG */
#define NEMO_IND_PP 1          /* Persson-Peraire on rho*p */
#define NEMO_IND_MLP 2         /* learned MLP indicator, Persson-Peraire as judge and fallback
                                  (Euler, n_poly = 3, EC volume flux, HLLC or Rusanov) */
/* time integration (timedisc/rungekutta/...) */
#define NEMO_TD_LSRK54 0       /* low-storage/4-5 */
#define NEMO_TD_SSPRK54 1      /* ssp/4-5 */

/* Compile-time configuration of the reference (preproc.py flags and
 * setup_config_mod constants) as a runtime struct.  Fortran: a bind(c)
 * derived type with the same members. */
typedef struct nemo_gpu_config {
    int32_t n_poly;        /* N; N_NODES = N+1 (PP_N_NODES), 2..4 */
    int32_t equations;     /* NEMO_EQ_* */
    int32_t volume;        /* NEMO_VOL_* */
    int32_t surface;       /* NEMO_SURF_* */
    int32_t indicator;     /* NEMO_IND_* */
    int32_t timedisc;      /* NEMO_TD_* */
    int32_t repair;        /* 1: kernel_repair after every stage */
    int32_t precision;     /* 0: FP64 operator; 1: FP32 operator (fluxes, Riemann
                              solvers, indicator, accumulation in FP32; state,
                              RK registers, CFL and repair in FP64) */
    double gamma;          /* setup_config_mod::kappa */
    double alpha_max;      /* blending factor clamp */
    double alpha_min;      /* blending factors below this become 0 */
    double cfl;            /* setup_config_mod::CFL */
    double mu0;            /* GLM-MHD only: setup_config_mod::mu0 */
} nemo_gpu_config;

/* Create / destroy a solver (replaces hook_init_400 operator setup,
 * kernel/dgsem/weakform/2d/kernel_share_mod.f90:56-151). */
int nemo_gpu_create(const nemo_gpu_config *cfg, void **handle);
int nemo_gpu_destroy(void *handle);

/* Upload mesh connectivity once per remesh (replaces the quad/side tables
 * of mesh_alloc, mesh/p4est/mesh_utils_mod.f90:160-441): nbr[nelem][6]
 * element across each face, delta[nelem][3] element edge lengths
 * (quad%delta).  Conforming, aligned faces. */
int nemo_gpu_upload_mesh(void *handle, int64_t nelem, const int32_t *nbr, const double *delta);

/* Adaptive (2:1 balanced) mesh: as nemo_gpu_upload_mesh plus the sides with
 * nquads = 4 of mesh_alloc (mesh/p4est/mesh_utils_mod.f90:340-376, side_t in
 * mesh/p4est/mesh_types_mod.f90:56-78).  mortars[m] = (big element, its face,
 * small_0..small_3); small_k covers half (k & 1) of the big face along the
 * face's first tangential axis and half (k >> 1) along the second (axes in
 * increasing order).  Faces on mortar m carry nbr = -(5m+1) (big) and
 * -(5m+2+k) (small_k).  Their fluxes follow the reference's rfaceMat /
 * cfaceMat sequence (kernel/dgsem/weakform/2d/kernel_flux_mod.f90:58-147);
 * for GLM-MHD with each side's Powell/GLM term added before the projection
 * (kernel/mcfv/3d/kernel_flux_mod.f90:229-268). */
int nemo_gpu_upload_mesh_amr(void *handle, int64_t nelem, const int32_t *nbr, const double *delta,
                             int64_t nmortar, const int32_t *mortars);

/* Non-periodic mesh: as nemo_gpu_upload_mesh_amr plus the boundary sides of
 * mesh_alloc (side%boundary > 0, mesh/p4est/mesh_utils_mod.f90:340-376) and
 * the BC hook of the face pass (kernel/mcfv/3d/kernel_flux_mod.f90:122-226).
 * bfaces[b] = (element, face, kind); boundary face b carries
 * nbr = -(5 nmortar + b + 1).  kind 0 = transmissive (the outer state mirrors
 * the inner one, :123-125, and the configured Riemann solver gives the flux;
 * GLM-MHD adds the inner state's non-conservative term), 1 = slip wall (the
 * reference's eulerian-slip-wall boundary solver,
 * equations/euler/riemann/eulerian-slip-wall/riemann_mod.f90:8-200; Euler
 * only). */
#define NEMO_BC_TRANSMISSIVE 0
#define NEMO_BC_SLIP_WALL 1
int nemo_gpu_upload_mesh_bc(void *handle, int64_t nelem, const int32_t *nbr, const double *delta,
                            int64_t nmortar, const int32_t *mortars, int64_t nbound, const int32_t *bfaces);

/* Mesh adaptation: replaces the state-transfer half of mesh_adapt_wrapped
 * (mesh/p4est/mesh_utils_mod.f90:966-1000) with kernel_amr_refine /
 * kernel_amr_coarsen (kernel/mcfv/3d/kernel_amr_mod.f90:12-70).  The new
 * mesh is given as for nemo_gpu_upload_mesh_amr; units[u] = (kind, first old
 * element, first new element) in element order, kind 0 = copy, 1 = refine
 * (one old into the 8 new children, child c = x + 2y + 4z), 2 = coarsen (8
 * old siblings into one).  A unit whose interpolated or projected state has
 * a nonpositive density or pressure (GLM-MHD: gas pressure) takes the
 * (conserved) element mean instead; their number is returned in *fallbacks.
 * RK registers are cleared: call between time steps. */
int nemo_gpu_remesh(void *handle, int64_t nelem, const int32_t *nbr, const double *delta, int64_t nmortar,
                    const int32_t *mortars, int64_t nunits, const int32_t *units, int64_t *fallbacks);

/* State upload/download (payload quad%pld%hydro%state of every quad). */
int nemo_gpu_upload_state(void *handle, const double *u);
int nemo_gpu_download_state(void *handle, double *u);

/* RHS of the current state, du/dt per node (replaces kernel_calc_flux,
 * kernel/dgsem/weakform/2d/kernel_flux_mod.f90:197-277, after the face pass
 * hook_pre_rk_timestep_300).  Parity entry point. */
int nemo_gpu_rhs(void *handle, double *rhs);

/* Blending factor per element for the current state (indicator only).
 * nemo_gpu_rhs, nemo_gpu_indicator and nemo_gpu_indicator_features are
 * diagnostics: with the learned indicator they evaluate a fresh factor but
 * leave the factor of the current time step (used by stages 2..5) and the
 * judge's counters untouched, so they may be called between stages. */
int nemo_gpu_indicator(void *handle, double *alpha);

/* Volume integrals of the current state with the LGL quadrature volumes
 * (the sums of hook_analyze, setups/euler/2d/_common_/setup_analyze_mod.f90:
 * 34-164, and setups/euler/3d/convergence-test/setup_analyze_mod.f90:
 * 40-151): sums[0..V-1] the conserved variables, [9] kinetic, [10] internal,
 * [11] magnetic energy, [12] volume, [13] min rho, [14] min p, [15] max |v|.
 * Deterministic (fixed reduction order).  Synchronises. */
#define NEMO_ANALYZE_NSUMS 16
int nemo_gpu_analyze(void *handle, double *sums);

/* Allocation-free nemo_gpu_analyze.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Enqueues the same kernels on the handle's
 * stream and returns without synchronising; sums_dev is a device buffer of
 * NEMO_ANALYZE_NSUMS doubles owned by the caller, filled when the stream
 * reaches the call (nemo_gpu_synchronize waits for it).  The workgroup
 * partials live in a fixed array of the library (8 MiB per GPU, held until
 * the process ends) instead of memory allocated on first use, so the call
 * allocates nothing; meshes of more than NEMO_ANALYZE_NOALLOC_MAX_NODES nodes
 * per variable (64^3 elements at n_poly = 3) are rejected with
 * NEMO_ERR_ARG.  All calls on one GPU share that array: calls whose work
 * overlaps on one GPU (several handles, several streams or host threads)
 * race and must be ordered by the caller. */
#define NEMO_ANALYZE_NOALLOC_MAX_NODES 16777216
int nemo_gpu_analyze_noalloc(void *handle, double *sums_dev);

/* CFL time step of the current state (replaces hook_calc_timestep_200,
 * kernel/dgsem/weakform/2d/kernel_timestep_mod.f90:19-63): dt =
 * 0.5*CFL/(2*N_NODES-1)*min(delta/lambda); lmax and vmax are the global
 * maxima used for GLM cleaning.  Synchronises. */
int nemo_gpu_calc_timestep(void *handle, double *dt, double *lmax, double *vmax);

/* GLM divergence cleaning (GLM-MHD): the cleaning speed c_h and damping
 * alpha that the operator uses, the reference's module variables
 * equations_share_mod::glm_ch / glm_alpha.  nemo_gpu_calc_timestep and
 * nemo_gpu_run set them from the CFL reduction as the reference's timestep
 * hook does (c_h = 2|lmax - vmax|, alpha = c_h/0.18,
 * kernel/mcfv/3d/kernel_timestep_mod.f90:53-56); nemo_gpu_set_glm assigns
 * them directly. */
int nemo_gpu_set_glm(void *handle, double ch, double alpha);
int nemo_gpu_get_glm(void *handle, double *ch, double *alpha);

/* One RK stage s = 1..5 with time step dt (replaces hook_mid_rk_timestep_500
 * -> rkstep_cb -> kernel_evolve + timedisc_probe + kernel_repair,
 * timedisc/rungekutta/low-storage/4-5/timedisc_hooks_mod.f90:19-56). */
int nemo_gpu_stage(void *handle, int32_t stage, double dt);

/* All stages of one time step (hook_evolve_500,
 * timedisc/rungekutta/timedisc_evolve_mod.f90:20-84), asynchronous. */
int nemo_gpu_step(void *handle, double dt);

/* nsteps steps with the CFL time step recomputed on the device before each
 * step (driver_mod.f90:62-150 without checkpoint/analysis), no host
 * synchronisation between steps.  dt_hist (may be NULL) receives the time
 * step of every step; this call synchronises. */
int nemo_gpu_run(void *handle, int32_t nsteps, double *dt_hist);

/* Wait for all queued work; returns NEMO_ERR_STATE if a repair failed
 * since the last call (rt%stat%ok check, timedisc_evolve_mod.f90:68-79). */
int nemo_gpu_synchronize(void *handle);
int nemo_gpu_get_status(void *handle, int32_t *code);

/* Kernel timing with HIP events on the solver's stream: enable, then read
 * the number of stage launches and their summed duration in milliseconds
 * (reading synchronises and resets the counters). */
int nemo_gpu_set_timing(void *handle, int32_t enable);
/* Measurement aid: one plain device copy of nbytes (reads and writes exactly
 * nbytes), the known-size kernel that scales hardware traffic counters. */
int nemo_gpu_copy_calibration(void *handle, int64_t nbytes);
/* Diagnostic, no counterpart in the reference: n generated operand pairs
 * (random exponents in [-60, 60], half of them nearly equal) through the
 * kernels' FP64 division (div_rn) and reciprocal (rcp_nr); counts the
 * results whose bits differ from the IEEE-rounded quotient / reciprocal.
 * The log-mean's ratio and beta rely on div_rn matching IEEE bit for bit. */
int nemo_gpu_check_arithmetic(void *handle, int64_t n, uint64_t seed, int64_t *div_mismatches,
                              int64_t *rcp_mismatches);
int nemo_gpu_read_timing(void *handle, int64_t *launches, double *total_ms);

/* ---- learned troubled-cell indicator --------------------------------------
 * No counterpart in the reference, whose DG/FV blending survives only as
 * traces (setups/mhd/3d/_common_/setup_checkpoint_mod.f90:395-450).
 * Network: 4 inputs -> 3 hidden layers of 64 (GELU) -> sigmoid alpha.
 * Inputs per element: 0.1 ln(f + 1e-12) for the modal energy fractions
 * f = E(top shell)/E, E(top two shells)/E of rho and of p.
 * Parameters (FP32, NEMO_MLP_NPARAMS): W1[64][4], b1[64], W2[64][64], b2[64],
 * W3[64][64], b3[64], w4[64], b4, W[j][k] from input k to unit j.
 * With NEMO_IND_MLP the factor is computed at the first stage of every time
 * step and kept for the step; the Persson-Peraire indicator judges a sampled
 * fraction of the elements each step and takes over where the prediction is
 * short by more than the tolerance, and where rho or p is non-positive. */
#define NEMO_MLP_NPARAMS 8705
int nemo_gpu_mlp_set_params(void *handle, const float *params);
int nemo_gpu_mlp_get_params(void *handle, float *params);
/* judge: sampled fraction of elements per step (default 0.05) and tolerance (0.05) */
int nemo_gpu_mlp_set_policy(void *handle, double sample_fraction, double tolerance);
/* judged, overridden and fallback element counts since the last call */
int nemo_gpu_mlp_stats(void *handle, int64_t *judged, int64_t *overridden, int64_t *fallback);
/* Training data of the current state: features[nelem][4] and the
 * Persson-Peraire factor alpha_pp[nelem] (any Euler scheme, n_poly = 3). */
int nemo_gpu_indicator_features(void *handle, float *features, double *alpha_pp);
/* Network output for n feature rows. */
int nemo_gpu_mlp_forward(void *handle, int64_t n, const float *features, float *y);
/* Loss 1/(2n) sum (y - target)^2 and its gradient (canonical layout). */
int nemo_gpu_mlp_grad(void *handle, int64_t n, const float *features, const float *target, float *loss,
                      float *grad);
/* Adam training on the device from the current parameters: `epochs` passes
 * over shuffled minibatches of `batch` rows; loss_per_epoch (may be NULL)
 * receives the mean minibatch loss of each epoch. */
int nemo_gpu_mlp_train(void *handle, int64_t n, const float *features, const float *target, int32_t epochs,
                       int32_t batch, float lr, uint64_t seed, float *loss_per_epoch);

/* ---- batched ensemble ----------------------------------------------------
 * Independent runs of one scheme advanced together (the reference runs one
 * mesh per process; this has no counterpart there).  Every run has its own
 * mesh, state, n_poly (2..4) and CFL time step; by default one launch per
 * RK stage covers the element blocks of every run and degree, and pairs of
 * time steps are replayed as a captured HIP graph.  Built for Euler with the
 * EC volume flux, HLLC, no indicator or Persson-Peraire; cfg->n_poly is not
 * used. */
int nemo_gpu_batch_create(const nemo_gpu_config *cfg, int32_t nruns, const int32_t *n_poly,
                          const int64_t *nelem, void **handle);
int nemo_gpu_batch_destroy(void *handle);
/* Mesh (as nemo_gpu_upload_mesh, run-local element ids) and state of a run. */
int nemo_gpu_batch_upload(void *handle, int32_t run, const int32_t *nbr, const double *delta, const double *u);
int nemo_gpu_batch_download(void *handle, int32_t run, double *u);
/* nsteps CFL-limited steps of every run; dt_hist[nsteps][nruns] may be NULL.
 * Synchronises; NEMO_ERR_STATE if a repair failed. */
int nemo_gpu_batch_run(void *handle, int32_t nsteps, int32_t use_graph, double *dt_hist);
/* Launch geometry: one_launch = 1 runs each RK stage as one launch over the
 * blocks of all degrees (order 0: degree 4, 3, 2 blocks in that order, the
 * costliest first; 1: 2, 3, 4), 0 as one launch per degree; grid_per_cu
 * workgroups per CU walking the blocks persistently (0: one workgroup per
 * block of elements) and the XCD-contiguous block map on or off.  Default:
 * one launch per stage, one workgroup per block, order 0. */
int nemo_gpu_batch_set_launch(void *handle, int32_t grid_per_cu, int32_t xcd_map, int32_t one_launch, int32_t order);
/* Times `steps` steps of every candidate geometry on the current states
 * (which are restored afterwards), keeps the fastest and reports it. */
int nemo_gpu_batch_autotune(void *handle, int32_t steps, int32_t *grid_per_cu, int32_t *xcd_map, int32_t *one_launch,
                            int32_t *order, double *best_ms_per_step);
/* Timings of the last autotuning: *rows candidates; for the first max_rows,
 * geometry[4*i..4*i+3] = (one_launch, order, grid_per_cu, xcd_map) and
 * ms[i] its time per step. */
int nemo_gpu_batch_tune_table(void *handle, int32_t max_rows, int32_t *rows, int32_t *geometry, double *ms);

/* Operators for polynomial degree N (host only, no device needed):
 * LGL nodes x[M], weights w[M] (utils/interpol_mod.f90:513-582),
 * D[M][M] with D[i][j] = l_j'(x_i) (interpol_mod.f90:38-80), and the
 * inverse normalised-Legendre Vandermonde Vinv[M][M] (interpol_mod.f90:
 * 584-601). */
int nemo_gpu_operators(int32_t n_poly, double *x, double *w, double *D, double *Vinv);

/* Last error message on this thread (NUL-terminated, truncated to len). */
int nemo_gpu_error_message(char *buf, int32_t len);

#ifdef __cplusplus
}
#endif
#endif
