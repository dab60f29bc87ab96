/*
 * nemo_oracle.h -- FP64 CPU restatement of nemo's spatial operator and RK
 * stage update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (nemo_amd/, libnemo_gpu.so) never links, loads or calls it.
 *
 * Layout shared with the product's C-ABI: a state array is
 *     u[elem][var][node],  node = i + M*j + M*M*k  (x fastest, M = N+1),
 * i.e. the Fortran array state(i,j,k,var,elem) of the reference's per-quad
 * payload (kernel/dgsem/weakform/2d/kernel_types_mod.f90:6-19, extended to
 * 3-D and stacked over elements).
 *
 * Face ids follow the reference (utils/constants_mod.f90:34-40) minus one:
 *     0 = x- (NOR), 1 = x+ (SOU), 2 = y- (WES), 3 = y+ (EAS),
 *     4 = z- (FRO), 5 = z+ (BAC).
 */
#ifndef NEMO_ORACLE_H
#define NEMO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_EQ_EULER = 0, ORC_EQ_MHD = 1 };
enum { ORC_VOL_WEAK = 0, ORC_VOL_CENTRAL = 1, ORC_VOL_EC = 2 };
enum { ORC_SURF_RUSANOV = 0, ORC_SURF_HLLC = 1, ORC_SURF_HLLD = 2,
       ORC_SURF_EC_TEST = 3 /* Chandrashekar flux at faces: entropy-conservation test only */ };
enum { ORC_TD_LSRK54 = 0, ORC_TD_SSPRK54 = 1 };
enum { ORC_IND_NONE = 0, ORC_IND_PP = 1 };

typedef struct {
    int32_t N;          /* polynomial degree, M = N+1 LGL nodes per direction */
    int32_t eq;         /* ORC_EQ_* */
    int32_t volume;     /* ORC_VOL_* */
    int32_t surface;    /* ORC_SURF_* */
    int32_t indicator;  /* ORC_IND_*: none = pure DG, PP = DG/FV blending */
    int32_t timedisc;   /* ORC_TD_* */
    int32_t repair;     /* 1 = positivity repair after every stage */
    int32_t pad;
    double gamma;       /* ratio of specific heats (kappa in the reference) */
    double alpha_max;   /* blending factor clamp */
    double alpha_min;   /* blending factors below this are set to zero */
    double cfl;
    double mu0;         /* MHD only */
    double glm_ch;      /* MHD only: GLM cleaning speed */
    double glm_alpha;   /* MHD only: GLM damping */
    int64_t nmortar;    /* non-conforming (2:1) interfaces */
    const int32_t *mortars; /* [nmortar][6]: big element, its face, small_0..small_3;
                               faces on mortar m carry nbr = -(5m+1) (big) and
                               -(5m+2+k) (small_k); small_k covers half k&1 of the
                               big face along its first tangential axis and half
                               k>>1 along the second (axes in increasing order) */
    int64_t nbound;     /* boundary faces (non-periodic meshes) */
    const int32_t *bfaces;  /* [nbound][3]: element, face, ORC_BC_* kind; face b
                               carries nbr = -(5 nmortar + b + 1) */
} orc_cfg;

/* Boundary kinds: transmissive (outer state = inner state, the configured
 * Riemann solver) and the Eulerian slip wall (Euler only). */
enum { ORC_BC_TRANSMISSIVE = 0, ORC_BC_SLIP_WALL = 1 };

/* ---- operators (utils/interpol_mod.f90, utils/linalg_mod.f90) ---------- */
void orc_lgl_nodes_weights(int N, double *x, double *w);
void orc_diff_matrix(int N, const double *x, double *D);      /* D[i*M+j] = l_j'(x_i) */
void orc_legendre(int n, double x, double *L, double *dL);     /* normalised */
void orc_vandermonde(int N, const double *x, double *V, double *Vinv);
void orc_invert(int n, const double *a, double *c);

/* ---- pointwise physics (Euler) ---------------------------------------- */
double orc_pressure(const orc_cfg *c, const double *u);
void orc_flux(const orc_cfg *c, int dir, const double *u, double *f);
double orc_lambda_max(const orc_cfg *c, int dir, const double *u);
void orc_riemann(const orc_cfg *c, int kind, int dir, const double *uL,
                 const double *uR, double *f);
void orc_two_point(const orc_cfg *c, int kind, int dir, const double *u,
                   const double *v, double *f);
double orc_logmean(double a, double b);
void orc_cons2evec(const orc_cfg *c, const double *u, double *w);
double orc_entropy(const orc_cfg *c, const double *u);
int orc_isvalid(const orc_cfg *c, const double *u);

/* ---- pointwise physics, dispatched on cfg->eq (Euler or GLM-MHD) ------ */
/* GLM-MHD state: rho, rho*v[3], E, B[3], psi (equations/mhd/equations_const_mod.f90). */
double orc_mhd_pressure(const orc_cfg *c, const double *u);
double orc_mhd_fastspeed(const orc_cfg *c, const double *u);   /* x direction */
void orc_phys_flux(const orc_cfg *c, int dir, const double *u, double *f);
void orc_phys_riemann(const orc_cfg *c, int kind, int dir, const double *uL,
                      const double *uR, double *f);
void orc_phys_two_point(const orc_cfg *c, int kind, int dir, const double *u,
                        const double *v, double *f);
/* Powell/GLM non-conservative two-point term Phi(uI) * {{B_n, psi}}(uL, uR)
 * (equations/mhd/non_conservative_flux_mod.f90); zero for Euler. */
void orc_phys_nc(const orc_cfg *c, int dir, const double *uI, const double *uL,
                 const double *uR, double *f);
double orc_phys_lambda_max(const orc_cfg *c, int dir, const double *u);
int orc_phys_isvalid(const orc_cfg *c, const double *u);

/* ---- 2:1 operators (rfaceMat/cfaceMat, refineMat/coarseMat) ------------- */
/* P[(c*M+i)*M+p] = l_p((x_i + 2c - 1)/2), R[(c*M+p)*M+i] = w_i P[c][i][p] / (2 w_p) */
void orc_mortar_ops(int N, double *P, double *R);
/* Face fluxes of every mortar slot: flux[slot][var][M*M]; GLM-MHD slots
 * include their side's non-conservative term. */
void orc_mortar_fluxes(const orc_cfg *c, int64_t nelem, const double *u, double *flux);
/* Face fluxes of every boundary slot (slots 5 nmortar + b of the same array). */
void orc_boundary_fluxes(const orc_cfg *c, int64_t nelem, const double *u, double *flux);
/* Star pressure of the slip wall for inner density rho, pressure p and
 * velocity v toward the wall. */
double orc_slip_wall_pressure(const orc_cfg *c, double rho, double v, double p);

/* ---- element operators -------------------------------------------------- */
int orc_nvars(const orc_cfg *c);
void orc_indicator(const orc_cfg *c, int64_t nelem, const double *u,
                   double *alpha, double *energy);
void orc_rhs(const orc_cfg *c, int64_t nelem, const int32_t *nbr,
             const double *delta, const double *u, const double *alpha,
             double *rhs);
double orc_calc_timestep(const orc_cfg *c, int64_t nelem, const double *delta,
                         const double *u, double *lmax, double *vmax);
int orc_repair(const orc_cfg *c, double *ue);

/* One RK stage s = 1..5 (RHS of the stage input, probe, repair, register
 * update); returns nonzero if a repair failed. */
int orc_stage(const orc_cfg *c, int64_t nelem, const int32_t *nbr,
              const double *delta, double *u, double *r0, double *r1,
              int s, double dt);
/* One full time step (all stages) with a given dt; registers r0/r1 sized
 * like u.  Returns 0 on success, nonzero if a repair failed. */
int orc_step(const orc_cfg *c, int64_t nelem, const int32_t *nbr,
             const double *delta, double *u, double *r0, double *r1,
             double dt);
/* nsteps steps, dt from the CFL rule before each step (written to
 * dt_hist[nsteps] if non-null).  Returns 0 on success. */
int orc_run(const orc_cfg *c, int64_t nelem, const int32_t *nbr,
            const double *delta, double *u, int nsteps, double *dt_hist);

/* LSRK 5-4 and SSPRK(5,4) coefficients
 * (timedisc/rungekutta/low-storage/4-5/timedisc_const_mod.f90:5-29,
 *  timedisc/rungekutta/ssp/4-5/timedisc_const_mod.f90:5-36). */
void orc_lsrk54_coeffs(double *a, double *b, double *cc);
void orc_ssprk54_coeffs(double *coef /* 23 values A0..E6 */);

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
