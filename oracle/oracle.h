/*
 * oracle.h -- CPU restatement of Photon's gas-phase pressure projection path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library (photon_amd/) may
 * include, link or call this code; it is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).
 *
 * Restated from the reference (paths relative to the reference repo root):
 *   src/sim/pressure_mg.cpp:28-191   buildPressureLevels (connected aggregation)
 *   src/sim/pressure_mg.cpp:195-300  gsColor / residualGrid / jacobi /
 *                                    residualSparse / gsSparseSweep / V-cycle
 *   src/sim/gasphase.cpp:505-677     project (RHS, CG, zone mean, correction)
 *   src/sim/gasphase.cpp:251-271     energyAndExpansion
 *   src/sim/gasphase.cpp:685-784     advectScalars
 *   src/sim/gasphase.cpp:111-158     computeZones
 *   src/sim/gasphase.cpp:41-49       sumSlices (per-k FP64 partials, in order)
 *
 * Conventions (same as the reference's Solver): cell c = i + nx*(j + ny*k);
 * x faces (nx+1)*ny*nz, y faces nx*(ny+1)*nz, z faces nx*ny*(nz+1).
 * "wind_closed[4]" flags the x-low, x-high, y-low, y-high domain sides on
 * which windAt() returns an inflow (those boundary faces leave the matrix,
 * gasphase.cpp:510,518-527; pressure_mg.cpp:31-46).  All zero = no wind.
 */
#ifndef PHOTON_ORACLE_H
#define PHOTON_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LEVELS 12

typedef struct {
    int n;
    int nx, ny, nz;               /* level 0 only */
    float *ax, *ay, *az;          /* level 0 only: boundary-masked face coefficients */
    int *rowStart, *col;          /* coarse levels: off-diagonal couplings (positive) */
    float *val;
    int nnz;
    int *bi, *bj, *bk;            /* coarse levels: block coordinate of each unknown */
    int *agg;                     /* unknown -> next level unknown; NULL on the coarsest */
    float *diag;
    uint8_t *leak;                /* unknown touches an open (Dirichlet) boundary face */
} orc_level;

typedef struct orc_mg {
    int nl;
    orc_level L[ORC_MAX_LEVELS];
    /* exact coarsest solve (FP64 Cholesky of the grounded coarsest operator) */
    double *chol;                 /* n_c*n_c lower factor, NULL until first use */
    uint8_t *cmask;               /* coarsest rows forced to zero (diag <= 0) */
    int *ccomp;                   /* coarsest rows: closed-component id, -1 if open or masked */
    int nccomp;
} orc_mg;

enum { ORC_COARSE_GS20 = 0, ORC_COARSE_EXACT = 1 };
enum { ORC_STOP_MAXNORM = 0, ORC_STOP_REL_L2 = 1 };

/* ---- hierarchy ---------------------------------------------------------- */
orc_mg *orc_mg_build(int nx, int ny, int nz, const float *ax_, const float *ay_, const float *az_,
                     const uint8_t *blocked, const uint8_t *wind_closed);
void orc_mg_free(orc_mg *mg);
int orc_mg_nlevels(const orc_mg *mg);
int orc_mg_level_n(const orc_mg *mg, int l);
int orc_mg_level_nnz(const orc_mg *mg, int l);
/* copies out level data; any pointer may be NULL */
void orc_mg_level_get(const orc_mg *mg, int l, int *rowStart, int *col, float *val, float *diag, int *agg,
                      int *bi, int *bj, int *bk, uint8_t *leak);

/* ---- V-cycle: z = M^-1 r  (FP32 as the reference, or FP64) -------------- */
void orc_precondition_f32(orc_mg *mg, const float *r, float *z, int coarse_mode);
void orc_precondition_f64(orc_mg *mg, const double *r, double *z, int coarse_mode);

/* ---- y = A x on level 0, returns sum y*x in per-slice FP64 (gasphase.cpp:551-570) */
double orc_apply_f32(const orc_mg *mg, const float *x, float *y);
void orc_coarse_solve_f64(orc_mg *mg, const double *b, double *x);

/* ---- PCG on a given right-hand side (the CG loop of project(), gasphase.cpp:571-627)
 * x is the warm start on input and the solution on output.  Stop rules:
 *   ORC_STOP_MAXNORM: max|r| / inv <= tol (the reference's rule, inv = dx^2/dt)
 *   ORC_STOP_REL_L2:  ||r||_2 / ||b||_2 <= tol  (recursive residual, FP64 norms)
 * Returns iterations; *res_out = the final value of the stop quantity.  */
int orc_pcg_f32(orc_mg *mg, const float *b, float *x, float inv, float tol, int max_iters, int stop_mode,
                int coarse_mode, double *res_out);
int orc_pcg_f64(orc_mg *mg, const double *b, double *x, double tol, int max_iters, int coarse_mode,
                double *res_out);
/* true relative residual ||b - A x|| / ||b|| with all arithmetic in FP64 */
double orc_true_relres_f32(const orc_mg *mg, const float *b, const float *x);
double orc_true_relres_f64(const orc_mg *mg, const double *b, const double *x);

/* ---- zones (gasphase.cpp:111-158): returns zone count; zone[c] = -1 on blocked */
int orc_compute_zones(int nx, int ny, int nz, const float *ax_, const float *ay_, const float *az_,
                      const uint8_t *blocked, int *zone, uint8_t *zone_open, int max_zones);

/* ---- project(dt) as a whole (gasphase.cpp:505-677) ------------------------ */
typedef struct {
    int nx, ny, nz;
    float dx, dt;
    const float *fx, *fy, *fz;     /* face open fractions */
    const float *ax, *ay, *az;     /* matrix coefficients (unmasked, as ax_/ay_/az_) */
    const uint8_t *blocked;
    const int *zone;               /* per cell, -1 solid */
    const uint8_t *zone_open;      /* per zone */
    const uint8_t *wind_closed;    /* 4 sides, may be NULL */
    float tol;
    int max_iters;
    int stop_mode, coarse_mode;
} orc_project_args;

typedef struct {
    int iters;
    float residual;                /* max-norm divergence error, 1/s (pressureResidual) */
    float max_vel;
    double relres;                 /* final recursive ||r||/||b|| */
} orc_project_out;

/* u,v,w,p in/out; S in; b_out/diag_out (may be NULL) receive the RHS and diagonal */
void orc_project(orc_mg *mg, const orc_project_args *a, float *u, float *v, float *w, float *p, const float *S,
                 float *b_out, float *diag_out, orc_project_out *out);
/* the three pieces of project() separately */
void orc_rhs(const orc_project_args *a, const float *u, const float *v, const float *w, const float *S, float *b,
             float *diag, float *p);
void orc_zone_mean(const orc_project_args *a, int nzones, float *p);
float orc_correct(const orc_project_args *a, const float *p, float *u, float *v, float *w);

/* ---- energyAndExpansion (gasphase.cpp:251-271); returns maxDiv ------------ */
float orc_energy_expansion(int n, float dx, float dt, const uint8_t *blocked, const float *Pb, const float *Q,
                           const float *massSrc, float *T, float *S);

/* ---- advectScalars (gasphase.cpp:685-784) -------------------------------- */
/* fields: 10 arrays in the order T, Yo, Yf, Ys, Ya, Yp, Yc, Yw, Yco, Yai (in), out: 10 arrays */
void orc_advect_scalars(int nx, int ny, int nz, float dx, float dt, const float *u, const float *v, const float *w,
                        const float *fx, const float *fy, const float *fz, const uint8_t *blocked, const float *Pb,
                        const float *nu, const float *const *fields, float *const *out);

/* ---- the rest of the gas step (oracle_step.c) ------------------------------ */
typedef struct {
    int nx, ny, nz;
    float dx;
    const float *fx, *fy, *fz;     /* face open fractions */
    const uint8_t *blocked;
    const uint8_t *lx, *ly, *lz;   /* crack (leak) faces, or NULL */
} orc_grid;

typedef struct {
    float speed, from_deg, exponent;   /* Solver::wind (solver.h:106-111) */
    float oz;                          /* height of the grid's floor, m */
} orc_wind;

typedef struct {
    float *T, *Pb;
    float *Yo, *Yf, *Ys, *Ya, *Yp, *Yc, *Yw, *Yco, *Yai;
    float *massSrc;
} orc_species;

float orc_wind_at(const orc_wind *wd, int axis, int side, float z);
float orc_stable_dt(float dt_max, float profile_dt_max, float cfl, float dx, float max_vel, float max_nu,
                    float max_div);
float orc_turbulence(const orc_grid *g, const float *u, const float *v, const float *w, const float *T, float *nu);
void orc_heat_diffusion(const orc_grid *g, float dt, const float *T, const float *nu, const float *Pb, float *Q);
void orc_momentum(const orc_grid *g, const orc_wind *wd, float dt, const float *u, const float *v, const float *w,
                  const float *T, const float *Pb, const float *nu, const float *p, float *u2, float *v2, float *w2);
void orc_add_gas(int ncell, const int *cells, float dx, float fuel_gen, float aerosol, float soot_pot, float ait,
                 int set_pilot, const uint8_t *blocked, const orc_species *sp, uint8_t *pilot);
double orc_combustion(const orc_grid *g, float dt, int radiation_on, const orc_species *sp, float *Q, float *Qr,
                      float *hrr, const uint8_t *pilot, uint8_t *lit);

/* leakage (gasphase.cpp:402-498): crack flow, closed zones' background
 * pressure, zoneDP and the shift of S; returns the number of crack faces */
int orc_leakage(const orc_grid *g, float dt, const int *zone, int nzones, const uint8_t *zone_open,
                const float *zone_vol, const float *zoneP, const float *T, const float *Pb, const float *p, float *u,
                float *v, float *w, float *S, float *zoneDP);

#ifdef __cplusplus
}
#endif
#endif
