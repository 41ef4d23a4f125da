/*
 * photon_gas.h -- C ABI of the MI355X (gfx950) gas-phase pressure projection.
 *
 * The reference has no plugin or FFI layer: the path is a set of private
 * members of photon::Solver (src/sim/solver.h:290-336), called only from
 * Solver::step (src/sim/solver.cpp:226-269).  These entry points are what a
 * Solver built on this library binds in place of those members; each
 * declaration cites the reference function it replaces.
 *
 * Conventions
 *   - cell c = i + nx*(j + ny*k); x faces (nx+1)*ny*nz, y faces nx*(ny+1)*nz,
 *     z faces nx*ny*(nz+1)  (src/sim/solver.h:216-220).
 *   - Geometry arguments (phg_grid_create) are HOST pointers; field and
 *     right-hand-side arguments of the compute entry points are DEVICE
 *     pointers (from phg_malloc or any HIP allocation on the context's
 *     device).  Work is queued on the context's stream; entry points that
 *     return statistics synchronise that stream before returning.
 *   - Every entry point returns 0 on success or a negative PHG_E* code;
 *     phg_last_error() then holds a message (per context, not thread-safe
 *     across threads sharing one context).  Contexts are independent: the
 *     library's only global mutable state is the per-device workspace of
 *     phg_zone_mean_noalloc.
 */
#ifndef PHOTON_GAS_H
#define PHOTON_GAS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PHG_OK = 0,
    PHG_EINVAL = -1,   /* bad argument or shape */
    PHG_EHIP = -2,     /* a HIP call failed */
    PHG_ENOMEM = -3,   /* device allocation failed */
    PHG_ESTATE = -4    /* unknown handle */
};

/* storage precision of the multigrid preconditioner */
enum {
    PHG_PREC_FP32 = 0,       /* all levels FP32 (the reference's precision) */
    PHG_PREC_MIXED_F16 = 1,  /* V-cycle levels and level-0 coefficients FP16, compute FP32 */
    PHG_PREC_MIXED_BF16 = 2  /* same with BF16 storage */
};

/* stop rules */
enum {
    PHG_STOP_REL_L2 = 0,   /* ||r||_2 / ||b||_2 <= tol (FP64 norms of the recursive residual) */
    PHG_STOP_MAXNORM = 1   /* max|r| / inv <= tol in 1/s, the reference's rule (gasphase.cpp:592-605) */
};

/* solve status: breakdown = d.Ad <= 0 (gasphase.cpp:607) or r.z <= 0 (:605);
 * nonfinite = the residual norm became NaN or infinite (a non-finite
 * right-hand side or warm start, or an overflow): the system stops there and
 * is never restarted.  The reference's loop test `res > tol` also ends on a
 * NaN residual, but reports nothing. */
enum {
    PHG_RUNNING = 0,
    PHG_CONVERGED = 1,
    PHG_MAXITER = 2,
    PHG_BREAKDOWN = 3,
    PHG_BREAKDOWN_RZ = 4,
    PHG_NONFINITE = 5
};

typedef struct phg_ctx phg_ctx;

typedef struct {
    double tol;          /* stop tolerance (see stop rule) */
    int max_iters;       /* iteration cap (FidelityProfile::pressureMaxIters, solver.h:89) */
    int stop_rule;       /* PHG_STOP_* */
    float inv;           /* dx^2/dt, for PHG_STOP_MAXNORM (must then be positive and finite) */
    int demean;          /* remove the mean of x on closed zones at the end (gasphase.cpp:628-639) */
    int check_every;     /* iterations between host convergence checks (0 = library default) */
    int use_graph;       /* 0: launch directly; 1: replay one captured CG iteration; 2: the set-up and
                            the first check_every iterations as one captured graph (one launch per
                            solve when every system stops within them), then as 1 with a check every
                            max(4, check_every / 4) iterations (at most check_every) */
    int zero_x;          /* start from x = 0 (x is written, not read) instead of warm-starting */
    int refine;          /* restarts allowed from the recomputed residual b - A x: in FP32 the
                            recursive residual can reach tol while the true one stalls higher; a
                            system whose true residual misses tol (and stands clear of FP32 noise)
                            is restarted from its x (0 = none), and so is one that broke down
                            (d.Ad <= 0 or r.z <= 0); it keeps PHG_BREAKDOWN* unless a restart
                            converges or finds its residual at the FP32 noise floor.  With
                            refine > 0 every residual formed from x, a warm start's included, is
                            accumulated in FP64 */
    float coarse_weight; /* weight of every coarse-grid correction of the V-cycle (the prolongation
                            adds w times the coarse level's correction), 0 <= w < 2; 0 or 1 keeps
                            the reference's cycle (pressure_mg.cpp:293-295) bit for bit.  Piecewise-
                            constant aggregation under-corrects smooth errors: w = 1.8 cuts the
                            iterations to 1e-6 on walled grids by a third or more (29.5 -> 19 on the
                            80x50x40 house, 38.6 -> 19.6 on the 160x100x80 one).  Not the reference's method, so its iteration counts differ;
                            the solution meets the same stop rule */
} phg_solve_params;

typedef struct {
    int iters;           /* CG iterations (Solver::pressureIters, solver.h:277) */
    int status;          /* PHG_CONVERGED / PHG_MAXITER / PHG_BREAKDOWN* / PHG_NONFINITE */
    double relres;       /* final recursive ||r||/||b|| */
    double maxres;       /* final max|r| / inv (Solver::pressureResidual, solver.h:278) */
} phg_solve_stats;

/* ---- context --------------------------------------------------------- */
/* Replaces the JobPool the reference Solver owns (solver.h:291): one HIP
 * stream on `device`. */
int phg_create(phg_ctx **out, int device);
void phg_destroy(phg_ctx *ctx);
const char *phg_last_error(phg_ctx *ctx);
int phg_synchronize(phg_ctx *ctx);
void *phg_stream(phg_ctx *ctx);   /* the hipStream_t work is queued on */

/* ---- device memory helpers (for callers without a HIP runtime of their own) */
int phg_malloc(phg_ctx *ctx, void **dptr, size_t bytes);
int phg_free(phg_ctx *ctx, void *dptr);
int phg_memcpy_h2d(phg_ctx *ctx, void *dst, const void *src, size_t bytes);
int phg_memcpy_d2h(phg_ctx *ctx, void *dst, const void *src, size_t bytes);
int phg_memset(phg_ctx *ctx, void *dst, int value, size_t bytes);

/* ---- geometry: one grid and its multigrid hierarchy -------------------
 * Replaces Solver::buildPressureLevels (pressure_mg.cpp:28-191, run when
 * mgDirty_ is set) and computeZones (gasphase.cpp:111-158).  ax/ay/az are the
 * matrix coefficients (Solver::ax_/ay_/az_, solver.h:296), fx/fy/fz the face
 * open fractions, blocked the solid mask; wind_closed (4 bytes: x-low,
 * x-high, y-low, y-high, or NULL) flags sides where windAt() prescribes
 * inflow (gasphase.cpp:52-61).  precision: PHG_PREC_*.  */
int phg_grid_create(phg_ctx *ctx, int nx, int ny, int nz, float dx, const float *fx, const float *fy,
                    const float *fz, const float *ax, const float *ay, const float *az, const uint8_t *blocked,
                    const uint8_t *wind_closed, int precision, int *grid_out);
int phg_grid_destroy(phg_ctx *ctx, int grid);
/* The same set-up in two halves, so a server can build a grid's host half
 * while the device runs other work (no counterpart in the reference, whose
 * job runner builds one Solver at a time, src/app/cli.cpp:23-101):
 * phg_grid_plan_build runs the host work of phg_grid_create (the hierarchy,
 * its dense coarsest inverse and the zones) with no context and no device;
 * calls on distinct plans may run on several threads at once.
 * phg_grid_create_planned is phg_grid_create taking that work from the plan
 * (which must come from the same nx, ny, nz, ax, ay, az, blocked and
 * wind_closed; a plan serves one grid), and yields the same grid.  A plan is
 * freed with phg_grid_plan_free, used or not. */
typedef struct phg_grid_plan phg_grid_plan;
int phg_grid_plan_build(int nx, int ny, int nz, const float *ax, const float *ay, const float *az,
                        const uint8_t *blocked, const uint8_t *wind_closed, phg_grid_plan **plan_out);
void phg_grid_plan_free(phg_grid_plan *plan);
int phg_grid_create_planned(phg_ctx *ctx, int nx, int ny, int nz, float dx, const float *fx, const float *fy,
                            const float *fz, const float *ax, const float *ay, const float *az,
                            const uint8_t *blocked, const uint8_t *wind_closed, int precision,
                            phg_grid_plan *plan, int *grid_out);
/* How the V-cycle solves its coarsest level (pressure_mg.cpp:277-281).
 * PHG_COARSE_EXACT (the default where it fits): the dense FP64 projected
 * inverse, formed once per grid; a better preconditioner than the
 * reference's, so CG needs fewer iterations but its iterates differ.
 * PHG_COARSE_GS20: the reference's own 20 forward/backward Gauss-Seidel
 * sweeps from zero, so the iterates, pressureIters and pressureResidual
 * follow the reference's to rounding; also the only mode of grids whose
 * coarsening stalls above 2048 coarsest unknowns (dense inverse refused;
 * up to 16384).  The mode applies to batches created after the call; the
 * grid's own cached one-system batch (phg_solve, phg_project, ...) follows
 * it at once. */
enum { PHG_COARSE_EXACT = 0, PHG_COARSE_GS20 = 1 };
int phg_grid_set_coarse(phg_ctx *ctx, int grid, int mode);
int phg_grid_get_coarse(phg_ctx *ctx, int grid, int *mode);
/* hierarchy shape: level count, sizes (nlevels entries), zones */
int phg_grid_info(phg_ctx *ctx, int grid, int *nlevels, int *level_n, int *nzones, int *ref_levels);
/* zone labels (N ints, -1 solid) and open flags (nzones bytes), host out */
int phg_grid_zones(phg_ctx *ctx, int grid, int *zone, uint8_t *zone_open);
/* Host-only: the hierarchy phg_grid_create would build (no device needed).
 * level_n: up to 12 entries; agg0 (N ints, may be NULL): the level-0 -> 1 map. */
int phg_hierarchy_shape(int nx, int ny, int nz, const float *ax, const float *ay, const float *az,
                        const uint8_t *blocked, const uint8_t *wind_closed, int *nlevels, int *level_n,
                        int *agg0);

/* ---- level-0 operator pieces (tests and callers of single kernels) ----
 * y = A x with the fused FP64 dot x.y  (the apply lambda, gasphase.cpp:551-570) */
int phg_apply(phg_ctx *ctx, int grid, const float *x, float *y, double *dot);
/* z = M^-1 r, one symmetric V-cycle (Solver::pressurePrecondition, pressure_mg.cpp:270-300) */
int phg_precondition(phg_ctx *ctx, int grid, const float *r, float *z);

/* ---- batched MG-PCG ------------------------------------------------------
 * A batch is a fixed set of systems (each on a grid, grids may differ or
 * repeat) that advance together, one kernel launch per step for all of
 * them.  The batch owns every work vector; the caller owns b and x. */
int phg_batch_create(phg_ctx *ctx, int nsys, const int *grids, int *batch_out);
/* tile shape of a batch (chosen per size class by the autotuner) */
typedef struct {
    int kchunk;          /* level-0 planes per block along z (0 = from level0_blocks) */
    int level0_blocks;   /* level-0 blocks per launch to aim for when kchunk = 0 (0 = 4096, with chunks
                            of at least 32 planes when the batch's 256-column tiles alone number 512
                            or more) */
    int unfused;         /* 1: level-0 smoothing as separate half sweeps instead of the two fused
                            passes (same results bit for bit; for comparison) */
    int sweeps0;         /* level-0 red-black sweeps before and after the coarse correction
                            (0 = 1, the reference's gsColor pair, pressure_mg.cpp:276,296; at most 4) */
    const int *sys_kchunk;   /* per system (nsys entries, 0 = batch value) or NULL */
    const int *sys_sweeps0;  /* per system (nsys entries, 0 = batch value) or NULL */
    int fused_threads;   /* threads (and own columns) per block of the level-0 passes that stage
                            halo rows (the fused smoothing passes, the direction update):
                            0 = 512 (1024 for the smoothing of grids wider than 128 cells), or
                            256, 512 or 1024 for every such pass (1024: the direction update
                            keeps 512); every choice gives the same results bit for bit */
    int coarse_lds;      /* coarse levels in LDS: from its first coarse level whose vectors fit
                            one block's LDS, a system's coarse levels run as one down and one up
                            launch, one block per system, instead of several launches per level
                            (same results bit for bit).  0: when at least two such blocks per CU
                            would run (fewer leave the chip idle); 1: never; 2: always */
} phg_batch_opts;
int phg_batch_create_ex(phg_ctx *ctx, int nsys, const int *grids, const phg_batch_opts *opts, int *batch_out);
int phg_batch_destroy(phg_ctx *ctx, int batch);
/* The CG loop of Solver::project (gasphase.cpp:571-627) for every system:
 * x (warm start in, solution out) and b are arrays of nsys device pointers.
 * stats: nsys entries (may be NULL). */
int phg_batch_solve(phg_ctx *ctx, int batch, const float *const *b, float *const *x, const phg_solve_params *params,
                    phg_solve_stats *stats);
/* Captures (without launching) the whole-solve graph phg_batch_solve replays
 * for use_graph = 2 (the set-up and `iterations` CG iterations, refine as in
 * phg_solve_params) and checks every node against the device's launch
 * limits: a kernel node, a non-empty grid, threads per block, LDS, grid
 * size, no scratch.  nodes_out: the node count; msg: a summary (may be NULL).
 * Returns PHG_EINVAL naming the first node outside the limits.  No
 * counterpart in the reference (a check on this library's own launches). */
int phg_batch_graph_check(phg_ctx *ctx, int batch, int iterations, int refine, int *nodes_out, char *msg,
                          size_t msg_len);
/* one system: convenience over a cached one-system batch of `grid` */
int phg_solve(phg_ctx *ctx, int grid, const float *b, float *x, const phg_solve_params *params,
              phg_solve_stats *stats);

/* ---- the projection step (Solver::project, gasphase.cpp:505-677) ------
 * u,v,w (faces) and p (cells) in/out, S in; all device pointers.  Builds the
 * RHS (:528-549), solves (:571-627, warm start from p), removes closed-zone
 * means (:628-639) and corrects the face velocities (:640-676).
 * max_vel out: Solver::maxVel.  */
int phg_project(phg_ctx *ctx, int grid, float dt, float *u, float *v, float *w, float *p, const float *S,
                const phg_solve_params *params, phg_solve_stats *stats, float *max_vel);
/* the pieces of phg_project: RHS + diagonal (b, diag device out; p zeroed in solids) */
int phg_rhs(phg_ctx *ctx, int grid, float dt, const float *u, const float *v, const float *w, const float *S,
            float *b, float *diag, float *p);
int phg_zone_mean(phg_ctx *ctx, int grid, float *p);
/* phg_zone_mean without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same result bit for bit as phg_zone_mean;
 * p is a device pointer, in/out.  The work is queued on the context's stream
 * and the call returns without waiting.  The partial sums and zone means go
 * to a fixed workspace of 2048 doubles (16 KiB) that the library holds once
 * per device: a grid needs one entry per closed-zone task (8192 cells) plus
 * one per closed zone, and a larger need returns PHG_EINVAL.  Every call on a
 * device shares that workspace, so calls that overlap on one device (two
 * contexts' streams, or two threads) race. */
int phg_zone_mean_noalloc(phg_ctx *ctx, int grid, float *p);
/* The compatible divergence source of Solver::leakage (gasphase.cpp:487-498)
 * for grids without crack faces: on each closed zone S -= (integral of S) /
 * zone volume, i.e. its mean, so the pressure equation of a sealed room is
 * solvable (the zone's pressure rise rate is gamma P times that shift).  S:
 * device, in/out.  shift (host, one float per zone, 0 on open zones) may be
 * NULL; when given, the call waits for the device. */
int phg_leakage(phg_ctx *ctx, int grid, float *S, float *shift);
int phg_correct(phg_ctx *ctx, int grid, float dt, const float *p, float *u, float *v, float *w, float *max_vel);

/* ---- divergence source (Solver::energyAndExpansion, gasphase.cpp:251-271)
 * T in/out, S out, Pb/Q/massSrc in (cells, device). max_div out: Solver::maxDiv. */
int phg_energy_expansion(phg_ctx *ctx, int grid, float dt, float *T, float *S, const float *Pb, const float *Q,
                         const float *massSrc, float *max_div);

/* ---- scalar transport (Solver::advectScalars, gasphase.cpp:685-784) ----
 * fields_in / fields_out: 10 device pointers each, in the order
 * T, Yo, Yf, Ys, Ya, Yp, Yc, Yw, Yco, Yai (gasphase.cpp:687).  Out must not
 * alias in (the reference writes tmp_ and swaps). */
int phg_advect_scalars(phg_ctx *ctx, int grid, float dt, const float *u, const float *v, const float *w,
                       const float *Pb, const float *nu, const float *const *fields_in, float *const *fields_out);

/* ---- the rest of the gas step (Solver::step, solver.cpp:226-269) ----------
 * The phases around the projection, each the device form of the reference
 * member it names; fields are device pointers on the grid's cells or faces.
 * Solver::step's order is: phg_step_clear, burners (phg_add_gas),
 * phg_combustion, phg_turbulence, phg_heat_diffusion, phg_energy_expansion,
 * phg_momentum, leakage (phg_zone_integrals + phg_zone_values), phg_project,
 * phg_advect_scalars, then the closed zones' background pressure
 * (phg_zone_values into Pb); photon_amd/host/gas_phase.hpp runs them. */

/* Wind at the open boundaries (Solver::wind, solver.h:106-111) and the height
 * of the grid's floor (Solver::oz); speed <= 0 means no wind.  The sides it
 * enters through must match the wind_closed flags of the grid. */
typedef struct {
    float speed;      /* m/s at 10 m height */
    float from_deg;   /* direction it blows from, degrees clockwise from north (+y) */
    float exponent;   /* power-law profile exponent */
    float oz;         /* m */
} phg_wind;

/* The gas state of one grid's cells that the source phases update (device
 * pointers; Solver's members of the same names, solver.h:229-247). */
typedef struct {
    float *T, *Yo, *Yf, *Ys, *Ya, *Yp, *Yc, *Yw, *Yco, *Yai;
    float *Pb;        /* background pressure of the cell's zone, Pa */
    float *massSrc;   /* kg of gas added this step */
    float *Q, *Qr;    /* gas heat and flame radiant energy this step, J */
    float *hrr;       /* heat release rate, W/m3 */
    uint8_t *pilot, *lit;
} phg_gas;

/* stableDt (gasphase.cpp:63-70): the step size from the last step's maxVel,
 * maxNu and maxDiv and the profile's dtMax and cfl (host only). */
int phg_stable_dt(float dt_max, float profile_dt_max, float cfl, float dx, float max_vel, float max_nu, float max_div,
                  float *dt_out);
/* Solver::turbulence (gasphase.cpp:163-218): Deardorff eddy viscosity nu from
 * the face velocities and T; max_nu (host, may be NULL): Solver::maxNu. */
int phg_turbulence(phg_ctx *ctx, int grid, const float *u, const float *v, const float *w, const float *T, float *nu,
                   float *max_nu);
/* Solver::heatDiffusion (gasphase.cpp:221-248): conduction into Q. */
int phg_heat_diffusion(phg_ctx *ctx, int grid, float dt, const float *T, const float *nu, const float *Pb, float *Q);
/* Solver::momentum (gasphase.cpp:275-399): new face velocities into
 * u_out/v_out/w_out (must not alias u/v/w; the reference swaps u2_ in).
 * wind may be NULL.  Leak faces keep their velocity. */
int phg_momentum(phg_ctx *ctx, int grid, float dt, const phg_wind *wind, const float *u, const float *v, const float *w,
                 const float *T, const float *Pb, const float *nu, const float *p, float *u_out, float *v_out,
                 float *w_out);
/* The start of Solver::step (solver.cpp:231-234): Q, Qr, pilot, massSrc cleared. */
int phg_step_clear(phg_ctx *ctx, int grid, const phg_gas *gas);
/* Solver::addGasSpecies (solver.cpp:213-224) on each of ncells device-listed
 * cells (distinct; solid ones skipped), pilot set when set_pilot: the
 * burners' injection (injectBurners, gasphase.cpp:92-105). */
int phg_add_gas(phg_ctx *ctx, int grid, int ncells, const int *cells, float fuel_gen, float aerosol, float soot_pot,
                float ait, int set_pilot, const phg_gas *gas);
/* Solver::combustion (solver.cpp:284-344); hrr_total (host, may be NULL):
 * the summed heat release rate, W (FP64 over fixed-order block sums). */
int phg_combustion(phg_ctx *ctx, int grid, float dt, int radiation_on, const phg_gas *gas, double *hrr_total);
/* The maxima the last phg_correct / phg_project, phg_turbulence and
 * phg_energy_expansion calls on the grid left on the device (maxVel, maxNu,
 * maxDiv), read in one wait, for callers that queue those calls without their
 * host outputs.  Any pointer may be NULL. */
int phg_grid_stats(phg_ctx *ctx, int grid, float *max_vel, float *max_nu, float *max_div);
/* Closed zones' integral of S dV (leakage, gasphase.cpp:405-409), FP64, into
 * dint (host, one per zone; 0 for open zones). */
int phg_zone_integrals(phg_ctx *ctx, int grid, const float *S, double *dint);
/* A value per zone (host, one per zone) applied to the cells of the closed
 * zones: x -= value (subtract = 1, leakage's shift of S, gasphase.cpp:494-498)
 * or x = value (subtract = 0, Pb from zoneP, solver.cpp:262-263). */
int phg_zone_values(phg_ctx *ctx, int grid, const float *values, float *x, int subtract);
/* The grid's crack faces: Solver::lx_, ly_, lz_ (solver.h:297), which
 * Solver::updateFace (solver.cpp:93-100) sets where an intact barrier is open
 * by less than 0.05 (such a face has matrix coefficient 0).  One byte per
 * face of each axis (host; NULL: none on that axis); all three NULL clears
 * them.  A crack with a solid side is ignored (gasphase.cpp:424).  Momentum
 * keeps crack velocities, which phg_zone_leakage drives. */
int phg_grid_set_cracks(phg_ctx *ctx, int grid, const uint8_t *lx, const uint8_t *ly, const uint8_t *lz);
/* Face changes of a grid (barrier failure, a door opened or closed): the
 * results of Solver::updateFace (solver.cpp:66-101) for n faces (host
 * arrays): axis (0..2), face index, open fraction f, matrix coefficient a,
 * crack flag.  Velocities (device, may be NULL) of faces that close (f = 0)
 * are zeroed, as updateFace does.  When a matrix coefficient changed, the
 * zones and the multigrid hierarchy are set up again under the same id,
 * precision, wind sides and coarsest-solve mode (computeZones and
 * buildPressureLevels on the reference's zonesDirty_ / mgDirty_,
 * gasphase.cpp:111-158, pressure_mg.cpp:28-191); batches made earlier over
 * this grid are stale then and their solves return PHG_ESTATE.  Otherwise
 * only the open fractions and the crack faces change.  A rebuild forms
 * levels 0 and 1 of the hierarchy (the work that grows with the cells) on the
 * device, bit for bit as the host set-up does, and the coarser levels, the
 * coarsest solve and the zones on the host.  rebuilt (may be NULL): 0 when
 * the set-up stood, 2 when it was redone with the device-formed levels, 1
 * when it was redone wholly on the host (a grid beyond the device build's
 * limits). */
int phg_grid_update_faces(phg_ctx *ctx, int grid, int n, const int *axis, const int *face, const float *f,
                          const float *a, const uint8_t *crack, float *u, float *v, float *w, int *rebuilt);
/* Solver::leakage (gasphase.cpp:402-498): each closed zone's end-of-step
 * background pressure against its cracks (orifice flow, 40-step bisection),
 * every crack velocity relaxed halfway to its orifice velocity, each closed
 * zone's rise rate and the matching shift of S.  zoneP (host, one per zone)
 * the zones' background pressures; zoneDP (host, one per zone, may be NULL)
 * receives the rise rates (0 for open zones); leak_faces (may be NULL) the
 * number of crack faces (Solver::leakFaces).  T, Pb, p read; u, v, w (crack
 * faces) and S updated.  Without cracks this is the closed zones' mean
 * shift of S. */
int phg_zone_leakage(phg_ctx *ctx, int grid, float dt, const float *zoneP, const float *T, const float *Pb,
                     const float *p, float *u, float *v, float *w, float *S, float *zoneDP, int *leak_faces);
/* Zone background pressures kept on the device.  The reference keeps
 * Solver::zoneP / zoneDP_ in host vectors (solver.h:243,294) and the calls
 * above take them from the host, which makes every step wait for the device
 * twice; a caller that keeps a scenario's whole state on the device (the
 * serving loop) keeps them in the grid instead, and these calls queue work
 * without waiting.  phg_zone_pressures_set loads them (host, one per zone;
 * rise rates reset to 0) and phg_zone_pressures_get reads them back (and the
 * rise rates, 0 for open zones; zoneDP may be NULL), waiting.
 * phg_zone_leakage_resident is phg_zone_leakage on them (the rise rates stay
 * in the grid); phg_zone_pressure_resident is the end of Solver::step
 * (solver.cpp:258-263): each closed zone's pressure follows its rise rate,
 * clamped to [P0 - 20 kPa, P0 + 50 kPa], and Pb of its cells takes it.  The
 * results equal the host path's bit for bit.  phg_zone_leakage (host
 * pressures) and a face change that rebuilds the zones replace them. */
int phg_zone_pressures_set(phg_ctx *ctx, int grid, const float *zoneP);
int phg_zone_pressures_get(phg_ctx *ctx, int grid, float *zoneP, float *zoneDP);
int phg_zone_leakage_resident(phg_ctx *ctx, int grid, float dt, const float *T, const float *Pb, const float *p,
                              float *u, float *v, float *w, float *S);
int phg_zone_pressure_resident(phg_ctx *ctx, int grid, float dt, float *Pb);

/* ---- measurement aid: a plain device copy (16-byte vector loads/stores),
 * the calibration kernel for HBM byte counters; bytes must be a multiple of 16 */
int phg_copy(phg_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- profiling: HIP-event timing of each kernel class on the context stream */
int phg_profile_enable(phg_ctx *ctx, int on);
/* time only the kernel classes whose bit (their index below) is set; all by
 * default.  Each timed launch is bracketed by two events on the stream, so a
 * run that times only the classes it reports stays close to an untimed one. */
int phg_profile_select(phg_ctx *ctx, unsigned long long class_mask);
/* kernel class names are returned by index; returns PHG_EINVAL past the last */
int phg_profile_get(phg_ctx *ctx, int index, const char **name, double *total_ms, long long *launches,
                    double *bytes);
int phg_profile_reset(phg_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
