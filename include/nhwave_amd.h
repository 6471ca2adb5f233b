/*
 * nhwave_amd -- MI355X (gfx950) implementation of the NHWAVE time-stepping hot
 * path behind a plain C ABI.
 *
 * The reference's hot path is a set of argument-less Fortran subroutines that
 * mutate `module GLOBAL` in place (src/mod_global.F:1-370).  This ABI keeps one
 * device-resident copy of that state per context; host arrays cross the
 * boundary only through nhw_upload/nhw_download, in the reference's own
 * Fortran layout (1-based (i,j,k), i fastest, Nghost = 2 ghost cells):
 *
 *   2-D cell    (Mloc ,Nloc )          Mloc  = Mglob+4, Nloc = Nglob+4
 *   x-face 2-D  (Mloc1,Nloc )          Mloc1 = Mloc+1
 *   y-face 2-D  (Mloc ,Nloc1)          Nloc1 = Nloc+1
 *   3-D cell    (Mloc ,Nloc ,Kloc )    Kloc  = Kglob+4
 *   z-face 3-D  (Mloc ,Nloc ,Kloc1)    Kloc1 = Kloc+1   (P, Omega)
 *   (src/initialize.F:1381-1386, 1423-1504)
 *
 * Every entry point returns 0 on success and a negative NHW_E* code on
 * failure; nhw_last_error() then holds a message.  Calls on one context are
 * not thread-safe; different contexts are independent.
 */
#ifndef NHWAVE_AMD_H
#define NHWAVE_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  NHW_OK = 0,
  NHW_EINVAL = -1,      /* bad argument or unsupported option */
  NHW_EHIP = -2,        /* HIP runtime error */
  NHW_EDT = -3,         /* dt < DT_MIN (reference: `stop`, src/nhwave.F:3840-3850) */
  NHW_ESOLVER = -4,     /* Krylov breakdown or non-finite residual */
  NHW_ENOMEM = -5,      /* device allocation failed / admission refused */
  NHW_ENOTFINITE = -6   /* non-finite state detected */
};

enum { NHW_FP64 = 0, NHW_MIXED = 1 };

/* Poisson solver of the non-hydrostatic step.  The reference uses NSPCG
 * IC(0)+GMRES (src/pressure.F:86-131); both choices here solve the same
 * system to the same relative residual TOL.
 *  NHW_SOLVER_MG:   BiCGStab, matrix-free operator, geometric multigrid
 *                   V(1,1) preconditioner with z-line smoothing (Kglob <= 16)
 *  NHW_SOLVER_LINE: BiCGStab on the assembled 15-diagonal matrix with a
 *                   vertical-line (tridiagonal) preconditioner */
enum { NHW_SOLVER_MG = 0, NHW_SOLVER_LINE = 1 };

/* WAVEMAKER values on the GPU path: the west-boundary wavemakers of
 * update_wave_bc (src/nhwave.F:5514-5562) -- solitary_wave_left_boundary,
 * linear_wave_left_boundary, stokes_wave_left_boundary and
 * cnoidal_wave_left_boundary (src/waves.F:869, 808, 518, 92); each needs
 * BC_X0 = 3 (inflow) -- and the internal irregular wavemaker INT_IRR, a
 * mass source over a box (src/source_terms.F:194-305), whose box and
 * spectrum nhw_set_internal_wavemaker supplies -- and FLUX_LR
 * (flux_left_right_boundary, src/nhwave.F:5565-5712): a logarithmic inflow
 * profile at the west and an open east boundary with Orlanski radiation;
 * it needs BC_X0 = BC_Xn = 3 (FP64, single contexts). */
enum {
  NHW_WAVE_NONE = 0, NHW_WAVE_LEF_SOL = 1, NHW_WAVE_LEF_LIN = 2, NHW_WAVE_LEF_STK = 3, NHW_WAVE_LEF_CON = 4,
  NHW_WAVE_INT_IRR = 5, NHW_WAVE_FLUX_LR = 6
};

/* Run-time configuration: the `input.txt` keys of the hot path
 * (src/initialize.F:166-555).  Units and meaning as in the reference. */
typedef struct nhw_config {
  int32_t mglob, nglob, kglob;     /* Mglob, Nglob, Kglob */
  double dx, dy;                   /* DX, DY */
  double cfl;                      /* CFL */
  double min_dep;                  /* MinDep */
  double dt_ini, dt_min, dt_max;   /* DT_INI, DT_MIN, DT_MAX */
  double froude_cap;               /* FROUDE_CAP */
  int32_t froude_cap_on;           /* compiled with -DFROUDE_CAP */
  double visc;                     /* VISCOSITY (the surface stress BC; the laminar viscosity of
                                      the viscous branch) */
  double fcor;                     /* Coriolis parameter as the reference derives it from slat */
  double wind_u, wind_v;           /* WindU, WindV (Iws = 1) */
  int32_t time_order;              /* It_Order: 1, 2 (SECOND) or 3 (THIRD) */
  int32_t non_hydro;               /* NON_HYDRO */
  int32_t hllc;                    /* HLLC (only F is supported) */
  int32_t bc_x0, bc_xn, bc_y0, bc_yn, bc_z0; /* BC_X0..BC_Z0 (lateral 1/2; 3 at X0 with a
                                                west wavemaker or FLUX_LR, at Xn with
                                                FLUX_LR; bottom 1 free slip, 2 no slip,
                                                5 friction law) */
  int32_t itmax;                   /* ITMAX */
  double tol;                      /* TOL: relative residual |b-Ax|/|b| */
  int32_t precision;               /* NHW_FP64 or NHW_MIXED */
  int32_t solver;                  /* NHW_SOLVER_MG (default) or NHW_SOLVER_LINE */
  /* boundary forcing (src/initialize.F:537, 601-613, 946-957) */
  int32_t wavemaker;               /* WAVEMAKER: NHW_WAVE_* */
  double amp_wave, per_wave, dep_wave, theta_wave;  /* AMP, PER, DEP, THETA */
  double cur_wave, tramp;          /* CUR, TRAMP (the sponge's current and its ramp) */
  int32_t sponge_on;               /* SPONGE_ON */
  double sponge_w, sponge_e, sponge_s, sponge_n;  /* Sponge_{West,East,South,North}_Width */
  /* viscous branch (src/initialize.F:357, 385-393): VISCOUS_FLOW = T with
   * IVTURB = IHTURB = 1 (constant viscosities) or the k-epsilon closure
   * (`turb` below); FP64 or mixed (single contexts and batch members), Kglob <= 32 */
  int32_t viscous;                 /* VISCOUS_FLOW */
  int32_t rng;                     /* RNG (coefficient of the initial turbulent viscosity) */
  double cvs, chs, schmidt, visc_number;  /* Cvs, Chs, Schmidt, VISCOUS_NUMBER */
  /* bottom friction of BC_Z0 = 5 (src/initialize.F:466-468) */
  int32_t ibot;                    /* Ibot: 1 constant Cd0, otherwise the log law with Zob */
  double cd0, zob;                 /* Cd0, Zob */
  /* turbulence model of the viscous branch (src/nhwave.F:6544-6782): 1 (or 0)
   * IVTURB = IHTURB = 1, constant Cvs and Chs; 10 IVTURB = 10 with IHTURB >= 10,
   * the 3-D k-epsilon model kepsilon_3D (:10030-11114) with CmuHt = CmuVt,
   * computed as the reference's serial build executes it (Kglob >= 3) */
  int32_t turb;
} nhw_config;

/* Host views of GLOBAL arrays (any pointer may be NULL = skip). */
typedef struct nhw_fields {
  double *Eta, *Eta0, *D, *D0, *Hc, *DelxH, *DelyH, *DelxEta, *DelyEta;
  int32_t *Mask, *Mask_Struct, *Mask9;
  double *Hfx, *Hfx0, *EtaxL, *EtaxR;          /* (Mloc1,Nloc) */
  double *Hfy, *Hfy0, *EtayL, *EtayR;          /* (Mloc,Nloc1) */
  double *U, *V, *W, *DU, *DV, *DW, *DU0, *DV0, *DW0;
  double *P, *Omega;                           /* (Mloc,Nloc,Kloc1) */
  double *SourceX, *SourceY, *Wsx, *Wsy;
  double *Ex, *Ey, *Fx, *Fy, *Gx, *Gy, *Hx, *Hy; /* face fluxes */
  double *Fz, *Gz, *Hz;                        /* (Mloc,Nloc,Kloc1) */
  /* k-epsilon state (turb = 10), (Mloc,Nloc,Kloc): the eddy viscosity CmuVt
   * (CmuHt equals it), Tke, Eps and DTke = D Tke, DEps = D Eps; a context
   * without the k-epsilon model holds none of them and refuses a non-NULL
   * pointer */
  double *CmuVt, *Tke, *Eps, *DTke, *DEps;
} nhw_fields;

typedef struct nhw_solver_stats {
  int32_t last_iters;      /* Krylov iterations of the last Poisson solve */
  double last_relres;      /* |b-Ax|/|b| at exit of the last solve */
  int64_t total_iters;     /* since creation */
  int64_t solves;          /* Poisson solves since creation */
  int32_t max_iters;       /* largest iteration count of any solve */
  int32_t not_converged;   /* solves that hit ITMAX */
} nhw_stats;

typedef struct nhw_ctx nhw_ctx;

const char *nhw_last_error(void);
/* Device-resident state for one scenario; binds to the current HIP device.
 * Replaces allocate_variables for the hot-path arrays
 * (src/initialize.F:1423-1643). */
int nhw_create(const nhw_config *cfg, nhw_ctx **out);
int nhw_destroy(nhw_ctx *ctx);
/* device bytes nhw_create would allocate for cfg (no allocation; no
 * reference counterpart: the admission test of a serving loop) */
int nhw_footprint(const nhw_config *cfg, int64_t *bytes);
/* Vertical grid from generate_grid (src/initialize.F:1700-1727):
 * dsig[Kloc], sig[Kloc1], sigc[Kloc]. */
int nhw_set_grid(nhw_ctx *ctx, const double *dsig, const double *sig, const double *sigc);
int nhw_upload(nhw_ctx *ctx, const nhw_fields *host);
int nhw_download(nhw_ctx *ctx, const nhw_fields *host);
int nhw_set_time(nhw_ctx *ctx, double time, double dt, int32_t run_step);
int nhw_get_time(nhw_ctx *ctx, double *time, double *dt, int32_t *run_step);

/* Tail of `initial` (src/initialize.F:3118-3170): Mask, Eta0, Mask9, D,
 * vel_bc, DU/DV/DW, from uploaded Hc/Eta/U/V/W (Hc ghosts filled). */
int nhw_initial(nhw_ctx *ctx);

/* The reference's hot-path subroutines (argument-less Fortran externals). */
int nhw_estimate_dt(nhw_ctx *ctx);          /* estimate_dt,  src/nhwave.F:3694-3874 */
int nhw_update_mask(nhw_ctx *ctx);          /* update_mask,  src/nhwave.F:5403-5456 */
int nhw_update_wind(nhw_ctx *ctx);          /* update_wind,  src/nhwave.F:5333-5400 */
int nhw_update_vars(nhw_ctx *ctx);          /* update_vars,  src/nhwave.F:5459-5511 */
int nhw_source_terms(nhw_ctx *ctx);         /* source_terms, src/source_terms.F:1-456 */
int nhw_fluxes(nhw_ctx *ctx);               /* fluxes,       src/nhwave.F:1686-1714 */
int nhw_eval_duvw(nhw_ctx *ctx, int32_t istep); /* eval_duvw(ISTEP), src/nhwave.F:539-1300 */
int nhw_poisson_solver(nhw_ctx *ctx);     /* poisson_solver, src/pressure.F:63-245 */
/* generate_coef_rhs (src/pressure.F:475-840) on the current state, into host
 * arrays: coef[15*neq] diagonal-major in the order of JCoef (:817-831), rhs[neq],
 * rows i fastest, then j, then k over the interior (neq = Mglob*Nglob*Kglob);
 * the reference's Coef(1:neq,1:15) and Rhs.  Any context. */
int nhw_generate_coef_rhs(nhw_ctx *ctx, double *coef, double *rhs);
/* y = A x with the multigrid solver's matrix-free FP64 operator on the current
 * state (the product NSPCG forms with Coef, src/nspcg.f), host vectors in the
 * row order above; NON_HYDRO contexts of NHW_SOLVER_MG and the mixed form. */
int nhw_poisson_apply(nhw_ctx *ctx, const double *x, double *y);
/* update_wave_bc (src/nhwave.F:5514-5562): the west inflow values of the
 * configured wavemaker at the device TIME; a no-op without a west wavemaker. */
int nhw_update_wave_bc(nhw_ctx *ctx);
/* The internal irregular wavemaker (WAVEMAKER = INT_IRR) of a context: the
 * source box Xsource_West/East, Ysource_Suth/Nrth (src/initialize.F:629-634)
 * and the spectrum the reference reads from irr2d.txt (:835-878): nfreq
 * frequencies (Hz), ndir directions (degrees), Wave_Spc2d and Random_Phs as
 * (ndir, nfreq) arrays, direction fastest.  It replaces the set-up that
 * source_terms (src/source_terms.F:194-305) repeats at every call; the
 * device then evaluates the mass source SourceC of every stage.  Required
 * before the first step of an INT_IRR context (a batch member: before the
 * batch's next step); copied, the arrays may be freed on return. */
int nhw_set_internal_wavemaker(nhw_ctx *ctx, double xs_west, double xs_east, double ys_south, double ys_north,
                               int32_t nfreq, int32_t ndir, const double *freq, const double *dire,
                               const double *spc2d, const double *phs);
/* sponge_damping (src/nhwave.F:6462-6502) with the Sponge table of
 * calculate_sponge (:6405-6459); a no-op with SPONGE_ON = F. */
int nhw_sponge_damping(nhw_ctx *ctx);
/* eval_turb (src/nhwave.F:6544-6782) with IVTURB = IHTURB = 1: from this call
 * on the vertical and horizontal turbulent viscosities are Cvs and Chs (before
 * it they hold the turbulence seed of src/initialize.F:3172-3196).  The driver
 * calls it after each stage (src/nhwave.F:211); nhw_step does so itself.  A
 * no-op when VISCOUS_FLOW is off. */
int nhw_eval_turb(nhw_ctx *ctx);        /* eval_turb,    src/nhwave.F:6544-6782 */

/* The driver's time loop body (src/nhwave.F:127-258) run `nsteps` times on
 * the device: estimate_dt, update_mask, update_wind, update_vars, then for
 * each RK stage source_terms, fluxes, eval_duvw. */
int nhw_step(nhw_ctx *ctx, int32_t nsteps);
/* The driver's loop itself, `do while (TIME<TOTAL_TIME .and.
 * RUN_STEP<SIM_STEPS)` (src/nhwave.F:119): steps until the device TIME
 * reaches total_time or RUN_STEP reaches sim_steps; *steps_taken (may be
 * NULL) gets the number of steps run by this call. */
int nhw_run(nhw_ctx *ctx, double total_time, int32_t sim_steps, int32_t *steps_taken);
int nhw_solver_stats(nhw_ctx *ctx, nhw_stats *out);
/* Blocks until all work queued on the context's stream is done. */
int nhw_sync(nhw_ctx *ctx);

/* Per-kernel timing: when enabled, every launch is bracketed by HIP events on
 * the context's stream; nhw_kernel_stats returns, for kernel id 0.. until it
 * fails with NHW_EINVAL, its name, launch count, summed device time and its
 * algorithmic HBM bytes per launch.  No reference counterpart (measurement). */
int nhw_kernel_timing(nhw_ctx *ctx, int32_t enable);
int nhw_kernel_stats(nhw_ctx *ctx, int32_t id, const char **name, int64_t *launches,
                     double *total_ms, double *bytes_per_launch);
/* Calibration: a plain copy kernel reading and writing `nbytes` per launch,
 * `reps` launches on the null stream; average ms per launch.  Used to scale
 * HBM byte counters.  No reference counterpart (measurement). */
int nhw_bench_copy(int64_t nbytes, int32_t reps, double *avg_ms);

/* ---- Output and restart -------------------------------------------------
 * statistics, src/nhwave.F:3119-3208: the quantities the reference prints,
 * reduced on the device from the resident state (sums in a fixed tree
 * order, so they match the reference's serial sums to rounding). */
typedef struct nhw_stat_values {
  double mass_volume;  /* MassVolume = sum D dx dy */
  double energy;       /* Energy = sum m g (D sigc - Hc) + m (u^2+v^2+w^2)/2, m = Rho0 dsig D dx dy */
  double max_eta, min_eta;
  double max_u, max_v, max_w;  /* max |U|, |V|, |W| over the interior */
} nhw_stat_values;
int nhw_statistics(nhw_ctx *ctx, nhw_stat_values *out);
/* Snapshot of the non-NULL members of `host`, taken after the work queued so
 * far and drained to the host while the caller keeps stepping; the arrays
 * hold it after nhw_snapshot_wait (a new begin completes a pending one
 * first).  No counterpart in the reference, whose preview/putfile
 * (src/nhwave.F:3291-3690) stop the run to write text. */
int nhw_snapshot_begin(nhw_ctx *ctx, const nhw_fields *host);
int nhw_snapshot_wait(nhw_ctx *ctx);
/* The same snapshot for callers that cannot allocate (no counterpart in the
 * reference): on success it calls no allocating routine -- no hipMalloc,
 * pinned host memory, stream, event or host heap.  The fields are staged,
 * in stream order, in one fixed device array of the library (128 MiB per
 * GPU, held until the process ends) and copied from there into the
 * caller's arrays; the call returns when they hold the snapshot.  A request
 * whose FP64/int32 bytes (each field rounded up to 256) exceed 128 MiB is
 * refused with NHW_EINVAL.  All calls on one GPU share that array: calls
 * that overlap on one GPU race and must be serialised by the caller. */
int nhw_snapshot_static(nhw_ctx *ctx, const nhw_fields *host);
/* Full-precision restart file: every array of nhw_fields, the sigma grid,
 * TIME, dt, dt_old and RUN_STEP; a run continued from it is bit-identical to
 * the uninterrupted run.  The role of the reference's hot start
 * (src/nhwave.F:13338-13692), in a binary format of this library.  Loading
 * needs a context of the same grid size, precision and NON_HYDRO. */
int nhw_checkpoint_save(nhw_ctx *ctx, const char *path);
int nhw_checkpoint_load(nhw_ctx *ctx, const char *path);

/* ---- Batched form (ensemble of scenarios) -------------------------------
 * No counterpart in the reference, which runs one scenario per process: a
 * batch advances many scenarios of uneven sizes with one kernel launch per
 * step of the time step for the whole ensemble (the same per-cell code as
 * the single form).  Members are ordinary contexts: fill each through
 * nhw_set_grid / nhw_upload / nhw_initial (nhw_batch_member gives the
 * handle), then nhw_batch_prepare, then nhw_batch_step; results come back
 * per member through nhw_download / nhw_get_time / nhw_solver_stats.
 * Members share precision, TIME_ORDER and NON_HYDRO; each keeps its own dt,
 * mask, solver tolerance and convergence. */
typedef struct nhw_batch nhw_batch;
int nhw_batch_create(const nhw_config *cfgs, int32_t n, nhw_batch **out);  /* n >= 0 */
int nhw_batch_member(nhw_batch *b, int32_t i, nhw_ctx **ctx);
/* continuous batching: members join and leave between steps; either call
 * requires nhw_batch_prepare before the next nhw_batch_step.  Member
 * indices after a removal shift down by one. */
int nhw_batch_add(nhw_batch *b, const nhw_config *cfg, nhw_ctx **ctx);
int nhw_batch_remove(nhw_batch *b, int32_t i);
int nhw_batch_size(nhw_batch *b, int32_t *n);
/* builds the device tables; autotune != 0 times 64/128/256 threads per
 * block for the face and column sweeps and keeps the fastest */
int nhw_batch_prepare(nhw_batch *b, int32_t autotune);
int nhw_batch_step(nhw_batch *b, int32_t nsteps);
/* enable != 0 (the default): the launches between Poisson solves are
 * recorded once per prepare into HIP graphs and replayed every step (run
 * eagerly while per-launch kernel timing is on); 0: always eager */
int nhw_batch_graphs(nhw_batch *b, int32_t enable);
/* autotuner result for kernel kind 0..6: name, threads per block and the
 * mean time (ms) of the 64/128/256 candidates */
int nhw_batch_tuning(nhw_batch *b, int32_t kind, const char **name, int32_t *threads, double *ms3);
int nhw_batch_destroy(nhw_batch *b);

#ifdef __cplusplus
}
#endif
#endif
