/*
 * ORACLE -- test infrastructure only.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker; the product path never links it.
 *
 * Plain-C FP64 restatement of the NHWAVE time-stepping hot path (serial,
 * -DDOUBLE_PRECISION -DFROUDE_CAP build of jtfrey/NHWAVE), routine by
 * routine, loop order for loop order, with the reference's expression order
 * and its single-precision-rounded literals.  Compile with
 * -ffp-contract=off and no -march (no FMA), as the reference build is.
 *
 * Arrays use the reference's Fortran layout (1-based indices, i fastest):
 *   2-D      (Mloc ,Nloc )          x-face 2-D (Mloc1,Nloc)   y-face (Mloc,Nloc1)
 *   3-D      (Mloc ,Nloc ,Kloc )    x-face 3-D (Mloc1,Nloc,Kloc) ...
 *   z-face   (Mloc ,Nloc ,Kloc1)
 * (src/initialize.F:1423-1504).
 */
#ifndef NHW_ORACLE_H
#define NHW_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cfg {
  int mglob, nglob, kglob;           /* Mglob, Nglob, Kglob */
  double dx, dy;
  double cfl, min_dep, dt_ini, dt_min, dt_max;
  double froude_cap, visc, fcor;     /* fcor as the reference computes it */
  double tol;                        /* oracle solves directly; kept for API parity */
  int itmax;
  int it_order;                      /* 1,2,3 */
  int non_hydro;                     /* 0/1 */
  int froude_cap_on;                 /* -DFROUDE_CAP */
  int bc_x0, bc_xn, bc_y0, bc_yn, bc_z0;  /* lateral 1/2 (3 at x0 with a wavemaker); bottom 1, 2, 5 */
  /* boundary forcing: WAVEMAKER (0 NONE, 1 LEF_SOL, 2 LEF_LIN, 3 LEF_STK, 4 LEF_CON),
   * AMP, PER, DEP, THETA, CUR, TRAMP; SPONGE_ON and the four widths
   * (West, East, South, North) */
  int wavemaker;
  double amp_wave, per_wave, dep_wave, theta_wave, cur_wave, tramp;
  int sponge_on;
  double sponge_w, sponge_e, sponge_s, sponge_n;
  /* viscous branch with constant viscosities (VISCOUS_FLOW = T, IVTURB =
   * IHTURB = 1): Cvs, Chs, Schmidt, VISCOUS_NUMBER; VISCOSITY is visc above;
   * RNG picks the coefficient of the initial turbulent viscosity */
  int viscous, rng;
  double cvs, chs, schmidt, visc_number;
  /* bottom friction for BC_Z0 = 5: Ibot (1: constant Cd0, else log law with
   * roughness Zob) */
  int ibot;
  double cd0, zob;
  /* turbulence model of the viscous branch: IVTURB = IHTURB = 1 (constant
   * Cvs, Chs) or IVTURB = IHTURB = 10 (the 3-D k-epsilon model, kepsilon_3D,
   * as the reference build executes it; Kglob >= 3) */
  int ivturb;
  /* internal irregular wavemaker (WAVEMAKER = INT_IRR, 5): the source box
   * Xsource_West/East, Ysource_Suth/Nrth and the spectrum of irr2d.txt
   * (src/initialize.F:835-878): NumFreq frequencies, NumDir directions
   * (degrees), Wave_Spc2d and Random_Phs as (NumDir, NumFreq) arrays,
   * direction fastest */
  double xs_w, xs_e, ys_s, ys_n;
  int nfreq, ndir;
  const double *freq, *dire, *spc, *phs;
} orc_cfg;

typedef struct orc orc;

/* lifetime */
orc *orc_create(const orc_cfg *cfg);
void orc_destroy(orc *o);
const char *orc_last_error(void);

/* raw access to a named array (Fortran layout, see header comment); returns
 * NULL for an unknown name.  *count receives the element count, *is_int 1
 * for int32 arrays (Mask, Mask_Struct, Mask9). */
void *orc_array(orc *o, const char *name, long *count, int *is_int);
/* scalars */
void orc_get_time(const orc *o, double *time, double *dt, int *run_step);
void orc_set_time(orc *o, double time, double dt, int run_step);

/* reference routines (each cites its reference lines in nhw_oracle.c) */
void orc_initial_tail(orc *o);      /* src/initialize.F:3118-3170 */
int  orc_estimate_dt(orc *o);       /* src/nhwave.F:3694-3874, 0 ok, -1 dt<dt_min */
void orc_update_mask(orc *o);       /* src/nhwave.F:5403-5456 */
void orc_update_vars(orc *o);       /* src/nhwave.F:5459-5511 */
void orc_source_terms(orc *o);      /* src/source_terms.F:446-453 */
void orc_fluxes(orc *o);            /* src/nhwave.F:1686-1714 */
int  orc_eval_duvw(orc *o, int istep); /* src/nhwave.F:539-1300 */
int  orc_poisson_solver(orc *o);    /* src/pressure.F:63-245 (direct solve) */
void orc_generate_coef_rhs(orc *o); /* src/pressure.F:475-840 */
void orc_interpolate_velocity_to_faces(orc *o); /* src/nhwave.F:1332-1408 (Uf, Vf, Wf) */
void orc_vel_bc(orc *o);            /* src/nhwave.F:4416-5066 */
void orc_update_wave_bc(orc *o);    /* src/nhwave.F:5514-5562, src/waves.F:92-170, 518-903 */
/* the cnoidal helpers of src/waves.F:184-361 (cnoidal_cn, cnoidal_ck,
 * cnoidal_ce), exposed for tests, and the wavemaker's set-up: out[0..4] =
 * wave length, celerity, trough level, modulus, zero-crossing start; -1 if
 * the start search does not converge */
double orc_cnoidal_cn(double u, double m);
double orc_cnoidal_ck(double m);
double orc_cnoidal_ce(double m);
int orc_cnoidal_setup(double height, double depth, double period, double *out);
void orc_sponge_damping(orc *o);    /* src/nhwave.F:6462-6502 */
void orc_eval_turb(orc *o);         /* src/nhwave.F:6544-6782 (IVturb = IHturb = 1) */
/* generate_grid + read_bathymetry (CELL_CENTER) + initial with
 * INITIAL_EUVW; hcg/etag (Mglob,Nglob), ug/vg/wg (Mglob,Nglob,Kglob) or NULL */
void orc_setup(orc *o, const double *hcg, const double *etag, const double *ug,
               const double *vg, const double *wg, int ivgrd, double grd_r,
               double wind_u, double wind_v);
void orc_set_wind(orc *o, double wind_u, double wind_v);
/* one full time step of the driver loop, src/nhwave.F:119-258 */
int  orc_step(orc *o);
/* Poisson system as assembled: coef[15*neqns] (diag-major, row index
 * i+Mg*(j-1)+Mg*Ng*(k-1) over the interior) and rhs[neqns]. */
long orc_neqns(const orc *o);
double *orc_coef(orc *o);
double *orc_rhs(orc *o);
/* dense/banded helpers exposed for tests */
int orc_solve_banded(orc *o, const double *coef, const double *rhs, double *x);

#ifdef __cplusplus
}
#endif
#endif
