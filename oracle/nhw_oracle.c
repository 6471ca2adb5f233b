/*
 * ORACLE -- test infrastructure only (see nhw_oracle.h).  Never linked into
 * the product; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, as the checker.
 *
 * A plain-C FP64 restatement of the jtfrey/NHWAVE hot path, following the
 * reference routines loop by loop.  Every function cites the reference lines
 * it restates.  Expression order follows the Fortran source left to right;
 * default-REAL literals that are inexact in binary32 are widened from their
 * binary32 value, exactly as the -DDOUBLE_PRECISION reference build does
 * (src/mod_global.F:21-32).  Build with -O2 -ffp-contract=off.
 *
 * The one deliberate algorithmic difference: poisson_solver's linear solve.
 * The reference runs NSPCG IC(0)+truncated GMRES to its stopping test
 * (src/pressure.F:86-131); the oracle solves the identical assembled system
 * directly (banded LU with partial pivoting), i.e. the limit that iteration
 * converges to.
 */
#include "nhw_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- constants with the reference's binary32 rounding ------------------ */
#define F32(x) ((double)(float)(x))
static const double GRAV = F32(9.81);          /* src/mod_global.F:24 */
static const double SMALL = F32(1.0e-16);      /* src/mod_global.F:22 */
static const double LARGE = 10000000.0;        /* src/mod_global.F:23 */
static const double RHO0 = 1000.0;             /* src/mod_global.F:27 */
static const double ZERO = 0.0;
static const double KAPPA = F32(0.41);         /* src/mod_global.F:30 */

static char g_err[256];
const char *orc_last_error(void) { return g_err; }

struct orc {
  orc_cfg c;
  int Mloc, Nloc, Kloc, Mloc1, Nloc1, Kloc1;
  int Ibeg, Iend, Iend1, Jbeg, Jend, Jend1, Kbeg, Kend, Kend1;
  long n2, nx2, ny2, n3, nx3, ny3, nz3, neqns;
  double alpha[3], beta[3];
  double dt, dt_old, time;
  int run_step;
  double *dsig, *sig, *sigc;
  /* 2-D */
  double *Hc, *DelxH, *DelyH, *DeltH, *Eta, *Eta0, *D, *D0, *DelxEta, *DelyEta;
  double *SourceX, *SourceY, *SourceC, *Wsx, *Wsy, *R1;
  double *irr;  /* the internal wavemaker's spectrum (copied from the configuration) */
  int *Mask, *Mask_Struct, *Mask9;
  double *Hfx, *Hfx0, *EtaxL, *EtaxR, *DxL, *DxR;   /* (Mloc1,Nloc) */
  double *Hfy, *Hfy0, *EtayL, *EtayR, *DyL, *DyR;   /* (Mloc,Nloc1) */
  /* 3-D cell */
  double *U, *V, *W, *DU, *DV, *DW, *DU0, *DV0, *DW0;
  double *DelxU, *DelxV, *DelxW, *DelxDU, *DelxDV, *DelxDW;
  double *DelyU, *DelyV, *DelyW, *DelyDU, *DelyDV, *DelyDW;
  double *DelzU, *DelzV, *DelzW, *DelxSc, *DelySc, *R2, *R3, *R4;
  /* viscous branch: sigma slopes at the sigma levels (z-face), the stress
   * terms of diffusion (cell), the turbulent viscosities CmuHt, CmuVt (cell;
   * the seed Cmut_min until the first eval_turb) and, with the k-epsilon
   * model, Tke, Eps and their depth-integrated forms DTke, DEps (cell) */
  double *DelxSl, *DelySl, *Diffxx, *Diffxy, *Diffyx, *Diffyy, *Diffzx, *Diffzy;
  double *CmuHt, *CmuVt, *Tke, *Eps, *DTke, *DEps;
  /* 3-D z-face (Mloc,Nloc,Kloc1) */
  double *Omega, *P, *Uf, *Vf, *Wf, *UzL, *UzR, *VzL, *VzR, *WzL, *WzR;
  double *Fz, *Gz, *Hz;
  /* 3-D x-face (Mloc1,Nloc,Kloc) */
  double *UxL, *UxR, *VxL, *VxR, *WxL, *WxR, *DUxL, *DUxR, *DVxL, *DVxR, *DWxL, *DWxR;
  double *SxL, *SxR, *SxS, *ExL, *ExR, *FxL, *FxR, *GxL, *GxR, *HxL, *HxR;
  double *Ex, *Fx, *Gx, *Hx;
  /* 3-D y-face (Mloc,Nloc1,Kloc) */
  double *UyL, *UyR, *VyL, *VyR, *WyL, *WyR, *DUyL, *DUyR, *DVyL, *DVyR, *DWyL, *DWyR;
  double *SyL, *SyR, *SyS, *EyL, *EyR, *FyL, *FyR, *GyL, *GyR, *HyL, *HyR;
  double *Ey, *Fy, *Gy, *Hy;
  /* Poisson system (src/pressure.F:475-840) */
  double *Coef, *Rhs;
  /* constant-wind settings for update_wind (src/nhwave.F:5333-5400) */
  double wind_u, wind_v;
  /* boundary forcing: west inflow values (src/initialize.F:1442, 1449) with
   * Uin/Vin/Win (Nloc,Kloc), j fastest; cell centres; sponge (Mloc,Nloc) */
  double *Ein_X0, *Din_X0, *Uin_X0, *Vin_X0, *Win_X0;
  /* east open boundary of FLUX_LR (BC_Xn = 3): Ein_Xn, Din_Xn (Nloc), Uin_Xn,
   * Vin_Xn, Win_Xn and the outflow velocity Uin_Xni with its previous value
   * Uin_Xni0 (Nloc,Kloc); the Orlanski radiation reads the previous steps'
   * Eta00 (Mloc,Nloc) and U0, V0, W0, U00, V00, W00 (Mloc,Nloc,Kloc) */
  double *Ein_Xn, *Din_Xn, *Uin_Xn, *Vin_Xn, *Win_Xn, *Uin_Xni, *Uin_Xni0;
  double *Eta00, *U0, *V0, *W0, *U00, *V00, *W00;
  double *xc, *yc, *Sponge;
  double wave_number;                /* wave_dispersion result at set-up */
};

/* ---- indexing (Fortran 1-based, i fastest) ------------------------------ */
#define I2(i, j) ((long)((j)-1) * o->Mloc + (i)-1)
#define IX2(i, j) ((long)((j)-1) * o->Mloc1 + (i)-1)
#define IY2(i, j) ((long)((j)-1) * o->Mloc + (i)-1)
#define I3(i, j, k) (((long)((k)-1) * o->Nloc + (j)-1) * o->Mloc + (i)-1)
#define IZ3(i, j, k) I3(i, j, k)
#define IX3(i, j, k) (((long)((k)-1) * o->Nloc + (j)-1) * o->Mloc1 + (i)-1)
#define IY3(i, j, k) (((long)((k)-1) * o->Nloc1 + (j)-1) * o->Mloc + (i)-1)
#define IB(j, k) ((long)((k)-1) * o->Nloc + (j)-1)

static double *zd(long n) { return (double *)calloc((size_t)n, sizeof(double)); }
static int *zi(long n) { return (int *)calloc((size_t)n, sizeof(int)); }
static double dmin(double a, double b) { return b < a ? b : a; }
static double dmax(double a, double b) { return b > a ? b : a; }

/* k-epsilon floors, src/initialize.F:3178-3179 (REAL(4) literals) */
#define KE_TKE_MIN F32(1.e-9)
#define KE_EPS_MIN F32(1.e-9)
/* Cmut_min = c Tke_min**2/Eps_min, c = 8.5e-2 with RNG else 9.0e-2
 * (src/initialize.F:3180-3184) */
static double ke_cmut_min(const orc *o) {
  return (o->c.rng ? F32(8.5e-2) : F32(9.0e-2)) * (KE_TKE_MIN * KE_TKE_MIN) / KE_EPS_MIN;
}

static int cnoidal_start(double A, double Dp, double Per, double *L, double *yt, double *m, double *ck,
                         double *xstart);

orc *orc_create(const orc_cfg *cfg) {
  if (cfg->mglob < 2 || cfg->nglob < 1 || cfg->kglob < 2) {
    snprintf(g_err, sizeof g_err, "orc_create: unsupported grid %d x %d x %d",
             cfg->mglob, cfg->nglob, cfg->kglob);
    return NULL;
  }
  const int bcs[4] = {cfg->bc_x0, cfg->bc_xn, cfg->bc_y0, cfg->bc_yn};
  for (int q = 0; q < 4; ++q)
    if (bcs[q] != 1 && bcs[q] != 2 && !(bcs[q] == 5 && cfg->viscous) &&
        !(q == 0 && bcs[q] == 3 && ((cfg->wavemaker >= 1 && cfg->wavemaker <= 4) || cfg->wavemaker == 6)) &&
        !(q == 1 && bcs[q] == 3 && cfg->wavemaker == 6)) {
      snprintf(g_err, sizeof g_err,
               "orc_create: lateral BC type %d unsupported (1, 2; 5 in viscous runs; 3 at x0 with a wavemaker)",
               bcs[q]);
      return NULL;
    }
  if (cfg->wavemaker < 0 || cfg->wavemaker > 6 || (cfg->wavemaker == 6 && (cfg->bc_x0 != 3 || cfg->bc_xn != 3))) {
    snprintf(g_err, sizeof g_err, "orc_create: wavemaker %d unsupported", cfg->wavemaker);
    return NULL;
  }
  if (cfg->wavemaker == 4) {
    double L, yt, m, ck, xs;
    if (cnoidal_start(cfg->amp_wave, cfg->dep_wave, cfg->per_wave, &L, &yt, &m, &ck, &xs)) {
      snprintf(g_err, sizeof g_err, "orc_create: cnoidal wave: too many iterations for its zero crossing");
      return NULL;
    }
  }
  if (cfg->viscous && cfg->kglob > 64) {
    snprintf(g_err, sizeof g_err, "orc_create: VISCOUS_FLOW takes Kglob <= 64");
    return NULL;
  }
  if (cfg->viscous && !(cfg->schmidt != 0.0)) {
    snprintf(g_err, sizeof g_err, "orc_create: VISCOUS_FLOW needs Schmidt != 0");
    return NULL;
  }
  if (cfg->viscous && cfg->ivturb != 1 && cfg->ivturb != 10) {
    snprintf(g_err, sizeof g_err, "orc_create: IVTURB=%d unsupported (1 constant, 10 k-epsilon)", cfg->ivturb);
    return NULL;
  }
  if (cfg->viscous && cfg->ivturb == 10 && cfg->kglob < 3) {
    snprintf(g_err, sizeof g_err, "orc_create: the k-epsilon model takes Kglob >= 3");
    return NULL;
  }
  if (cfg->bc_z0 != 1 && cfg->bc_z0 != 2 && cfg->bc_z0 != 5) {
    snprintf(g_err, sizeof g_err, "orc_create: BC_Z0=%d unsupported (1, 2, 5)", cfg->bc_z0);
    return NULL;
  }
  if (cfg->wavemaker == 5 && (cfg->nfreq < 1 || cfg->ndir < 1 || !cfg->freq || !cfg->dire || !cfg->spc || !cfg->phs)) {
    snprintf(g_err, sizeof g_err, "orc_create: INT_IRR needs its spectrum (NumFreq, NumDir >= 1 and the arrays)");
    return NULL;
  }
  orc *o = (orc *)calloc(1, sizeof(orc));
  o->c = *cfg;
  if (cfg->wavemaker == 5) {  /* keep a copy of the spectrum */
    const long nf = cfg->nfreq, nd = cfg->ndir;
    o->irr = (double *)malloc(sizeof(double) * (nf + nd + 2 * nf * nd));
    memcpy(o->irr, cfg->freq, sizeof(double) * nf);
    memcpy(o->irr + nf, cfg->dire, sizeof(double) * nd);
    memcpy(o->irr + nf + nd, cfg->spc, sizeof(double) * nf * nd);
    memcpy(o->irr + nf + nd + nf * nd, cfg->phs, sizeof(double) * nf * nd);
    o->c.freq = o->irr;
    o->c.dire = o->irr + nf;
    o->c.spc = o->irr + nf + nd;
    o->c.phs = o->irr + nf + nd + nf * nd;
  } else {
    o->c.nfreq = o->c.ndir = 0;
    o->c.freq = o->c.dire = o->c.spc = o->c.phs = NULL;
  }
  /* src/initialize.F:1381-1396 (serial: PX=PY=1, Nghost=2) */
  o->Mloc = cfg->mglob + 4; o->Nloc = cfg->nglob + 4; o->Kloc = cfg->kglob + 4;
  o->Mloc1 = o->Mloc + 1; o->Nloc1 = o->Nloc + 1; o->Kloc1 = o->Kloc + 1;
  o->Ibeg = 3; o->Iend = o->Mloc - 2; o->Iend1 = o->Mloc1 - 2;
  o->Jbeg = 3; o->Jend = o->Nloc - 2; o->Jend1 = o->Nloc1 - 2;
  o->Kbeg = 3; o->Kend = o->Kloc - 2; o->Kend1 = o->Kloc1 - 2;
  o->n2 = (long)o->Mloc * o->Nloc;
  o->nx2 = (long)o->Mloc1 * o->Nloc;
  o->ny2 = (long)o->Mloc * o->Nloc1;
  o->n3 = o->n2 * o->Kloc;
  o->nx3 = o->nx2 * o->Kloc;
  o->ny3 = o->ny2 * o->Kloc;
  o->nz3 = o->n2 * o->Kloc1;
  o->neqns = (long)cfg->mglob * cfg->nglob * cfg->kglob;
  o->dsig = zd(o->Kloc); o->sig = zd(o->Kloc1); o->sigc = zd(o->Kloc);
#define A2D(x) o->x = zd(o->n2)
  A2D(Hc); A2D(DelxH); A2D(DelyH); A2D(DeltH); A2D(Eta); A2D(Eta0); A2D(D); A2D(D0);
  A2D(DelxEta); A2D(DelyEta); A2D(SourceX); A2D(SourceY); A2D(SourceC); A2D(Wsx); A2D(Wsy); A2D(R1);
  o->Mask = zi(o->n2); o->Mask_Struct = zi(o->n2); o->Mask9 = zi(o->n2);
#define AX2D(x) o->x = zd(o->nx2)
#define AY2D(x) o->x = zd(o->ny2)
  AX2D(Hfx); AX2D(Hfx0); AX2D(EtaxL); AX2D(EtaxR); AX2D(DxL); AX2D(DxR);
  AY2D(Hfy); AY2D(Hfy0); AY2D(EtayL); AY2D(EtayR); AY2D(DyL); AY2D(DyR);
#define A3D(x) o->x = zd(o->n3)
  A3D(U); A3D(V); A3D(W); A3D(DU); A3D(DV); A3D(DW); A3D(DU0); A3D(DV0); A3D(DW0);
  A3D(DelxU); A3D(DelxV); A3D(DelxW); A3D(DelxDU); A3D(DelxDV); A3D(DelxDW);
  A3D(DelyU); A3D(DelyV); A3D(DelyW); A3D(DelyDU); A3D(DelyDV); A3D(DelyDW);
  A3D(DelzU); A3D(DelzV); A3D(DelzW); A3D(DelxSc); A3D(DelySc); A3D(R2); A3D(R3); A3D(R4);
  A3D(Diffxx); A3D(Diffxy); A3D(Diffyx); A3D(Diffyy); A3D(Diffzx); A3D(Diffzy);
  A3D(CmuHt); A3D(CmuVt); A3D(Tke); A3D(Eps); A3D(DTke); A3D(DEps);
#define AZ3D(x) o->x = zd(o->nz3)
  AZ3D(Omega); AZ3D(P); AZ3D(Uf); AZ3D(Vf); AZ3D(Wf); AZ3D(UzL); AZ3D(UzR); AZ3D(VzL); AZ3D(VzR);
  AZ3D(WzL); AZ3D(WzR); AZ3D(Fz); AZ3D(Gz); AZ3D(Hz); AZ3D(DelxSl); AZ3D(DelySl);
#define AX3D(x) o->x = zd(o->nx3)
  AX3D(UxL); AX3D(UxR); AX3D(VxL); AX3D(VxR); AX3D(WxL); AX3D(WxR); AX3D(DUxL); AX3D(DUxR);
  AX3D(DVxL); AX3D(DVxR); AX3D(DWxL); AX3D(DWxR); AX3D(SxL); AX3D(SxR); AX3D(SxS);
  AX3D(ExL); AX3D(ExR); AX3D(FxL); AX3D(FxR); AX3D(GxL); AX3D(GxR); AX3D(HxL); AX3D(HxR);
  AX3D(Ex); AX3D(Fx); AX3D(Gx); AX3D(Hx);
#define AY3D(x) o->x = zd(o->ny3)
  AY3D(UyL); AY3D(UyR); AY3D(VyL); AY3D(VyR); AY3D(WyL); AY3D(WyR); AY3D(DUyL); AY3D(DUyR);
  AY3D(DVyL); AY3D(DVyR); AY3D(DWyL); AY3D(DWyR); AY3D(SyL); AY3D(SyR); AY3D(SyS);
  AY3D(EyL); AY3D(EyR); AY3D(FyL); AY3D(FyR); AY3D(GyL); AY3D(GyR); AY3D(HyL); AY3D(HyR);
  AY3D(Ey); AY3D(Fy); AY3D(Gy); AY3D(Hy);
  o->Coef = zd(15 * o->neqns);
  o->Rhs = zd(o->neqns);
  o->Ein_X0 = zd(o->Nloc); o->Din_X0 = zd(o->Nloc);
  o->Ein_Xn = zd(o->Nloc); o->Din_Xn = zd(o->Nloc);
  o->Uin_Xn = zd((long)o->Nloc * o->Kloc); o->Vin_Xn = zd((long)o->Nloc * o->Kloc);
  o->Win_Xn = zd((long)o->Nloc * o->Kloc); o->Uin_Xni = zd((long)o->Nloc * o->Kloc);
  o->Uin_Xni0 = zd((long)o->Nloc * o->Kloc);
  o->Eta00 = zd(o->n2); o->U0 = zd(o->n3); o->V0 = zd(o->n3); o->W0 = zd(o->n3);
  o->U00 = zd(o->n3); o->V00 = zd(o->n3); o->W00 = zd(o->n3);
  o->Uin_X0 = zd((long)o->Nloc * o->Kloc); o->Vin_X0 = zd((long)o->Nloc * o->Kloc);
  o->Win_X0 = zd((long)o->Nloc * o->Kloc);
  o->xc = zd(o->Mloc); o->yc = zd(o->Nloc); o->Sponge = zd(o->n2);
  /* SSP-RK coefficients, src/initialize.F:3382-3400: 1/3, 2/3 are REAL(4). */
  if (cfg->it_order == 3) {
    o->alpha[0] = 0.0; o->alpha[1] = 0.75; o->alpha[2] = F32(1.0f / 3.0f);
    o->beta[0] = 1.0; o->beta[1] = 0.25; o->beta[2] = F32(2.0f / 3.0f);
  } else if (cfg->it_order == 2) {
    o->alpha[0] = 0.0; o->alpha[1] = 0.5; o->beta[0] = 1.0; o->beta[1] = 0.5;
  } else {
    o->alpha[0] = 0.0; o->beta[0] = 1.0;
  }
  o->dt = cfg->dt_ini;
  /* turbulence seeding, src/initialize.F:3172-3196: CmuHt = CmuVt = Cmut_min
   * = c Tke_min**2/Eps_min with Tke_min = Eps_min = 1.e-9 and c = 8.5e-2
   * (RNG) or 9.0e-2, all REAL(4) literals; Tke, Eps at their minima (DTke,
   * DEps follow once D and Mask are known, orc_initial_tail) */
  if (cfg->viscous) {
    const double cmut_min = ke_cmut_min(o);
    for (long q = 0; q < o->n3; ++q) {
      o->CmuHt[q] = cmut_min;
      o->CmuVt[q] = cmut_min;
      o->Tke[q] = KE_TKE_MIN;
      o->Eps[q] = KE_EPS_MIN;
    }
  }
  return o;
}

void orc_destroy(orc *o) {
  if (!o) return;
  free(o->irr);
  double **dp[] = {
      &o->dsig, &o->sig, &o->sigc, &o->Hc, &o->DelxH, &o->DelyH, &o->DeltH, &o->Eta, &o->Eta0,
      &o->D, &o->D0, &o->DelxEta, &o->DelyEta, &o->SourceX, &o->SourceY, &o->SourceC, &o->Wsx,
      &o->Wsy, &o->R1, &o->Hfx, &o->Hfx0, &o->EtaxL, &o->EtaxR, &o->DxL, &o->DxR, &o->Hfy,
      &o->Hfy0, &o->EtayL, &o->EtayR, &o->DyL, &o->DyR, &o->U, &o->V, &o->W, &o->DU, &o->DV,
      &o->DW, &o->DU0, &o->DV0, &o->DW0, &o->DelxU, &o->DelxV, &o->DelxW, &o->DelxDU,
      &o->DelxDV, &o->DelxDW, &o->DelyU, &o->DelyV, &o->DelyW, &o->DelyDU, &o->DelyDV,
      &o->DelyDW, &o->DelzU, &o->DelzV, &o->DelzW, &o->DelxSc, &o->DelySc, &o->R2, &o->R3,
      &o->R4, &o->Omega, &o->P, &o->Uf, &o->Vf, &o->Wf, &o->UzL, &o->UzR, &o->VzL, &o->VzR,
      &o->WzL, &o->WzR, &o->Fz, &o->Gz, &o->Hz, &o->UxL, &o->UxR, &o->VxL, &o->VxR, &o->WxL,
      &o->WxR, &o->DUxL, &o->DUxR, &o->DVxL, &o->DVxR, &o->DWxL, &o->DWxR, &o->SxL, &o->SxR,
      &o->SxS, &o->ExL, &o->ExR, &o->FxL, &o->FxR, &o->GxL, &o->GxR, &o->HxL, &o->HxR, &o->Ex,
      &o->Fx, &o->Gx, &o->Hx, &o->UyL, &o->UyR, &o->VyL, &o->VyR, &o->WyL, &o->WyR, &o->DUyL,
      &o->DUyR, &o->DVyL, &o->DVyR, &o->DWyL, &o->DWyR, &o->SyL, &o->SyR, &o->SyS, &o->EyL,
      &o->EyR, &o->FyL, &o->FyR, &o->GyL, &o->GyR, &o->HyL, &o->HyR, &o->Ey, &o->Fy, &o->Gy,
      &o->Hy, &o->Coef, &o->Rhs, &o->Ein_X0, &o->Din_X0, &o->Uin_X0, &o->Vin_X0, &o->Win_X0,
      &o->Ein_Xn, &o->Din_Xn, &o->Uin_Xn, &o->Vin_Xn, &o->Win_Xn, &o->Uin_Xni, &o->Uin_Xni0, &o->Eta00,
      &o->U0, &o->V0, &o->W0, &o->U00, &o->V00, &o->W00,
      &o->xc, &o->yc, &o->Sponge, &o->DelxSl, &o->DelySl, &o->Diffxx, &o->Diffxy, &o->Diffyx,
      &o->Diffyy, &o->Diffzx, &o->Diffzy, &o->CmuHt, &o->CmuVt, &o->Tke, &o->Eps, &o->DTke, &o->DEps};
  for (size_t q = 0; q < sizeof dp / sizeof dp[0]; ++q) free(*dp[q]);
  free(o->Mask); free(o->Mask_Struct); free(o->Mask9);
  free(o);
}

void *orc_array(orc *o, const char *name, long *count, int *is_int) {
  struct { const char *n; void *p; long c; int ii; } t[] = {
      {"dsig", o->dsig, o->Kloc, 0}, {"sig", o->sig, o->Kloc1, 0}, {"sigc", o->sigc, o->Kloc, 0},
      {"Hc", o->Hc, o->n2, 0}, {"DelxH", o->DelxH, o->n2, 0}, {"DelyH", o->DelyH, o->n2, 0},
      {"Eta", o->Eta, o->n2, 0}, {"Eta0", o->Eta0, o->n2, 0}, {"D", o->D, o->n2, 0},
      {"D0", o->D0, o->n2, 0}, {"DelxEta", o->DelxEta, o->n2, 0}, {"DelyEta", o->DelyEta, o->n2, 0},
      {"SourceX", o->SourceX, o->n2, 0}, {"SourceY", o->SourceY, o->n2, 0}, {"SourceC", o->SourceC, o->n2, 0},
      {"Wsx", o->Wsx, o->n2, 0}, {"Wsy", o->Wsy, o->n2, 0}, {"R1", o->R1, o->n2, 0},
      {"Mask", o->Mask, o->n2, 1}, {"Mask_Struct", o->Mask_Struct, o->n2, 1},
      {"Mask9", o->Mask9, o->n2, 1},
      {"Hfx", o->Hfx, o->nx2, 0}, {"Hfx0", o->Hfx0, o->nx2, 0}, {"EtaxL", o->EtaxL, o->nx2, 0},
      {"EtaxR", o->EtaxR, o->nx2, 0}, {"DxL", o->DxL, o->nx2, 0}, {"DxR", o->DxR, o->nx2, 0},
      {"Hfy", o->Hfy, o->ny2, 0}, {"Hfy0", o->Hfy0, o->ny2, 0}, {"EtayL", o->EtayL, o->ny2, 0},
      {"EtayR", o->EtayR, o->ny2, 0}, {"DyL", o->DyL, o->ny2, 0}, {"DyR", o->DyR, o->ny2, 0},
      {"U", o->U, o->n3, 0}, {"V", o->V, o->n3, 0}, {"W", o->W, o->n3, 0},
      {"DU", o->DU, o->n3, 0}, {"DV", o->DV, o->n3, 0}, {"DW", o->DW, o->n3, 0},
      {"DU0", o->DU0, o->n3, 0}, {"DV0", o->DV0, o->n3, 0}, {"DW0", o->DW0, o->n3, 0},
      {"DelxSc", o->DelxSc, o->n3, 0}, {"DelySc", o->DelySc, o->n3, 0},
      {"Omega", o->Omega, o->nz3, 0}, {"P", o->P, o->nz3, 0}, {"Uf", o->Uf, o->nz3, 0},
      {"Vf", o->Vf, o->nz3, 0}, {"Wf", o->Wf, o->nz3, 0},
      {"Ex", o->Ex, o->nx3, 0}, {"Fx", o->Fx, o->nx3, 0}, {"Gx", o->Gx, o->nx3, 0},
      {"Hx", o->Hx, o->nx3, 0}, {"Ey", o->Ey, o->ny3, 0}, {"Fy", o->Fy, o->ny3, 0},
      {"Gy", o->Gy, o->ny3, 0}, {"Hy", o->Hy, o->ny3, 0}, {"Fz", o->Fz, o->nz3, 0},
      {"Gz", o->Gz, o->nz3, 0}, {"Hz", o->Hz, o->nz3, 0},
      {"Coef", o->Coef, 15 * o->neqns, 0}, {"Rhs", o->Rhs, o->neqns, 0},
      {"Ein_X0", o->Ein_X0, o->Nloc, 0}, {"Din_X0", o->Din_X0, o->Nloc, 0},
      {"Ein_Xn", o->Ein_Xn, o->Nloc, 0}, {"Din_Xn", o->Din_Xn, o->Nloc, 0}, {"Eta00", o->Eta00, o->n2, 0},
      {"U0", o->U0, o->n3, 0}, {"V0", o->V0, o->n3, 0}, {"W0", o->W0, o->n3, 0},
      {"Uin_X0", o->Uin_X0, (long)o->Nloc * o->Kloc, 0}, {"Vin_X0", o->Vin_X0, (long)o->Nloc * o->Kloc, 0},
      {"Win_X0", o->Win_X0, (long)o->Nloc * o->Kloc, 0}, {"xc", o->xc, o->Mloc, 0},
      {"yc", o->yc, o->Nloc, 0}, {"Sponge", o->Sponge, o->n2, 0},
      {"DelxSl", o->DelxSl, o->nz3, 0}, {"DelySl", o->DelySl, o->nz3, 0},
      {"Diffxx", o->Diffxx, o->n3, 0}, {"Diffxy", o->Diffxy, o->n3, 0},
      {"Diffyx", o->Diffyx, o->n3, 0}, {"Diffyy", o->Diffyy, o->n3, 0},
      {"Diffzx", o->Diffzx, o->n3, 0}, {"Diffzy", o->Diffzy, o->n3, 0},
      {"CmuHt", o->CmuHt, o->n3, 0}, {"CmuVt", o->CmuVt, o->n3, 0}, {"Tke", o->Tke, o->n3, 0},
      {"Eps", o->Eps, o->n3, 0}, {"DTke", o->DTke, o->n3, 0}, {"DEps", o->DEps, o->n3, 0},
  };
  for (size_t q = 0; q < sizeof t / sizeof t[0]; ++q)
    if (strcmp(t[q].n, name) == 0) {
      if (count) *count = t[q].c;
      if (is_int) *is_int = t[q].ii;
      return t[q].p;
    }
  return NULL;
}

void orc_get_time(const orc *o, double *time, double *dt, int *run_step) {
  if (time) *time = o->time;
  if (dt) *dt = o->dt;
  if (run_step) *run_step = o->run_step;
}
void orc_set_time(orc *o, double time, double dt, int run_step) {
  o->time = time; o->dt = dt; o->run_step = run_step;
}
void orc_set_wind(orc *o, double wu, double wv) { o->wind_u = wu; o->wind_v = wv; }
long orc_neqns(const orc *o) { return o->neqns; }
double *orc_coef(orc *o) { return o->Coef; }
double *orc_rhs(orc *o) { return o->Rhs; }

/* ======================================================================== */
/* Slopes and reconstruction, src/nhwave.F:1890-2351, 2889-2906             */
/* ======================================================================== */

/* LIMITER (van Leer), src/nhwave.F:2889-2906 */
static double limiter(double a, double b) {
  return (a * fabs(b) + fabs(a) * b) / (fabs(a) + fabs(b));
}

/* delxFun_2D, src/nhwave.F:2060-2094 */
static void delx_2d(orc *o, const double *Din, double *Dout) {
  const double dx = o->c.dx;
  for (int i = 2; i <= o->Mloc - 1; ++i)
    for (int j = 1; j <= o->Nloc; ++j) {
      if (o->Mask[I2(i, j)] == 0) {
        Dout[I2(i, j)] = ZERO;
      } else {
        double t1 = (Din[I2(i + 1, j)] - Din[I2(i, j)]) / dx;
        double t2 = (Din[I2(i, j)] - Din[I2(i - 1, j)]) / dx;
        if ((fabs(t1) + fabs(t2)) < SMALL) Dout[I2(i, j)] = ZERO;
        else Dout[I2(i, j)] = limiter(t1, t2);
      }
    }
  for (int j = 1; j <= o->Nloc; ++j) {
    Dout[I2(1, j)] = (Din[I2(2, j)] - Din[I2(1, j)]) / dx;
    Dout[I2(o->Mloc, j)] = (Din[I2(o->Mloc, j)] - Din[I2(o->Mloc - 1, j)]) / dx;
  }
}

/* delyFun_2D, src/nhwave.F:2182-2216 */
static void dely_2d(orc *o, const double *Din, double *Dout) {
  const double dy = o->c.dy;
  for (int i = 1; i <= o->Mloc; ++i)
    for (int j = 2; j <= o->Nloc - 1; ++j) {
      if (o->Mask[I2(i, j)] == 0) {
        Dout[I2(i, j)] = ZERO;
      } else {
        double t1 = (Din[I2(i, j + 1)] - Din[I2(i, j)]) / dy;
        double t2 = (Din[I2(i, j)] - Din[I2(i, j - 1)]) / dy;
        if ((fabs(t1) + fabs(t2)) < SMALL) Dout[I2(i, j)] = ZERO;
        else Dout[I2(i, j)] = limiter(t1, t2);
      }
    }
  for (int i = 1; i <= o->Mloc; ++i) {
    Dout[I2(i, 1)] = (Din[I2(i, 2)] - Din[I2(i, 1)]) / dy;
    Dout[I2(i, o->Nloc)] = (Din[I2(i, o->Nloc)] - Din[I2(i, o->Nloc - 1)]) / dy;
  }
}

/* delxFun_3D / delxFun1_3D (nk = Kloc or Kloc1), src/nhwave.F:2097-2179 */
static void delx_3d(orc *o, const double *Din, double *Dout, int nk) {
  const double dx = o->c.dx;
  for (int i = 2; i <= o->Mloc - 1; ++i)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int k = 1; k <= nk; ++k) {
        if (o->Mask[I2(i, j)] == 0) {
          Dout[I3(i, j, k)] = ZERO;
        } else {
          double t1 = (Din[I3(i + 1, j, k)] - Din[I3(i, j, k)]) / dx;
          double t2 = (Din[I3(i, j, k)] - Din[I3(i - 1, j, k)]) / dx;
          if ((fabs(t1) + fabs(t2)) < SMALL) Dout[I3(i, j, k)] = ZERO;
          else Dout[I3(i, j, k)] = limiter(t1, t2);
        }
      }
  for (int j = 1; j <= o->Nloc; ++j)
    for (int k = 1; k <= nk; ++k) {
      Dout[I3(1, j, k)] = (Din[I3(2, j, k)] - Din[I3(1, j, k)]) / dx;
      Dout[I3(o->Mloc, j, k)] = (Din[I3(o->Mloc, j, k)] - Din[I3(o->Mloc - 1, j, k)]) / dx;
    }
}

/* delyFun_3D / delyFun1_3D, src/nhwave.F:2219-2311 */
static void dely_3d(orc *o, const double *Din, double *Dout, int nk) {
  const double dy = o->c.dy;
  for (int i = 1; i <= o->Mloc; ++i)
    for (int j = 2; j <= o->Nloc - 1; ++j)
      for (int k = 1; k <= nk; ++k) {
        if (o->Mask[I2(i, j)] == 0) {
          Dout[I3(i, j, k)] = ZERO;
        } else {
          double t1 = (Din[I3(i, j + 1, k)] - Din[I3(i, j, k)]) / dy;
          double t2 = (Din[I3(i, j, k)] - Din[I3(i, j - 1, k)]) / dy;
          if ((fabs(t1) + fabs(t2)) < SMALL) Dout[I3(i, j, k)] = ZERO;
          else Dout[I3(i, j, k)] = limiter(t1, t2);
        }
      }
  for (int i = 1; i <= o->Mloc; ++i)
    for (int k = 1; k <= nk; ++k) {
      Dout[I3(i, 1, k)] = (Din[I3(i, 2, k)] - Din[I3(i, 1, k)]) / dy;
      Dout[I3(i, o->Nloc, k)] = (Din[I3(i, o->Nloc, k)] - Din[I3(i, o->Nloc - 1, k)]) / dy;
    }
}

/* delzFun_3D, src/nhwave.F:2314-2351 (no mask) */
static void delz_3d(orc *o, const double *Din, double *Dout) {
  const double *sigc = o->sigc - 1, *dsig = o->dsig - 1;
  for (int i = 1; i <= o->Mloc; ++i)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int k = 2; k <= o->Kloc - 1; ++k) {
        double t1 = (Din[I3(i, j, k + 1)] - Din[I3(i, j, k)]) / (sigc[k + 1] - sigc[k]);
        double t2 = (Din[I3(i, j, k)] - Din[I3(i, j, k - 1)]) / (sigc[k] - sigc[k - 1]);
        if ((fabs(t1) + fabs(t2)) < SMALL) Dout[I3(i, j, k)] = ZERO;
        else Dout[I3(i, j, k)] = limiter(t1, t2);
      }
  for (int i = 1; i <= o->Mloc; ++i)
    for (int j = 1; j <= o->Nloc; ++j) {
      Dout[I3(i, j, 1)] = (Din[I3(i, j, 2)] - Din[I3(i, j, 1)]) / (0.5 * (dsig[1] + dsig[2]));
      int K = o->Kloc;
      Dout[I3(i, j, K)] = (Din[I3(i, j, K)] - Din[I3(i, j, K - 1)]) / (0.5 * (dsig[K - 1] + dsig[K]));
    }
}

/* construct_2D_x, src/nhwave.F:1890-1913 */
static void construct_2d_x(orc *o, const double *Vin, const double *Din, double *L, double *R) {
  memset(L, 0, sizeof(double) * o->nx2);
  memset(R, 0, sizeof(double) * o->nx2);
  for (int i = o->Ibeg; i <= o->Iend1; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      L[IX2(i, j)] = Vin[I2(i - 1, j)] + 0.5 * o->c.dx * Din[I2(i - 1, j)];
      R[IX2(i, j)] = Vin[I2(i, j)] - 0.5 * o->c.dx * Din[I2(i, j)];
    }
}
/* construct_3D_x, src/nhwave.F:1916-1941 */
static void construct_3d_x(orc *o, const double *Vin, const double *Din, double *L, double *R) {
  memset(L, 0, sizeof(double) * o->nx3);
  memset(R, 0, sizeof(double) * o->nx3);
  for (int i = o->Ibeg; i <= o->Iend1; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int k = o->Kbeg; k <= o->Kend; ++k) {
        L[IX3(i, j, k)] = Vin[I3(i - 1, j, k)] + 0.5 * o->c.dx * Din[I3(i - 1, j, k)];
        R[IX3(i, j, k)] = Vin[I3(i, j, k)] - 0.5 * o->c.dx * Din[I3(i, j, k)];
      }
}
/* construct_2D_y, src/nhwave.F:1944-1967 */
static void construct_2d_y(orc *o, const double *Vin, const double *Din, double *L, double *R) {
  memset(L, 0, sizeof(double) * o->ny2);
  memset(R, 0, sizeof(double) * o->ny2);
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend1; ++j) {
      L[IY2(i, j)] = Vin[I2(i, j - 1)] + 0.5 * o->c.dy * Din[I2(i, j - 1)];
      R[IY2(i, j)] = Vin[I2(i, j)] - 0.5 * o->c.dy * Din[I2(i, j)];
    }
}
/* construct_3D_y, src/nhwave.F:1970-1995 */
static void construct_3d_y(orc *o, const double *Vin, const double *Din, double *L, double *R) {
  memset(L, 0, sizeof(double) * o->ny3);
  memset(R, 0, sizeof(double) * o->ny3);
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend1; ++j)
      for (int k = o->Kbeg; k <= o->Kend; ++k) {
        L[IY3(i, j, k)] = Vin[I3(i, j - 1, k)] + 0.5 * o->c.dy * Din[I3(i, j - 1, k)];
        R[IY3(i, j, k)] = Vin[I3(i, j, k)] - 0.5 * o->c.dy * Din[I3(i, j, k)];
      }
}
/* construct_3D_z, src/nhwave.F:1998-2023 */
static void construct_3d_z(orc *o, const double *Vin, const double *Din, double *L, double *R) {
  const double *dsig = o->dsig - 1;
  memset(L, 0, sizeof(double) * o->nz3);
  memset(R, 0, sizeof(double) * o->nz3);
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int k = o->Kbeg; k <= o->Kend1; ++k) {
        L[IZ3(i, j, k)] = Vin[I3(i, j, k - 1)] + 0.5 * dsig[k - 1] * Din[I3(i, j, k - 1)];
        R[IZ3(i, j, k)] = Vin[I3(i, j, k)] - 0.5 * dsig[k] * Din[I3(i, j, k)];
      }
}

/* KBC_surface, src/nhwave.F:5069-5088 */
static void kbc_surface(orc *o) {
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      o->WzL[IZ3(i, j, o->Kend1)] = (o->Eta[I2(i, j)] - o->Eta0[I2(i, j)]) / o->dt +
                                    o->DelxEta[I2(i, j)] * o->UzL[IZ3(i, j, o->Kend1)] +
                                    o->DelyEta[I2(i, j)] * o->VzL[IZ3(i, j, o->Kend1)];
}
/* KBC_bottom, src/nhwave.F:5090-5116 (Bc_Z0 = 1) */
static void kbc_bottom(orc *o) {
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      if (o->c.bc_z0 == 2) {  /* no-slip: horizontal bottom-face states vanish */
        o->UzR[IZ3(i, j, o->Kbeg)] = ZERO; o->VzR[IZ3(i, j, o->Kbeg)] = ZERO;
        o->UzL[IZ3(i, j, o->Kbeg)] = ZERO; o->VzL[IZ3(i, j, o->Kbeg)] = ZERO;
      }
      o->WzR[IZ3(i, j, o->Kbeg)] = -o->DeltH[I2(i, j)] - o->DelxH[I2(i, j)] * o->UzR[IZ3(i, j, o->Kbeg)] -
                                   o->DelyH[I2(i, j)] * o->VzR[IZ3(i, j, o->Kbeg)];
    }
}

/* ======================================================================== */
/* fluxes, src/nhwave.F:1686-1714                                           */
/* ======================================================================== */

/* delxyzFun, src/nhwave.F:2026-2057 */
static void delxyzfun(orc *o) {
  delx_2d(o, o->Eta, o->DelxEta);
  delx_3d(o, o->U, o->DelxU, o->Kloc);
  delx_3d(o, o->V, o->DelxV, o->Kloc);
  delx_3d(o, o->W, o->DelxW, o->Kloc);
  delx_3d(o, o->DU, o->DelxDU, o->Kloc);
  delx_3d(o, o->DV, o->DelxDV, o->Kloc);
  delx_3d(o, o->DW, o->DelxDW, o->Kloc);
  dely_2d(o, o->Eta, o->DelyEta);
  dely_3d(o, o->U, o->DelyU, o->Kloc);
  dely_3d(o, o->V, o->DelyV, o->Kloc);
  dely_3d(o, o->W, o->DelyW, o->Kloc);
  dely_3d(o, o->DU, o->DelyDU, o->Kloc);
  dely_3d(o, o->DV, o->DelyDV, o->Kloc);
  dely_3d(o, o->DW, o->DelyDW, o->Kloc);
  delz_3d(o, o->U, o->DelzU);
  delz_3d(o, o->V, o->DelzV);
  delz_3d(o, o->W, o->DelzW);
}

/* construction, src/nhwave.F:1717-1887 */
static void construction(orc *o) {
  const double MinDep = o->c.min_dep, g = GRAV;
  construct_2d_x(o, o->Eta, o->DelxEta, o->EtaxL, o->EtaxR);
  construct_3d_x(o, o->U, o->DelxU, o->UxL, o->UxR);
  construct_3d_x(o, o->V, o->DelxV, o->VxL, o->VxR);
  construct_3d_x(o, o->W, o->DelxW, o->WxL, o->WxR);
  construct_3d_x(o, o->DU, o->DelxDU, o->DUxL, o->DUxR);
  construct_3d_x(o, o->DV, o->DelxDV, o->DVxL, o->DVxR);
  construct_3d_x(o, o->DW, o->DelxDW, o->DWxL, o->DWxR);
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend1; ++i) {
      o->DxL[IX2(i, j)] = o->EtaxL[IX2(i, j)] + o->Hfx[IX2(i, j)];
      o->DxR[IX2(i, j)] = o->EtaxR[IX2(i, j)] + o->Hfx[IX2(i, j)];
    }
  for (long q = 0; q < o->nx2; ++q) {
    o->DxL[q] = dmax(o->DxL[q], MinDep);
    o->DxR[q] = dmax(o->DxR[q], MinDep);
  }
  /* wet-dry front overrides, :1745-1788 */
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        int m = o->Mask[I2(i, j)], mp = o->Mask[I2(i + 1, j)], mm = o->Mask[I2(i - 1, j)];
        if (m == 1 && mp == 0) {
          o->EtaxL[IX2(i + 1, j)] = o->Eta[I2(i, j)];
          o->DxL[IX2(i + 1, j)] = o->D[I2(i, j)];
          o->Hfx0[IX2(i + 1, j)] = o->Hc[I2(i, j)];
          o->DUxL[IX3(i + 1, j, k)] = ZERO; o->DVxL[IX3(i + 1, j, k)] = ZERO;
          o->DWxL[IX3(i + 1, j, k)] = ZERO; o->UxL[IX3(i + 1, j, k)] = ZERO;
          o->VxL[IX3(i + 1, j, k)] = ZERO; o->WxL[IX3(i + 1, j, k)] = ZERO;
        } else if (m == 1 && mm == 0) {
          o->EtaxR[IX2(i, j)] = o->Eta[I2(i, j)];
          o->DxR[IX2(i, j)] = o->D[I2(i, j)];
          o->Hfx0[IX2(i, j)] = o->Hc[I2(i, j)];
          o->DUxR[IX3(i, j, k)] = ZERO; o->DVxR[IX3(i, j, k)] = ZERO;
          o->DWxR[IX3(i, j, k)] = ZERO; o->UxR[IX3(i, j, k)] = ZERO;
          o->VxR[IX3(i, j, k)] = ZERO; o->WxR[IX3(i, j, k)] = ZERO;
        } else if (m == 0) {
          o->EtaxL[IX2(i + 1, j)] = o->Eta[I2(i, j)];
          o->DxL[IX2(i + 1, j)] = o->D[I2(i, j)];
          o->DUxL[IX3(i + 1, j, k)] = ZERO; o->DVxL[IX3(i + 1, j, k)] = ZERO;
          o->DWxL[IX3(i + 1, j, k)] = ZERO; o->UxL[IX3(i + 1, j, k)] = ZERO;
          o->VxL[IX3(i + 1, j, k)] = ZERO; o->WxL[IX3(i + 1, j, k)] = ZERO;
          o->EtaxR[IX2(i, j)] = o->Eta[I2(i, j)];
          o->DxR[IX2(i, j)] = o->D[I2(i, j)];
          o->DUxR[IX3(i, j, k)] = ZERO; o->DVxR[IX3(i, j, k)] = ZERO;
          o->DWxR[IX3(i, j, k)] = ZERO; o->UxR[IX3(i, j, k)] = ZERO;
          o->VxR[IX3(i, j, k)] = ZERO; o->WxR[IX3(i, j, k)] = ZERO;
        }
      }
  /* physical fluxes, :1789-1802 */
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend1; ++i) {
        long f = IX3(i, j, k), f2 = IX2(i, j);
        double eL = o->EtaxL[f2], eR = o->EtaxR[f2], h0 = o->Hfx0[f2];
        o->ExL[f] = o->DUxL[f];
        o->ExR[f] = o->DUxR[f];
        o->FxL[f] = o->DUxL[f] * o->UxL[f] + 0.5 * g * (eL * eL + 2.0 * eL * h0);
        o->FxR[f] = o->DUxR[f] * o->UxR[f] + 0.5 * g * (eR * eR + 2.0 * eR * h0);
        o->GxL[f] = o->DxL[f2] * o->UxL[f] * o->VxL[f];
        o->GxR[f] = o->DxR[f2] * o->UxR[f] * o->VxR[f];
        o->HxL[f] = o->DxL[f2] * o->UxL[f] * o->WxL[f];
        o->HxR[f] = o->DxR[f2] * o->UxR[f] * o->WxR[f];
      }

  construct_2d_y(o, o->Eta, o->DelyEta, o->EtayL, o->EtayR);
  construct_3d_y(o, o->U, o->DelyU, o->UyL, o->UyR);
  construct_3d_y(o, o->V, o->DelyV, o->VyL, o->VyR);
  construct_3d_y(o, o->W, o->DelyW, o->WyL, o->WyR);
  construct_3d_y(o, o->DU, o->DelyDU, o->DUyL, o->DUyR);
  construct_3d_y(o, o->DV, o->DelyDV, o->DVyL, o->DVyR);
  construct_3d_y(o, o->DW, o->DelyDW, o->DWyL, o->DWyR);
  for (int j = o->Jbeg; j <= o->Jend1; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      o->DyL[IY2(i, j)] = o->EtayL[IY2(i, j)] + o->Hfy[IY2(i, j)];
      o->DyR[IY2(i, j)] = o->EtayR[IY2(i, j)] + o->Hfy[IY2(i, j)];
    }
  for (long q = 0; q < o->ny2; ++q) {
    o->DyL[q] = dmax(o->DyL[q], MinDep);
    o->DyR[q] = dmax(o->DyR[q], MinDep);
  }
  /* :1821-1864 */
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        int m = o->Mask[I2(i, j)], mp = o->Mask[I2(i, j + 1)], mm = o->Mask[I2(i, j - 1)];
        if (m == 1 && mp == 0) {
          o->EtayL[IY2(i, j + 1)] = o->Eta[I2(i, j)];
          o->DyL[IY2(i, j + 1)] = o->D[I2(i, j)];
          o->Hfy0[IY2(i, j + 1)] = o->Hc[I2(i, j)];
          o->DUyL[IY3(i, j + 1, k)] = ZERO; o->DVyL[IY3(i, j + 1, k)] = ZERO;
          o->DWyL[IY3(i, j + 1, k)] = ZERO; o->UyL[IY3(i, j + 1, k)] = ZERO;
          o->VyL[IY3(i, j + 1, k)] = ZERO; o->WyL[IY3(i, j + 1, k)] = ZERO;
        } else if (m == 1 && mm == 0) {
          o->EtayR[IY2(i, j)] = o->Eta[I2(i, j)];
          o->DyR[IY2(i, j)] = o->D[I2(i, j)];
          o->Hfy0[IY2(i, j)] = o->Hc[I2(i, j)];
          o->DUyR[IY3(i, j, k)] = ZERO; o->DVyR[IY3(i, j, k)] = ZERO;
          o->DWyR[IY3(i, j, k)] = ZERO; o->UyR[IY3(i, j, k)] = ZERO;
          o->VyR[IY3(i, j, k)] = ZERO; o->WyR[IY3(i, j, k)] = ZERO;
        } else if (m == 0) {
          o->EtayL[IY2(i, j + 1)] = o->Eta[I2(i, j)];
          o->DyL[IY2(i, j + 1)] = o->D[I2(i, j)];
          o->DUyL[IY3(i, j + 1, k)] = ZERO; o->DVyL[IY3(i, j + 1, k)] = ZERO;
          o->DWyL[IY3(i, j + 1, k)] = ZERO; o->UyL[IY3(i, j + 1, k)] = ZERO;
          o->VyL[IY3(i, j + 1, k)] = ZERO; o->WyL[IY3(i, j + 1, k)] = ZERO;
          o->EtayR[IY2(i, j)] = o->Eta[I2(i, j)];
          o->DyR[IY2(i, j)] = o->D[I2(i, j)];
          o->DUyR[IY3(i, j, k)] = ZERO; o->DVyR[IY3(i, j, k)] = ZERO;
          o->DWyR[IY3(i, j, k)] = ZERO; o->UyR[IY3(i, j, k)] = ZERO;
          o->VyR[IY3(i, j, k)] = ZERO; o->WyR[IY3(i, j, k)] = ZERO;
        }
      }
  /* :1865-1878 */
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend1; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        long f = IY3(i, j, k), f2 = IY2(i, j);
        double eL = o->EtayL[f2], eR = o->EtayR[f2], h0 = o->Hfy0[f2];
        o->EyL[f] = o->DVyL[f];
        o->EyR[f] = o->DVyR[f];
        o->FyL[f] = o->DyL[f2] * o->UyL[f] * o->VyL[f];
        o->FyR[f] = o->DyR[f2] * o->UyR[f] * o->VyR[f];
        o->GyL[f] = o->DVyL[f] * o->VyL[f] + 0.5 * g * (eL * eL + 2.0 * eL * h0);
        o->GyR[f] = o->DVyR[f] * o->VyR[f] + 0.5 * g * (eR * eR + 2.0 * eR * h0);
        o->HyL[f] = o->DyL[f2] * o->VyL[f] * o->WyL[f];
        o->HyR[f] = o->DyR[f2] * o->VyR[f] * o->WyR[f];
      }

  construct_3d_z(o, o->U, o->DelzU, o->UzL, o->UzR);
  construct_3d_z(o, o->V, o->DelzV, o->VzL, o->VzR);
  construct_3d_z(o, o->W, o->DelzW, o->WzL, o->WzR);
  kbc_surface(o);
  kbc_bottom(o);
}

/* wave_speed, src/nhwave.F:2813-2886 (speeds persist where both sides dry) */
static void wave_speed(orc *o) {
  const double g = GRAV;
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend1; ++i) {
        int ml = o->Mask[I2(i - 1, j)], mr = o->Mask[I2(i, j)];
        long f = IX3(i, j, k), f2 = IX2(i, j);
        if (ml == 1 && mr == 1) {
          double pl = sqrt(g * fabs(o->DxL[f2])), pr = sqrt(g * fabs(o->DxR[f2]));
          double ps = 0.5 * (pl + pr) + 0.25 * (o->UxL[f] - o->UxR[f]);
          double us = 0.5 * (o->UxL[f] + o->UxR[f]) + pl - pr;
          o->SxL[f] = dmin(o->UxL[f] - pl, us - ps);
          o->SxR[f] = dmax(o->UxR[f] + pr, us + ps);
          o->SxS[f] = us;
        } else if (ml == 0 && mr == 1) {
          double pr = sqrt(g * fabs(o->DxR[f2]));
          o->SxL[f] = o->UxR[f] - 2.0 * pr;
          o->SxR[f] = o->UxR[f] + pr;
          o->SxS[f] = o->SxL[f];
        } else if (ml == 1 && mr == 0) {
          double pl = sqrt(g * fabs(o->DxL[f2]));
          o->SxL[f] = o->UxL[f] - pl;
          o->SxR[f] = o->UxL[f] + 2.0 * pl;
          o->SxS[f] = o->SxR[f];
        }
      }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend1; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        int ml = o->Mask[I2(i, j - 1)], mr = o->Mask[I2(i, j)];
        long f = IY3(i, j, k), f2 = IY2(i, j);
        if (ml == 1 && mr == 1) {
          double pl = sqrt(g * fabs(o->DyL[f2])), pr = sqrt(g * fabs(o->DyR[f2]));
          double ps = 0.5 * (pl + pr) + 0.25 * (o->VyL[f] - o->VyR[f]);
          double us = 0.5 * (o->VyL[f] + o->VyR[f]) + pl - pr;
          o->SyL[f] = dmin(o->VyL[f] - pl, us - ps);
          o->SyR[f] = dmax(o->VyR[f] + pr, us + ps);
          o->SyS[f] = us;
        } else if (ml == 0 && mr == 1) {
          double pr = sqrt(g * fabs(o->DyR[f2]));
          o->SyL[f] = o->VyR[f] - 2.0 * pr;
          o->SyR[f] = o->VyR[f] + pr;
          o->SyS[f] = o->SyL[f];
        } else if (ml == 1 && mr == 0) {
          double pl = sqrt(g * fabs(o->DyL[f2]));
          o->SyL[f] = o->VyL[f] - pl;
          o->SyR[f] = o->VyL[f] + 2.0 * pl;
          o->SyS[f] = o->SyR[f];
        }
      }
}

/* HLL, src/nhwave.F:2741-2776 (whole array, UL/UR given per face) */
static double hll1(double sl, double sr, double fl, double fr, double ul, double ur) {
  if (sl >= ZERO) return fl;
  if (sr <= ZERO) return fr;
  double f = sr * fl - sl * fr + sl * sr * (ur - ul);
  if (fabs(sr - sl) < SMALL) return f / SMALL;
  return f / (sr - sl);
}
static void hll_x(orc *o, const double *FL, const double *FR, const double *UL3,
                  const double *UR3, const double *UL2, const double *UR2, double *out) {
  for (int k = 1; k <= o->Kloc; ++k)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int i = 1; i <= o->Mloc1; ++i) {
        long f = IX3(i, j, k);
        double ul = UL3 ? UL3[f] : UL2[IX2(i, j)];
        double ur = UR3 ? UR3[f] : UR2[IX2(i, j)];
        out[f] = hll1(o->SxL[f], o->SxR[f], FL[f], FR[f], ul, ur);
      }
}
static void hll_y(orc *o, const double *FL, const double *FR, const double *UL3,
                  const double *UR3, const double *UL2, const double *UR2, double *out) {
  for (int k = 1; k <= o->Kloc; ++k)
    for (int j = 1; j <= o->Nloc1; ++j)
      for (int i = 1; i <= o->Mloc; ++i) {
        long f = IY3(i, j, k);
        double ul = UL3 ? UL3[f] : UL2[IY2(i, j)];
        double ur = UR3 ? UR3[f] : UR2[IY2(i, j)];
        out[f] = hll1(o->SyL[f], o->SyR[f], FL[f], FR[f], ul, ur);
      }
}

/* fluxes_at_faces_HLL, src/nhwave.F:2673-2738 */
static void fluxes_at_faces_hll(orc *o) {
  hll_x(o, o->ExL, o->ExR, NULL, NULL, o->DxL, o->DxR, o->Ex);
  hll_y(o, o->EyL, o->EyR, NULL, NULL, o->DyL, o->DyR, o->Ey);
  hll_x(o, o->FxL, o->FxR, o->DUxL, o->DUxR, NULL, NULL, o->Fx);
  hll_y(o, o->FyL, o->FyR, o->DUyL, o->DUyR, NULL, NULL, o->Fy);
  hll_x(o, o->GxL, o->GxR, o->DVxL, o->DVxR, NULL, NULL, o->Gx);
  hll_y(o, o->GyL, o->GyR, o->DVyL, o->DVyR, NULL, NULL, o->Gy);
  hll_x(o, o->HxL, o->HxR, o->DWxL, o->DWxR, NULL, NULL, o->Hx);
  hll_y(o, o->HyL, o->HyR, o->DWyL, o->DWyR, NULL, NULL, o->Hy);
  for (int k = o->Kbeg + 1; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        long q = IZ3(i, j, k);
        double om = o->Omega[q];
        o->Fz[q] = 0.5 * (om * (o->UzL[q] + o->UzR[q]) - fabs(om) * (o->UzR[q] - o->UzL[q]));
        o->Gz[q] = 0.5 * (om * (o->VzL[q] + o->VzR[q]) - fabs(om) * (o->VzR[q] - o->VzL[q]));
        o->Hz[q] = 0.5 * (om * (o->WzL[q] + o->WzR[q]) - fabs(om) * (o->WzR[q] - o->WzL[q]));
      }
}

static int is_wall(int bc) { return bc == 1 || bc == 2 || bc == 5; }

/* flux_bc, src/nhwave.F:2354-2552 (serial, walls) */
static void flux_bc(orc *o) {
  const double g = GRAV;
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int k = o->Kbeg; k <= o->Kend; ++k)
      if (is_wall(o->c.bc_x0)) {
        long f = IX3(o->Ibeg, j, k);
        double e = o->EtaxR[IX2(o->Ibeg, j)];
        o->Ex[f] = ZERO;
        o->Fx[f] = 0.5 * g * (e * e + 2.0 * e * o->Hfx0[IX2(o->Ibeg, j)]);
        o->Gx[f] = ZERO;
        o->Hx[f] = ZERO;
      } else if (o->c.bc_x0 == 3) { /* inflow, :2381-2386 */
        long f = IX3(o->Ibeg, j, k);
        double din = o->Din_X0[j - 1], ein = o->Ein_X0[j - 1];
        double uin = o->Uin_X0[IB(j, k)], vin = o->Vin_X0[IB(j, k)], win = o->Win_X0[IB(j, k)];
        o->Ex[f] = din * uin;
        o->Fx[f] = din * uin * uin + 0.5 * g * (ein * ein + 2.0 * ein * o->Hfx0[IX2(o->Ibeg, j)]);
        o->Gx[f] = din * uin * vin;
        o->Hx[f] = din * uin * win;
      }
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int k = o->Kbeg; k <= o->Kend; ++k)
      if (is_wall(o->c.bc_xn)) {
        long f = IX3(o->Iend1, j, k);
        double e = o->EtaxL[IX2(o->Iend1, j)];
        o->Ex[f] = ZERO;
        o->Fx[f] = 0.5 * g * (e * e + 2.0 * e * o->Hfx0[IX2(o->Iend1, j)]);
        o->Gx[f] = ZERO;
        o->Hx[f] = ZERO;
      } else if (o->c.bc_xn == 3) { /* open east boundary, :2411-2416 */
        long f = IX3(o->Iend1, j, k);
        double din = o->Din_Xn[j - 1], ein = o->Ein_Xn[j - 1];
        double uin = o->Uin_Xn[IB(j, k)], vin = o->Vin_Xn[IB(j, k)], win = o->Win_Xn[IB(j, k)];
        o->Ex[f] = din * uin;
        o->Fx[f] = din * uin * uin + 0.5 * g * (ein * ein + 2.0 * ein * o->Hfx0[IX2(o->Iend1, j)]);
        o->Gx[f] = din * uin * vin;
        o->Hx[f] = din * uin * win;
      }
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int k = o->Kbeg; k <= o->Kend; ++k)
      if (is_wall(o->c.bc_y0)) {
        long f = IY3(i, o->Jbeg, k);
        double e = o->EtayR[IY2(i, o->Jbeg)];
        o->Ey[f] = ZERO;
        o->Fy[f] = ZERO;
        o->Gy[f] = 0.5 * g * (e * e + 2.0 * e * o->Hfy0[IY2(i, o->Jbeg)]);
        o->Hy[f] = ZERO;
      }
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int k = o->Kbeg; k <= o->Kend; ++k)
      if (is_wall(o->c.bc_yn)) {
        long f = IY3(i, o->Jend1, k);
        double e = o->EtayL[IY2(i, o->Jend1)];
        o->Ey[f] = ZERO;
        o->Fy[f] = ZERO;
        o->Gy[f] = 0.5 * g * (e * e + 2.0 * e * o->Hfy0[IY2(i, o->Jend1)]);
        o->Hy[f] = ZERO;
      }
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      o->Fz[IZ3(i, j, o->Kbeg)] = ZERO; o->Gz[IZ3(i, j, o->Kbeg)] = ZERO; o->Hz[IZ3(i, j, o->Kbeg)] = ZERO;
      o->Fz[IZ3(i, j, o->Kend1)] = ZERO; o->Gz[IZ3(i, j, o->Kend1)] = ZERO; o->Hz[IZ3(i, j, o->Kend1)] = ZERO;
    }
  /* dry cells, :2488-2550 */
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg - 1; j <= o->Jend + 1; ++j)
      for (int i = o->Ibeg - 1; i <= o->Iend + 1; ++i) {
        if (o->Mask[I2(i, j)] != 0) continue;
        o->Ex[IX3(i, j, k)] = ZERO;
        if (i == o->Ibeg) {
          o->Fx[IX3(i, j, k)] = ZERO;
        } else {
          double e = o->EtaxL[IX2(i, j)];
          o->Fx[IX3(i, j, k)] = 0.5 * g * (e * e + 2.0 * e * o->Hfx0[IX2(i, j)]) * o->Mask[I2(i - 1, j)];
        }
        o->Gx[IX3(i, j, k)] = ZERO;
        o->Hx[IX3(i, j, k)] = ZERO;
        o->Ex[IX3(i + 1, j, k)] = ZERO;
        if (i == o->Iend) {
          o->Fx[IX3(i + 1, j, k)] = ZERO;
        } else {
          double e = o->EtaxR[IX2(i + 1, j)];
          o->Fx[IX3(i + 1, j, k)] = 0.5 * g * (e * e + 2.0 * e * o->Hfx0[IX2(i + 1, j)]) * o->Mask[I2(i + 1, j)];
        }
        o->Gx[IX3(i + 1, j, k)] = ZERO;
        o->Hx[IX3(i + 1, j, k)] = ZERO;
        o->Ey[IY3(i, j, k)] = ZERO;
        o->Fy[IY3(i, j, k)] = ZERO;
        if (j == o->Jbeg) {
          o->Gy[IY3(i, j, k)] = ZERO;
        } else {
          double e = o->EtayL[IY2(i, j)];
          o->Gy[IY3(i, j, k)] = 0.5 * g * (e * e + 2.0 * e * o->Hfy0[IY2(i, j)]) * o->Mask[I2(i, j - 1)];
        }
        o->Hy[IY3(i, j, k)] = ZERO;
        o->Ey[IY3(i, j + 1, k)] = ZERO;
        o->Fy[IY3(i, j + 1, k)] = ZERO;
        if (j == o->Jend) {
          o->Gy[IY3(i, j + 1, k)] = ZERO;
        } else {
          double e = o->EtayR[IY2(i, j + 1)];
          o->Gy[IY3(i, j + 1, k)] = 0.5 * g * (e * e + 2.0 * e * o->Hfy0[IY2(i, j + 1)]) * o->Mask[I2(i, j + 1)];
        }
        o->Hy[IY3(i, j + 1, k)] = ZERO;
      }
}

void orc_fluxes(orc *o) {
  delxyzfun(o);
  construction(o);
  wave_speed(o);
  fluxes_at_faces_hll(o);
  flux_bc(o);
}

/* ======================================================================== */
/* boundary conditions                                                      */
/* ======================================================================== */

/* vel_bc, src/nhwave.F:4416-5066 (lateral BC 1/2, Bc_Z0 = 1) */
/* the wall friction boundary, code 5 (vel_bc :4486-4501, :4583-4598,
 * :4676-4691, :4737-4752): Cdrag = Cd0 (Ibot = 1) or the log law at half a
 * cell, 1/(1/Kappa log(30 h/2/Zob))**2, and the ghost factor 1 - Phi with
 * Phi = min(h Cdrag |u_t| / (Cmu + CmuHt), 2), |u_t| from the two velocity
 * components along the wall at the first interior cell c */
static double wall_phi(const orc *o, double h, double a, double b, long c) {
  double cdrag = o->c.cd0;
  if (o->c.ibot != 1) {
    const double t = 1. / KAPPA * log(30.0 * h / 2.0 / o->c.zob);
    cdrag = 1. / (t * t);
  }
  const double phi = h * cdrag * sqrt(a * a + b * b) / (o->c.visc + o->CmuHt[c]);
  return phi < 2.0 ? phi : 2.0;
}

void orc_vel_bc(orc *o) {
  const int Ib = o->Ibeg, Ie = o->Iend, Jb = o->Jbeg, Je = o->Jend, Kb = o->Kbeg, Ke = o->Kend;
  double *U = o->U, *V = o->V, *W = o->W, *DU = o->DU, *DV = o->DV, *DW = o->DW;
  const double MinDepBC = F32(0.01);           /* :4438 */
  /* west / east, :4440-4638 */
  for (int k = Kb; k <= Ke; ++k)
    for (int j = Jb; j <= Je; ++j)
      for (int i = 1; i <= 2; ++i) {
        long g = I3(Ib - i, j, k), s = I3(Ib + i - 1, j, k);
        if (o->c.bc_x0 == 3) { /* inflow, :4468-4476 */
          double din = o->Din_X0[j - 1];
          double uin = o->Uin_X0[IB(j, k)], vin = o->Vin_X0[IB(j, k)], win = o->Win_X0[IB(j, k)];
          U[g] = 2.0 * uin - U[s]; V[g] = 2.0 * vin - V[s]; W[g] = 2.0 * win - W[s];
          DU[g] = 2.0 * din * uin - DU[s]; DV[g] = 2.0 * din * vin - DV[s];
          DW[g] = 2.0 * din * win - DW[s];
          continue;
        }
        if (o->c.bc_x0 == 5) {  /* wall friction; the ghosts' D(i,j) is as written, :4494-4500 */
          const long b = I3(Ib, j, k);
          const double ph = wall_phi(o, o->c.dx, V[b], W[b], b), dg = o->D[I2(i, j)];
          U[g] = -U[s]; V[g] = (1.0 - ph) * V[s]; W[g] = (1.0 - ph) * W[s];
          DU[g] = dg * U[g]; DV[g] = dg * V[g]; DW[g] = dg * W[g];
          continue;
        }
        double sv = o->c.bc_x0 == 2 ? -1.0 : 1.0;
        U[g] = -U[s]; V[g] = sv == 1.0 ? V[s] : -V[s]; W[g] = sv == 1.0 ? W[s] : -W[s];
        DU[g] = -DU[s]; DV[g] = sv == 1.0 ? DV[s] : -DV[s]; DW[g] = sv == 1.0 ? DW[s] : -DW[s];
      }
  for (int k = Kb; k <= Ke; ++k)
    for (int j = Jb; j <= Je; ++j)
      for (int i = 1; i <= 2; ++i) {
        long g = I3(Ie + i, j, k), s = I3(Ie - i + 1, j, k);
        if (o->c.bc_xn == 3) {  /* open east boundary with Orlanski radiation, :4547-4573 */
          const long a = I3(Ie + i - 1, j, k), b = I3(Ie + i - 2, j, k);
          double cg = -(o->U0[a] - o->U00[a] + SMALL) / (o->U00[a] - o->U00[b] + SMALL);
          cg = fmax(fmin(cg, 1.0), 0.0);
          if (i == 1)
            U[g] = cg * U[a] + o->Uin_Xni0[IB(j, k)] - cg * o->Uin_Xni[IB(j, k)];
          else
            U[g] = cg * U[a] + (1.0 - cg) * U[g];
          cg = -(o->V0[a] - o->V00[a] + SMALL) / (o->V00[a] - o->V00[b] + SMALL);
          cg = fmax(fmin(cg, 1.0), 0.0);
          V[g] = cg * o->V0[a] + (1.0 - cg) * o->V0[g];
          cg = -(o->W0[a] - o->W00[a] + SMALL) / (o->W00[a] - o->W00[b] + SMALL);
          cg = fmax(fmin(cg, 1.0), 0.0);
          W[g] = cg * o->W0[a] + (1.0 - cg) * o->W0[g];
          const double din = o->Din_Xn[j - 1];
          DU[g] = 2.0 * din * o->Uin_Xn[IB(j, k)] - DU[s];
          DV[g] = 2.0 * din * o->Vin_Xn[IB(j, k)] - DV[s];
          DW[g] = 2.0 * din * o->Win_Xn[IB(j, k)] - DW[s];
          continue;
        }
        if (o->c.bc_xn == 5) {  /* wall friction, :4583-4598 (D(i,j) as written) */
          const long b = I3(Ie, j, k);
          const double ph = wall_phi(o, o->c.dx, V[b], W[b], b), dg = o->D[I2(i, j)];
          U[g] = -U[s]; V[g] = (1.0 - ph) * V[s]; W[g] = (1.0 - ph) * W[s];
          DU[g] = dg * U[g]; DV[g] = dg * V[g]; DW[g] = dg * W[g];
          continue;
        }
        int ns = o->c.bc_xn == 2;
        U[g] = -U[s]; V[g] = ns ? -V[s] : V[s]; W[g] = ns ? -W[s] : W[s];
        DU[g] = -DU[s]; DV[g] = ns ? -DV[s] : DV[s]; DW[g] = ns ? -DW[s] : DW[s];
      }
  /* south / north, :4640-4761 */
  for (int k = Kb; k <= Ke; ++k)
    for (int i = Ib; i <= Ie; ++i)
      for (int j = 1; j <= 2; ++j) {
        long g = I3(i, Jb - j, k), s = I3(i, Jb + j - 1, k);
        if (o->c.bc_y0 == 5) {  /* wall friction, :4676-4691 (D(i,j) as written) */
          const long b = I3(i, Jb, k);
          const double ph = wall_phi(o, o->c.dy, U[b], W[b], b), dg = o->D[I2(i, j)];
          U[g] = (1.0 - ph) * U[s]; V[g] = -V[s]; W[g] = (1.0 - ph) * W[s];
          DU[g] = dg * U[g]; DV[g] = dg * V[g]; DW[g] = dg * W[g];
          continue;
        }
        int ns = o->c.bc_y0 == 2;
        U[g] = ns ? -U[s] : U[s]; V[g] = -V[s]; W[g] = ns ? -W[s] : W[s];
        DU[g] = ns ? -DU[s] : DU[s]; DV[g] = -DV[s]; DW[g] = ns ? -DW[s] : DW[s];
      }
  for (int k = Kb; k <= Ke; ++k)
    for (int i = Ib; i <= Ie; ++i)
      for (int j = 1; j <= 2; ++j) {
        long g = I3(i, Je + j, k), s = I3(i, Je - j + 1, k);
        if (o->c.bc_yn == 5) {  /* wall friction, :4737-4752 (D(i,j) as written) */
          const long b = I3(i, Je, k);
          const double ph = wall_phi(o, o->c.dy, U[b], W[b], b), dg = o->D[I2(i, j)];
          U[g] = (1.0 - ph) * U[s]; V[g] = -V[s]; W[g] = (1.0 - ph) * W[s];
          DU[g] = dg * U[g]; DV[g] = dg * V[g]; DW[g] = dg * W[g];
          continue;
        }
        int ns = o->c.bc_yn == 2;
        U[g] = ns ? -U[s] : U[s]; V[g] = -V[s]; W[g] = ns ? -W[s] : W[s];
        DU[g] = ns ? -DU[s] : DU[s]; DV[g] = -DV[s]; DW[g] = ns ? -DW[s] : DW[s];
      }
  /* vertical ghosts, :4763-4956 */
  static double *dx_uzl, *dx_vzl, *dx_wzl, *dy_uzl, *dy_vzl, *dy_wzl;
  static double *dx_uzr, *dx_vzr, *dx_wzr, *dy_uzr, *dy_vzr, *dy_wzr;
  static long cap = 0;
  if (cap < o->nz3) {
    double **sl[] = {&dx_uzl, &dx_vzl, &dx_wzl, &dy_uzl, &dy_vzl, &dy_wzl,
                     &dx_uzr, &dx_vzr, &dx_wzr, &dy_uzr, &dy_vzr, &dy_wzr};
    for (int q = 0; q < 12; ++q) {
      free(*sl[q]);
      *sl[q] = zd(o->nz3);
    }
    cap = o->nz3;
  }
  delz_3d(o, U, o->DelzU);
  delz_3d(o, V, o->DelzV);
  delz_3d(o, W, o->DelzW);
  construct_3d_z(o, U, o->DelzU, o->UzL, o->UzR);
  construct_3d_z(o, V, o->DelzV, o->VzL, o->VzR);
  construct_3d_z(o, W, o->DelzW, o->WzL, o->WzR);
  kbc_surface(o);
  kbc_bottom(o);
  delx_3d(o, o->UzL, dx_uzl, o->Kloc1);
  delx_3d(o, o->VzL, dx_vzl, o->Kloc1);
  delx_3d(o, o->WzL, dx_wzl, o->Kloc1);
  dely_3d(o, o->UzL, dy_uzl, o->Kloc1);
  dely_3d(o, o->VzL, dy_vzl, o->Kloc1);
  dely_3d(o, o->WzL, dy_wzl, o->Kloc1);
  if (o->c.bc_z0 != 1) {  /* the *R slopes feed only the Bc_Z0 /= 1 branches */
    delx_3d(o, o->UzR, dx_uzr, o->Kloc1);
    delx_3d(o, o->VzR, dx_vzr, o->Kloc1);
    delx_3d(o, o->WzR, dx_wzr, o->Kloc1);
    dely_3d(o, o->UzR, dy_uzr, o->Kloc1);
    dely_3d(o, o->VzR, dy_vzr, o->Kloc1);
    dely_3d(o, o->WzR, dy_wzr, o->Kloc1);
  }
  const double *sigc = o->sigc - 1, *dsig1 = o->dsig - 1;
  const double visc = o->c.visc;
  const int Ke1 = o->Kend1;
  for (int j = Jb; j <= Je; ++j)
    for (int i = Ib; i <= Ie; ++i) {
      long c2 = I2(i, j);
      double dxe = o->DelxEta[c2], dye = o->DelyEta[c2], Dd = o->D[c2];
      double Atop = sqrt(1.0 + dxe * dxe + dye * dye);
      double fw1 = 1.0 / Atop * ((1.0 - dxe * dxe) * o->Wsx[c2] - dxe * dye * o->Wsy[c2]);
      double fw2 = 1.0 / Atop * ((1.0 - dye * dye) * o->Wsy[c2] - dxe * dye * o->Wsx[c2]);
      /* Cmu = Visc, CmuR = 0; CmuVt of the bottom row and of the two levels
       * around the surface (0 on the inviscid path) */
      const double cv = o->CmuVt[I3(i, j, Kb)];
      double cmu6 = 0.5 * (((((visc + 0.0) + o->CmuVt[I3(i, j, Ke)]) + visc) + 0.0) + o->CmuVt[I3(i, j, Ke + 1)]);
      long zt = IZ3(i, j, Ke1), zb = IZ3(i, j, Kb);
      /* bottom stress, :4800-4858 (Bc_Z0 = 2 no-slip, 5 friction law) */
      const double dxh = o->DelxH[c2], dyh = o->DelyH[c2];
      double fb1 = 0.0, fb2 = 0.0, Abot = 1.0;
      if (o->c.bc_z0 != 1) {
        Abot = sqrt(1.0 + dxh * dxh + dyh * dyh);
        const double Bbot = sqrt(1.0 + dxh * dxh);
        const long b0 = I3(i, j, Kb);
        const double dup = (DU[b0] - dxh * DW[b0]) / Bbot;
        /* `-DelxH*DelyH*DU* +(1.0+DelxH**2)*DV -DelyH*DW` as the reference's
         * compiler reads the unary plus after `*` */
        const double dvp = (-(dxh * dyh * DU[b0] * (1.0 + dxh * dxh) * DV[b0]) - dyh * DW[b0]) / Abot / Bbot;
        const double dumag = sqrt(dup * dup + dvp * dvp);
        const double Dz1 = 0.5 * Dd * dsig1[Kb] / Abot;
        const double Z0 = o->c.zob / 30.0;
        double cdrag;
        if (o->c.ibot == 1) cdrag = o->c.cd0;
        else {
          const double t = KAPPA / log(Dz1 / Z0);
          cdrag = t * t;
        }
        cdrag = dmin(cdrag, ((visc + 0.0) + cv) / Dz1 / (dumag + SMALL) * Dd);
        double tb1 = 0.0, tb2 = 0.0;
        if (o->c.bc_z0 == 2) {
          tb1 = dup / Dd / Dz1 * ((visc + 0.0) + cv);
          tb2 = dvp / Dd / Dz1 * ((visc + 0.0) + cv);
        } else {
          tb1 = cdrag * dup * dumag / (Dd * Dd);
          tb2 = cdrag * dvp * dumag / (Dd * Dd);
        }
        fb1 = Bbot * tb1;
        fb2 = (dxh * dyh * tb1 + Abot * tb2) / Bbot;
      }
      for (int k = 1; k <= 2; ++k) {
        long gb = I3(i, j, Kb - k), sb = I3(i, j, Kb + k - 1);
        DW[gb] = 2.0 * Dd * o->WzR[zb] - DW[sb];
        if (o->c.bc_z0 == 1) {
          DU[gb] = DU[sb];
          DV[gb] = DV[sb];
        } else {
          const double dsb = sigc[Kb + k - 1] - sigc[Kb - k];
          double t1 = fb1 * (Dd * Dd / Abot / (((SMALL + visc) + 0.0) + cv));
          double t2 = dxh * (DW[sb] - Dd * o->WzR[zb]) / (dsb / 2.0);
          double t3 = Dd * Dd / (Abot * Abot) *
                      (-2.0 * dxh * dx_uzr[zb] - (1.0 - dxh * dxh) * dx_wzr[zb] -
                       dyh * (dy_uzr[zb] + dx_vzr[zb] - dxh * dy_wzr[zb]));
          DU[gb] = DU[sb] - dsb * (t1 + t2 + t3);
          double old = 2.0 * Dd * o->UzR[zb] - DU[sb];
          if (old > 0.) DU[gb] = dmin(3.0 * old, dmax(-1.0 * old, DU[gb]));
          else if (old < 0.) DU[gb] = dmax(3.0 * old, dmin(-1.0 * old, DU[gb]));
          if (Dd < MinDepBC) DU[gb] = DU[sb];
          t1 = fb2 * (Dd * Dd / Abot / (((SMALL + visc) + 0.0) + cv));
          t2 = dyh * (DW[sb] - Dd * o->WzR[zb]) / (dsb / 2.0);
          t3 = Dd * Dd / (Abot * Abot) *
               (-2.0 * dyh * dy_vzr[zb] - (1.0 - dyh * dyh) * dy_wzr[zb] -
                dxh * (dx_vzr[zb] + dy_uzr[zb] - dyh * dx_wzr[zb]));
          DV[gb] = DV[sb] - dsb * (t1 + t2 + t3);
          old = 2.0 * Dd * o->VzR[zb] - DV[sb];
          if (old > 0.) DV[gb] = dmin(3.0 * old, dmax(-1.0 * old, DV[gb]));
          else if (old < 0.) DV[gb] = dmax(3.0 * old, dmin(-1.0 * old, DV[gb]));
          if (Dd < MinDepBC) DV[gb] = DV[sb];
        }

        long gt = I3(i, j, Ke + k), st = I3(i, j, Ke - k + 1);
        double dsg = sigc[Ke + k] - sigc[Ke - k + 1];
        DW[gt] = 2.0 * Dd * o->WzL[zt] - DW[st];
        double t1 = fw1 * (Dd * Dd / Atop / (SMALL + cmu6));
        double t2 = dxe * (Dd * o->WzL[zt] - DW[st]) / (dsg / 2.0);
        double t3 = Dd * Dd / (Atop * Atop) *
                    (+2.0 * dxe * dx_uzl[zt] - (1.0 - dxe * dxe) * dx_wzl[zt] +
                     dye * (dy_uzl[zt] + dx_vzl[zt] + dxe * dy_wzl[zt]));
        DU[gt] = DU[st] + dsg * (t1 - t2 + t3);
        double old = 2.0 * Dd * o->UzL[zt] - DU[st];
        if (old > 0.) DU[gt] = dmin(3.0 * old, dmax(-1.0 * old, DU[gt]));
        else if (old < 0.) DU[gt] = dmax(3.0 * old, dmin(-1.0 * old, DU[gt]));
        if (Dd < MinDepBC) DU[gt] = DU[st];

        t1 = fw2 * (Dd * Dd / Atop / (SMALL + cmu6));
        t2 = dye * (Dd * o->WzL[zt] - DW[st]) / (dsg / 2.0);
        t3 = Dd * Dd / (Atop * Atop) *
             (+2.0 * dye * dy_vzl[zt] - (1.0 - dye * dye) * dy_wzl[zt] +
              dxe * (dx_vzl[zt] + dy_uzl[zt] + dye * dx_wzl[zt]));
        DV[gt] = DV[st] + dsg * (t1 - t2 + t3);
        old = 2.0 * Dd * o->VzL[zt] - DV[st];
        if (old > 0.) DV[gt] = dmin(3.0 * old, dmax(-1.0 * old, DV[gt]));
        else if (old < 0.) DV[gt] = dmax(3.0 * old, dmin(-1.0 * old, DV[gt]));
        if (Dd < MinDepBC) DV[gt] = DV[st];

        U[gb] = DU[gb] / Dd; V[gb] = DV[gb] / Dd; W[gb] = DW[gb] / Dd;
        U[gt] = DU[gt] / Dd; V[gt] = DV[gt] / Dd; W[gt] = DW[gt] / Dd;
      }
    }
}

/* wl_bc, src/nhwave.F:5119-5262 (walls; inflow at the west, :5144-5153) */
static void wl_bc(orc *o) {
  for (int j = o->Jbeg; j <= o->Jend; ++j) {
    if (o->c.bc_x0 == 3 && o->c.wavemaker == 6)  /* FLUX_LR: the first column too, :5145-5148 */
      for (int i = 0; i <= 2; ++i) o->D[I2(o->Ibeg - i, j)] = o->Din_X0[j - 1];
    else
      for (int i = 1; i <= 2; ++i) {
        if (o->c.bc_x0 == 3)
          o->D[I2(o->Ibeg - i, j)] = 2.0 * o->Din_X0[j - 1] - o->D[I2(o->Ibeg + i - 1, j)];
        else
          o->D[I2(o->Ibeg - i, j)] = o->D[I2(o->Ibeg + i - 1, j)];
      }
  }
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = 1; i <= 2; ++i) {
      if (o->c.bc_xn == 3) {  /* Orlanski radiation of Eta, :5182-5193 */
        const long a = I2(o->Iend + i - 1, j), b = I2(o->Iend + i - 2, j), e = I2(o->Iend + i, j);
        double cg = -(o->Eta0[a] - o->Eta00[a] + SMALL) / (o->Eta00[a] - o->Eta00[b] + SMALL);
        cg = fmax(fmin(cg, 1.0), 0.0);
        o->Eta[e] = cg * o->Eta0[a] + (1.0 - cg) * o->Eta0[e];
        o->D[e] = o->Hc[e] + o->Eta[e];
      } else {
        o->D[I2(o->Iend + i, j)] = o->D[I2(o->Iend - i + 1, j)];
      }
    }
  for (int i = 1; i <= o->Mloc; ++i)
    for (int j = 1; j <= 2; ++j) o->D[I2(i, j)] = o->D[I2(i, o->Jbeg + 2 - j)];
  for (int i = 1; i <= o->Mloc; ++i)
    for (int j = 1; j <= 2; ++j) o->D[I2(i, o->Jend + j)] = o->D[I2(i, o->Jend - j + 1)];
}

/* phi_2D_coll, src/nhwave.F:5264-5330 */
static void phi_2d_coll(orc *o, double *phi) {
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = 1; i <= 2; ++i) phi[I2(i, j)] = phi[I2(o->Ibeg + 2 - i, j)];
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = 1; i <= 2; ++i) phi[I2(o->Iend + i, j)] = phi[I2(o->Iend - i + 1, j)];
  for (int i = 1; i <= o->Mloc; ++i)
    for (int j = 1; j <= 2; ++j) phi[I2(i, j)] = phi[I2(i, o->Jbeg + 2 - j)];
  for (int i = 1; i <= o->Mloc; ++i)
    for (int j = 1; j <= 2; ++j) phi[I2(i, o->Jend + j)] = phi[I2(i, o->Jend - j + 1)];
}

/* get_UVW, src/nhwave.F:1651-1683 */
static void get_uvw(orc *o) {
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i)
      for (int k = o->Kbeg; k <= o->Kend; ++k) {
        double d = o->D[I2(i, j)];
        o->U[I3(i, j, k)] = o->DU[I3(i, j, k)] / d;
        o->V[I3(i, j, k)] = o->DV[I3(i, j, k)] / d;
        o->W[I3(i, j, k)] = o->DW[I3(i, j, k)] / d;
      }
  orc_vel_bc(o);
}

/* ======================================================================== */
/* per-step routines                                                        */
/* ======================================================================== */

/* estimate_dt, src/nhwave.F:3694-3874 (inviscid, no slides) */
int orc_estimate_dt(orc *o) {
  o->dt_old = o->dt;
  double dt_growth = F32(1.05) * o->dt_old;
  double tmp2 = LARGE;
  const double *dsig = o->dsig - 1;
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        double d = o->D[I2(i, j)];
        double t1 = fabs(o->U[I3(i, j, k)]) + sqrt(GRAV * d);
        t1 = dmax(t1, SMALL);
        double v = o->c.dx / t1;
        if (v < tmp2) tmp2 = v;
        t1 = fabs(o->V[I3(i, j, k)]) + sqrt(GRAV * d);
        t1 = dmax(t1, SMALL);
        v = o->c.dy / t1;
        if (v < tmp2) tmp2 = v;
        t1 = dmax(fabs(o->W[I3(i, j, k)]), SMALL);
        v = dsig[k] * d / t1;
        if (v < tmp2) tmp2 = v;
      }
  double dt_courant = o->c.cfl * tmp2;
  /* explicit viscous stress limit, :3812-3832 (1.e-16 is a REAL(4) literal) */
  double dt_viscous = LARGE;
  if (o->c.viscous) {
    double t2 = LARGE;
    for (int k = o->Kbeg; k <= o->Kend; ++k)
      for (int j = o->Jbeg; j <= o->Jend; ++j)
        for (int i = o->Ibeg; i <= o->Iend; ++i) {
          const double ch = o->CmuHt[I3(i, j, k)];
          double t1 = o->c.dx * o->c.dx / (fabs(ch) + SMALL);
          if (t1 < t2) t2 = t1;
          t1 = o->c.dy * o->c.dy / (fabs(ch) + SMALL);
          if (t1 < t2) t2 = t1;
        }
    dt_viscous = o->c.visc_number * t2;
  }
  double dt = dmin(dmin(dmin(dt_growth, dt_courant), dt_viscous), o->c.dt_max);
  o->dt = dt;
  if (dt < o->c.dt_min) {
    snprintf(g_err, sizeof g_err, "time step too small !! %g %g", dt, dt_courant);
    return -1;
  }
  o->time = o->time + dt;
  o->run_step += 1;
  return 0;
}

/* update_mask, src/nhwave.F:5403-5456 */
void orc_update_mask(orc *o) {
  const double MinDep = o->c.min_dep, thr = F32(1.e-6);
  int *tmp = (int *)malloc(sizeof(int) * o->n2);
  memcpy(tmp, o->Mask, sizeof(int) * o->n2);
  const int *M = o->Mask;
  double *Eta = o->Eta, *D = o->D;
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      if (o->Mask_Struct[I2(i, j)] == 0) continue;
      long c = I2(i, j);
      if (M[c] == 0) {
        if (M[I2(i - 1, j)] == 1 && Eta[I2(i - 1, j)] > Eta[c] && D[I2(i - 1, j)] - MinDep > thr) tmp[c] = 1;
        if (M[I2(i + 1, j)] == 1 && Eta[I2(i + 1, j)] > Eta[c] && D[I2(i + 1, j)] - MinDep > thr) tmp[c] = 1;
        if (M[I2(i, j - 1)] == 1 && Eta[I2(i, j - 1)] > Eta[c] && D[I2(i, j - 1)] - MinDep > thr) tmp[c] = 1;
        if (M[I2(i, j + 1)] == 1 && Eta[I2(i, j + 1)] > Eta[c] && D[I2(i, j + 1)] - MinDep > thr) tmp[c] = 1;
      } else {
        if (fabs(D[c] - MinDep) <= thr) {
          tmp[c] = 0;
          Eta[c] = MinDep - o->Hc[c];
          D[c] = Eta[c] + o->Hc[c];
        }
      }
    }
  for (long q = 0; q < o->n2; ++q) o->Mask[q] = tmp[q] * o->Mask_Struct[q];
  free(tmp);
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i)
      o->Mask9[I2(i, j)] = M[I2(i, j)] * M[I2(i - 1, j)] * M[I2(i + 1, j)] * M[I2(i + 1, j + 1)] *
                           M[I2(i, j + 1)] * M[I2(i - 1, j + 1)] * M[I2(i + 1, j - 1)] *
                           M[I2(i, j - 1)] * M[I2(i - 1, j - 1)];
}

/* update_vars, src/nhwave.F:5459-5511 (fields the hot path reads back) */
void orc_update_vars(orc *o) {
  memcpy(o->Eta00, o->Eta0, sizeof(double) * o->n2);  /* :5470-5473 */
  memcpy(o->U00, o->U0, sizeof(double) * o->n3);
  memcpy(o->V00, o->V0, sizeof(double) * o->n3);
  memcpy(o->W00, o->W0, sizeof(double) * o->n3);
  memcpy(o->U0, o->U, sizeof(double) * o->n3);
  memcpy(o->V0, o->V, sizeof(double) * o->n3);
  memcpy(o->W0, o->W, sizeof(double) * o->n3);
  memcpy(o->D0, o->D, sizeof(double) * o->n2);
  memcpy(o->Eta0, o->Eta, sizeof(double) * o->n2);
  memcpy(o->DU0, o->DU, sizeof(double) * o->n3);
  memcpy(o->DV0, o->DV, sizeof(double) * o->n3);
  memcpy(o->DW0, o->DW, sizeof(double) * o->n3);
}

/* update_wind (Iws = 1), src/nhwave.F:5333-5400 */
static void update_wind(orc *o) {
  /* constant expressions of default reals are folded in REAL(4) */
  float cmin = 1.e-3f * (0.61f + 0.063f * 6.0f);
  float cmax = 1.e-3f * (0.61f + 0.063f * 50.0f);
  double Cdsmin = (double)cmin, Cdsmax = (double)cmax;
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      double wu = o->wind_u, wv = o->wind_v;
      double wds = sqrt(wu * wu + wv * wv);
      double cds = F32(1.e-3) * (F32(0.61) + F32(0.063) * wds);
      cds = dmin(dmax(cds, Cdsmin), Cdsmax);
      o->Wsx[I2(i, j)] = F32(0.001293) * cds * wu * wds;
      o->Wsy[I2(i, j)] = F32(0.001293) * cds * wv * wds;
    }
}

/* source_terms (bed slope), src/source_terms.F:446-453 */
static void int_irr_source(orc *o);

/* source_terms, src/source_terms.F:1-456: the internal wavemaker's mass
 * source SourceC (INT_IRR, :194-305), then the bed-slope terms (:446-453) */
void orc_source_terms(orc *o) {
  if (o->c.wavemaker == 5) int_irr_source(o);
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      o->SourceX[I2(i, j)] = GRAV * 0.5 * (o->EtaxL[IX2(i + 1, j)] + o->EtaxR[IX2(i, j)]) *
                             (o->Hfx0[IX2(i + 1, j)] - o->Hfx0[IX2(i, j)]) / o->c.dx *
                             o->Mask[I2(i, j)];
      o->SourceY[I2(i, j)] = GRAV * 0.5 * (o->EtayL[IY2(i, j + 1)] + o->EtayR[IY2(i, j)]) *
                             (o->Hfy0[IY2(i, j + 1)] - o->Hfy0[IY2(i, j)]) / o->c.dy *
                             o->Mask[I2(i, j)];
    }
}

/* sigma_transform (DelxSc/DelySc), src/nhwave.F:6505-6541 */
static void sigma_transform(orc *o) {
  const double *sigc = o->sigc - 1;
  for (int k = 1; k <= o->Kloc; ++k)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int i = 1; i <= o->Mloc; ++i) {
        long c = I2(i, j);
        o->DelxSc[I3(i, j, k)] = (1.0 - sigc[k]) / o->D[c] * o->DelxH[c] - sigc[k] / o->D[c] * o->DelxEta[c];
        o->DelySc[I3(i, j, k)] = (1.0 - sigc[k]) / o->D[c] * o->DelyH[c] - sigc[k] / o->D[c] * o->DelyEta[c];
      }
  const double *sig = o->sig - 1;
  for (int k = 1; k <= o->Kloc1; ++k)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int i = 1; i <= o->Mloc; ++i) {
        long c = I2(i, j);
        o->DelxSl[IZ3(i, j, k)] = (1. - sig[k]) / o->D[c] * o->DelxH[c] - sig[k] / o->D[c] * o->DelxEta[c];
        o->DelySl[IZ3(i, j, k)] = (1. - sig[k]) / o->D[c] * o->DelyH[c] - sig[k] / o->D[c] * o->DelyEta[c];
      }
}

/* interpolate_velocity_to_faces (linear), src/nhwave.F:1332-1408 */
static void interpolate_velocity_to_faces(orc *o) {
  const double *dsig = o->dsig - 1;
  for (int k = 2; k <= o->Kloc; ++k)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int i = 1; i <= o->Mloc; ++i) {
        double f = dsig[k] / (dsig[k] + dsig[k - 1]);
        o->Uf[IZ3(i, j, k)] = (1.0 - f) * o->U[I3(i, j, k)] + f * o->U[I3(i, j, k - 1)];
        o->Vf[IZ3(i, j, k)] = (1.0 - f) * o->V[I3(i, j, k)] + f * o->V[I3(i, j, k - 1)];
        o->Wf[IZ3(i, j, k)] = (1.0 - f) * o->W[I3(i, j, k)] + f * o->W[I3(i, j, k - 1)];
      }
  for (int j = 1; j <= o->Nloc; ++j)
    for (int i = 1; i <= o->Mloc; ++i) {
      o->Uf[IZ3(i, j, 1)] = o->U[I3(i, j, 1)];
      o->Vf[IZ3(i, j, 1)] = o->V[I3(i, j, 1)];
      o->Wf[IZ3(i, j, 1)] = o->W[I3(i, j, 1)];
      o->Uf[IZ3(i, j, o->Kloc1)] = o->U[I3(i, j, o->Kloc)];
      o->Vf[IZ3(i, j, o->Kloc1)] = o->V[I3(i, j, o->Kloc)];
      o->Wf[IZ3(i, j, o->Kloc1)] = o->W[I3(i, j, o->Kloc)];
    }
}

void orc_interpolate_velocity_to_faces(orc *o) { interpolate_velocity_to_faces(o); }

/* ======================================================================== */
/* Poisson, src/pressure.F:63-245, 475-840                                  */
/* ======================================================================== */

#define ROW(i, j, k) ((long)((i)-o->Ibeg) + (long)o->c.mglob * (((j)-o->Jbeg) + (long)o->c.nglob * ((k)-o->Kbeg)))
#define CO(d, r) o->Coef[(long)((d)-1) * o->neqns + (r)]

void orc_generate_coef_rhs(orc *o) {
  const double dx = o->c.dx, dy = o->c.dy, dt = o->dt;
  const double *dsig = o->dsig - 1, *sig = o->sig - 1;
  double *DelxS = zd(o->nz3), *DelyS = zd(o->nz3), *DelzS = zd(o->nz3), *A1 = zd(o->nz3);
  for (int k = o->Kbeg; k <= o->Kend1; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        long c = I2(i, j), q = IZ3(i, j, k);
        DelxS[q] = (1. - sig[k]) / o->D[c] * o->DelxH[c] - sig[k] / o->D[c] * o->DelxEta[c];
        DelyS[q] = (1. - sig[k]) / o->D[c] * o->DelyH[c] - sig[k] / o->D[c] * o->DelyEta[c];
        DelzS[q] = 1. / o->D[c];
        A1[q] = DelxS[q] * DelxS[q] + DelyS[q] * DelyS[q] + DelzS[q] * DelzS[q];
      }
  /* source term, :517-535 */
  memset(o->Rhs, 0, sizeof(double) * o->neqns);
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        long q = IZ3(i, j, k);
        double hs = 0.5 * (dsig[k] + dsig[k - 1]);
        o->Rhs[ROW(i, j, k)] =
            -((o->Uf[IZ3(i + 1, j, k)] - o->Uf[IZ3(i - 1, j, k)]) / (2.0 * dx) +
              (o->U[I3(i, j, k)] - o->U[I3(i, j, k - 1)]) / hs * DelxS[q] +
              (o->Vf[IZ3(i, j + 1, k)] - o->Vf[IZ3(i, j - 1, k)]) / (2.0 * dy) +
              (o->V[I3(i, j, k)] - o->V[I3(i, j, k - 1)]) / hs * DelyS[q] +
              (o->W[I3(i, j, k)] - o->W[I3(i, j, k - 1)]) / hs * DelzS[q] - o->SourceC[I2(i, j)]) *
            RHO0 / dt;
      }
  memset(o->Coef, 0, sizeof(double) * 15 * o->neqns);
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        long r = ROW(i, j, k), q = IZ3(i, j, k);
        double s = dsig[k] + dsig[k - 1];
        CO(1, r) = (2. / (dx * dx) + 2. / (dy * dy) + A1[q] / (0.5 * s * dsig[k]) + A1[q] / (0.5 * s * dsig[k - 1]));
        CO(2, r) = -1. / (dx * dx);
        CO(3, r) = -1. / (dy * dy);
        CO(4, r) = (DelyS[IZ3(i, j - 1, k)] / (2. * dy * s) + DelyS[q] / (2. * dy * s));
        CO(5, r) = (DelxS[IZ3(i - 1, j, k)] / (2. * dx * s) + DelxS[q] / (2. * dx * s));
        CO(6, r) = -A1[q] / (0.5 * s * dsig[k]);
        CO(7, r) = -(DelxS[IZ3(i + 1, j, k)] / (2. * dx * s) + DelxS[q] / (2. * dx * s));
        CO(8, r) = -(DelyS[IZ3(i, j + 1, k)] / (2. * dy * s) + DelyS[q] / (2. * dy * s));
        CO(9, r) = -1. / (dx * dx);
        CO(10, r) = -1. / (dy * dy);
        CO(11, r) = (DelyS[IZ3(i, j + 1, k)] / (2. * dy * s) + DelyS[q] / (2. * dy * s));
        CO(12, r) = (DelxS[IZ3(i + 1, j, k)] / (2. * dx * s) + DelxS[q] / (2. * dx * s));
        CO(13, r) = -A1[q] / (0.5 * s * dsig[k - 1]);
        CO(14, r) = -(DelxS[IZ3(i - 1, j, k)] / (2. * dx * s) + DelxS[q] / (2. * dx * s));
        CO(15, r) = -(DelyS[IZ3(i, j - 1, k)] / (2. * dy * s) + DelyS[q] / (2. * dy * s));
      }
  /* boundary folds, :649-814 */
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      long r = ROW(o->Ibeg, j, k);
      CO(1, r) = CO(1, r) + CO(9, r); CO(6, r) = CO(6, r) + CO(5, r); CO(13, r) = CO(13, r) + CO(14, r);
      CO(9, r) = ZERO; CO(5, r) = ZERO; CO(14, r) = ZERO;
    }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      long r = ROW(o->Iend, j, k);
      CO(1, r) = CO(1, r) + CO(2, r); CO(6, r) = CO(6, r) + CO(7, r); CO(13, r) = CO(13, r) + CO(12, r);
      CO(2, r) = ZERO; CO(7, r) = ZERO; CO(12, r) = ZERO;
    }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      long r = ROW(i, o->Jbeg, k);
      CO(1, r) = CO(1, r) + CO(10, r); CO(6, r) = CO(6, r) + CO(4, r); CO(13, r) = CO(13, r) + CO(15, r);
      CO(10, r) = ZERO; CO(4, r) = ZERO; CO(15, r) = ZERO;
    }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      long r = ROW(i, o->Jend, k);
      CO(1, r) = CO(1, r) + CO(3, r); CO(6, r) = CO(6, r) + CO(8, r); CO(13, r) = CO(13, r) + CO(11, r);
      CO(3, r) = ZERO; CO(8, r) = ZERO; CO(11, r) = ZERO;
    }
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      long r = ROW(i, j, o->Kbeg);
      CO(6, r) = CO(6, r) + CO(13, r); CO(7, r) = CO(7, r) + CO(12, r); CO(8, r) = CO(8, r) + CO(11, r);
      CO(5, r) = CO(5, r) + CO(14, r); CO(4, r) = CO(4, r) + CO(15, r);
      CO(13, r) = ZERO; CO(12, r) = ZERO; CO(11, r) = ZERO; CO(14, r) = ZERO; CO(15, r) = ZERO;
    }
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      long r = ROW(i, j, o->Kend);
      CO(4, r) = ZERO; CO(5, r) = ZERO; CO(6, r) = ZERO; CO(7, r) = ZERO; CO(8, r) = ZERO;
    }
  free(DelxS); free(DelyS); free(DelzS); free(A1);
}

/* 15-point stencil offsets (di,dj,dk), JCoef order, src/pressure.F:817-831 */
static const int OFF[15][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},   {0, -1, 1}, {-1, 0, 1},
                               {0, 0, 1},  {1, 0, 1},  {0, 1, 1},   {-1, 0, 0}, {0, -1, 0},
                               {0, 1, -1}, {1, 0, -1}, {0, 0, -1},  {-1, 0, -1}, {0, -1, -1}};

/* Banded LU with partial pivoting (the classical unblocked band algorithm),
 * on the unknowns reordered k-fastest so the half bandwidth is
 * K*min(M,N)+1.  Solves coef*x = rhs; coef is diag-major as assembled. */
int orc_solve_banded(orc *o, const double *coef, const double *rhs, double *x) {
  const int M = o->c.mglob, N = o->c.nglob, K = o->c.kglob;
  const long n = o->neqns;
  int islow = M >= N; /* slowest index: i if M >= N else j */
  const int S = islow ? N : M;
  const int kl = K * S + 1, ku = kl, kv = kl + ku, ldab = 2 * kl + ku + 1;
  double *ab = zd((long)ldab * n);
  long *piv = (long *)malloc(sizeof(long) * n);
  double *b = zd(n);
  if (!ab || !piv || !b) { free(ab); free(piv); free(b); snprintf(g_err, sizeof g_err, "banded: oom"); return -1; }
#define BQ(i0, j0, k0) (islow ? ((long)(k0) + (long)K * ((j0) + (long)N * (i0))) \
                              : ((long)(k0) + (long)K * ((i0) + (long)M * (j0))))
  for (int k0 = 0; k0 < K; ++k0)
    for (int j0 = 0; j0 < N; ++j0)
      for (int i0 = 0; i0 < M; ++i0) {
        long r = i0 + (long)M * (j0 + (long)N * k0);
        long qr = BQ(i0, j0, k0);
        b[qr] = rhs[r];
        for (int d = 0; d < 15; ++d) {
          double a = coef[(long)d * n + r];
          if (a == 0.0) continue;
          int i1 = i0 + OFF[d][0], j1 = j0 + OFF[d][1], k1 = k0 + OFF[d][2];
          if (i1 < 0 || i1 >= M || j1 < 0 || j1 >= N || k1 < 0 || k1 >= K) {
            snprintf(g_err, sizeof g_err, "banded: nonzero coefficient %d leaves the domain at row %ld", d + 1, r);
            free(ab); free(piv); free(b);
            return -2;
          }
          long qc = BQ(i1, j1, k1);
          ab[(long)(kv + qr - qc) + (long)ldab * qc] += a;
        }
      }
#undef BQ
  /* factor */
  long ju = 0;
  for (long j = 0; j < n; ++j) {
    long km = kl < n - 1 - j ? kl : n - 1 - j;
    long jp = 0;
    double amax = fabs(ab[kv + ldab * j]);
    for (long i = 1; i <= km; ++i) {
      double v = fabs(ab[kv + i + ldab * j]);
      if (v > amax) { amax = v; jp = i; }
    }
    piv[j] = j + jp;
    if (ab[kv + jp + ldab * j] == 0.0) {
      snprintf(g_err, sizeof g_err, "banded: singular at %ld", j);
      free(ab); free(piv); free(b);
      return -3;
    }
    long jue = j + ku + jp < n - 1 ? j + ku + jp : n - 1;
    if (jue > ju) ju = jue;
    if (jp != 0)
      for (long c = j; c <= ju; ++c) {
        double *p1 = &ab[kv + jp + (j - c) + ldab * c], *p2 = &ab[kv + (j - c) + ldab * c];
        double t = *p1; *p1 = *p2; *p2 = t;
      }
    if (km > 0) {
      double inv = 1.0 / ab[kv + ldab * j];
      for (long i = 1; i <= km; ++i) ab[kv + i + ldab * j] *= inv;
      for (long c = j + 1; c <= ju; ++c) {
        double u = ab[kv + (j - c) + ldab * c];
        if (u == 0.0) continue;
        for (long i = 1; i <= km; ++i) ab[kv + i + (j - c) + ldab * c] -= ab[kv + i + ldab * j] * u;
      }
    }
  }
  /* solve L */
  for (long j = 0; j < n; ++j) {
    long lm = kl < n - 1 - j ? kl : n - 1 - j;
    long p = piv[j];
    if (p != j) { double t = b[p]; b[p] = b[j]; b[j] = t; }
    for (long i = 1; i <= lm; ++i) b[j + i] -= ab[kv + i + ldab * j] * b[j];
  }
  /* solve U */
  for (long j = n - 1; j >= 0; --j) {
    b[j] /= ab[kv + ldab * j];
    long lim = kv < j ? kv : j;
    for (long i = 1; i <= lim; ++i) b[j - i] -= ab[kv - i + ldab * j] * b[j];
  }
  for (int k0 = 0; k0 < K; ++k0)
    for (int j0 = 0; j0 < N; ++j0)
      for (int i0 = 0; i0 < M; ++i0) {
        long r = i0 + (long)M * (j0 + (long)N * k0);
        long qr = islow ? ((long)k0 + (long)K * (j0 + (long)N * i0)) : ((long)k0 + (long)K * (i0 + (long)M * j0));
        x[r] = b[qr];
      }
  free(ab); free(piv); free(b);
  return 0;
}

/* poisson_solver, src/pressure.F:63-245 (serial) */
int orc_poisson_solver(orc *o) {
  orc_generate_coef_rhs(o);
  double *phi = zd(o->neqns);
  int rc = orc_solve_banded(o, o->Coef, o->Rhs, phi);
  if (rc != 0) { free(phi); return rc; }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) o->P[IZ3(i, j, k)] = phi[ROW(i, j, k)];
  free(phi);
  /* ghost cells, :185-239 */
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = 1; i <= 2; ++i) o->P[IZ3(o->Ibeg - i, j, k)] = o->P[IZ3(o->Ibeg + i - 1, j, k)];
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = 1; i <= 2; ++i) o->P[IZ3(o->Iend + i, j, k)] = o->P[IZ3(o->Iend - i + 1, j, k)];
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int i = o->Ibeg; i <= o->Iend; ++i)
      for (int j = 1; j <= 2; ++j) o->P[IZ3(i, o->Jbeg - j, k)] = o->P[IZ3(i, o->Jbeg + j - 1, k)];
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int i = o->Ibeg; i <= o->Iend; ++i)
      for (int j = 1; j <= 2; ++j) o->P[IZ3(i, o->Jend + j, k)] = o->P[IZ3(i, o->Jend - j + 1, k)];
  return 0;
}

/* projection_corrector, src/nhwave.F:487-536 */
static void projection_corrector(orc *o) {
  const double dx = o->c.dx, dy = o->c.dy, dt = o->dt;
  const double *dsig = o->dsig - 1;
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        if (o->Mask[I2(i, j)] == 0) continue;
        const double *P = o->P;
        double dxp = 0.5 * ((P[IZ3(i + 1, j, k)] - P[IZ3(i - 1, j, k)]) / (2. * dx) +
                            (P[IZ3(i + 1, j, k + 1)] - P[IZ3(i - 1, j, k + 1)]) / (2. * dx));
        double dyp = 0.5 * ((P[IZ3(i, j + 1, k)] - P[IZ3(i, j - 1, k)]) / (2. * dy) +
                            (P[IZ3(i, j + 1, k + 1)] - P[IZ3(i, j - 1, k + 1)]) / (2. * dy));
        double dzp = (P[IZ3(i, j, k + 1)] - P[IZ3(i, j, k)]) / dsig[k];
        long c = I3(i, j, k);
        double d = o->D[I2(i, j)];
        o->DU[c] = o->DU[c] - d * dt / RHO0 * (dxp + dzp * o->DelxSc[c]);
        o->DV[c] = o->DV[c] - d * dt / RHO0 * (dyp + dzp * o->DelySc[c]);
        o->DW[c] = o->DW[c] - dt / RHO0 * dzp;
      }
}

/* get_Omega, src/nhwave.F:1411-1648 */
static void get_omega(orc *o) {
  const double MinDep = o->c.min_dep;
  delx_2d(o, o->Eta, o->DelxEta);
  dely_2d(o, o->Eta, o->DelyEta);
  delx_3d(o, o->U, o->DelxU, o->Kloc);
  delx_3d(o, o->V, o->DelxV, o->Kloc);
  dely_3d(o, o->U, o->DelyU, o->Kloc);
  dely_3d(o, o->V, o->DelyV, o->Kloc);
  delx_3d(o, o->DU, o->DelxDU, o->Kloc);
  delx_3d(o, o->DV, o->DelxDV, o->Kloc);
  dely_3d(o, o->DU, o->DelyDU, o->Kloc);
  dely_3d(o, o->DV, o->DelyDV, o->Kloc);
  construct_2d_x(o, o->Eta, o->DelxEta, o->EtaxL, o->EtaxR);
  construct_2d_y(o, o->Eta, o->DelyEta, o->EtayL, o->EtayR);
  construct_3d_x(o, o->U, o->DelxU, o->UxL, o->UxR);
  construct_3d_x(o, o->V, o->DelxV, o->VxL, o->VxR);
  construct_3d_y(o, o->U, o->DelyU, o->UyL, o->UyR);
  construct_3d_y(o, o->V, o->DelyV, o->VyL, o->VyR);
  construct_3d_x(o, o->DU, o->DelxDU, o->DUxL, o->DUxR);
  construct_3d_x(o, o->DV, o->DelxDV, o->DVxL, o->DVxR);
  construct_3d_y(o, o->DU, o->DelyDU, o->DUyL, o->DUyR);
  construct_3d_y(o, o->DV, o->DelyDV, o->DVyL, o->DVyR);
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend1; ++i) {
      o->DxL[IX2(i, j)] = o->EtaxL[IX2(i, j)] + o->Hfx[IX2(i, j)];
      o->DxR[IX2(i, j)] = o->EtaxR[IX2(i, j)] + o->Hfx[IX2(i, j)];
    }
  for (long q = 0; q < o->nx2; ++q) { o->DxL[q] = dmax(o->DxL[q], MinDep); o->DxR[q] = dmax(o->DxR[q], MinDep); }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        int m = o->Mask[I2(i, j)], mp = o->Mask[I2(i + 1, j)], mm = o->Mask[I2(i - 1, j)];
        if (m == 1 && mp == 0) {
          o->EtaxL[IX2(i + 1, j)] = o->Eta[I2(i, j)];
          o->DxL[IX2(i + 1, j)] = o->D[I2(i, j)];
          o->DUxL[IX3(i + 1, j, k)] = ZERO;
        } else if (m == 1 && mm == 0) {
          o->EtaxR[IX2(i, j)] = o->Eta[I2(i, j)];
          o->DxR[IX2(i, j)] = o->D[I2(i, j)];
          o->DUxR[IX3(i, j, k)] = ZERO;
        } else if (m == 0) {
          o->EtaxL[IX2(i + 1, j)] = o->Eta[I2(i, j)];
          o->DxL[IX2(i + 1, j)] = o->D[I2(i, j)];
          o->DUxL[IX3(i + 1, j, k)] = ZERO;
          o->DVxL[IX3(i + 1, j, k)] = ZERO;
          o->EtaxR[IX2(i, j)] = o->Eta[I2(i, j)];
          o->DxR[IX2(i, j)] = o->D[I2(i, j)];
          o->DUxR[IX3(i, j, k)] = ZERO;
        }
      }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend1; ++i) {
        o->ExL[IX3(i, j, k)] = o->DUxL[IX3(i, j, k)];
        o->ExR[IX3(i, j, k)] = o->DUxR[IX3(i, j, k)];
      }
  for (int j = o->Jbeg; j <= o->Jend1; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      o->DyL[IY2(i, j)] = o->EtayL[IY2(i, j)] + o->Hfy[IY2(i, j)];
      o->DyR[IY2(i, j)] = o->EtayR[IY2(i, j)] + o->Hfy[IY2(i, j)];
    }
  for (long q = 0; q < o->ny2; ++q) { o->DyL[q] = dmax(o->DyL[q], MinDep); o->DyR[q] = dmax(o->DyR[q], MinDep); }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        int m = o->Mask[I2(i, j)], mp = o->Mask[I2(i, j + 1)], mm = o->Mask[I2(i, j - 1)];
        if (m == 1 && mp == 0) {
          o->EtayL[IY2(i, j + 1)] = o->Eta[I2(i, j)];
          o->DyL[IY2(i, j + 1)] = o->D[I2(i, j)];
          o->DVyL[IY3(i, j + 1, k)] = ZERO;
        } else if (m == 1 && mm == 0) {
          o->EtayR[IY2(i, j)] = o->Eta[I2(i, j)];
          o->DyR[IY2(i, j)] = o->D[I2(i, j)];
          o->DVyR[IY3(i, j, k)] = ZERO;
        } else if (m == 0) {
          o->EtayL[IY2(i, j + 1)] = o->Eta[I2(i, j)];
          o->DyL[IY2(i, j + 1)] = o->D[I2(i, j)];
          o->DVyL[IY3(i, j + 1, k)] = ZERO;
          o->EtayR[IY2(i, j)] = o->Eta[I2(i, j)];
          o->DyR[IY2(i, j)] = o->D[I2(i, j)];
          o->DVyR[IY3(i, j, k)] = ZERO;
        }
      }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend1; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        o->EyL[IY3(i, j, k)] = o->DVyL[IY3(i, j, k)];
        o->EyR[IY3(i, j, k)] = o->DVyR[IY3(i, j, k)];
      }
  wave_speed(o);
  hll_x(o, o->ExL, o->ExR, NULL, NULL, o->DxL, o->DxR, o->Ex);
  hll_y(o, o->EyL, o->EyR, NULL, NULL, o->DyL, o->DyR, o->Ey);
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int k = o->Kbeg; k <= o->Kend; ++k) {
      if (is_wall(o->c.bc_x0)) o->Ex[IX3(o->Ibeg, j, k)] = ZERO;
      else if (o->c.bc_x0 == 3) o->Ex[IX3(o->Ibeg, j, k)] = o->Din_X0[j - 1] * o->Uin_X0[IB(j, k)]; /* :1564-1565 */
      if (is_wall(o->c.bc_xn)) o->Ex[IX3(o->Iend1, j, k)] = ZERO;
      else if (o->c.bc_xn == 3) o->Ex[IX3(o->Iend1, j, k)] = o->Din_Xn[j - 1] * o->Uin_Xn[IB(j, k)]; /* :1580-1581 */
    }
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int k = o->Kbeg; k <= o->Kend; ++k) {
      if (is_wall(o->c.bc_y0)) o->Ey[IY3(i, o->Jbeg, k)] = ZERO;
      if (is_wall(o->c.bc_yn)) o->Ey[IY3(i, o->Jend1, k)] = ZERO;
    }
  memset(o->Omega, 0, sizeof(double) * o->nz3);
  const double *dsig = o->dsig - 1;
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int k = o->Kbeg + 1; k <= o->Kend1; ++k) {
        o->Omega[IZ3(i, j, k)] =
            o->Omega[IZ3(i, j, k - 1)] -
            dsig[k - 1] * (o->R1[I2(i, j)] + (o->Ex[IX3(i + 1, j, k - 1)] - o->Ex[IX3(i, j, k - 1)]) / o->c.dx +
                           (o->Ey[IY3(i, j + 1, k - 1)] - o->Ey[IY3(i, j, k - 1)]) / o->c.dy);
        if (o->Mask[I2(i, j)] == 0) o->Omega[IZ3(i, j, k)] = ZERO;
      }
}

/* eval_duvw(ISTEP), src/nhwave.F:539-1300 (inviscid, barotropic) */
/* ======================================================================== */
/* Viscous branch with constant viscosities (IVTURB = IHTURB = 1)           */
/* ======================================================================== */
/* kepsilon_3D, src/nhwave.F:10030-11114, as the reference build executes it
 * (Kglob >= 3, lateral boundaries 1/2/3, no sediment, vegetation or porous
 * media).
 *
 * The subroutine reads its local variables alpha_c, beta_c and gamma_c
 * (declared :10058-10059, never assigned in it) in the interior-row
 * coefficients of both tridiagonal systems (Acoef, Bcoef, Ccoef for
 * Kbeg < k < Kend, :10468-10541 and :10669-10742).  The reference build
 * (amdflang -O2, oracle/build_ref.sh) folds those reads: its LLVM IR stores
 * the constants +NaN (Bcoef) and -NaN (Acoef, Ccoef) for every interior
 * row.  With Kglob >= 3 every column has such a row, so trig (:13182-13224)
 * returns NaN for all of it; the blend ALPHA*DEps0 + BETA*DEps stays NaN
 * (BETA > 0); and the clamp max(DEps, D*Eps_min), compiled as
 * (a > b ? a : b), yields D*Eps_min, likewise D*Tke_min for DTke, in every
 * column with D >= Dmin and Mask = 1.  The production terms (:10102-10340),
 * the HLPA advection with horizontal diffusion (adv_scalar_hlpa, :7604-7806)
 * and the vertical diffusion coefficients enter only those systems, so they
 * leave no trace in the state.  What does, restated here in the reference's
 * order: the NaN of the solves with its blend and clamp, the bottom wall
 * function (:10766-10836), the free-surface gradient condition of the ghost
 * levels (:10838-10862) built from the old Tke/Eps (:10400-10419), the
 * lateral ghost copies (:10869-10996), the new Tke, Eps, CmuVt
 * (:10999-11030) and the wavemaker-region reset (:11055-11071; the scenarios
 * here leave its box at the origin, where no cell centre lies). */
/* kepsilon_3D's wall function at a wall friction boundary (code 5,
 * :10874-10892 and its east, south, north twins): from the speed along the
 * wall |u_t| (below 1.e-6, a REAL(4) literal: Tke_min, Eps_min) the friction
 * velocity sqrt(Cd0)|u_t| (Ibot = 1) or |u_t|/(1/Kappa log(30 h/2/Zob)), then
 * Tke = u*^2/sqrt(cmiu), Eps = u*^3/(Kappa h/2) at the cell c (column c2) */
static void ke_wall(orc *o, double h, double a, double b, long c, long c2, double cmiu) {
  const double umag = sqrt(a * a + b * b);
  double tkeb = KE_TKE_MIN, epsb = KE_EPS_MIN;
  if (!(umag < F32(1.e-6))) {
    const double fricu = o->c.ibot == 1 ? sqrt(o->c.cd0) * umag : umag / (1. / KAPPA * log(30. * h / 2.0 / o->c.zob));
    tkeb = fricu * fricu / sqrt(cmiu);
    epsb = fricu * fricu * fricu / (KAPPA * h / 2.0);
  }
  o->DTke[c] = o->D[c2] * tkeb;
  o->DEps[c] = o->D[c2] * epsb;
}

static void kepsilon_3d(orc *o) {
  const double *dsig = o->dsig - 1, *sigc = o->sigc - 1;
  const int Kb = o->Kbeg, Ke = o->Kend;
  const double Dmin = 2.0 * o->c.min_dep;               /* :10087 */
  const double cmiu = o->c.rng ? F32(0.085) : F32(0.09); /* :10091-10110 */
  const double tke_min = KE_TKE_MIN, eps_min = KE_EPS_MIN, cmut_min = ke_cmut_min(o);
  double *Tke_old = zd(o->n3), *Eps_old = zd(o->n3);
  memcpy(Tke_old, o->Tke, sizeof(double) * o->n3);
  memcpy(Eps_old, o->Eps, sizeof(double) * o->n3);
  /* reconstructed values at the sigma levels and their x/y slopes (:10400-10419) */
  double *DelzTke = zd(o->n3), *DelzEps = zd(o->n3);
  double *TkezL = zd(o->nz3), *TkezR = zd(o->nz3), *EpszL = zd(o->nz3), *EpszR = zd(o->nz3);
  double *DxTkezL = zd(o->nz3), *DxEpszL = zd(o->nz3), *DyTkezL = zd(o->nz3), *DyEpszL = zd(o->nz3);
  delz_3d(o, Tke_old, DelzTke);
  delz_3d(o, Eps_old, DelzEps);
  construct_3d_z(o, Tke_old, DelzTke, TkezL, TkezR);
  construct_3d_z(o, Eps_old, DelzEps, EpszL, EpszR);
  for (long q = 0; q < o->nz3; ++q) {
    TkezL[q] = dmax(TkezL[q], tke_min);
    TkezR[q] = dmax(TkezR[q], tke_min);
    EpszL[q] = dmax(EpszL[q], eps_min);
    EpszR[q] = dmax(EpszR[q], eps_min);
  }
  delx_3d(o, TkezL, DxTkezL, o->Kloc1);
  delx_3d(o, EpszL, DxEpszL, o->Kloc1);
  dely_3d(o, TkezL, DyTkezL, o->Kloc1);
  dely_3d(o, EpszL, DyEpszL, o->Kloc1);
  const double nan = NAN;  /* what trig returns for a column with a NaN row */
  /* epsilon: solve (:10421-10548), blend and clamp (:10550-10558); then k
   * likewise (:10562-10759) */
  for (int var = 0; var < 2; ++var) {
    double *DQ = var == 0 ? o->DEps : o->DTke;
    const double qmin = var == 0 ? eps_min : tke_min;
    for (int i = o->Ibeg; i <= o->Iend; ++i)
      for (int j = o->Jbeg; j <= o->Jend; ++j) {
        const long c2 = I2(i, j);
        if (o->D[c2] < Dmin && o->Mask[c2] == 0) continue;
        for (int k = Kb; k <= Ke; ++k) DQ[I3(i, j, k)] = nan;
      }
    for (int i = o->Ibeg; i <= o->Iend; ++i)
      for (int j = o->Jbeg; j <= o->Jend; ++j)
        for (int k = Kb; k <= Ke; ++k) {
          const long c2 = I2(i, j), c = I3(i, j, k);
          if (o->D[c2] >= Dmin && o->Mask[c2] == 1) {
            const double fl = o->D[c2] * qmin;
            DQ[c] = DQ[c] > fl ? DQ[c] : fl;
          }
        }
  }
  /* bottom wall function, :10763-10836 */
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      const long c2 = I2(i, j), b0 = I3(i, j, Kb);
      const double dxh = o->DelxH[c2], dyh = o->DelyH[c2], Dd = o->D[c2];
      const double Abot = sqrt(1.0 + dxh * dxh + dyh * dyh);
      const double Bbot = sqrt(1.0 + dxh * dxh);
      const double dup = (o->DU[b0] - dxh * o->DW[b0]) / Bbot;
      const double dvp = (-(dxh * dyh * o->DU[b0]) + (1.0 + dxh * dxh) * o->DV[b0] - dyh * o->DW[b0]) / Abot / Bbot;
      const double dumag = sqrt(dup * dup + dvp * dvp);
      if (Dd < Dmin || o->Mask[c2] == 0) continue;
      double tkeb, epsb;
      if (dumag / Dd < F32(1.e-6)) {
        tkeb = tke_min;
        epsb = eps_min;
      } else {
        const double Dz1 = 0.5 * Dd * dsig[Kb] / Abot;
        const double Z0 = o->c.zob / 30.0;
        double cdrag;
        if (o->c.ibot == 1) cdrag = o->c.cd0;
        else {
          const double t = KAPPA / log(Dz1 / Z0);
          cdrag = t * t;
        }
        cdrag = dmin(cdrag, ((o->c.visc + 0.0) + o->CmuVt[b0]) / Dz1 / (dumag + SMALL) * Dd);
        const double ustar = sqrt(cdrag) * dumag / Dd;
        tkeb = dmax(ustar * ustar / sqrt(cmiu), tke_min);
        epsb = dmax(ustar * ustar * ustar / (KAPPA * Dz1), eps_min);
      }
      o->DTke[b0] = Dd * tkeb;
      o->DEps[b0] = Dd * epsb;
      for (int k = 1; k <= 2; ++k) {
        o->DTke[I3(i, j, Kb - k)] = o->DTke[b0];
        o->DEps[I3(i, j, Kb - k)] = o->DEps[b0];
      }
    }
  /* free surface, :10838-10862 */
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      const long c2 = I2(i, j), zt = IZ3(i, j, Ke + 1);
      if (o->Mask[c2] == 0 || o->D[c2] < Dmin) continue;
      const double dxe = o->DelxEta[c2], dye = o->DelyEta[c2], Dd = o->D[c2];
      for (int k = 1; k <= 2; ++k) {
        const double Atop = sqrt(1.0 + dxe * dxe + dye * dye);
        const double ds = sigc[Ke + k] - sigc[Ke - k + 1];
        o->DTke[I3(i, j, Ke + k)] =
            o->DTke[I3(i, j, Ke - k + 1)] + ds * (Dd * Dd) / (Atop * Atop) * (dxe * DxTkezL[zt] + dye * DyTkezL[zt]);
        o->DEps[I3(i, j, Ke + k)] =
            o->DEps[I3(i, j, Ke - k + 1)] + ds * (Dd * Dd) / (Atop * Atop) * (dxe * DxEpszL[zt] + dye * DyEpszL[zt]);
      }
    }
  /* lateral ghosts, west, east, south, north (:10869-10996), each wall
   * friction boundary (code 5) first setting its first interior cell by the
   * wall function (ke_wall) at every level; the south and north passes run
   * over every column, ghost columns included */
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int k = 1; k <= o->Kloc; ++k) {
      if (o->c.bc_x0 == 5) {
        const long b = I3(o->Ibeg, j, k);
        ke_wall(o, o->c.dx, o->V[b], o->W[b], b, I2(o->Ibeg, j), cmiu);
      }
      for (int i = 1; i <= 2; ++i) {
        o->DTke[I3(o->Ibeg - i, j, k)] = o->DTke[I3(o->Ibeg + i - 1, j, k)];
        o->DEps[I3(o->Ibeg - i, j, k)] = o->DEps[I3(o->Ibeg + i - 1, j, k)];
      }
    }
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int k = 1; k <= o->Kloc; ++k) {
      if (o->c.bc_xn == 5) {
        const long b = I3(o->Iend, j, k);
        ke_wall(o, o->c.dx, o->V[b], o->W[b], b, I2(o->Iend, j), cmiu);
      }
      for (int i = 1; i <= 2; ++i) {
        o->DTke[I3(o->Iend + i, j, k)] = o->DTke[I3(o->Iend - i + 1, j, k)];
        o->DEps[I3(o->Iend + i, j, k)] = o->DEps[I3(o->Iend - i + 1, j, k)];
      }
    }
  for (int i = 1; i <= o->Mloc; ++i)
    for (int k = 1; k <= o->Kloc; ++k) {
      if (o->c.bc_y0 == 5) {
        const long b = I3(i, o->Jbeg, k);
        ke_wall(o, o->c.dy, o->U[b], o->W[b], b, I2(i, o->Jbeg), cmiu);
      }
      for (int j = 1; j <= 2; ++j) {
        o->DTke[I3(i, o->Jbeg - j, k)] = o->DTke[I3(i, o->Jbeg + j - 1, k)];
        o->DEps[I3(i, o->Jbeg - j, k)] = o->DEps[I3(i, o->Jbeg + j - 1, k)];
      }
    }
  for (int i = 1; i <= o->Mloc; ++i)
    for (int k = 1; k <= o->Kloc; ++k) {
      if (o->c.bc_yn == 5) {
        const long b = I3(i, o->Jend, k);
        ke_wall(o, o->c.dy, o->U[b], o->W[b], b, I2(i, o->Jend), cmiu);
      }
      for (int j = 1; j <= 2; ++j) {
        o->DTke[I3(i, o->Jend + j, k)] = o->DTke[I3(i, o->Jend - j + 1, k)];
        o->DEps[I3(i, o->Jend + j, k)] = o->DEps[I3(i, o->Jend - j + 1, k)];
      }
    }
  /* Tke, Eps and the eddy viscosity, :10999-11030 (1.e-6 is REAL(4)) */
  for (int i = 1; i <= o->Mloc; ++i)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int k = 1; k <= o->Kloc; ++k) {
        const long c2 = I2(i, j), c = I3(i, j, k);
        if (o->D[c2] >= Dmin && o->Mask[c2] == 1) {
          o->Tke[c] = o->DTke[c] / o->D[c2];
          o->Eps[c] = o->DEps[c] / o->D[c2];
          if (o->Tke[c] < F32(1.e-6) || o->Eps[c] < F32(1.e-6)) o->CmuVt[c] = cmut_min;
          else o->CmuVt[c] = cmiu * (o->Tke[c] * o->Tke[c]) / o->Eps[c];
        } else {
          o->Tke[c] = tke_min;
          o->Eps[c] = eps_min;
          o->DTke[c] = o->D[c2] * tke_min;
          o->DEps[c] = o->D[c2] * eps_min;
          o->CmuVt[c] = cmut_min;
        }
      }
  for (int i = 1; i <= o->Mloc; ++i)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int k = 1; k <= 2; ++k) {
        o->CmuVt[I3(i, j, Kb - k)] = o->CmuVt[I3(i, j, Kb)];
        o->CmuVt[I3(i, j, Ke + k)] = o->CmuVt[I3(i, j, Ke)];
      }
  /* no turbulence in the wavemaker region (:11055-11071): the internal
   * wavemaker's source box; the west wavemakers never read the box keys
   * (src/initialize.F:629-634), which stay 0: a box at the origin */
  const int internal = o->c.wavemaker == 5;
  const double bw = internal ? o->c.xs_w : 0.0, be = internal ? o->c.xs_e : 0.0;
  const double bs = internal ? o->c.ys_s : 0.0, bn = internal ? o->c.ys_n : 0.0;
  if (o->c.wavemaker != 0)
    for (int k = 1; k <= o->Kloc; ++k)
      for (int j = 1; j <= o->Nloc; ++j)
        for (int i = 1; i <= o->Mloc; ++i)
          if (o->xc[i - 1] >= bw && o->xc[i - 1] <= be && o->yc[j - 1] >= bs && o->yc[j - 1] <= bn) {
            const long c = I3(i, j, k), c2 = I2(i, j);
            o->Tke[c] = tke_min;
            o->Eps[c] = eps_min;
            o->DTke[c] = o->D[c2] * tke_min;
            o->DEps[c] = o->D[c2] * eps_min;
            o->CmuVt[c] = cmut_min;
          }
  double *tmp[] = {Tke_old, Eps_old, DelzTke, DelzEps, TkezL, TkezR, EpszL, EpszR, DxTkezL, DxEpszL, DyTkezL, DyEpszL};
  for (size_t q = 0; q < sizeof tmp / sizeof tmp[0]; ++q) free(tmp[q]);
}

/* eval_turb, src/nhwave.F:6544-6782: Cmu = Visc in every cell
 * (src/initialize.F:2666, :6559-6576), CmuR = 0.  IVturb = IHturb = 1:
 * CmuVt = Cvs (:6650-6652), CmuHt = Chs (:6683-6685), in every cell.
 * IVturb = IHturb = 10: kepsilon_3D (:6667-6674), then CmuHt = CmuVt
 * (:6773-6778).  Before the first call CmuHt and CmuVt hold the turbulence
 * seed Cmut_min (src/initialize.F:3172-3196). */
void orc_eval_turb(orc *o) {
  if (!o->c.viscous) return;
  if (o->c.ivturb == 10) {
    kepsilon_3d(o);
    memcpy(o->CmuHt, o->CmuVt, sizeof(double) * o->n3);
    return;
  }
  for (long q = 0; q < o->n3; ++q) {
    o->CmuVt[q] = o->c.cvs;
    o->CmuHt[q] = o->c.chs;
  }
}

/* diffusion, src/nhwave.F:6785-7141.  Cmu = Visc in every cell and CmuR =
 * 0, so they enter as scalars inside the reference's own expressions; CmuHt
 * is read per cell.  Diffxz and Diffyz feed no momentum equation (:643-683)
 * and are not formed.  DelxSc/DelySc/DelxSl/DelySl are those of the last
 * sigma_transform. */
static void diffusion(orc *o) {
  const double dx = o->c.dx, dy = o->c.dy, S = o->c.schmidt;
  const double cm = o->c.visc;
  const double *dsig = o->dsig - 1;
  double *out[6] = {o->Diffxx, o->Diffxy, o->Diffyx, o->Diffyy, o->Diffzx, o->Diffzy};
  for (int q = 0; q < 6; ++q) memset(out[q], 0, sizeof(double) * o->n3);
#define DU_(i, j, k) o->DU[I3(i, j, k)]
#define DV_(i, j, k) o->DV[I3(i, j, k)]
#define DW_(i, j, k) o->DW[I3(i, j, k)]
#define SX(i, j, k) o->DelxSc[I3(i, j, k)]
#define SY(i, j, k) o->DelySc[I3(i, j, k)]
#define LX(i, j, k) o->DelxSl[IZ3(i, j, k)]
#define LY(i, j, k) o->DelySl[IZ3(i, j, k)]
  for (int k = o->Kbeg; k <= o->Kend; ++k) {
    /* Lagrange weights at k+-1/2 and the non-uniform sigma derivative
     * weights, :6813-6827 */
    const double L1top = dsig[k + 1] / (dsig[k] + dsig[k + 1]);
    const double L2top = dsig[k] / (dsig[k] + dsig[k + 1]);
    const double L1bot = dsig[k] / (dsig[k] + dsig[k - 1]);
    const double L2bot = dsig[k - 1] / (dsig[k] + dsig[k - 1]);
    const double dsigck = (dsig[k] + dsig[k + 1]) / 2.0;
    const double dsigck1 = (dsig[k - 1] + dsig[k]) / 2.0;
    const double alpha_c = -dsigck / (dsigck + dsigck1) / dsigck1;
    const double beta_c = (dsigck - dsigck1) / (dsigck * dsigck1);
    const double gamma_c = dsigck1 / (dsigck + dsigck1) / dsigck;
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        if (o->Mask[I2(i, j)] == 0) continue;
        const double Dc = o->D[I2(i, j)];
#define CH(i, j, k) o->CmuHt[I3(i, j, k)]
        const double nuH_top = (L1top * cm + L2top * cm) + (L1top * CH(i, j, k) + L2top * CH(i, j, k + 1)) / S;
        const double nuH_bot = (L1bot * cm + L2bot * cm) + (L1bot * CH(i, j, k - 1) + L2bot * CH(i, j, k)) / S;
        /* face averages at i+-1/2, j+-1/2 */
        const double hfe = 0.5 * (cm + cm) + 0.5 * (CH(i + 1, j, k) + CH(i, j, k)) / S;
        const double hfw = 0.5 * (cm + cm) + 0.5 * (CH(i, j, k) + CH(i - 1, j, k)) / S;
        const double hfn = 0.5 * (cm + cm) + 0.5 * (CH(i, j + 1, k) + CH(i, j, k)) / S;
        const double hfs = 0.5 * (cm + cm) + 0.5 * (CH(i, j, k) + CH(i, j - 1, k)) / S;
        /* cell values Cmu + CmuHt/Schmidt at the neighbours that enter */
        const double hE = cm + CH(i + 1, j, k) / S, hW = cm + CH(i - 1, j, k) / S;
        const double hN = cm + CH(i, j + 1, k) / S, hS = cm + CH(i, j - 1, k) / S;
        const double hB = cm + CH(i, j, k - 1) / S, hC = cm + CH(i, j, k) / S, hT = cm + CH(i, j, k + 1) / S;
#undef CH
        const double lxd = nuH_top * LX(i, j, k + 1) - nuH_bot * LX(i, j, k);
        const double lxs = nuH_top * LX(i, j, k + 1) + nuH_bot * LX(i, j, k);
        const double lyd = nuH_top * LY(i, j, k + 1) - nuH_bot * LY(i, j, k);
        const double lys = nuH_top * LY(i, j, k + 1) + nuH_bot * LY(i, j, k);
        const long c = I3(i, j, k);
        /* :6841-6866 */
        o->Diffxx[c] =
            (2.0 * hfe * (DU_(i + 1, j, k) - DU_(i, j, k)) / dx - 2.0 * hfw * (DU_(i, j, k) - DU_(i - 1, j, k)) / dx) / dx +
            (2.0 * hE * (alpha_c * DU_(i + 1, j, k - 1) * SX(i + 1, j, k - 1) + beta_c * DU_(i + 1, j, k) * SX(i + 1, j, k) +
                         gamma_c * DU_(i + 1, j, k + 1) * SX(i + 1, j, k + 1)) -
             2.0 * hW * (alpha_c * DU_(i - 1, j, k - 1) * SX(i - 1, j, k - 1) + beta_c * DU_(i - 1, j, k) * SX(i - 1, j, k) +
                         gamma_c * DU_(i - 1, j, k + 1) * SX(i - 1, j, k + 1))) / 2.0 / dx +
            (alpha_c * 2.0 * hB * SX(i, j, k - 1) * (DU_(i + 1, j, k - 1) - DU_(i - 1, j, k - 1)) / 2.0 / dx +
             beta_c * 2.0 * hC * SX(i, j, k) * (DU_(i + 1, j, k) - DU_(i - 1, j, k)) / 2.0 / dx +
             gamma_c * 2.0 * hT * SX(i, j, k + 1) * (DU_(i + 1, j, k + 1) - DU_(i - 1, j, k + 1)) / 2.0 / dx);
        /* :6877-6929 */
        o->Diffxy[c] =
            (hfn * (DU_(i, j + 1, k) - DU_(i, j, k)) / dy - hfs * (DU_(i, j, k) - DU_(i, j - 1, k)) / dy) / dy +
            (hN * ((DV_(i + 1, j + 1, k) - DV_(i - 1, j + 1, k)) / 2.0 / dx +
                   alpha_c * (SX(i, j + 1, k - 1) * DV_(i, j + 1, k - 1) + SY(i, j + 1, k - 1) * DU_(i, j + 1, k - 1)) +
                   beta_c * (SX(i, j + 1, k) * DV_(i, j + 1, k) + SY(i, j + 1, k) * DU_(i, j + 1, k)) +
                   gamma_c * (SX(i, j + 1, k + 1) * DV_(i, j + 1, k + 1) + SY(i, j + 1, k + 1) * DU_(i, j + 1, k + 1))) -
             hS * ((DV_(i + 1, j - 1, k) - DV_(i - 1, j - 1, k)) / 2.0 / dx +
                   alpha_c * (SX(i, j - 1, k - 1) * DV_(i, j - 1, k - 1) + SY(i, j - 1, k - 1) * DU_(i, j - 1, k - 1)) +
                   beta_c * (SX(i, j - 1, k) * DV_(i, j - 1, k) + SY(i, j - 1, k) * DU_(i, j - 1, k)) +
                   gamma_c * (SX(i, j - 1, k + 1) * DV_(i, j - 1, k + 1) + SY(i, j - 1, k + 1) * DU_(i, j - 1, k + 1)))) /
                2.0 / dy +
            (alpha_c * hB * SY(i, j, k - 1) *
                 ((DU_(i, j + 1, k - 1) - DU_(i, j - 1, k - 1)) / 2.0 / dy + (DV_(i + 1, j, k - 1) - DV_(i - 1, j, k - 1)) / 2.0 / dx) +
             beta_c * hC * SY(i, j, k) *
                 ((DU_(i, j + 1, k) - DU_(i, j - 1, k)) / 2.0 / dy + (DV_(i + 1, j, k) - DV_(i - 1, j, k)) / 2.0 / dx) +
             gamma_c * hT * SY(i, j, k + 1) *
                 ((DU_(i, j + 1, k + 1) - DU_(i, j - 1, k + 1)) / 2.0 / dy + (DV_(i + 1, j, k + 1) - DV_(i - 1, j, k + 1)) / 2.0 / dx)) +
            ((lyd * alpha_c / dsig[k] + lys / dsigck1 / (dsigck + dsigck1)) * SX(i, j, k - 1) * DV_(i, j, k - 1) +
             (lyd * beta_c / dsig[k] - lys / dsigck1 / dsigck) * SX(i, j, k) * DV_(i, j, k) +
             (lyd * gamma_c / dsig[k] + lys / dsigck / (dsigck + dsigck1)) * SX(i, j, k + 1) * DV_(i, j, k + 1));
        /* :6960-7001 */
        o->Diffyx[c] =
            (hfe * (DV_(i + 1, j, k) - DV_(i, j, k)) / dx - hfw * (DV_(i, j, k) - DV_(i - 1, j, k)) / dx) / dx +
            (hE * ((DU_(i + 1, j + 1, k) - DU_(i + 1, j - 1, k)) / 2.0 / dy +
                   alpha_c * (SX(i + 1, j, k - 1) * DV_(i + 1, j, k - 1) + SY(i + 1, j, k - 1) * DU_(i + 1, j, k - 1)) +
                   beta_c * (SX(i + 1, j, k) * DV_(i + 1, j, k) + SY(i + 1, j, k) * DU_(i + 1, j, k)) +
                   gamma_c * (SX(i + 1, j, k + 1) * DV_(i + 1, j, k + 1) + SY(i + 1, j, k + 1) * DU_(i + 1, j, k + 1))) -
             hW * ((DU_(i - 1, j + 1, k) - DU_(i - 1, j - 1, k)) / 2.0 / dy +
                   alpha_c * (SX(i - 1, j, k - 1) * DV_(i - 1, j, k - 1) + SY(i - 1, j, k - 1) * DU_(i - 1, j, k - 1)) +
                   beta_c * (SX(i - 1, j, k) * DV_(i - 1, j, k) + SY(i - 1, j, k) * DU_(i - 1, j, k)) +
                   gamma_c * (SX(i - 1, j, k + 1) * DV_(i - 1, j, k + 1) + SY(i - 1, j, k + 1) * DU_(i - 1, j, k + 1)))) /
                2.0 / dx +
            (alpha_c * hB * SX(i, j, k - 1) *
                 ((DU_(i, j + 1, k - 1) - DU_(i, j - 1, k - 1)) / 2.0 / dy + (DV_(i + 1, j, k - 1) - DV_(i - 1, j, k - 1)) / 2.0 / dx) +
             beta_c * hC * SX(i, j, k) *
                 ((DU_(i, j + 1, k) - DU_(i, j - 1, k)) / 2.0 / dy + (DV_(i + 1, j, k) - DV_(i - 1, j, k)) / 2.0 / dx) +
             gamma_c * hT * SX(i, j, k + 1) *
                 ((DU_(i, j + 1, k + 1) - DU_(i, j - 1, k + 1)) / 2.0 / dy + (DV_(i + 1, j, k + 1) - DV_(i - 1, j, k + 1)) / 2.0 / dx)) +
            ((lxd * alpha_c / dsig[k] + lxs / dsigck1 / (dsigck + dsigck1)) * SY(i, j, k - 1) * DU_(i, j, k - 1) +
             (lxd * beta_c / dsig[k] - lxs / dsigck1 / dsigck) * SY(i, j, k) * DU_(i, j, k) +
             (lxd * gamma_c / dsig[k] + lxs / dsigck / (dsigck + dsigck1)) * SY(i, j, k + 1) * DU_(i, j, k + 1));
        /* :7008-7032 */
        o->Diffyy[c] =
            (2.0 * hfn * (DV_(i, j + 1, k) - DV_(i, j, k)) / dy - 2.0 * hfs * (DV_(i, j, k) - DV_(i, j - 1, k)) / dy) / dy +
            (2.0 * hN * (alpha_c * DV_(i, j + 1, k - 1) * SY(i, j + 1, k - 1) + beta_c * DV_(i, j + 1, k) * SY(i, j + 1, k) +
                         gamma_c * DV_(i, j + 1, k + 1) * SY(i, j + 1, k + 1)) -
             2.0 * hS * (alpha_c * DV_(i, j - 1, k - 1) * SY(i, j - 1, k - 1) + beta_c * DV_(i, j - 1, k) * SY(i, j - 1, k) +
                         gamma_c * DV_(i, j - 1, k + 1) * SY(i, j - 1, k + 1))) / 2.0 / dy +
            (alpha_c * 2.0 * hB * SY(i, j, k - 1) * (DV_(i, j + 1, k - 1) - DV_(i, j - 1, k - 1)) / 2.0 / dy +
             beta_c * 2.0 * hC * SY(i, j, k) * (DV_(i, j + 1, k) - DV_(i, j - 1, k)) / 2.0 / dy +
             gamma_c * 2.0 * hT * SY(i, j, k + 1) * (DV_(i, j + 1, k + 1) - DV_(i, j - 1, k + 1)) / 2.0 / dy);
        /* :7063-7099 */
        o->Diffzx[c] =
            (hfe * (DW_(i + 1, j, k) - DW_(i, j, k)) / dx - hfw * (DW_(i, j, k) - DW_(i - 1, j, k)) / dx) / dx +
            (hE * (alpha_c * (DW_(i + 1, j, k - 1) * SX(i + 1, j, k - 1) + DU_(i + 1, j, k - 1) / Dc) +
                   beta_c * (DW_(i + 1, j, k) * SX(i + 1, j, k) + DU_(i + 1, j, k) / Dc) +
                   gamma_c * (DW_(i + 1, j, k + 1) * SX(i + 1, j, k + 1) + DU_(i + 1, j, k + 1) / Dc)) -
             hW * (alpha_c * (DW_(i - 1, j, k - 1) * SX(i - 1, j, k - 1) + DU_(i - 1, j, k - 1) / Dc) +
                   beta_c * (DW_(i - 1, j, k) * SX(i - 1, j, k) + DU_(i - 1, j, k) / Dc) +
                   gamma_c * (DW_(i - 1, j, k + 1) * SX(i - 1, j, k + 1) + DU_(i - 1, j, k + 1) / Dc))) /
                2.0 / dx +
            (alpha_c * hB * SX(i, j, k - 1) * (DW_(i + 1, j, k - 1) - DW_(i - 1, j, k - 1)) / 2.0 / dx +
             beta_c * hC * SX(i, j, k) * (DW_(i + 1, j, k) - DW_(i - 1, j, k)) / 2.0 / dx +
             gamma_c * hT * SX(i, j, k + 1) * (DW_(i + 1, j, k + 1) - DW_(i - 1, j, k + 1)) / 2.0 / dx) +
            ((lxd * alpha_c / dsig[k] + lxs / dsigck1 / (dsigck + dsigck1)) * DU_(i, j, k - 1) / Dc +
             (lxd * beta_c / dsig[k] - lxs / dsigck1 / dsigck) * DU_(i, j, k) / Dc +
             (lxd * gamma_c / dsig[k] + lxs / dsigck / (dsigck + dsigck1)) * DU_(i, j, k + 1) / Dc);
        /* :7100-7136 */
        o->Diffzy[c] =
            (hfn * (DW_(i, j + 1, k) - DW_(i, j, k)) / dy - hfs * (DW_(i, j, k) - DW_(i, j - 1, k)) / dy) / dy +
            (hN * (alpha_c * (DW_(i, j + 1, k - 1) * SY(i, j + 1, k - 1) + DV_(i, j + 1, k - 1) / Dc) +
                   beta_c * (DW_(i, j + 1, k) * SY(i, j + 1, k) + DV_(i, j + 1, k) / Dc) +
                   gamma_c * (DW_(i, j + 1, k + 1) * SY(i, j + 1, k + 1) + DV_(i, j + 1, k + 1) / Dc)) -
             hS * (alpha_c * (DW_(i, j - 1, k - 1) * SY(i, j - 1, k - 1) + DV_(i, j - 1, k - 1) / Dc) +
                   beta_c * (DW_(i, j - 1, k) * SY(i, j - 1, k) + DV_(i, j - 1, k) / Dc) +
                   gamma_c * (DW_(i, j - 1, k + 1) * SY(i, j - 1, k + 1) + DV_(i, j - 1, k + 1) / Dc))) /
                2.0 / dy +
            (alpha_c * hB * SY(i, j, k - 1) * (DW_(i, j + 1, k - 1) - DW_(i, j - 1, k - 1)) / 2.0 / dy +
             beta_c * hC * SY(i, j, k) * (DW_(i, j + 1, k) - DW_(i, j - 1, k)) / 2.0 / dy +
             gamma_c * hT * SY(i, j, k + 1) * (DW_(i, j + 1, k + 1) - DW_(i, j - 1, k + 1)) / 2.0 / dy) +
            ((lyd * alpha_c / dsig[k] + lys / dsigck1 / (dsigck + dsigck1)) * DV_(i, j, k - 1) / Dc +
             (lyd * beta_c / dsig[k] - lys / dsigck1 / dsigck) * DV_(i, j, k) / Dc +
             (lyd * gamma_c / dsig[k] + lys / dsigck / (dsigck + dsigck1)) * DV_(i, j, k + 1) / Dc);
      }
  }
#undef DU_
#undef DV_
#undef DW_
#undef SX
#undef SY
#undef LX
#undef LY
}

/* trig, src/nhwave.F:13182-13224: Thomas algorithm, a/b/c the sub-, main and
 * super-diagonal (a[0], c[n-1] unused) */
static void trig(const double *a, const double *b, const double *c, const double *r, double *x, int n) {
  double bp[64], rp[64];
  bp[0] = b[0];
  rp[0] = r[0];
  for (int q = 1; q < n; ++q) {
    double coeff = a[q] / bp[q - 1];
    bp[q] = b[q] - coeff * c[q - 1];
    rp[q] = r[q] - coeff * rp[q - 1];
  }
  x[n - 1] = rp[n - 1] / bp[n - 1];
  for (int q = n - 2; q >= 0; --q) x[q] = (rp[q] - c[q] * x[q + 1]) / bp[q];
}

/* implicit vertical viscous solves of eval_duvw for DU, DV and DW
 * (src/nhwave.F:693-821, 852-985, 1015-1144) with Bc_Z0 = 1: per wet column
 * a tridiagonal system in k; the surface rows of DU/DV take the wind stress,
 * the bottom and surface rows of DW the kinematic W (Wbot, Wtop). */
static void viscous_solves(orc *o) {
  const double dt = o->dt, S = o->c.schmidt, cm = o->c.visc;
  const double *dsig = o->dsig - 1;
  const int n = o->Kend - o->Kbeg + 1;
  double A[64], B[64], Cc[64], R[64], X[64];
  for (int var = 0; var < 3; ++var) {
    double *Q = var == 0 ? o->DU : var == 1 ? o->DV : o->DW;
    const double *RQ = var == 0 ? o->R2 : var == 1 ? o->R3 : o->R4;
    for (int i = o->Ibeg; i <= o->Iend; ++i)
      for (int j = o->Jbeg; j <= o->Jend; ++j) {
        const long c2 = I2(i, j);
        if (o->Mask[c2] == 0) continue;
        const double Dc = o->D[c2];
        for (int k = o->Kbeg; k <= o->Kend; ++k) {
          const int q = k - o->Kbeg;
          const long c = I3(i, j, k);
          /* viscosity at the sigma levels below and above level k */
          const double cvm = o->CmuVt[I3(i, j, k - 1)], cv = o->CmuVt[c], cvp = o->CmuVt[I3(i, j, k + 1)];
          const double nub = 0.5 * (cm + cm + 0.0 + 0.0) + 0.5 * (cvm + cv) / S;
          const double nut = 0.5 * (cm + cm + 0.0 + 0.0) + 0.5 * (cv + cvp) / S;
          A[q] = k == o->Kbeg ? 0.0 : -dt / (Dc * Dc) * nub / (0.5 * dsig[k] * (dsig[k] + dsig[k - 1]));
          Cc[q] = k == o->Kend ? 0.0 : -dt / (Dc * Dc) * nut / (0.5 * dsig[k] * (dsig[k] + dsig[k + 1]));
          if (var < 2) {
            if (k == o->Kbeg && o->c.bc_z0 == 2)  /* no-slip bottom row */
              B[q] = 1.0 - Cc[q] + dt / (Dc * Dc) * (cm + 0.0 + cv / S) / (0.5 * dsig[k] * dsig[k]);
            else
              B[q] = 1.0 - A[q] - Cc[q];
            if (k == o->Kbeg && o->c.bc_z0 == 5) {  /* friction law */
              const double Dz1 = 0.5 * Dc * dsig[o->Kbeg];
              double cdrag;
              if (o->c.ibot == 1) cdrag = o->c.cd0;
              else {
                const double t = 1. / KAPPA * log(30.0 * Dz1 / o->c.zob);
                cdrag = 1. / (t * t);
              }
              const double u = o->U[c], v = o->V[c];
              const double ustar2 = cdrag * sqrt(u * u + v * v) * (var == 0 ? u : v);
              R[q] = Q[c] + dt * RQ[c] - dt * ustar2 / dsig[k];
            } else if (k == o->Kend)
              R[q] = Q[c] + dt * RQ[c] + dt * (var == 0 ? o->Wsx : o->Wsy)[c2] / dsig[k];
            else
              R[q] = Q[c] + dt * RQ[c];
          } else if (k == o->Kbeg) {
            B[q] = 1.0 - Cc[q] + dt / (Dc * Dc) * nub / (0.25 * dsig[k] * (dsig[k] + dsig[k - 1]));
            const double Wbot = -o->DeltH[c2] - o->U[I3(i, j, o->Kbeg)] * o->DelxH[c2] -
                                o->V[I3(i, j, o->Kbeg)] * o->DelyH[c2];
            R[q] = Q[c] + dt * RQ[c] + dt / Dc * nub / (0.25 * dsig[k] * (dsig[k] + dsig[k - 1])) * Wbot;
          } else if (k == o->Kend) {
            B[q] = 1.0 - A[q] + dt / (Dc * Dc) * nut / (0.25 * dsig[k] * (dsig[k] + dsig[k + 1]));
            const double Wtop = (o->Eta[c2] - o->Eta0[c2]) / dt + o->U[I3(i, j, o->Kend)] * o->DelxEta[c2] +
                                o->V[I3(i, j, o->Kend)] * o->DelyEta[c2];
            R[q] = Q[c] + dt * RQ[c] + dt / Dc * nut / (0.25 * dsig[k] * (dsig[k] + dsig[k + 1])) * Wtop;
          } else {
            B[q] = 1.0 - A[q] - Cc[q];
            R[q] = Q[c] + dt * RQ[c];
          }
        }
        trig(A, B, Cc, R, X, n);
        for (int k = o->Kbeg; k <= o->Kend; ++k) Q[I3(i, j, k)] = X[k - o->Kbeg];
      }
  }
}

int orc_eval_duvw(orc *o, int istep) {
  const double dx = o->c.dx, dy = o->c.dy, dt = o->dt;
  const double *dsig = o->dsig - 1;
  const double A = o->alpha[istep - 1], B = o->beta[istep - 1];
  /* horizontal and cross stress terms, :575-576 */
  if (o->c.viscous) diffusion(o);
  /* mass, :611-628 */
  memset(o->R1, 0, sizeof(double) * o->n2);
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      long c = I2(i, j);
      if (o->Mask[c] == 0) continue;
      for (int k = o->Kbeg; k <= o->Kend; ++k)
        o->R1[c] = o->R1[c] - 1.0 / dx * (o->Ex[IX3(i + 1, j, k)] - o->Ex[IX3(i, j, k)]) * dsig[k] -
                   1.0 / dy * (o->Ey[IY3(i, j + 1, k)] - o->Ey[IY3(i, j, k)]) * dsig[k];
      o->R1[c] = o->R1[c] + o->D[c] * o->SourceC[c];
      o->D[c] = A * o->D0[c] + B * (o->D[c] + dt * o->R1[c]);
    }
  for (long q = 0; q < o->n2; ++q) o->D[q] = dmax(o->D[q], o->c.min_dep);
  wl_bc(o);
  for (long q = 0; q < o->n2; ++q) o->Eta[q] = o->D[q] - o->Hc[q];
  delx_2d(o, o->Eta, o->DelxEta);
  dely_2d(o, o->Eta, o->DelyEta);
  sigma_transform(o);
  /* momentum right-hand sides, :637-688; the baroclinic, diffusion and
   * external-forcing arrays are identically zero on this path and are added
   * as 0.0 in their places. */
  memset(o->R2, 0, sizeof(double) * o->n3);
  memset(o->R3, 0, sizeof(double) * o->n3);
  memset(o->R4, 0, sizeof(double) * o->n3);
  const double fcor = o->c.fcor;
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      if (o->Mask[I2(i, j)] == 0) continue;
      for (int k = o->Kbeg; k <= o->Kend; ++k) {
        long c = I3(i, j, k);
        o->R2[c] = -1.0 / dx * (o->Fx[IX3(i + 1, j, k)] - o->Fx[IX3(i, j, k)]) -
                   1.0 / dy * (o->Fy[IY3(i, j + 1, k)] - o->Fy[IY3(i, j, k)]) -
                   1.0 / dsig[k] * (o->Fz[IZ3(i, j, k + 1)] - o->Fz[IZ3(i, j, k)]) + fcor * o->DV[c] +
                   0.0 + o->SourceX[I2(i, j)] + o->Diffxx[c] + o->Diffxy[c] + 0.0; /* DRhoX, ExtForceX */
        o->R3[c] = -1.0 / dx * (o->Gx[IX3(i + 1, j, k)] - o->Gx[IX3(i, j, k)]) -
                   1.0 / dy * (o->Gy[IY3(i, j + 1, k)] - o->Gy[IY3(i, j, k)]) -
                   1.0 / dsig[k] * (o->Gz[IZ3(i, j, k + 1)] - o->Gz[IZ3(i, j, k)]) - fcor * o->DU[c] +
                   0.0 + o->SourceY[I2(i, j)] + o->Diffyx[c] + o->Diffyy[c] + 0.0; /* DRhoY, ExtForceY */
        o->R4[c] = -1.0 / dx * (o->Hx[IX3(i + 1, j, k)] - o->Hx[IX3(i, j, k)]) -
                   1.0 / dy * (o->Hy[IY3(i, j + 1, k)] - o->Hy[IY3(i, j, k)]) -
                   1.0 / dsig[k] * (o->Hz[IZ3(i, j, k + 1)] - o->Hz[IZ3(i, j, k)]) + o->Diffzx[c] + o->Diffzy[c];
      }
    }
  /* explicit update, :841-845, 1004-1008, 1163-1167, or the implicit
   * vertical viscous solves */
  if (o->c.viscous) viscous_solves(o);
  else for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      if (o->Mask[I2(i, j)] == 0) continue;
      for (int k = o->Kbeg; k <= o->Kend; ++k) {
        long c = I3(i, j, k);
        o->DU[c] = o->DU[c] + dt * o->R2[c];
      }
    }
  if (!o->c.viscous) for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      if (o->Mask[I2(i, j)] == 0) continue;
      for (int k = o->Kbeg; k <= o->Kend; ++k) {
        long c = I3(i, j, k);
        o->DV[c] = o->DV[c] + dt * o->R3[c];
      }
    }
  if (!o->c.viscous) for (int i = o->Ibeg; i <= o->Iend; ++i)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      if (o->Mask[I2(i, j)] == 0) continue;
      for (int k = o->Kbeg; k <= o->Kend; ++k) {
        long c = I3(i, j, k);
        o->DW[c] = o->DW[c] + dt * o->R4[c];
      }
    }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i)
        if (o->Mask[I2(i, j)] == 0) {
          long c = I3(i, j, k);
          o->DU[c] = ZERO; o->DV[c] = ZERO; o->DW[c] = ZERO;
        }
  /* non-hydrostatic block, :1186-1198 */
  if (o->c.non_hydro) {
    get_uvw(o);
    interpolate_velocity_to_faces(o);
    int rc = orc_poisson_solver(o);
    if (rc != 0) return rc;
    projection_corrector(o);
  }
  /* SSP-RK blend, :1227-1241 */
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        long c = I3(i, j, k);
        o->DU[c] = A * o->DU0[c] + B * o->DU[c];
        o->DV[c] = A * o->DV0[c] + B * o->DV[c];
        o->DW[c] = A * o->DW0[c] + B * o->DW[c];
        if (o->Mask[I2(i, j)] == 0) { o->DU[c] = ZERO; o->DV[c] = ZERO; o->DW[c] = ZERO; }
      }
  /* Froude cap, :1244-1261 */
  if (o->c.froude_cap_on)
    for (int j = o->Jbeg; j <= o->Jend; ++j)
      for (int i = o->Ibeg; i <= o->Iend; ++i) {
        double d = o->D[I2(i, j)];
        double fu = sqrt(GRAV * d) * o->c.froude_cap * d;
        if (o->Mask[I2(i, j)] > 0)
          for (int k = o->Kbeg; k <= o->Kend; ++k) {
            long c = I3(i, j, k);
            double duu = sqrt(o->DU[c] * o->DU[c] + o->DV[c] * o->DV[c] + o->DW[c] * o->DW[c]);
            if (duu > fu) {
              double ang = atan2(o->DV[c], o->DU[c]);
              o->DU[c] = fu * cos(ang);
              o->DV[c] = fu * sin(ang);
              o->DW[c] = ZERO;
            }
          }
      }
  get_uvw(o);
  get_omega(o);
  return 0;
}

/* one time step, src/nhwave.F:119-258 */
int orc_step(orc *o) {
  int rc = orc_estimate_dt(o);
  if (rc) return rc;
  orc_update_wave_bc(o);
  orc_update_mask(o);
  update_wind(o);
  orc_update_vars(o);
  for (int s = 1; s <= o->c.it_order; ++s) {
    orc_source_terms(o);
    orc_fluxes(o);
    rc = orc_eval_duvw(o, s);
    if (rc) return rc;
    if (o->c.sponge_on) orc_sponge_damping(o);
    if (o->c.viscous) orc_eval_turb(o);
  }
  return 0;
}

/* ======================================================================== */
/* Boundary forcing                                                         */
/* ======================================================================== */

static const double PI = F32(3.141592653);     /* src/mod_global.F:21 */

/* wave_dispersion, src/waves.F:61-80: Newton iteration for the linear wave
 * number from Eckart's guess.  Integer powers are products (x**2 = x*x). */
static double wave_dispersion(double segma, double dep) {
  double wk = (segma * segma) / (GRAV * sqrt(tanh(segma * segma * dep / GRAV)));
  double err = 100.0;
  while (err >= F32(1.0e-8)) {
    double th = tanh(wk * dep), ch = cosh(wk * dep);
    double fk = segma * segma - GRAV * wk * th;
    double fkdif = -GRAV * (th + (wk * dep) / (ch * ch));
    double wn = wk - fk / fkdif;
    err = fabs((wk - wn) / wk);
    wk = wn;
  }
  return wk;
}

/* The internal irregular wavemaker (WAVEMAKER = INT_IRR), source_terms,
 * src/source_terms.F:194-305: inside the source box every interior cell
 * gets the sum over the spectrum's components (frequency outer, direction
 * inner) of Celerity*Amp/Source_Area*cos(pi/2 - Segma*TIME + Phs_lag +
 * Random_Phs), the first frequency's terms in the form
 * 2*Celerity/Source_Area*(Amp/2*cos(...) + eta_mean); eta_mean is never
 * assigned in the reference (a zero module variable).  Source_Area sums
 * dx*D(i,Jbeg+1) over the box's columns.  A direction off the x axis is
 * moved to the nearest one a periodic y boundary admits (:230-275), as
 * written there (for negative angles the candidate of the smaller wave
 * number keeps its positive sign). */
static void int_irr_source(orc *o) {
  const double dx = o->c.dx, dy = o->c.dy, T = o->time, Dp = o->c.dep_wave;
  const int nf = o->c.nfreq, nd = o->c.ndir, ng = o->c.nglob;
  double area = 0.0;
  for (int i = o->Ibeg; i <= o->Iend; ++i)
    if (o->xc[i - 1] >= o->c.xs_w && o->xc[i - 1] <= o->c.xs_e) area = area + dx * o->D[I2(i, o->Jbeg + 1)];
  /* the per-component values do not depend on the cell: the same
   * expressions are formed once per call here (the reference forms them
   * again in every cell of the box) */
  const long n = (long)nf * nd;
  double *segma = (double *)malloc(sizeof(double) * 4 * n), *cel = segma + n, *wny = cel + n, *amp = wny + n;
  for (int f = 1; f <= nf; ++f)
    for (int dd = 1; dd <= nd; ++dd) {
      const long q = (long)(f - 1) * nd + (dd - 1);
      const double per = 1.0 / o->c.freq[f - 1];
      segma[q] = 2.0 * PI / per;
      const double wk = wave_dispersion(segma[q], Dp);
      const double wl = 2.0 * PI / wk;
      cel[q] = wl / per;
      double angle = o->c.dire[dd - 1] * PI / 180.0;
      if (angle > ZERO) {
        double t3 = ZERO;
        const double t1 = wk;
        int nk = 0;
        while (t3 < angle) {
          ++nk;
          const double t2 = nk * 2.0 * PI / (ng * dy);
          t3 = t2 >= t1 ? 0.5 * PI - SMALL : asin(t2 / t1);
        }
        const double t4 = asin((nk - 1) * 2.0 * PI / (ng * dy) / t1);
        angle = fabs(t4 - angle) < fabs(angle - t3) ? t4 : t3;
      } else if (angle < ZERO) {
        double t3 = ZERO;
        const double t1 = wk;
        int nk = 0;
        while (t3 > angle) {
          ++nk;
          const double t2 = nk * 2.0 * PI / (ng * dy);
          t3 = t2 >= t1 ? -0.5 * PI + SMALL : -asin(t2 / t1);
        }
        const double t4 = asin((nk - 1) * 2.0 * PI / (ng * dy) / t1);
        angle = fabs(t4 - angle) < fabs(angle - t3) ? t4 : t3;
      } else {
        angle = 0.0;
      }
      wny[q] = wk * sin(angle);
      amp[q] = 2.0 * o->c.spc[q];
    }
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      double sc = ZERO;
      if (o->xc[i - 1] >= o->c.xs_w && o->xc[i - 1] <= o->c.xs_e && o->yc[j - 1] >= o->c.ys_s &&
          o->yc[j - 1] <= o->c.ys_n)
        for (long q = 0; q < n; ++q) {
          const double lag = (dy / 2.0 + o->yc[j - 1]) * wny[q];
          const double arg = PI / 2.0 - segma[q] * T + lag + o->c.phs[q];
          if (q < nd)  /* the first frequency */
            sc = sc + 2.0 * cel[q] / area * (amp[q] / 2.0 * cos(arg) + 0.0);
          else
            sc = sc + cel[q] * amp[q] / area * cos(arg);
        }
      o->SourceC[I2(i, j)] = sc;
    }
  free(segma);
}

/* The arithmetic-geometric mean of the cnoidal helpers (src/waves.F:
 * 225-361): a(1) = 1, b(1) = sqrt(1-m), c(1) = sqrt(m), iterated until
 * |c| < 1.0e-15 (a default-REAL literal) or n > 1000; a[], c[] 1-based.
 * Returns n. */
#define CN_MAXN 1001
static int cn_agm(double m, double *a, double *c) {
  double a0 = 1.0, b0 = sqrt(1.0 - m), c0 = sqrt(m);
  int n = 1;
  a[1] = a0;
  c[1] = c0;
  while (!(fabs(c0) < F32(1.0e-15f) || n > 1000)) {
    ++n;
    const double a1 = (a0 + b0) / 2.0, b1 = sqrt(a0 * b0), c1 = (a0 - b0) / 2.0;
    a0 = a1;
    b0 = b1;
    c0 = c1;
    a[n] = a0;
    c[n] = c0;
  }
  return n;
}

/* cnoidal_ck, src/waves.F:275-311: K(m) = pi/2/a(n); the pi literal is a
 * default REAL */
static double cn_ck(double m) {
  double a[CN_MAXN + 1], c[CN_MAXN + 1];
  const int n = cn_agm(m, a, c);
  return F32(3.1415926535897932384626f) / 2.0 / a[n];
}

/* cnoidal_ce, src/waves.F:315-361: E(m) = K(m) (1 - sum 2**(k-2) c(k)**2) */
static double cn_ce(double m) {
  double a[CN_MAXN + 1], c[CN_MAXN + 1];
  const int n = cn_agm(m, a, c);
  const double ck = F32(3.1415926535897932384626f) / 2.0 / a[n];
  double sum = 0.0;
  for (int k = 1; k <= n; ++k) sum = sum + ldexp(1.0, k - 2) * (c[k] * c[k]);
  return ck * (1.0 - sum);
}

/* cnoidal_cn, src/waves.F:225-271: the Jacobi cn(u|m) by the descending
 * Landen recurrence y(n) = 2**(n-1) a(n) u, y(i) = (y(i+1) + asin(c(i+1)/
 * a(i+1) sin y(i+1)))/2 */
static double cn_cn(double u, double m) {
  double a[CN_MAXN + 1], c[CN_MAXN + 1];
  const int n = cn_agm(m, a, c);
  double y = ldexp(1.0, n - 1) * a[n] * u;
  for (int i = n - 1; i >= 1; --i) y = (y + asin(c[i + 1] / a[i + 1] * sin(y))) / 2.0;
  return cos(y);
}

/* cnoidal, src/waves.F:184-220: Newton iteration on the modulus from
 * m = 0.99999999 until |xa| <= 1.0e-8 (a default-REAL literal) or n > 1000,
 * then the wave length, celerity and trough level */
static void cnoidal(double height, double depth, double period, double *l, double *cel, double *x2,
                    double *mod) {
  double m = 0.99999999;
  int n = 0;
  for (;;) {
    ++n;
    const double ck = cn_ck(m), ce = cn_ce(m);
    const double d3 = depth * depth * depth, p2 = period * period, ck2 = ck * ck;
    const double xa = m * depth + 2.0 * height - m * height - 3.0 * height * ce / ck -
                      16.0 * d3 * (m * m) * ck2 / 3.0 / GRAV / height / p2;
    if (fabs(xa) <= F32(1.0e-8f) || n > 1000) break;
    const double xb = depth - height +
                      3.0 * height / 2.0 / m / (1.0 - m) / ck2 *
                          ((1.0 - m) * ck2 + ce * ce - 2.0 * (1.0 - m) * ck * ce) -
                      16.0 * d3 * m * ck / 3.0 / GRAV / (1.0 - m) / height / p2 * ((1.0 - m) * ck + ce);
    m = m - xa / xb;
  }
  const double ck = cn_ck(m), ce = cn_ce(m);
  *l = 4.0 * ck * depth * sqrt(m * depth / height / 3.0);
  const double xtemp = -m + 2.0 - 3.0 * ce / ck;
  *cel = sqrt(GRAV * depth * (1.0 + height / depth / m * xtemp));
  *x2 = height / m * (1.0 - m - ce / ck);
  *mod = m;
}

/* cnoidal_wave_left_boundary, src/waves.F:92-170: the surface starts at its
 * zero crossing (bisection on [0, 1] to |Eta| <= 1.0e-6 within 1000 steps),
 * then Eta and the velocity profiles of the cnoidal wave less its mean
 * mass transport */
static int cnoidal_start(double A, double Dp, double Per, double *L, double *yt, double *m, double *ck,
                         double *xstart) {
  double cel;
  cnoidal(A, Dp, Per, L, &cel, yt, m);
  *ck = cn_ck(*m);
  double zup = 1.0, zlow = 0.0, zmid = (zup + zlow) / 2.0;
  for (int nn = 1;; ++nn) {
    const double cn = cn_cn(zmid * 0.5 * *ck, *m);
    const double zero1 = *yt + A * (cn * cn);
    if (fabs(zero1) <= F32(1.0e-6f)) break;
    if (nn > 1000) return -1;  /* the reference stops: 'too many iterations' */
    if (zero1 < 0.0)
      zup = zmid;
    else
      zlow = zmid;
    zmid = (zup + zlow) / 2.0;
  }
  *xstart = zmid;
  return 0;
}

double orc_cnoidal_cn(double u, double m) { return cn_cn(u, m); }
double orc_cnoidal_ck(double m) { return cn_ck(m); }
double orc_cnoidal_ce(double m) { return cn_ce(m); }
int orc_cnoidal_setup(double height, double depth, double period, double *out) {
  double cel, ck;
  cnoidal(height, depth, period, &out[0], &cel, &out[2], &out[3]);
  out[1] = cel;
  return cnoidal_start(height, depth, period, &out[0], &out[2], &out[3], &ck, &out[4]);
}

static void cnoidal_left(orc *o) {
  const double A = o->c.amp_wave, Dp = o->c.dep_wave, Per = o->c.per_wave, T = o->time;
  const double *sigc = o->sigc - 1;
  double L, yt, m, ck, xstart;
  (void)cnoidal_start(A, Dp, Per, &L, &yt, &m, &ck, &xstart);  /* checked at orc_create */
  const double u = xstart * 0.5 * ck + 2.0 * ck * (-(T / Per));
  const double xcn = cn_cn(u, m);
  for (int j = 1; j <= o->Nloc; ++j) {
    o->Ein_X0[j - 1] = yt + A * (xcn * xcn);
    o->Din_X0[j - 1] = o->Ein_X0[j - 1] + o->Hfx[IX2(o->Ibeg, j)];
  }
  const double drift = GRAV * (A * A) / (L / Per) / Dp / 8.0;
  const double fact = 1.0;
  const double sgd = sqrt(GRAV * Dp);
  const double c2 = xcn * xcn, c4 = c2 * xcn * xcn;
  const double xsn = sqrt(1.0 - c2), xdn = sqrt(1.0 - m * (1.0 - c2));
  const double s2 = xsn * xsn, d2 = xdn * xdn, m2 = m * m;
  const double dy = Dp + yt;
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      const double z = sigc[k] * o->Din_X0[j - 1];
      o->Uin_X0[IB(j, k)] =
          sgd * (-1.25 + 3.0 * dy / 2.0 / Dp - (dy * dy) / 4.0 / (Dp * Dp) +
                 (3.0 * A / 2.0 / Dp - dy * A / 2.0 / (Dp * Dp)) * c2 - (A * A) / 4.0 / (Dp * Dp) * c4 -
                 8.0 * A * (ck * ck) / (L * L) * (Dp / 3.0 - (z * z) / 2.0 / Dp) *
                     (-(m2 * s2 * c2) + c2 * d2 - s2 * d2)) -
          drift * fact;
      o->Win_X0[IB(j, k)] = sgd * z * 2.0 * A * ck / L / Dp *
                            (1.0 + dy / Dp + A / Dp * c2 +
                             32.0 * (ck * ck) / 3.0 / (L * L) * (Dp * Dp - (z * z) / 2.0) *
                                 (m2 * s2 - m2 * c2 - d2)) *
                            xsn * xcn * xdn;
      o->Vin_X0[IB(j, k)] = 0.0;
    }
}

/* flux_left_right_boundary, src/nhwave.F:5565-5712 (WAVEMAKER = FLUX_LR):
 * at the west a still surface (Ein_X0 = 0, Din_X0 = Hfx(Ibeg)) with a
 * logarithmic velocity profile of mean 0.115 m/s over the roughness Zob,
 * ramped by tanh(TIME/TRAMP), and its discharge FluxL (k outer, j inner);
 * at the east Orlanski radiation of Eta, V and W from the two previous
 * steps, and an outflow velocity Uin_Xni = FluxL/(Din_Xn dy Nglob) that
 * carries FluxL out.  The literals 0.115, 30. and 1.e-16 are default REALs. */
static void flux_left_right(orc *o) {
  const double uavg = F32(0.115), kappa = F32(0.41), zob = o->c.zob, dy = o->c.dy;
  const double ramp = o->c.tramp > 0.0 ? tanh(o->time / o->c.tramp) : 1.0;
  const double *sigc = o->sigc - 1, *dsig = o->dsig - 1;
  double fluxl = 0.0;
  for (int j = 1; j <= o->Nloc; ++j) {
    o->Ein_X0[j - 1] = 0.0;
    o->Din_X0[j - 1] = o->Ein_X0[j - 1] + o->Hfx[IX2(o->Ibeg, j)];
  }
  const double ufric = uavg * kappa / log(30.0 * o->Din_X0[o->Jbeg - 1] / zob);
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      const double zlev1 = sigc[k] * o->Din_X0[j - 1];
      o->Uin_X0[IB(j, k)] = ufric / kappa * log(30.0 * zlev1 / zob) * ramp;
      fluxl = fluxl + dsig[k] * o->Din_X0[j - 1] * dy * o->Uin_X0[IB(j, k)];
      o->Win_X0[IB(j, k)] = 0.0;
      o->Vin_X0[IB(j, k)] = 0.0;
    }
  const int Ie = o->Iend;
  for (int j = 1; j <= o->Nloc; ++j) {
    const long a = I2(Ie, j), b = I2(Ie - 1, j);
    double cg = -(o->Eta0[a] - o->Eta00[a] + SMALL) / (o->Eta00[a] - o->Eta00[b] + SMALL);
    cg = fmax(fmin(cg, 1.0), 0.0);
    o->Ein_Xn[j - 1] = (cg * o->Eta0[a] + (1.0 - cg) * o->Eta0[I2(Ie + 1, j)] + o->Eta[a]) / 2.0;
    o->Din_Xn[j - 1] = o->Hfx[IX2(Ie + 1, j)] + o->Ein_Xn[j - 1];
  }
  memcpy(o->Uin_Xni0, o->Uin_Xni, sizeof(double) * o->Nloc * o->Kloc);
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      const long q = IB(j, k), a = I3(Ie, j, k), b = I3(Ie - 1, j, k), e = I3(Ie + 1, j, k);
      o->Uin_Xni[q] = fluxl / (o->Din_Xn[j - 1] * dy * (double)o->c.nglob);
      double cg = -(o->U0[a] - o->U00[a] + SMALL) / (o->U00[a] - o->U00[b] + SMALL);
      cg = fmax(fmin(cg, 1.0), 0.0);
      o->Uin_Xn[q] = (cg * o->U[a] + o->Uin_Xni0[q] - cg * o->Uin_Xni[q] + o->U[a]) / 2.0;
      cg = -(o->V0[a] - o->V00[a] + SMALL) / (o->V00[a] - o->V00[b] + SMALL);
      cg = fmax(fmin(cg, 1.0), 0.0);
      o->Vin_Xn[q] = (cg * o->V0[a] + (1.0 - cg) * o->V0[e] + o->V[a]) / 2.0;
      cg = -(o->W0[a] - o->W00[a] + SMALL) / (o->W00[a] - o->W00[b] + SMALL);
      cg = fmax(fmin(cg, 1.0), 0.0);
      o->Win_Xn[q] = (cg * o->W0[a] + (1.0 - cg) * o->W0[e] + o->W[a]) / 2.0;
    }
}

/* update_wave_bc, src/nhwave.F:5514-5562, for the west wavemakers
 * solitary_wave_left_boundary (src/waves.F:869-903),
 * linear_wave_left_boundary (:808-855), cnoidal_wave_left_boundary
 * (:92-170) and stokes_wave_left_boundary (:518-560).  x**n with constant
 * n is evaluated as the reference build evaluates it: n-1 multiplications
 * from the left, ((x*x)*x)*... */
void orc_update_wave_bc(orc *o) {
  const int wm = o->c.wavemaker;
  if (wm == 0 || wm == 5) return;  /* none; the internal wavemaker is a source term */
  if (wm == 4) {
    cnoidal_left(o);
    return;
  }
  if (wm == 6) {
    flux_left_right(o);
    return;
  }
  const double A = o->c.amp_wave, Dp = o->c.dep_wave, T = o->time;
  const double *sigc = o->sigc - 1;
  const double *hfx = o->Hfx;
  if (wm == 1) {
    for (long q = 0; q < (long)o->Nloc * o->Kloc; ++q) {
      o->Uin_X0[q] = ZERO; o->Vin_X0[q] = ZERO; o->Win_X0[q] = ZERO;
    }
    const double cel = sqrt(GRAV * Dp * (1.0 + A / Dp));
    for (int j = 1; j <= o->Nloc; ++j) {
      const double at = sqrt(0.75 * A / (Dp * Dp * Dp));
      const double xs = 4.0 * Dp / sqrt(A / Dp);
      const double xc = xs - cel * T;
      const double ch = cosh(at * xc), sh = sinh(at * xc);
      const double ein = A / (ch * ch);
      const double din = ein + hfx[IX2(o->Ibeg, j)];
      o->Ein_X0[j - 1] = ein;
      o->Din_X0[j - 1] = din;
      const double c2 = sqrt(GRAV * Dp);
      const double ch2 = ch * ch, ch3 = ch2 * ch, ch4 = ch3 * ch, ch5 = ch4 * ch;
      const double d2 = 2.0 * A * (at * at) * (2.0 * ch2 - 3.0) / ch4;
      const double d1 = -2.0 * A * sh * at / ch3;
      const double d3 = -8.0 * A * sh * (at * at * at) * (ch2 - 3.0) / ch5;
      for (int k = o->Kbeg; k <= o->Kend; ++k) {
        const double z = sigc[k] * din;
        o->Uin_X0[IB(j, k)] = c2 * ein / Dp *
                              (1.0 - 0.25 * ein / Dp + Dp / 3.0 * (Dp / ein) * (1.0 - 1.5 * (z * z) / (Dp * Dp)) * d2);
        o->Win_X0[IB(j, k)] = -c2 * z / Dp *
                              ((1.0 - 0.5 * ein / Dp) * d1 +
                               F32(1.0f / 3.0f) * (Dp * Dp) * (1.0 - 0.5 * (z * z) / (Dp * Dp)) * d3);
        o->Vin_X0[IB(j, k)] = ZERO;
      }
    }
    return;
  }
  const double segma = 2.0 * PI / o->c.per_wave;
  const double wk = o->wave_number;
  const double ph = PI / 2.0 - segma * T;
  const double shd = sinh(wk * Dp);
  if (wm == 2) {
    const double wl = 2.0 * PI / wk;
    const double drift = GRAV * (A * A) / (wl / o->c.per_wave) / Dp / 8.0;
    const double fact = 1.0;
    const double th = o->c.theta_wave * PI / 180.0;
    for (int j = 1; j <= o->Nloc; ++j) {
      const double lag = o->yc[j - 1] * sin(th) * wk;
      o->Ein_X0[j - 1] = 0.5 * A * cos(ph + lag);
      o->Din_X0[j - 1] = o->Ein_X0[j - 1] + hfx[IX2(o->Ibeg, j)];
    }
    for (int k = o->Kbeg; k <= o->Kend; ++k)
      for (int j = o->Jbeg; j <= o->Jend; ++j) {
        const double lag = o->yc[j - 1] * sin(th) * wk;
        const double z = sigc[k] * o->Din_X0[j - 1];
        const double c = cos(ph + lag), sn = sin(ph + lag);
        o->Uin_X0[IB(j, k)] = (0.5 * A * segma * cosh(wk * z) / shd * c - drift * fact) * cos(th);
        o->Win_X0[IB(j, k)] = 0.5 * A * segma * sinh(wk * z) / shd * sn;
        o->Vin_X0[IB(j, k)] = (0.5 * A * segma * cosh(wk * z) / shd * c - drift * fact) * sin(th);
      }
    return;
  }
  /* wm == 3: second-order Stokes */
  const double chd = cosh(wk * Dp), ch2d = cosh(2.0 * wk * Dp);
  const double sh3 = shd * shd * shd, sh4 = sh3 * shd;
  for (int j = 1; j <= o->Nloc; ++j) {
    o->Ein_X0[j - 1] = 0.5 * A * cos(ph) + (A * A) * wk / 16.0 * chd / sh3 * (2.0 + ch2d) * cos(2.0 * ph);
    o->Din_X0[j - 1] = o->Ein_X0[j - 1] + hfx[IX2(o->Ibeg, j)];
  }
  for (int k = o->Kbeg; k <= o->Kend; ++k)
    for (int j = o->Jbeg; j <= o->Jend; ++j) {
      const double z = sigc[k] * o->Din_X0[j - 1];
      o->Uin_X0[IB(j, k)] = 0.5 * A * segma * cosh(wk * z) / shd * cos(ph) +
                            0.1875 * (A * A) * segma * wk * cosh(2.0 * wk * z) / sh4 * cos(2.0 * ph);
      o->Win_X0[IB(j, k)] = 0.5 * A * segma * sinh(wk * z) / shd * sin(ph) +
                            0.1875 * (A * A) * segma * wk * sinh(2.0 * wk * z) / sh4 * sin(2.0 * ph);
      o->Vin_X0[IB(j, k)] = ZERO;
    }
}

/* sponge_damping, src/nhwave.F:6462-6502 (every cell, ghosts included) */
void orc_sponge_damping(orc *o) {
  const double ramp = o->c.tramp > 0.0 ? tanh(o->time / o->c.tramp) : 1.0;
  const double ucur = ramp * o->c.cur_wave;
  for (int j = 1; j <= o->Nloc; ++j)
    for (int i = 1; i <= o->Mloc; ++i) {
      const long c = I2(i, j);
      if (o->Mask[c] != 1) continue;
      const double sp = o->Sponge[c];
      o->Eta[c] = o->Eta[c] * sp;
      o->D[c] = o->Eta[c] + o->Hc[c];
      for (int k = 1; k <= o->Kloc; ++k) {
        const long q = I3(i, j, k);
        o->U[q] = (o->U[q] - ucur) * sp + ucur;
        o->V[q] = o->V[q] * sp;
        o->DU[q] = o->D[c] * o->U[q];
        o->DV[q] = o->D[c] * o->V[q];
      }
    }
}

/* horizontal grid of generate_grid (serial), src/initialize.F:1666-1698:
 * face coordinates by repeated addition, centres half a cell in */
static void horizontal_grid(orc *o) {
  double *x = zd(o->Mloc1 + 1), *y = zd(o->Nloc1 + 1);  /* 1-based */
  double *xc = o->xc - 1, *yc = o->yc - 1;
  const double dx = o->c.dx, dy = o->c.dy;
  x[o->Ibeg] = ZERO;
  for (int i = o->Ibeg + 1; i <= o->Mloc1; ++i) {
    x[i] = x[i - 1] + dx;
    xc[i - 1] = x[i - 1] + 0.5 * dx;
  }
  for (int i = o->Ibeg - 1; i >= 1; --i) {
    x[i] = x[i + 1] - dx;
    xc[i] = x[i + 1] - 0.5 * dx;
  }
  y[o->Jbeg] = ZERO;
  for (int j = o->Jbeg + 1; j <= o->Nloc1; ++j) {
    y[j] = y[j - 1] + dy;
    yc[j - 1] = y[j - 1] + 0.5 * dy;
  }
  for (int j = o->Jbeg - 1; j >= 1; --j) {
    y[j] = y[j + 1] - dy;
    yc[j] = y[j + 1] - 0.5 * dy;
  }
  free(x);
  free(y);
}

/* calculate_sponge, src/nhwave.F:6405-6459 (Sponge = 1 elsewhere,
 * src/initialize.F:2665); a later side overrides an earlier one */
static void calculate_sponge(orc *o) {
  for (long q = 0; q < o->n2; ++q) o->Sponge[q] = 1.0;
  if (!o->c.sponge_on) return;
  const double W = o->c.sponge_w, E = o->c.sponge_e, S = o->c.sponge_s, N = o->c.sponge_n;
  const double xe = o->c.mglob * o->c.dx, yn = o->c.nglob * o->c.dy;
  if (W > ZERO)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int i = 1; i <= o->Mloc; ++i)
        if (o->xc[i - 1] <= W) {
          double r = (o->xc[i - 1] - W) / W;
          o->Sponge[I2(i, j)] = sqrt(1.0 - dmin(r * r, 1.0));
        }
  if (E > ZERO)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int i = 1; i <= o->Mloc; ++i)
        if (o->xc[i - 1] >= xe - E) {
          double r = (o->xc[i - 1] - (xe - E)) / E;
          o->Sponge[I2(i, j)] = sqrt(1.0 - dmin(r * r, 1.0));
        }
  if (S > ZERO)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int i = 1; i <= o->Mloc; ++i)
        if (o->yc[j - 1] <= S) {
          double r = (o->yc[j - 1] - S) / S;
          o->Sponge[I2(i, j)] = sqrt(1.0 - dmin(r * r, 1.0));
        }
  if (N > ZERO)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int i = 1; i <= o->Mloc; ++i)
        if (o->yc[j - 1] >= yn - N) {
          double r = (o->yc[j - 1] - (yn - N)) / N;
          o->Sponge[I2(i, j)] = sqrt(1.0 - dmin(r * r, 1.0));
        }
}

/* tail of `initial`, src/initialize.F:3118-3170, on Hc/Eta/U/V/W already
 * loaded (interior) with Hc ghosts filled by phi_2D_coll */
void orc_initial_tail(orc *o) {
  const double MinDep = o->c.min_dep;
  for (int j = 1; j <= o->Nloc; ++j)
    for (int i = 1; i <= o->Mloc; ++i) {
      long c = I2(i, j);
      if ((o->Eta[c] + o->Hc[c]) <= MinDep) {
        o->Mask[c] = 0;
        o->Eta[c] = MinDep - o->Hc[c];
      } else {
        o->Mask[c] = 1;
      }
    }
  for (long q = 0; q < o->n2; ++q) o->Mask[q] = o->Mask[q] * o->Mask_Struct[q];
  phi_2d_coll(o, o->Eta);
  memcpy(o->Eta0, o->Eta, sizeof(double) * o->n2);
  const int *M = o->Mask;
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i)
      o->Mask9[I2(i, j)] = M[I2(i, j)] * M[I2(i - 1, j)] * M[I2(i + 1, j)] * M[I2(i + 1, j + 1)] *
                           M[I2(i, j + 1)] * M[I2(i - 1, j + 1)] * M[I2(i + 1, j - 1)] *
                           M[I2(i, j - 1)] * M[I2(i - 1, j - 1)];
  for (long q = 0; q < o->n2; ++q) o->D[q] = dmax(o->Hc[q] + o->Eta[q], MinDep);
  /* this vel_bc (:3155) runs before the turbulence seed, on the viscosities
   * as allocated: CmuHt = CmuVt = Zero (:2667-2668) */
  if (o->c.viscous) {
    memset(o->CmuHt, 0, sizeof(double) * o->n3);
    memset(o->CmuVt, 0, sizeof(double) * o->n3);
  }
  orc_vel_bc(o);
  for (int k = 1; k <= o->Kloc; ++k)
    for (int j = 1; j <= o->Nloc; ++j)
      for (int i = 1; i <= o->Mloc; ++i) {
        long c = I3(i, j, k), c2 = I2(i, j);
        o->DU[c] = o->D[c2] * o->U[c] * o->Mask[c2];
        o->DV[c] = o->D[c2] * o->V[c] * o->Mask[c2];
        o->DW[c] = o->D[c2] * o->W[c] * o->Mask[c2];
      }
  /* turbulence seed, src/initialize.F:3185-3196 */
  if (o->c.viscous) {
    const double cmut_min = ke_cmut_min(o);
    for (int k = 1; k <= o->Kloc; ++k)
      for (int j = 1; j <= o->Nloc; ++j)
        for (int i = 1; i <= o->Mloc; ++i) {
          long c = I3(i, j, k), c2 = I2(i, j);
          o->Tke[c] = KE_TKE_MIN;
          o->Eps[c] = KE_EPS_MIN;
          o->CmuHt[c] = cmut_min;
          o->CmuVt[c] = cmut_min;
          o->DTke[c] = o->D[c2] * o->Tke[c] * o->Mask[c2];
          o->DEps[c] = o->D[c2] * o->Eps[c] * o->Mask[c2];
        }
  }
}

/* generate_grid (vertical part), src/initialize.F:1700-1727 */
static void generate_grid(orc *o, int ivgrd, double grd_r) {
  double *dsig = o->dsig - 1, *sig = o->sig - 1, *sigc = o->sigc - 1;
  const int Kb = o->Kbeg, Ke = o->Kend, Kglob = o->c.kglob;
  if (ivgrd == 1) {
    for (int k = 1; k <= o->Kloc; ++k) dsig[k] = (double)(1.0f / (float)Kglob);
  } else {
    dsig[Kb] = (grd_r - 1.0) / (pow(grd_r, (double)(float)Kglob) - 1.0);
    for (int k = Kb + 1; k <= Ke; ++k) dsig[k] = dsig[k - 1] * grd_r;
    for (int k = 1; k <= 2; ++k) dsig[Kb - k] = dsig[Kb + k - 1];
    for (int k = 1; k <= 2; ++k) dsig[Ke + k] = dsig[Ke - k + 1];
  }
  sig[Kb] = ZERO;
  for (int k = Kb + 1; k <= o->Kloc1; ++k) {
    sig[k] = sig[k - 1] + dsig[k - 1];
    sigc[k - 1] = sig[k - 1] + 0.5 * dsig[k - 1];
  }
  for (int k = Kb - 1; k >= 1; --k) {
    sig[k] = sig[k + 1] - dsig[k];
    sigc[k] = sig[k + 1] - 0.5 * dsig[k];
  }
}

/* generate_grid + read_bathymetry (CELL_CENTER) + initial (INITIAL_EUVW),
 * src/initialize.F:1647-1729, 1734-2175, 2502-3170.  hcg/etag are
 * (Mglob,Nglob), ug/vg/wg (Mglob,Nglob,Kglob) or NULL for rest. */
void orc_setup(orc *o, const double *hcg, const double *etag, const double *ug,
               const double *vg, const double *wg, int ivgrd, double grd_r,
               double wind_u, double wind_v) {
  const int Mg = o->c.mglob, Ng = o->c.nglob;
  generate_grid(o, ivgrd, grd_r);
  memset(o->Hc, 0, sizeof(double) * o->n2);
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i)
      o->Hc[I2(i, j)] = hcg[(long)(j - 3) * Mg + (i - 3)];
  phi_2d_coll(o, o->Hc);
  for (long q = 0; q < o->n2; ++q) o->Mask_Struct[q] = 1;
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i)
      if (o->Hc[I2(i, j)] < -1000.0) o->Mask_Struct[I2(i, j)] = 0;
  for (int j = 1; j <= o->Nloc; ++j) {
    for (int i = 2; i <= o->Mloc; ++i) o->Hfx[IX2(i, j)] = 0.5 * (o->Hc[I2(i - 1, j)] + o->Hc[I2(i, j)]);
    o->Hfx[IX2(1, j)] = o->Hc[I2(1, j)];
    o->Hfx[IX2(o->Mloc1, j)] = o->Hc[I2(o->Mloc, j)];
  }
  memcpy(o->Hfx0, o->Hfx, sizeof(double) * o->nx2);
  for (int i = 1; i <= o->Mloc; ++i) {
    for (int j = 2; j <= o->Nloc; ++j) o->Hfy[IY2(i, j)] = 0.5 * (o->Hc[I2(i, j - 1)] + o->Hc[I2(i, j)]);
    o->Hfy[IY2(i, 1)] = o->Hc[I2(i, 1)];
    o->Hfy[IY2(i, o->Nloc1)] = o->Hc[I2(i, o->Nloc)];
  }
  memcpy(o->Hfy0, o->Hfy, sizeof(double) * o->ny2);
  for (int j = 1; j <= o->Nloc; ++j)
    for (int i = 1; i <= o->Mloc; ++i) {
      o->DelxH[I2(i, j)] = (o->Hfx[IX2(i + 1, j)] - o->Hfx[IX2(i, j)]) / o->c.dx;
      o->DelyH[I2(i, j)] = (o->Hfy[IY2(i, j + 1)] - o->Hfy[IY2(i, j)]) / o->c.dy;
    }
  /* initial: everything zero, then the loaded fields (src/initialize.F:2991-3045) */
  o->time = 0.0; o->run_step = 0; o->dt = o->c.dt_ini;
  memset(o->Eta, 0, sizeof(double) * o->n2);
  memset(o->U, 0, sizeof(double) * o->n3);
  memset(o->V, 0, sizeof(double) * o->n3);
  memset(o->W, 0, sizeof(double) * o->n3);
  for (int j = o->Jbeg; j <= o->Jend; ++j)
    for (int i = o->Ibeg; i <= o->Iend; ++i) {
      long g = (long)(j - 3) * Mg + (i - 3);
      if (etag) o->Eta[I2(i, j)] = etag[g];
      for (int k = o->Kbeg; k <= o->Kend; ++k) {
        long g3 = g + (long)Mg * Ng * (k - 3);
        if (ug) o->U[I3(i, j, k)] = ug[g3];
        if (vg) o->V[I3(i, j, k)] = vg[g3];
        if (wg) o->W[I3(i, j, k)] = wg[g3];
      }
    }
  o->wind_u = wind_u; o->wind_v = wind_v;
  horizontal_grid(o);
  calculate_sponge(o);
  if (o->c.wavemaker == 2 || o->c.wavemaker == 3)
    o->wave_number = wave_dispersion(2.0 * PI / o->c.per_wave, o->c.dep_wave);
  orc_initial_tail(o);
}
