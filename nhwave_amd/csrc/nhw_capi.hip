// C ABI of libnhwave_amd: context lifetime, host<->device transfers of the
// reference's GLOBAL arrays, and the hot-path subroutines as entry points.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "nhw_impl.h"

namespace nhw {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return NHW_EHIP;
}

int launch_p_from_x(nhw_ctx *c, const double *x);

void field_table_into(nhw_ctx *c, const nhw_fields *h, Xfer *out) {
  const Geo &g = c->g;
  Dev &d = c->d;
  const size_t b2 = sizeof(double) * g.n2(), bx2 = sizeof(double) * g.nx2(), by2 = sizeof(double) * g.ny2();
  const size_t b3 = sizeof(double) * g.n3(), bz3 = sizeof(double) * g.nz3();
  const size_t bx3 = sizeof(double) * g.nx3(), by3 = sizeof(double) * g.ny3(), bi2 = sizeof(int) * g.n2();
  const Xfer t[] = {
      {h->Eta, d.Eta, b2},      {h->Eta0, d.Eta0, b2},     {h->D, d.D, b2},           {h->D0, d.D0, b2},
      {h->Hc, d.Hc, b2},        {h->DelxH, d.DelxH, b2},   {h->DelyH, d.DelyH, b2},   {h->DelxEta, d.DelxEta, b2},
      {h->DelyEta, d.DelyEta, b2}, {h->Mask, d.Mask, bi2}, {h->Mask_Struct, d.Mask_Struct, bi2},
      {h->Mask9, d.Mask9, bi2}, {h->Hfx, d.Hfx, bx2},      {h->Hfx0, d.Hfx0, bx2},    {h->EtaxL, d.EtaxL, bx2},
      {h->EtaxR, d.EtaxR, bx2}, {h->Hfy, d.Hfy, by2},      {h->Hfy0, d.Hfy0, by2},    {h->EtayL, d.EtayL, by2},
      {h->EtayR, d.EtayR, by2}, {h->U, d.U, b3, true},     {h->V, d.V, b3, true},     {h->W, d.W, b3, true},
      {h->DU, d.DU, b3, true},  {h->DV, d.DV, b3, true},   {h->DW, d.DW, b3, true},   {h->DU0, d.DU0, b3, true},
      {h->DV0, d.DV0, b3, true}, {h->DW0, d.DW0, b3, true}, {h->P, d.P, bz3},         {h->Omega, d.Omega, bz3, true},
      {h->SourceX, d.SourceX, b2}, {h->SourceY, d.SourceY, b2}, {h->Wsx, d.Wsx, b2}, {h->Wsy, d.Wsy, b2},
      {h->Ex, d.Ex, bx3, true}, {h->Ey, d.Ey, by3, true},  {h->Fx, d.Fx, bx3, true},  {h->Fy, d.Fy, by3, true},
      {h->Gx, d.Gx, bx3, true}, {h->Gy, d.Gy, by3, true},  {h->Hx, d.Hx, bx3, true},  {h->Hy, d.Hy, by3, true},
      {h->Fz, d.Fz, bz3, true}, {h->Gz, d.Gz, bz3, true},  {h->Hz, d.Hz, bz3, true},
      // k-epsilon state: absent (device null, 0 bytes) without the model
      {h->CmuVt, d.CmuVt, d.CmuVt ? b3 : 0}, {h->Tke, d.Tke, d.Tke ? b3 : 0}, {h->Eps, d.Eps, d.Eps ? b3 : 0},
      {h->DTke, d.DTke, d.DTke ? b3 : 0},    {h->DEps, d.DEps, d.DEps ? b3 : 0},
  };
  static_assert(sizeof t / sizeof t[0] == kFieldSlots, "field table out of step with nhw_fields");
  for (size_t q = 0; q < kFieldSlots; ++q) out[q] = t[q];
}

std::vector<Xfer> field_table(nhw_ctx *c, const nhw_fields *h) {
  std::vector<Xfer> t(kFieldSlots);
  field_table_into(c, h, t.data());
  return t;
}

int check_present(const Xfer *t, size_t n, const char *who) {
  for (size_t q = 0; q < n; ++q)
    if (t[q].host && !t[q].dev) {
      set_error(std::string(who) + ": an array was requested that this context does not hold (the k-epsilon "
                                   "state exists only with the k-epsilon model)");
      return NHW_EINVAL;
    }
  return 0;
}

static int validate(const nhw_config *c) {
  char buf[256];
  auto bad = [&](const char *m) {
    set_error(m);
    return NHW_EINVAL;
  };
  if (c->mglob < 2 || c->nglob < 1 || c->kglob < 2) {
    snprintf(buf, sizeof buf, "nhw_create: grid %d x %d x %d unsupported (Mglob, Kglob >= 2; Nglob >= 1)",
             c->mglob, c->nglob, c->kglob);
    return bad(buf);
  }
  if (!(c->dx > 0) || !(c->dy > 0)) return bad("nhw_create: DX and DY must be positive");
  if (c->time_order < 1 || c->time_order > 3) return bad("nhw_create: TIME_ORDER must be 1, 2 or 3");
  if (c->hllc) return bad("nhw_create: HLLC = T is not supported (HLL only)");
  if (c->wavemaker < NHW_WAVE_NONE || c->wavemaker > NHW_WAVE_FLUX_LR)
    return bad("nhw_create: WAVEMAKER unsupported (NONE, LEF_SOL, LEF_LIN, LEF_STK, LEF_CON, INT_IRR, FLUX_LR)");
  const bool flux_lr = c->wavemaker == NHW_WAVE_FLUX_LR;
  const int lat[4] = {c->bc_x0, c->bc_xn, c->bc_y0, c->bc_yn};
  for (int q = 0; q < 4; ++q)
    if (lat[q] != 1 && lat[q] != 2 && !(lat[q] == 5 && c->viscous) &&
        !(q == 0 && lat[q] == 3 && (west_wavemaker(c->wavemaker) || flux_lr)) &&
        !(q == 1 && lat[q] == 3 && flux_lr)) {
      snprintf(buf, sizeof buf,
               "nhw_create: lateral BC type %d unsupported (1 free-slip, 2 no-slip, 5 wall friction in viscous "
               "runs; 3 inflow at X0 with a west wavemaker, open at X0 and Xn with FLUX_LR)", lat[q]);
      return bad(buf);
    }
  if (flux_lr && (c->bc_x0 != 3 || c->bc_xn != 3)) return bad("nhw_create: FLUX_LR needs BC_X0 = BC_Xn = 3");
  if (flux_lr && c->precision != NHW_FP64) return bad("nhw_create: FLUX_LR runs in FP64 only");
  if (flux_lr && !(c->zob > 0)) return bad("nhw_create: FLUX_LR needs Zob > 0 (its logarithmic inflow profile)");
  if (c->wavemaker == NHW_WAVE_INT_IRR && !(c->dep_wave > 0))
    return bad("nhw_create: the internal wavemaker needs DEP > 0");
  if (west_wavemaker(c->wavemaker)) {
    if (!(c->dep_wave > 0)) return bad("nhw_create: the wavemaker needs DEP > 0");
    if ((c->wavemaker == NHW_WAVE_LEF_SOL || c->wavemaker == NHW_WAVE_LEF_CON) && !(c->amp_wave > 0))
      return bad("nhw_create: the solitary and cnoidal wavemakers need AMP > 0");
    if (c->wavemaker != NHW_WAVE_LEF_SOL && !(c->per_wave > 0))
      return bad("nhw_create: the periodic wavemakers need PER > 0");
  }
  if (c->sponge_on && (c->sponge_w < 0 || c->sponge_e < 0 || c->sponge_s < 0 || c->sponge_n < 0))
    return bad("nhw_create: sponge widths must be >= 0");
  if (c->bc_z0 != 1 && c->bc_z0 != 2 && c->bc_z0 != 5)
    return bad("nhw_create: BC_Z0 must be 1 (free slip), 2 (no slip) or 5 (friction law)");
  if (c->bc_z0 == 5 && (c->ibot == 1 ? !(c->cd0 >= 0) : !(c->zob > 0)))
    return bad("nhw_create: BC_Z0 = 5 needs Cd0 >= 0 (Ibot = 1) or Zob > 0");
  const bool wall5 = c->bc_x0 == 5 || c->bc_xn == 5 || c->bc_y0 == 5 || c->bc_yn == 5;
  if (wall5 && (c->ibot == 1 ? !(c->cd0 >= 0) : !(c->zob > 0)))
    return bad("nhw_create: wall friction side boundaries (BC 5) need Cd0 >= 0 (Ibot = 1) or Zob > 0");
  if (wall5 && !(c->visc > 0)) return bad("nhw_create: wall friction side boundaries (BC 5) need VISCOSITY > 0");
  if ((c->bc_y0 == 5 || c->bc_yn == 5) && c->nglob < 2)
    return bad("nhw_create: wall friction at Y0/Yn (BC 5) needs Nglob >= 2");
  if (c->precision != NHW_FP64 && c->precision != NHW_MIXED) return bad("nhw_create: unknown precision");
  if (c->solver != NHW_SOLVER_MG && c->solver != NHW_SOLVER_LINE) return bad("nhw_create: unknown solver");
  if (c->precision == NHW_MIXED && c->non_hydro && c->solver != NHW_SOLVER_MG)
    return bad("nhw_create: the mixed form runs the multigrid solver only");
  if (c->non_hydro && c->solver == NHW_SOLVER_MG && c->kglob > kMgKMax)
    return bad("nhw_create: the multigrid solver takes Kglob <= 16; use NHW_SOLVER_LINE");
  if (c->non_hydro && (c->itmax < 1 || !(c->tol > 0))) return bad("nhw_create: ITMAX >= 1 and TOL > 0 required");
  if (c->viscous) {
    if (c->kglob > kViscKMax) return bad("nhw_create: VISCOUS_FLOW takes Kglob <= 32");
    if (!(c->schmidt != 0.0)) return bad("nhw_create: VISCOUS_FLOW needs Schmidt != 0");
    if (c->turb != 0 && c->turb != 1 && c->turb != 10)
      return bad("nhw_create: turbulence model unsupported (1 constant viscosities, 10 k-epsilon)");
    if (c->turb == 10 && c->kglob < 3) return bad("nhw_create: the k-epsilon model takes Kglob >= 3");
  }
  return 0;
}

// the k-epsilon model is on (IVTURB = 10 with IHTURB >= 10)
static inline bool kepsilon(const nhw_config *c) { return c->viscous && c->turb == 10; }

}  // namespace nhw

using namespace nhw;

extern "C" {

const char *nhw_last_error(void) { return g_err.c_str(); }

int nhw_create(const nhw_config *cfg, nhw_ctx **out) {
  if (!cfg || !out) {
    set_error("nhw_create: null argument");
    return NHW_EINVAL;
  }
  *out = nullptr;
  int rc = validate(cfg);
  if (rc) return rc;
  nhw_ctx *c = new nhw_ctx();
  c->cfg = *cfg;
  c->mixed = cfg->precision == NHW_MIXED;
  Geo &g = c->g;
  g.M = cfg->mglob; g.N = cfg->nglob; g.K = cfg->kglob;
  g.Mloc = g.M + 4; g.Nloc = g.N + 4; g.Kloc = g.K + 4;
  g.Mloc1 = g.Mloc + 1; g.Nloc1 = g.Nloc + 1; g.Kloc1 = g.Kloc + 1;
  g.Ibeg = 3; g.Iend = g.Mloc - 2; g.Iend1 = g.Mloc1 - 2;
  g.Jbeg = 3; g.Jend = g.Nloc - 2; g.Jend1 = g.Nloc1 - 2;
  g.Kbeg = 3; g.Kend = g.Kloc - 2; g.Kend1 = g.Kloc1 - 2;
  g.dx = cfg->dx; g.dy = cfg->dy;
  Par &p = c->p;
  p.min_dep = cfg->min_dep; p.cfl = cfg->cfl; p.dt_min = cfg->dt_min; p.dt_max = cfg->dt_max;
  p.froude_cap = cfg->froude_cap; p.visc = cfg->visc; p.fcor = cfg->fcor;
  p.wind_u = cfg->wind_u; p.wind_v = cfg->wind_v;
  p.bc_x0 = cfg->bc_x0; p.bc_xn = cfg->bc_xn; p.bc_y0 = cfg->bc_y0; p.bc_yn = cfg->bc_yn; p.bc_z0 = cfg->bc_z0;
  p.froude_on = cfg->froude_cap_on; p.non_hydro = cfg->non_hydro; p.it_order = cfg->time_order;
  p.viscous = cfg->viscous; p.schmidt = cfg->schmidt; p.visc_number = cfg->visc_number;
  p.turb = kepsilon(cfg) ? 10 : 1;
  p.rng_on = cfg->rng;
  p.ibot = cfg->ibot; p.cd0 = cfg->cd0; p.zob = cfg->zob;
  // SSP-RK coefficients, src/initialize.F:3382-3400 (1/3 and 2/3 are REAL(4))
  if (cfg->time_order == 3) {
    p.alpha[0] = 0.0; p.alpha[1] = 0.75; p.alpha[2] = (double)(1.0f / 3.0f);
    p.beta[0] = 1.0; p.beta[1] = 0.25; p.beta[2] = (double)(2.0f / 3.0f);
  } else if (cfg->time_order == 2) {
    p.alpha[0] = 0.0; p.alpha[1] = 0.5; p.alpha[2] = 0.0;
    p.beta[0] = 1.0; p.beta[1] = 0.5; p.beta[2] = 0.0;
  } else {
    p.alpha[0] = 0.0; p.beta[0] = 1.0;
    p.alpha[1] = p.alpha[2] = p.beta[1] = p.beta[2] = 0.0;
  }
  if (hip_check(hipGetDevice(&c->device), "hipGetDevice") ||
      hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate") ||
      hip_check(hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault), "hipHostMalloc")) {
    nhw_destroy(c);
    return NHW_EHIP;
  }
  Dev &d = c->d;
  bool ok = true;
  auto ad = [&](double **ptr, long n) {
    if (!ok) return;
    if (hipMalloc(ptr, sizeof(double) * std::max<long>(n, 1)) != hipSuccess ||
        hipMemset(*ptr, 0, sizeof(double) * std::max<long>(n, 1)) != hipSuccess)
      ok = false;
  };
  const size_t r3 = c->mixed ? sizeof(float) : sizeof(double);
  auto ar = [&](double **ptr, long n) {  // 3-D field in the storage type
    if (!ok) return;
    if (hipMalloc(ptr, r3 * std::max<long>(n, 1)) != hipSuccess || hipMemset(*ptr, 0, r3 * std::max<long>(n, 1)) != hipSuccess)
      ok = false;
  };
  auto ai = [&](int **ptr, long n) {
    if (!ok) return;
    if (hipMalloc(ptr, sizeof(int) * n) != hipSuccess || hipMemset(*ptr, 0, sizeof(int) * n) != hipSuccess) ok = false;
  };
  double *dsig, *sig, *sigc;
  ad(&dsig, g.Kloc); ad(&sig, g.Kloc1); ad(&sigc, g.Kloc);
  d.dsig = dsig; d.sig = sig; d.sigc = sigc;
  for (double **f : {&d.Hc, &d.DelxH, &d.DelyH, &d.Eta, &d.Eta0, &d.D, &d.D0, &d.DelxEta, &d.DelyEta, &d.SourceX,
                     &d.SourceY, &d.Wsx, &d.Wsy, &d.R1, &d.ztU, &d.ztV, &d.ztW, &d.zbU, &d.zbV, &d.zbW})
    ad(f, g.n2());
  ai(&d.Mask, g.n2()); ai(&d.MaskNew, g.n2()); ai(&d.Mask_Struct, g.n2()); ai(&d.Mask9, g.n2());
  for (double **f : {&d.Hfx, &d.Hfx0, &d.EtaxL, &d.EtaxR}) ad(f, g.nx2());
  for (double **f : {&d.Hfy, &d.Hfy0, &d.EtayL, &d.EtayR}) ad(f, g.ny2());
  for (double **f : {&d.U, &d.V, &d.W, &d.DU, &d.DV, &d.DW, &d.DU0, &d.DV0, &d.DW0}) ar(f, g.n3());
  for (double **f : {&d.Ex, &d.Fx, &d.Gx, &d.Hx}) ar(f, g.nx3());
  for (double **f : {&d.Ey, &d.Fy, &d.Gy, &d.Hy}) ar(f, g.ny3());
  for (double **f : {&d.Fz, &d.Gz, &d.Hz, &d.Omega}) ar(f, g.nz3());
  ad(&d.P, g.nz3());
  ad(&d.scal, 8);
  ai(&d.iscal, 8);
  ad(&d.red, 4096);
  if (!ok) {
    set_error("nhw_create: device allocation failed");
    nhw_destroy(c);
    return NHW_ENOMEM;
  }
  rc = forcing_setup(c);
  if (!rc) rc = viscous_setup(c);
  if (!rc) rc = turb_seed(c);
  if (rc) {
    nhw_destroy(c);
    return rc;
  }
  if (cfg->non_hydro) {
    rc = poisson_alloc(c);
    if (rc) {
      nhw_destroy(c);
      return rc;
    }
  }
  rc = nhw_set_time(c, 0.0, cfg->dt_ini, 0);
  if (rc) {
    nhw_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

// Device bytes nhw_create allocates for a configuration (same arithmetic as
// the allocations in nhw_create, poisson_alloc and mg_alloc), without
// allocating: the admission test of a serving loop.
int nhw_footprint(const nhw_config *cfg, int64_t *bytes) {
  if (!cfg || !bytes) {
    set_error("nhw_footprint: null argument");
    return NHW_EINVAL;
  }
  int rc = validate(cfg);
  if (rc) return rc;
  const long M = cfg->mglob, N = cfg->nglob, K = cfg->kglob;
  const long Ml = M + 4, Nl = N + 4, Kl = K + 4;
  const double n2 = (double)Ml * Nl, nx2 = (double)(Ml + 1) * Nl, ny2 = (double)Ml * (Nl + 1);
  const double n3 = n2 * Kl, nx3 = nx2 * Kl, ny3 = ny2 * Kl, nz3 = n2 * (Kl + 1);
  const bool mixed = cfg->precision == NHW_MIXED;
  const double r3 = mixed ? 4.0 : 8.0;
  double b = 8.0 * (3 * Kl + 1) + 8.0 * 20 * n2 + 4.0 * 4 * n2 + 8.0 * 4 * (nx2 + ny2);
  b += r3 * (9 * n3 + 4 * nx3 + 4 * ny3 + 4 * nz3) + 8.0 * nz3 + 8.0 * (8 + 4096) + 4.0 * 8;
  if (west_wavemaker(cfg->wavemaker) || cfg->wavemaker == NHW_WAVE_FLUX_LR) b += 8.0 * (3.0 * Nl + 3.0 * Nl * Kl);
  if (cfg->wavemaker == NHW_WAVE_FLUX_LR) b += 8.0 * (2.0 * Nl + 5.0 * Nl * Kl + n2 + 6.0 * n3);
  if (cfg->wavemaker == NHW_WAVE_INT_IRR) b += 8.0 * n2;  // SourceC (the spectrum table is extra)
  if (cfg->sponge_on) b += 8.0 * n2;
  if (cfg->viscous) b += 8.0 * (8 * n3 + 2 * nz3);
  if (kepsilon(cfg)) b += 8.0 * 5 * n3;  // CmuVt (= CmuHt), Tke, Eps, DTke, DEps
  if (cfg->non_hydro) {
    const double n = (double)M * N * K;
    const double pad = (double)(((long)M * N + M + 1 + 63) / 64 * 64);
    b += 8.0 * (n + (n + 2 * pad) + 3 * 8192 + 16) + 16;
    if (cfg->solver == NHW_SOLVER_LINE) b += 8.0 * 17 * n;
    if (mixed) b += 4.0 * 9 * n;
    else b += 8.0 * (6 * n + 2 * (n + 2 * pad));
    if (cfg->solver == NHW_SOLVER_MG) {
      // hierarchy levels: 5 column fields each, and x, b below level 0
      auto levels = [&](double es, int maxlev) {
        double t = 0.0;
        long m = M, nn = N;
        for (int l = 0;; ++l) {
          t += es * 5 * (double)(m + 2) * (nn + 2) + es * 5 * kMgKMax;
          if (maxlev > 1 && l > 0) t += es * 2 * (double)m * nn * K;
          const int fx = m >= 4 ? 2 : 1, fy = nn >= 4 ? 2 : 1;
          if (l + 1 >= maxlev || (fx == 1 && fy == 1) || (m * nn <= kMgCoarseCols && std::max(m, nn) <= kMgCoarseSide))
            break;
          m = (m + fx - 1) / fx;
          nn = (nn + fy - 1) / fy;
        }
        return t;
      };
      b += mixed ? levels(8.0, 1) + levels(4.0, 64) : levels(8.0, 64);
    }
  }
  *bytes = (int64_t)b;
  return 0;
}

int nhw_destroy(nhw_ctx *c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  snap_free(c);
  forcing_free(c);
  viscous_free(c);
  Dev &d = c->d;
  void *ptrs[] = {(void *)d.dsig, (void *)d.sig, (void *)d.sigc, d.Hc, d.DelxH, d.DelyH, d.Eta, d.Eta0, d.D,
                  d.D0, d.DelxEta, d.DelyEta, d.SourceX, d.SourceY, d.Wsx, d.Wsy, d.R1, d.ztU, d.ztV, d.ztW,
                  d.zbU, d.zbV, d.zbW, d.Mask, d.MaskNew, d.Mask_Struct, d.Mask9, d.Hfx, d.Hfx0, d.EtaxL,
                  d.EtaxR, d.Hfy, d.Hfy0, d.EtayL, d.EtayR, d.U, d.V, d.W, d.DU, d.DV, d.DW, d.DU0, d.DV0,
                  d.DW0, d.Ex, d.Fx, d.Gx, d.Hx, d.Ey, d.Fy, d.Gy, d.Hy, d.Fz, d.Gz, d.Hz, d.Omega, d.P,
                  d.scal, d.iscal, d.red};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  poisson_free(c);
  for (auto &p : c->kt.pending) {
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  for (hipEvent_t e : c->kt.pool) (void)hipEventDestroy(e);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int nhw_set_grid(nhw_ctx *c, const double *dsig, const double *sig, const double *sigc) {
  if (!c || !dsig || !sig || !sigc) {
    set_error("nhw_set_grid: null argument");
    return NHW_EINVAL;
  }
  const Geo &g = c->g;
  int rc = hip_check(hipMemcpy((void *)c->d.dsig, dsig, sizeof(double) * g.Kloc, hipMemcpyHostToDevice), "set_grid");
  rc |= hip_check(hipMemcpy((void *)c->d.sig, sig, sizeof(double) * g.Kloc1, hipMemcpyHostToDevice), "set_grid");
  rc |= hip_check(hipMemcpy((void *)c->d.sigc, sigc, sizeof(double) * g.Kloc, hipMemcpyHostToDevice), "set_grid");
  if (rc) return NHW_EHIP;
  // 1-based host copies for the multigrid level constants
  c->mg.dsig_h.assign(g.Kloc + 1, 0.0);
  c->mg.sig_h.assign(g.Kloc1 + 1, 0.0);
  std::copy(dsig, dsig + g.Kloc, c->mg.dsig_h.begin() + 1);
  std::copy(sig, sig + g.Kloc1, c->mg.sig_h.begin() + 1);
  if ((rc = mg_upload_consts(c))) return rc;
  c->grid_set = true;
  return 0;
}

int nhw_upload(nhw_ctx *c, const nhw_fields *h) {
  if (!c || !h) {
    set_error("nhw_upload: null argument");
    return NHW_EINVAL;
  }
  if (hip_check(hipStreamSynchronize(c->stream), "nhw_upload")) return NHW_EHIP;
  const std::vector<Xfer> tab = field_table(c, h);
  if (int rc = check_present(tab.data(), tab.size(), "nhw_upload")) return rc;
  std::vector<float> tmp;
  for (const Xfer &x : tab) {
    if (!x.host) continue;
    if (x.r3 && c->mixed) {  // FP64 host array -> FP32 device storage
      const size_t n = x.bytes / sizeof(double);
      tmp.resize(n);
      const double *src = (const double *)x.host;
      for (size_t q = 0; q < n; ++q) tmp[q] = (float)src[q];
      if (hip_check(hipMemcpy(x.dev, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice), "nhw_upload"))
        return NHW_EHIP;
    } else if (hip_check(hipMemcpy(x.dev, x.host, x.bytes, hipMemcpyHostToDevice), "nhw_upload")) {
      return NHW_EHIP;
    }
  }
  return 0;
}

int nhw_download(nhw_ctx *c, const nhw_fields *h) {
  if (!c || !h) {
    set_error("nhw_download: null argument");
    return NHW_EINVAL;
  }
  if (hip_check(hipStreamSynchronize(c->stream), "nhw_download")) return NHW_EHIP;
  const std::vector<Xfer> tab = field_table(c, h);
  if (int rc = check_present(tab.data(), tab.size(), "nhw_download")) return rc;
  std::vector<float> tmp;
  for (const Xfer &x : tab) {
    if (!x.host) continue;
    if (x.r3 && c->mixed) {
      const size_t n = x.bytes / sizeof(double);
      tmp.resize(n);
      if (hip_check(hipMemcpy(tmp.data(), x.dev, n * sizeof(float), hipMemcpyDeviceToHost), "nhw_download"))
        return NHW_EHIP;
      double *dst = (double *)x.host;
      for (size_t q = 0; q < n; ++q) dst[q] = (double)tmp[q];
    } else if (hip_check(hipMemcpy(x.host, x.dev, x.bytes, hipMemcpyDeviceToHost), "nhw_download")) {
      return NHW_EHIP;
    }
  }
  return 0;
}

int nhw_set_time(nhw_ctx *c, double time, double dt, int32_t run_step) {
  if (!c) {
    set_error("nhw_set_time: null context");
    return NHW_EINVAL;
  }
  double s[4] = {dt, time, dt, 0.0};
  int is[2] = {run_step, 0};
  int rc = hip_check(hipMemcpyAsync(c->d.scal, s, sizeof s, hipMemcpyHostToDevice, c->stream), "set_time");
  rc |= hip_check(hipMemcpyAsync(c->d.iscal, is, sizeof is, hipMemcpyHostToDevice, c->stream), "set_time");
  rc |= hip_check(hipStreamSynchronize(c->stream), "set_time");
  return rc ? NHW_EHIP : 0;
}

int nhw_get_time(nhw_ctx *c, double *time, double *dt, int32_t *run_step) {
  if (!c) {
    set_error("nhw_get_time: null context");
    return NHW_EINVAL;
  }
  double *hs = (double *)((char *)c->pinned + 1024);
  int *hi = (int *)((char *)c->pinned + 1024 + 64);
  int rc = hip_check(hipMemcpyAsync(hs, c->d.scal, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream), "get_time");
  rc |= hip_check(hipMemcpyAsync(hi, c->d.iscal, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream), "get_time");
  rc |= hip_check(hipStreamSynchronize(c->stream), "get_time");
  if (rc) return NHW_EHIP;
  if (dt) *dt = hs[0];
  if (time) *time = hs[1];
  if (run_step) *run_step = hi[0];
  return 0;
}

static int need_grid(nhw_ctx *c, const char *who) {
  if (!c) {
    set_error(std::string(who) + ": null context");
    return NHW_EINVAL;
  }
  if (!c->grid_set) {
    set_error(std::string(who) + ": nhw_set_grid has not been called");
    return NHW_EINVAL;
  }
  return 0;
}

int nhw_initial(nhw_ctx *c) {
  int rc = need_grid(c, "nhw_initial");
  if (rc) return rc;
  // initial's vel_bc (src/initialize.F:3155) reads the viscosities as
  // allocated (Zero, :2667-2668); its turbulence seed (:3172-3196) follows,
  // the depth-integrated k and epsilon once D and Mask exist
  if ((rc = turb_zero(c))) return rc;
  if ((rc = launch_initial(c))) return rc;
  if ((rc = turb_seed(c))) return rc;
  return turb_seed_depth(c);
}

static int check_flags(nhw_ctx *c) {
  int *hi = (int *)((char *)c->pinned + 2048);
  if (hip_check(hipMemcpyAsync(hi, c->d.iscal, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream), "flags") ||
      hip_check(hipStreamSynchronize(c->stream), "flags"))
    return NHW_EHIP;
  ktimer_flush(c);
  if (hi[1] & 1) {
    set_error("estimate_dt: time step too small (dt < DT_MIN)");
    return NHW_EDT;
  }
  return 0;
}

int nhw_estimate_dt(nhw_ctx *c) {
  int rc = need_grid(c, "nhw_estimate_dt");
  if (rc) return rc;
  rc = launch_estimate_dt(c);
  return rc ? rc : check_flags(c);
}

int nhw_update_mask(nhw_ctx *c) {
  int rc = need_grid(c, "nhw_update_mask");
  return rc ? rc : launch_update_mask(c);
}

int nhw_update_wind(nhw_ctx *c) {
  int rc = need_grid(c, "nhw_update_wind");
  return rc ? rc : launch_update_wind(c);
}

int nhw_update_vars(nhw_ctx *c) {
  int rc = need_grid(c, "nhw_update_vars");
  return rc ? rc : launch_update_vars(c);
}

int nhw_source_terms(nhw_ctx *c) {
  int rc = need_grid(c, "nhw_source_terms");
  return rc ? rc : launch_source_terms(c);
}

int nhw_fluxes(nhw_ctx *c) {
  int rc = need_grid(c, "nhw_fluxes");
  return rc ? rc : launch_fluxes(c);
}

int nhw_poisson_solver(nhw_ctx *c) {
  int rc = need_grid(c, "nhw_poisson_solver");
  if (rc) return rc;
  if (!c->cfg.non_hydro) {
    set_error("nhw_poisson_solver: context was created with NON_HYDRO = F");
    return NHW_EINVAL;
  }
  return poisson_solve(c);
}

int nhw_generate_coef_rhs(nhw_ctx *c, double *coef, double *rhs) {
  int rc = need_grid(c, "nhw_generate_coef_rhs");
  if (rc) return rc;
  if (!coef || !rhs) {
    set_error("nhw_generate_coef_rhs: null argument");
    return NHW_EINVAL;
  }
  return poisson_system(c, coef, rhs);
}

int nhw_poisson_apply(nhw_ctx *c, const double *x, double *y) {
  int rc = need_grid(c, "nhw_poisson_apply");
  if (rc) return rc;
  if (!x || !y) {
    set_error("nhw_poisson_apply: null argument");
    return NHW_EINVAL;
  }
  return poisson_apply(c, x, y);
}

int nhw_eval_duvw(nhw_ctx *c, int32_t istep) {
  int rc = need_grid(c, "nhw_eval_duvw");
  if (rc) return rc;
  if (istep < 1 || istep > c->cfg.time_order) {
    set_error("nhw_eval_duvw: ISTEP out of range");
    return NHW_EINVAL;
  }
  if ((rc = launch_diffusion(c))) return rc;
  if ((rc = launch_eval_explicit(c, istep))) return rc;
  if ((rc = launch_visc_solve(c))) return rc;
  if (c->cfg.non_hydro) {
    if ((rc = launch_get_uvw(c))) return rc;
    if ((rc = poisson_solve(c))) return rc;
    if ((rc = launch_projection(c))) return rc;
  }
  if ((rc = launch_rk_blend(c, istep))) return rc;
  if ((rc = launch_get_uvw(c))) return rc;
  return launch_get_omega(c);
}

int nhw_step(nhw_ctx *c, int32_t nsteps) {
  int rc = need_grid(c, "nhw_step");
  if (rc) return rc;
  for (int s = 0; s < nsteps; ++s) {
    if ((rc = launch_estimate_dt(c))) return rc;
    if ((rc = launch_wave_bc(c))) return rc;
    if ((rc = launch_update_mask(c))) return rc;
    if ((rc = launch_update_wind(c))) return rc;
    if ((rc = launch_update_vars(c))) return rc;
    for (int st = 1; st <= c->cfg.time_order; ++st) {
      if ((rc = launch_source_terms(c))) return rc;
      if ((rc = launch_fluxes(c))) return rc;
      if ((rc = nhw_eval_duvw(c, st))) return rc;
      if ((rc = launch_sponge(c))) return rc;
      if ((rc = eval_turb(c))) return rc;
    }
    if ((rc = check_flags(c))) return rc;
  }
  return 0;
}

int nhw_eval_turb(nhw_ctx *c) {
  if (!c) {
    set_error("nhw_eval_turb: null argument");
    return NHW_EINVAL;
  }
  return eval_turb(c);
}

int nhw_run(nhw_ctx *c, double total_time, int32_t sim_steps, int32_t *steps_taken) {
  int rc = need_grid(c, "nhw_run");
  if (rc) return rc;
  if (steps_taken) *steps_taken = 0;
  double t;
  int32_t rs;
  if ((rc = nhw_get_time(c, &t, nullptr, &rs))) return rc;
  // nhw_step already waits for every step (its dt check); the TIME/RUN_STEP
  // copy rides on the same idle stream
  while (t < total_time && rs < sim_steps) {
    if ((rc = nhw_step(c, 1))) return rc;
    if ((rc = nhw_get_time(c, &t, nullptr, &rs))) return rc;
    if (steps_taken) ++*steps_taken;
  }
  return 0;
}

int nhw_solver_stats(nhw_ctx *c, nhw_stats *out) {
  if (!c || !out) {
    set_error("nhw_solver_stats: null argument");
    return NHW_EINVAL;
  }
  *out = c->stats;
  return 0;
}

int nhw_sync(nhw_ctx *c) {
  if (!c) {
    set_error("nhw_sync: null context");
    return NHW_EINVAL;
  }
  if (hip_check(hipStreamSynchronize(c->stream), "nhw_sync")) return NHW_EHIP;
  ktimer_flush(c);
  return 0;
}

int nhw_kernel_timing(nhw_ctx *c, int32_t enable) {
  if (!c) {
    set_error("nhw_kernel_timing: null context");
    return NHW_EINVAL;
  }
  if (hip_check(hipStreamSynchronize(c->stream), "nhw_kernel_timing")) return NHW_EHIP;
  ktimer_flush(c);
  for (int q = 0; q < KID_COUNT; ++q) {
    c->kt.total_ms[q] = 0.0;
    c->kt.total_bytes[q] = 0.0;
    c->kt.launches[q] = 0;
  }
  c->kt.on = enable != 0;
  return 0;
}

int nhw_kernel_stats(nhw_ctx *c, int32_t id, const char **name, int64_t *launches, double *total_ms,
                     double *bytes_per_launch) {
  if (!c) {
    set_error("nhw_kernel_stats: null context");
    return NHW_EINVAL;
  }
  if (id < 0 || id >= KID_COUNT) return NHW_EINVAL;
  if (hip_check(hipStreamSynchronize(c->stream), "nhw_kernel_stats")) return NHW_EHIP;
  ktimer_flush(c);
  if (name) *name = kernel_name(id);
  if (launches) *launches = c->kt.launches[id];
  if (total_ms) *total_ms = c->kt.total_ms[id];
  if (bytes_per_launch)
    *bytes_per_launch = c->kt.launches[id] ? c->kt.total_bytes[id] / c->kt.launches[id] : kernel_bytes(c, id);
  return 0;
}

}  // extern "C"
