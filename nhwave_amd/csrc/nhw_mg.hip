// Matrix-free sigma-coordinate Poisson operator and a geometric multigrid
// preconditioner for it, templated on the vector type V (double for the FP64
// solve, float for the inner solve of the mixed form).
//
// The operator is the one generate_coef_rhs assembles (src/pressure.F:475-840):
// per row, 15 coefficients built from A1 = DelxS^2 + DelyS^2 + DelzS^2 and the
// face slopes DelxS/DelyS, with Neumann folds on the lateral and bottom
// boundaries and the Dirichlet surface.  Instead of streaming 15 stored
// diagonals (120 B/row) it recomputes them from five 2-D fields per column
// (ax = DelxH/D, bx = (DelxH+DelxEta)/D, ay, by, rD = 1/D, so that
// DelxS = ax - sig_k*bx) and per-level constants: the operator pass then moves
// about three vectors per row.  The coefficients agree with the assembled ones
// to rounding (products by precomputed reciprocals instead of quotients).
//
// Multigrid: cell-centred x-y coarsening by 2 (K kept), coarse operators by
// rediscretisation on block-averaged column fields, zebra z-line Gauss-Seidel
// smoothing (columns coloured by (i+j) parity; each column's vertical
// tridiagonal -- c13, c1, c6 -- solved exactly, which absorbs the strong
// vertical coupling of shallow and dry columns), block-average restriction
// and bilinear prolongation.  One V(1,1)-cycle from a zero guess is a fixed
// linear operator, used as the right preconditioner of BiCGStab.  The
// 15-point stencil has no (+-1,+-1) x-y couplings, so columns of one colour
// are independent and a half sweep is one parallel launch.
#include <algorithm>
#include <vector>

#include "nhw_mg_dev.h"

namespace nhw {

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline unsigned nb(long n) { return (unsigned)((n + kBSm - 1) / kBSm); }

template <class V>
static LvlC<V> make_lvl(const nhw_ctx *c, const MGLevel &lv) {
  LvlC<V> L{};
  const double dx = lv.dx, dy = lv.dy;
  L.M = lv.M; L.N = lv.N; L.K = c->g.K; L.MN = (long)lv.M * lv.N;
  L.cxy = (V)(2. / (dx * dx) + 2. / (dy * dy));
  L.idx2 = (V)(1. / (dx * dx));
  L.idy2 = (V)(1. / (dy * dy));
  L.kc = (const LvlK<V> *)lv.kc;
  return L;
}

// the sigma-level constants of one level (the rows' coefficients,
// src/pressure.F:540-555), kMgKMax entries (unused ones zero)
template <class V>
static std::vector<LvlK<V>> level_table(const nhw_ctx *c, const MGLevel &lv) {
  std::vector<LvlK<V>> t(kMgKMax, LvlK<V>{V(0), V(0), V(0), V(0), V(0)});
  const std::vector<double> &dsig = c->mg.dsig_h, &sig = c->mg.sig_h;  // 1-based
  for (int kk = 0; kk < c->g.K; ++kk) {
    const int k = c->g.Kbeg + kk;
    const double s = dsig[k] + dsig[k - 1];
    t[kk].sig = (V)sig[k];
    t[kk].r6 = (V)(1. / (0.5 * s * dsig[k]));
    t[kk].r13 = (V)(1. / (0.5 * s * dsig[k - 1]));
    t[kk].rx = (V)(1. / (2. * lv.dx * s));
    t[kk].ry = (V)(1. / (2. * lv.dy * s));
  }
  return t;
}

// uploads every level's table (after the vertical grid is set)
int mg_upload_consts(nhw_ctx *c) {
  for (int h = 0; h < 2; ++h) {
    std::vector<MGLevel> &lvs = h == 0 ? c->mg.lv : c->mg.lvf;
    const bool f32 = h == 1;
    for (const MGLevel &lv : lvs) {
      if (!lv.kc) continue;
      hipError_t e;
      if (f32) {
        const std::vector<LvlK<float>> t = level_table<float>(c, lv);
        e = hipMemcpy(lv.kc, t.data(), sizeof(LvlK<float>) * t.size(), hipMemcpyHostToDevice);
      } else {
        const std::vector<LvlK<double>> t = level_table<double>(c, lv);
        e = hipMemcpy(lv.kc, t.data(), sizeof(LvlK<double>) * t.size(), hipMemcpyHostToDevice);
      }
      if (hip_check(e, "multigrid level constants")) return NHW_EHIP;
    }
  }
  return 0;
}

template <class V>
LvlC<V> mg_level_consts(const nhw_ctx *c, const MGLevel &lv) {
  return make_lvl<V>(c, lv);
}
template LvlC<double> mg_level_consts<double>(const nhw_ctx *, const MGLevel &);
template LvlC<float> mg_level_consts<float>(const nhw_ctx *, const MGLevel &);

template <class V>
static G2<V> g2(const MGLevel &lv) {
  return G2<V>{(const V *)lv.ax, (const V *)lv.bx, (const V *)lv.ay, (const V *)lv.by, (const V *)lv.rd, lv.M + 2};
}

// Levels of one hierarchy; elements are `esize` bytes (8: FP64, 4: FP32).
// maxlev = 1 keeps only the fine-level geometry (the FP64 operator of the
// mixed form's outer residual).
int mg_alloc(nhw_ctx *c, std::vector<MGLevel> &lvs, int esize, int maxlev) {
  const Geo &g = c->g;
  if (g.K > kMgKMax) {
    set_error("mg_alloc: Kglob above the multigrid limit");
    return NHW_EINVAL;
  }
  int M = g.M, N = g.N;
  double dx = g.dx, dy = g.dy;
  while (true) {
    MGLevel lv;
    lv.M = M; lv.N = N; lv.dx = dx; lv.dy = dy;
    const long n = (long)M * N * g.K, n2h = (long)(M + 2) * (N + 2);
    const bool fine = lvs.empty();
    bool ok = true;
    for (double **p : {&lv.ax, &lv.bx, &lv.ay, &lv.by, &lv.rd})
      ok = ok && hipMalloc(p, esize * n2h) == hipSuccess && hipMemset(*p, 0, esize * n2h) == hipSuccess;
    const size_t kcb = (size_t)5 * esize * kMgKMax;
    ok = ok && hipMalloc(&lv.kc, kcb) == hipSuccess && hipMemset(lv.kc, 0, kcb) == hipSuccess;
    if (maxlev > 1 && !fine)  // level 0 solves into the caller's vectors
      for (double **p : {&lv.x, &lv.b})
        ok = ok && hipMalloc(p, esize * n) == hipSuccess && hipMemset(*p, 0, esize * n) == hipSuccess;
    lvs.push_back(lv);
    if (!ok) {
      set_error("mg_alloc: hipMalloc failed");
      return NHW_ENOMEM;
    }
    const int fx = M >= 4 ? 2 : 1, fy = N >= 4 ? 2 : 1;
    if ((int)lvs.size() >= maxlev || (fx == 1 && fy == 1) ||
        ((long)M * N <= kMgCoarseCols && std::max(M, N) <= kMgCoarseSide))
      break;
    lvs.back().fx = fx;
    lvs.back().fy = fy;
    M = (M + fx - 1) / fx;
    N = (N + fy - 1) / fy;
    dx *= fx;
    dy *= fy;
  }
  return 0;
}

void mg_free(std::vector<MGLevel> &lvs) {
  for (MGLevel &lv : lvs) {
    for (double *p : {lv.ax, lv.bx, lv.ay, lv.by, lv.rd, lv.x, lv.b, lv.r})
      if (p) (void)hipFree(p);
    if (lv.kc) (void)hipFree(lv.kc);
  }
  lvs.clear();
}

// column fields of every level from the current D, DelxH/Eta, DelyH/Eta
template <class V>
int mg_setup(nhw_ctx *c, std::vector<MGLevel> &lvs) {
  const Geo &g = c->g;
  hipStream_t st = c->stream;
  if (c->mg.dsig_h.size() != (size_t)g.Kloc + 1) {
    set_error("poisson_solver: vertical grid not set (nhw_set_grid)");
    return NHW_EINVAL;
  }
  MGLevel &l0 = lvs[0];
  {
    Timed t(c, KID_MG_SETUP, 8.0 * 5 * g.M * g.N + sizeof(V) * 5.0 * g.M * g.N);
    hipLaunchKernelGGL(k_geom0<V>, dim3(nb((long)g.M * g.N)), dim3(kBSm), 0, st, g, c->d, (V *)l0.ax, (V *)l0.bx,
                       (V *)l0.ay, (V *)l0.by, (V *)l0.rd, l0.M + 2);
  }
  for (size_t l = 0; l + 1 < lvs.size(); ++l) {
    const MGLevel &f = lvs[l];
    MGLevel &q = lvs[l + 1];
    Timed t(c, KID_MG_SETUP, sizeof(V) * 5.0 * ((double)f.M * f.N + (double)q.M * q.N));
    hipLaunchKernelGGL(k_geom_coarsen<V>, dim3(nb((long)q.M * q.N)), dim3(kBSm), 0, st, f.M, f.N, f.M + 2, q.M,
                       q.N, q.M + 2, f.fx, f.fy, (const V *)f.ax, (const V *)f.bx, (const V *)f.ay,
                       (const V *)f.by, (const V *)f.rd, (V *)q.ax, (V *)q.bx, (V *)q.ay, (V *)q.by, (V *)q.rd);
  }
  return hip_check(hipGetLastError(), "mg_setup");
}
template int mg_setup<double>(nhw_ctx *, std::vector<MGLevel> &);
template int mg_setup<float>(nhw_ctx *, std::vector<MGLevel> &);

// one zebra half sweep; algorithmic bytes: b and x of the swept columns, x
// of the other colour (unless zero), five column fields
template <class V, int KM>
static void zebra_k(nhw_ctx *c, const LvlC<V> &L, const G2<V> &G, const V *b, V *x, int color, bool zero,
                    const int *stop, int kid) {
  const long nth = (long)((L.M + 1) / 2) * L.N;
  const double n = (double)L.MN * L.K;
  Timed t(c, kid, sizeof(V) * ((zero ? 1.0 : 1.5) * n + 5.0 * L.MN));
  if (zero)
    hipLaunchKernelGGL((k_zebra<V, KM, true>), dim3(nb(nth)), dim3(kBSm), 0, c->stream, L, G, b, x, color, stop);
  else
    hipLaunchKernelGGL((k_zebra<V, KM, false>), dim3(nb(nth)), dim3(kBSm), 0, c->stream, L, G, b, x, color, stop);
}

template <class V>
static void smooth(nhw_ctx *c, const LvlC<V> &L, const G2<V> &G, const V *b, V *x, int color, bool zero,
                   const int *stop, int kid) {
  if (L.K <= 4) zebra_k<V, 4>(c, L, G, b, x, color, zero, stop, kid);
  else if (L.K <= 8) zebra_k<V, 8>(c, L, G, b, x, color, zero, stop, kid);
  else if (L.K <= 12) zebra_k<V, 12>(c, L, G, b, x, color, zero, stop, kid);
  else zebra_k<V, 16>(c, L, G, b, x, color, zero, stop, kid);
}

// the coarsest level: coarse_sweeps zebra pairs in one single-workgroup
// launch; algorithmic bytes: b read and x written per half sweep pair
template <class V>
static void coarse(nhw_ctx *c, const LvlC<V> &L, const G2<V> &G, const V *b, V *x, const int *stop) {
  const int sw = c->mg.coarse_sweeps;
  if (L.MN > kMgCoarseCols) {  // coarsening stopped early: separate sweeps
    for (int s = 0; s < sw; ++s) {
      smooth(c, L, G, b, x, 0, s == 0, stop, KID_MG_COARSE);
      smooth(c, L, G, b, x, 1, false, stop, KID_MG_COARSE);
    }
    return;
  }
  Timed t(c, KID_MG_COARSE, sizeof(V) * (2.0 * L.MN * L.K + 5.0 * L.MN));
  if (L.K <= 4)
    hipLaunchKernelGGL((k_coarse<V, 4>), dim3(1), dim3(kCoarseBS), 0, c->stream, L, G, b, x, sw, stop);
  else if (L.K <= 8)
    hipLaunchKernelGGL((k_coarse<V, 8>), dim3(1), dim3(kCoarseBS), 0, c->stream, L, G, b, x, sw, stop);
  else if (L.K <= 12)
    hipLaunchKernelGGL((k_coarse<V, 12>), dim3(1), dim3(kCoarseBS), 0, c->stream, L, G, b, x, sw, stop);
  else
    hipLaunchKernelGGL((k_coarse<V, 16>), dim3(1), dim3(kCoarseBS), 0, c->stream, L, G, b, x, sw, stop);
}

// the 2x2 restriction on lane pairs (k_resid22_pair), with ZSWEEP also the
// next level's first half sweep
template <class V, int KM>
static void resid_pair_k(nhw_ctx *c, const LvlC<V> &L, const G2<V> &G, const V *x, const V *b,
                         const LvlC<V> &Lc, const G2<V> &Gc, V *bc, V *xc, bool zsweep, const int *stop) {
  const dim3 grid(nb(2 * Lc.MN));
  if (zsweep)
    hipLaunchKernelGGL((k_resid22_pair<V, KM, true>), grid, dim3(kBSm), 0, c->stream, L, G, x, b, Lc, Gc, bc, xc,
                       stop);
  else
    hipLaunchKernelGGL((k_resid22_pair<V, KM, false>), grid, dim3(kBSm), 0, c->stream, L, G, x, b, Lc, Gc, bc, xc,
                       stop);
}

// x = V(b): one V(1,1) cycle from a zero guess on level l; the coarsest
// level (at most kMgCoarseCols columns) gets coarse_sweeps zebra sweeps.
// With pre_done the first (zero-guess) half sweep has been done by the
// restriction from the level above.
template <class V>
static void vcycle(nhw_ctx *c, const std::vector<MGLevel> &lvs, size_t l, const V *b, V *x, const int *stop,
                   bool pre_done) {
  const MGLevel &lv = lvs[l];
  const LvlC<V> L = make_lvl<V>(c, lv);
  const G2<V> G = g2<V>(lv);
  hipStream_t st = c->stream;
  const double n = (double)L.MN * L.K;
  if (l + 1 == lvs.size()) {
    coarse(c, L, G, b, x, stop);
    return;
  }
  const int kid = KID_MG_SMOOTH;  // one kernel at every level above the coarsest
  if (!pre_done) smooth(c, L, G, b, x, 0, true, stop, kid);
  smooth(c, L, G, b, x, 1, false, stop, kid);
  const MGLevel &q = lvs[l + 1];
  const double nc = (double)q.M * q.N * L.K;
  // the 2x2 path on lane pairs starts the next level's V-cycle unless that
  // level is the coarsest (solved by k_coarse).  Up to kMgPairCols fine
  // columns, where the launches are short, the pairs halve each thread's
  // walk and the fused half sweep saves a launch; above it (measured at
  // 1024 x 1024 x 10: 53 us against 38 + 10 us) the single-lane restriction
  // and a separate half sweep move the bytes faster.
  const bool even_only = lv.fx == 2 && lv.fy == 2;
  const bool pair = mg_pair_restrict(lv);
  const bool zsweep = pair && l + 2 < lvs.size();
  {
    // algorithmic bytes: x, b (the even half only on the 2x2 path), the coarse
    // right side, five fields; with the coarse half sweep also its x and
    // five coarse fields
    Timed t(c, KID_MG_RESID, sizeof(V) * ((even_only ? 1.5 : 2.0) * n + nc + 5.0 * L.MN +
                                          (zsweep ? 0.5 * nc + 5.0 * q.M * q.N : 0.0)));
    const dim3 grid(nb((long)q.M * q.N));
    if (pair) {
      const LvlC<V> Lc = make_lvl<V>(c, q);
      const G2<V> Gc = g2<V>(q);
      if (L.K <= 4)
        resid_pair_k<V, 4>(c, L, G, (const V *)x, b, Lc, Gc, (V *)q.b, (V *)q.x, zsweep, stop);
      else if (L.K <= 8)
        resid_pair_k<V, 8>(c, L, G, (const V *)x, b, Lc, Gc, (V *)q.b, (V *)q.x, zsweep, stop);
      else if (L.K <= 12)
        resid_pair_k<V, 12>(c, L, G, (const V *)x, b, Lc, Gc, (V *)q.b, (V *)q.x, zsweep, stop);
      else
        resid_pair_k<V, 16>(c, L, G, (const V *)x, b, Lc, Gc, (V *)q.b, (V *)q.x, zsweep, stop);
    } else if (even_only) {
      const dim3 g22(nb((long)q.M * q.N));
      if (L.K <= 4)
        hipLaunchKernelGGL((k_resid_restrict22<V, 4>), g22, dim3(kBSm), 0, st, L, G, (const V *)x, b, q.M, q.N,
                           (V *)q.b, stop);
      else if (L.K <= 8)
        hipLaunchKernelGGL((k_resid_restrict22<V, 8>), g22, dim3(kBSm), 0, st, L, G, (const V *)x, b, q.M, q.N,
                           (V *)q.b, stop);
      else if (L.K <= 12)
        hipLaunchKernelGGL((k_resid_restrict22<V, 12>), g22, dim3(kBSm), 0, st, L, G, (const V *)x, b, q.M, q.N,
                           (V *)q.b, stop);
      else
        hipLaunchKernelGGL((k_resid_restrict22<V, 16>), g22, dim3(kBSm), 0, st, L, G, (const V *)x, b, q.M, q.N,
                           (V *)q.b, stop);
    } else if (L.K <= 4)
      hipLaunchKernelGGL((k_resid_restrict<V, 4>), grid, dim3(kBSm), 0, st, L, G, (const V *)x, b, q.M, q.N, lv.fx,
                         lv.fy, (V *)q.b, stop);
    else if (L.K <= 8)
      hipLaunchKernelGGL((k_resid_restrict<V, 8>), grid, dim3(kBSm), 0, st, L, G, (const V *)x, b, q.M, q.N, lv.fx,
                         lv.fy, (V *)q.b, stop);
    else
      hipLaunchKernelGGL((k_resid_restrict<V, 16>), grid, dim3(kBSm), 0, st, L, G, (const V *)x, b, q.M, q.N,
                         lv.fx, lv.fy, (V *)q.b, stop);
  }
  vcycle<V>(c, lvs, l + 1, (const V *)q.b, (V *)q.x, stop, zsweep);
  {
    Timed t(c, KID_MG_PROLONG, sizeof(V) * (n + nc));
    hipLaunchKernelGGL(k_prolong_add<V>, dim3(nb((long)((lv.M + 1) / 2) * lv.N)), dim3(kBSm), 0, st, lv.M, lv.N,
                       q.M, q.N, L.K, lv.fx, lv.fy, (const V *)q.x, x, stop);
  }
  smooth(c, L, G, b, x, 1, false, stop, kid);
  smooth(c, L, G, b, x, 0, false, stop, kid);
}

template <class V>
void mg_apply(nhw_ctx *c, const std::vector<MGLevel> &lvs, const V *b, V *x, const int *stop) {
  vcycle<V>(c, lvs, 0, b, x, stop, false);
}
template void mg_apply<double>(nhw_ctx *, const std::vector<MGLevel> &, const double *, double *, const int *);
template void mg_apply<float>(nhw_ctx *, const std::vector<MGLevel> &, const float *, float *, const int *);

// y = A x on the fine level, with ndot dot partials (y,d1) [and (y,y)]
template <class V>
void mf_apply(nhw_ctx *c, const std::vector<MGLevel> &lvs, const V *x, V *y, const V *d1, int ndot, double *part,
              const int *stop, int nblk, int kid) {
  const MGLevel &l0 = lvs[0];
  const LvlC<V> L = make_lvl<V>(c, l0);
  const G2<V> G = g2<V>(l0);
  const double n = (double)L.MN * L.K;
  Timed t(c, kid, sizeof(V) * 3.0 * n + sizeof(V) * 5.0 * L.MN);
  if (ndot == 0)
    hipLaunchKernelGGL((k_mf_apply<V, 0>), dim3(nblk), dim3(kBSm), 0, c->stream, L, G, x, y, d1, part, stop);
  else if (ndot == 1)
    hipLaunchKernelGGL((k_mf_apply<V, 1>), dim3(nblk), dim3(kBSm), 0, c->stream, L, G, x, y, d1, part, stop);
  else
    hipLaunchKernelGGL((k_mf_apply<V, 2>), dim3(nblk), dim3(kBSm), 0, c->stream, L, G, x, y, d1, part, stop);
}

template void mf_apply<double>(nhw_ctx *, const std::vector<MGLevel> &, const double *, double *, const double *,
                               int, double *, const int *, int, int);
template void mf_apply<float>(nhw_ctx *, const std::vector<MGLevel> &, const float *, float *, const float *, int,
                              double *, const int *, int, int);

// FP64 residual of x against b into the Krylov vectors of type R (see
// k_mf_init); geometry from the FP64 level 0 of `lvs`
template <class R>
void mf_init(nhw_ctx *c, const std::vector<MGLevel> &lvs, const double *x, const double *b, R *r, R *rh, R *p, R *v,
             R *e, double *part, int nblk) {
  const MGLevel &l0 = lvs[0];
  const LvlC<double> L = make_lvl<double>(c, l0);
  const double n = (double)L.MN * L.K;
  Timed t(c, KID_MF_RESID, 16.0 * n + sizeof(R) * (e ? 5.0 : 4.0) * n + 40.0 * L.MN);
  hipLaunchKernelGGL(k_mf_init<R>, dim3(nblk), dim3(kBSm), 0, c->stream, L, g2<double>(l0), x, b, r, rh, p, v, e,
                     part);
}
template void mf_init<double>(nhw_ctx *, const std::vector<MGLevel> &, const double *, const double *, double *,
                              double *, double *, double *, double *, double *, int);
template void mf_init<float>(nhw_ctx *, const std::vector<MGLevel> &, const double *, const double *, float *,
                             float *, float *, float *, float *, double *, int);

void mf_true_resid(nhw_ctx *c, const std::vector<MGLevel> &lvs, const double *x, const double *b, double *part,
                   int nblk) {
  const MGLevel &l0 = lvs[0];
  const LvlC<double> L = make_lvl<double>(c, l0);
  Timed t(c, KID_MF_RESID, 16.0 * L.MN * L.K + 40.0 * L.MN);
  hipLaunchKernelGGL(k_mf_true, dim3(nblk), dim3(kBSm), 0, c->stream, L, g2<double>(l0), x, b, part);
}

}  // namespace nhw
