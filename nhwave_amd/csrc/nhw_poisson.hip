// Non-hydrostatic pressure on gfx950: assembly of the 15-coefficient
// sigma-coordinate Poisson matrix (generate_coef_rhs, src/pressure.F:475-840)
// into a device-resident diagonal (SoA) layout, and its solve by a
// right-preconditioned BiCGStab with a vertical-line (tridiagonal, one Thomas
// solve per water column) preconditioner.  The reference solves the same
// system with NSPCG IC(0)+truncated GMRES (src/pressure.F:86-131).
//
// Reductions are deterministic: a fixed grid of blocks, fixed per-thread
// strides, a fixed-shape tree per block and a single-block final sum, so a
// solve gives bit-identical results run to run.
#include "nhw_batch.h"
#include "nhw_krylov_dev.h"
#include "nhw_mg_dev.h"

namespace nhw {

constexpr int kRedBlocks = 8192;

struct Off15 {
  long o[15];
};

// 15-point stencil, JCoef order (src/pressure.F:817-831), as row offsets.
static Off15 stencil(const Geo &g) {
  const int di[15] = {0, 1, 0, 0, -1, 0, 1, 0, -1, 0, 0, 1, 0, -1, 0};
  const int dj[15] = {0, 0, 1, -1, 0, 0, 0, 1, 0, -1, 1, 0, 0, 0, -1};
  const int dk[15] = {0, 0, 0, 1, 1, 1, 1, 1, 0, 0, -1, -1, -1, -1, -1};
  Off15 s;
  for (int q = 0; q < 15; ++q) s.o[q] = di[q] + (long)g.M * dj[q] + (long)g.M * g.N * dk[q];
  return s;
}

// ----------------------------------------------------------------------------
// generate_coef_rhs, src/pressure.F:475-840.  One thread per interior cell;
// DelxS/DelyS/DelzS/A1 at the cell's sigma face and Uf/Vf
// (interpolate_velocity_to_faces, src/nhwave.F:1346-1356) are formed in place.
// ----------------------------------------------------------------------------
template <class DV>
__device__ __forceinline__ double delxs(const Geo &g, const DV &d, int i, int j, double sk) {
  if (i < g.Ibeg || i > g.Iend || j < g.Jbeg || j > g.Jend) return 0.0;  // DelxS = Zero outside
  const long c = g.c2(i, j);
  return (1. - sk) / d.D[c] * d.DelxH[c] - sk / d.D[c] * d.DelxEta[c];
}
template <class DV>
__device__ __forceinline__ double delys(const Geo &g, const DV &d, int i, int j, double sk) {
  if (i < g.Ibeg || i > g.Iend || j < g.Jbeg || j > g.Jend) return 0.0;
  const long c = g.c2(i, j);
  return (1. - sk) / d.D[c] * d.DelyH[c] - sk / d.D[c] * d.DelyEta[c];
}

template <bool COEF, class R>
__device__ __forceinline__ void assemble_body(const Geo &g, const DevT<R> &d, double *__restrict__ coef,
                                              double *__restrict__ rhs, long r) {
  const long n = g.neq();
  if (r >= n) return;
  const int i = g.Ibeg + (int)(r % g.M);
  const long rr = r / g.M;
  const int j = g.Jbeg + (int)(rr % g.N), k = g.Kbeg + (int)(rr / g.N);
  const double dx = g.dx, dy = g.dy, dt = d.scal[0];
  const double *dsig = d.dsig - 1, *sig = d.sig - 1;
  const double sk = sig[k];
  const long c = g.c2(i, j);
  const double Dd = d.D[c];
  const double dxs = delxs(g, d, i, j, sk), dys = delys(g, d, i, j, sk), dzs = 1. / Dd;
  const double A1 = dxs * dxs + dys * dys + dzs * dzs;
  const double s = dsig[k] + dsig[k - 1];
  if (COEF) {
  const double dxs_m = delxs(g, d, i - 1, j, sk), dxs_p = delxs(g, d, i + 1, j, sk);
  const double dys_m = delys(g, d, i, j - 1, sk), dys_p = delys(g, d, i, j + 1, sk);
  double cf[15];
  cf[0] = (2. / (dx * dx) + 2. / (dy * dy) + A1 / (0.5 * s * dsig[k]) + A1 / (0.5 * s * dsig[k - 1]));
  cf[1] = -1. / (dx * dx);
  cf[2] = -1. / (dy * dy);
  cf[3] = (dys_m / (2. * dy * s) + dys / (2. * dy * s));
  cf[4] = (dxs_m / (2. * dx * s) + dxs / (2. * dx * s));
  cf[5] = -A1 / (0.5 * s * dsig[k]);
  cf[6] = -(dxs_p / (2. * dx * s) + dxs / (2. * dx * s));
  cf[7] = -(dys_p / (2. * dy * s) + dys / (2. * dy * s));
  cf[8] = -1. / (dx * dx);
  cf[9] = -1. / (dy * dy);
  cf[10] = (dys_p / (2. * dy * s) + dys / (2. * dy * s));
  cf[11] = (dxs_p / (2. * dx * s) + dxs / (2. * dx * s));
  cf[12] = -A1 / (0.5 * s * dsig[k - 1]);
  cf[13] = -(dxs_m / (2. * dx * s) + dxs / (2. * dx * s));
  cf[14] = -(dys_m / (2. * dy * s) + dys / (2. * dy * s));
  // boundary folds in the reference's order (west, east, south, north,
  // bottom, top), src/pressure.F:649-814 (1-based diagonal numbers)
#define C(q) cf[(q)-1]
  if (i == g.Ibeg) { C(1) = C(1) + C(9); C(6) = C(6) + C(5); C(13) = C(13) + C(14); C(9) = 0.0; C(5) = 0.0; C(14) = 0.0; }
  if (i == g.Iend) { C(1) = C(1) + C(2); C(6) = C(6) + C(7); C(13) = C(13) + C(12); C(2) = 0.0; C(7) = 0.0; C(12) = 0.0; }
  if (j == g.Jbeg) { C(1) = C(1) + C(10); C(6) = C(6) + C(4); C(13) = C(13) + C(15); C(10) = 0.0; C(4) = 0.0; C(15) = 0.0; }
  if (j == g.Jend) { C(1) = C(1) + C(3); C(6) = C(6) + C(8); C(13) = C(13) + C(11); C(3) = 0.0; C(8) = 0.0; C(11) = 0.0; }
  if (k == g.Kbeg) {
    C(6) = C(6) + C(13); C(7) = C(7) + C(12); C(8) = C(8) + C(11); C(5) = C(5) + C(14); C(4) = C(4) + C(15);
    C(13) = 0.0; C(12) = 0.0; C(11) = 0.0; C(14) = 0.0; C(15) = 0.0;
  }
  if (k == g.Kend) { C(4) = 0.0; C(5) = 0.0; C(6) = 0.0; C(7) = 0.0; C(8) = 0.0; }
#undef C
  for (int q = 0; q < 15; ++q) coef[(long)q * n + r] = cf[q];
  }
  // right-hand side, src/pressure.F:517-535
  const long sk3 = (long)g.Mloc * g.Nloc, sj = g.Mloc;
  const long q = g.c3(i, j, k);
  const double f = dsig[k] / (dsig[k] + dsig[k - 1]);
  const R *U = d.U, *V = d.V, *W = d.W;
  const double ufp = (1.0 - f) * U[q + 1] + f * U[q + 1 - sk3];
  const double ufm = (1.0 - f) * U[q - 1] + f * U[q - 1 - sk3];
  const double vfp = (1.0 - f) * V[q + sj] + f * V[q + sj - sk3];
  const double vfm = (1.0 - f) * V[q - sj] + f * V[q - sj - sk3];
  const double hs = 0.5 * (dsig[k] + dsig[k - 1]);
  // the multigrid solver (rhs only) keeps its vectors in the zebra layout
  const long ro = COEF ? r : (rr / g.N) * ((long)g.M * g.N) + zcol(i - g.Ibeg, j - g.Jbeg, g.M);
  rhs[ro] = -((ufp - ufm) / (2.0 * dx) + (U[q] - U[q - sk3]) / hs * dxs + (vfp - vfm) / (2.0 * dy) +
             (V[q] - V[q - sk3]) / hs * dys + (W[q] - W[q - sk3]) / hs * dzs - (d.SourceC ? d.SourceC[c] : 0.0)) *
           kRho0 / dt;
}

template <bool COEF, class R>
__global__ __launch_bounds__(256) void k_assemble(Geo g, DevT<R> d, double *__restrict__ coef,
                                                  double *__restrict__ rhs) {
  assemble_body<COEF, R>(g, d, coef, rhs, (long)xcd_block() * blockDim.x + threadIdx.x);
}

// ----------------------------------------------------------------------------
// Krylov kernels
// ----------------------------------------------------------------------------
// Krylov state in the vector type V; scalars, partial sums and flags are
// FP64/int whatever V is.  `inner` marks the mixed form's FP32 inner solve,
// which stops on |r| <= itol*|r0| instead of |r| <= tol*|b|.
template <class V>
struct KVT {
  long n, M, MN;
  int K;
  const double *coef;
  const double *b;
  V *x, *r, *rh, *p, *v, *s, *t, *ph, *sh;
  double *lu_c, *lu_d;
  double *part, *sc;
  int *flag;
  int itmax;
  double tol2, itol2;
  int inner;
  __host__ __device__ KScal scal() const { return KScal{sc, flag, itmax, tol2, itol2, inner}; }
};
using KV = KVT<double>;

__device__ __forceinline__ double spmv_row(const double *__restrict__ coef, const double *__restrict__ x,
                                           long n, long r, const Off15 &o) {
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 15; ++q) acc += coef[(long)q * n + r] * x[r + o.o[q]];
  return acc;
}

// r = b - A x, rh = r, p = v = 0; partials (r,r), (b,b)
__global__ __launch_bounds__(kBS) void k_resid0(KV k, Off15 o) {
  double acc[2] = {0.0, 0.0};
  for (long r = (long)blockIdx.x * kBS + threadIdx.x; r < k.n; r += (long)gridDim.x * kBS) {
    const double bv = k.b[r];
    const double rv = bv - spmv_row(k.coef, k.x, k.n, r, o);
    k.r[r] = rv; k.rh[r] = rv; k.p[r] = 0.0; k.v[r] = 0.0;
    acc[0] += rv * rv;
    acc[1] += bv * bv;
  }
  block_reduce<2>(acc, k.part);
}

template <class V>
__global__ void k_fin0(KVT<V> k, int nblk) {
  const double rr = final_sum(k.part, 0, 0, nblk);
  const double bb = final_sum(k.part, 1, 0, nblk);
  if (threadIdx.x == 0) fin0_logic(k.scal(), rr, bb);
}

// Thomas factors of the vertical tridiagonal (c13, c1, c6) per column
__global__ void k_line_factor(KV k) {
  long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= k.MN) return;
  const double *c1 = k.coef, *c6 = k.coef + 5 * k.n, *c13 = k.coef + 12 * k.n;
  double cp = 0.0;
  for (int kk = 0; kk < k.K; ++kk) {
    const long r = col + kk * k.MN;
    const double den = kk == 0 ? c1[r] : c1[r] - c13[r] * cp;
    const double inv = 1.0 / den;
    cp = c6[r] * inv;
    k.lu_d[r] = inv;
    k.lu_c[r] = cp;
  }
}

// z = M^-1 y on one column (y already stored in z)
__device__ __forceinline__ void line_solve(const KV &k, long col, double *z) {
  const double *c13 = k.coef + 12 * k.n;
  double gp = 0.0;
  for (int kk = 0; kk < k.K; ++kk) {
    const long r = col + kk * k.MN;
    const double gv = (kk == 0 ? z[r] : z[r] - c13[r] * gp) * k.lu_d[r];
    z[r] = gv;
    gp = gv;
  }
  double zn = gp;
  for (int kk = k.K - 2; kk >= 0; --kk) {
    const long r = col + kk * k.MN;
    zn = z[r] - k.lu_c[r] * zn;
    z[r] = zn;
  }
}

// p = r + beta*(p - omega*v); ph = M^-1 p
__global__ void k_prec_p(KV k) {
  if (k.flag[0]) return;
  long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= k.MN) return;
  const double beta = k.sc[S_BETA], omega = k.sc[S_OMEGA];
  for (int kk = 0; kk < k.K; ++kk) {
    const long r = col + kk * k.MN;
    const double pv = k.r[r] + beta * (k.p[r] - omega * k.v[r]);
    k.p[r] = pv;
    k.ph[r] = pv;
  }
  line_solve(k, col, k.ph);
}

// v = A ph; partial (rh, v)
__global__ __launch_bounds__(kBS) void k_spmv_v(KV k, Off15 o) {
  if (k.flag[0]) return;
  double acc[1] = {0.0};
  for (long r = (long)blockIdx.x * kBS + threadIdx.x; r < k.n; r += (long)gridDim.x * kBS) {
    const double vv = spmv_row(k.coef, k.ph, k.n, r, o);
    k.v[r] = vv;
    acc[0] += k.rh[r] * vv;
  }
  block_reduce<1>(acc, k.part);
}

template <class V>
__global__ void k_fin1(KVT<V> k, int nblk) {
  if (k.flag[0]) return;
  const double rhv = final_sum(k.part, 0, 0, nblk);
  if (threadIdx.x == 0) fin1_logic(k.scal(), rhv);
}

// s = r - alpha*v; sh = M^-1 s
__global__ void k_prec_s(KV k) {
  if (k.flag[0]) return;
  long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= k.MN) return;
  const double alpha = k.sc[S_ALPHA];
  for (int kk = 0; kk < k.K; ++kk) {
    const long r = col + kk * k.MN;
    const double sv = k.r[r] - alpha * k.v[r];
    k.s[r] = sv;
    k.sh[r] = sv;
  }
  line_solve(k, col, k.sh);
}

// t = A sh; partials (t,s), (t,t)
__global__ __launch_bounds__(kBS) void k_spmv_t(KV k, Off15 o) {
  if (k.flag[0]) return;
  double acc[2] = {0.0, 0.0};
  for (long r = (long)blockIdx.x * kBS + threadIdx.x; r < k.n; r += (long)gridDim.x * kBS) {
    const double tv = spmv_row(k.coef, k.sh, k.n, r, o);
    k.t[r] = tv;
    acc[0] += tv * k.s[r];
    acc[1] += tv * tv;
  }
  block_reduce<2>(acc, k.part);
}

template <class V>
__global__ void k_fin2(KVT<V> k, int nblk) {
  if (k.flag[0]) return;
  const double ts = final_sum(k.part, 0, 0, nblk);
  const double tt = final_sum(k.part, 1, 0, nblk);
  if (threadIdx.x == 0) fin2_logic(k.scal(), ts, tt);
}

// x += alpha*ph + omega*sh; r = s - omega*t; partials (rh,r), (r,r)
template <class V>
__global__ __launch_bounds__(kBS) void k_update(KVT<V> k) {
  if (k.flag[0]) return;
  const V alpha = (V)k.sc[S_ALPHA], omega = (V)k.sc[S_OMEGA];
  double acc[2] = {0.0, 0.0};
  for (long r = (long)blockIdx.x * kBS + threadIdx.x; r < k.n; r += (long)gridDim.x * kBS) {
    k.x[r] = k.x[r] + alpha * k.ph[r] + omega * k.sh[r];
    const V rv = k.s[r] - omega * k.t[r];
    k.r[r] = rv;
    acc[0] += (double)k.rh[r] * (double)rv;
    acc[1] += (double)rv * (double)rv;
  }
  block_reduce<2>(acc, k.part);
}

template <class V>
__global__ void k_fin3(KVT<V> k, int nblk) {
  if (k.flag[0]) return;
  const double rhr = final_sum(k.part, 0, 0, nblk);
  const double rr = final_sum(k.part, 1, 0, nblk);
  if (threadIdx.x == 0) fin3_logic(k.scal(), rhr, rr);
}

// p = r + beta*(p - omega*v)  (multigrid path: the preconditioner is a
// separate V-cycle)
template <class V>
__global__ __launch_bounds__(kBS) void k_pvec(KVT<V> k) {
  if (k.flag[0]) return;
  const V beta = (V)k.sc[S_BETA], omega = (V)k.sc[S_OMEGA];
  for (long r = (long)blockIdx.x * kBS + threadIdx.x; r < k.n; r += (long)gridDim.x * kBS)
    k.p[r] = k.r[r] + beta * (k.p[r] - omega * k.v[r]);
}
// s = r - alpha*v
template <class V>
__global__ __launch_bounds__(kBS) void k_svec(KVT<V> k) {
  if (k.flag[0]) return;
  const V alpha = (V)k.sc[S_ALPHA];
  for (long r = (long)blockIdx.x * kBS + threadIdx.x; r < k.n; r += (long)gridDim.x * kBS)
    k.s[r] = k.r[r] - alpha * k.v[r];
}

// true residual |b - A x|^2 for the solve report
__global__ __launch_bounds__(kBS) void k_true_res(KV k, Off15 o) {
  double acc[1] = {0.0};
  for (long r = (long)blockIdx.x * kBS + threadIdx.x; r < k.n; r += (long)gridDim.x * kBS) {
    const double rv = k.b[r] - spmv_row(k.coef, k.x, k.n, r, o);
    acc[0] += rv * rv;
  }
  block_reduce<1>(acc, k.part);
}
__global__ void k_fin_true(KV k, int nblk) {
  const double rr = final_sum(k.part, 0, 0, nblk);
  if (threadIdx.x == 0) k.sc[S_TRUE] = rr;
}

// initial guess x = P (interior), src/pressure.F:105-113; zero when the
// right-hand side vanishes.  zl: x in the zebra layout (multigrid solver)
__global__ void k_x_from_p(Geo g, Dev d, double *__restrict__ x, int zl) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.neq()) return;
  const int i = g.Ibeg + (int)(t % g.M);
  const long r = t / g.M;
  const int j = g.Jbeg + (int)(r % g.N), k = g.Kbeg + (int)(r / g.N);
  const long o = zl ? (r / g.N) * ((long)g.M * g.N) + zcol(i - g.Ibeg, j - g.Jbeg, g.M) : t;
  x[o] = d.P[g.c3(i, j, k)];
}
// x += e (the mixed form's FP32 correction, added in FP64)
__global__ void k_add_correction(double *__restrict__ x, const float *__restrict__ e, long n) {
  for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
    x[r] += (double)e[r];
}
__global__ void k_zero_if(double *__restrict__ x, long n, const int *flag, int val) {
  if (flag[0] != val) return;
  for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) x[r] = 0.0;
}

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------
int poisson_alloc(nhw_ctx *c) {
  Krylov &k = c->kr;
  const Geo &g = c->g;
  k.n = g.neq();
  k.pad = ((long)g.M * g.N + g.M + 1 + 63) / 64 * 64;
  const long nv = k.n + 2 * k.pad;
  auto al = [&](double **p, long cnt) {
    if (hipMalloc(p, sizeof(double) * cnt) != hipSuccess) return false;
    return hipMemset(*p, 0, sizeof(double) * cnt) == hipSuccess;
  };
  auto alf = [&](float **p, long cnt) {
    if (hipMalloc(p, sizeof(float) * cnt) != hipSuccess) return false;
    return hipMemset(*p, 0, sizeof(float) * cnt) == hipSuccess;
  };
  k.nblk = (int)std::min<long>(kRedBlocks, (k.n + kBS - 1) / kBS);
  const bool line = c->cfg.solver == NHW_SOLVER_LINE;
  bool ok = al(&k.rhs, k.n) && al(&k.x, nv) && al(&k.part, 3 * kRedBlocks) && al(&k.sc, 16);
  if (line)  // the assembled matrix and line factors exist only for NHW_SOLVER_LINE
    ok = ok && al(&k.coef, 15 * k.n) && al(&k.lu_c, k.n) && al(&k.lu_d, k.n);
  if (c->mixed)  // FP32 inner Krylov vectors; the FP64 ones are not needed
    ok = ok && alf(&k.fe, k.n) && alf(&k.fr, k.n) && alf(&k.frh, k.n) && alf(&k.fp, k.n) && alf(&k.fv, k.n) &&
         alf(&k.fs, k.n) && alf(&k.ft, k.n) && alf(&k.fph, k.n) && alf(&k.fsh, k.n);
  else
    ok = ok && al(&k.r, k.n) && al(&k.rh, k.n) && al(&k.p, k.n) && al(&k.v, k.n) && al(&k.s, k.n) &&
         al(&k.t, k.n) && al(&k.ph, nv) && al(&k.sh, nv);
  ok = ok && hipMalloc(&k.flag, sizeof(int) * 4) == hipSuccess && hipMemset(k.flag, 0, sizeof(int) * 4) == hipSuccess;
  if (!ok) {
    set_error("poisson_alloc: hipMalloc failed");
    return NHW_ENOMEM;
  }
  if (!line) {
    // FP64 hierarchy; in the mixed form only its fine-level geometry (the
    // outer residual) plus a full FP32 hierarchy for the inner solve
    int rc = mg_alloc(c, c->mg.lv, sizeof(double), c->mixed ? 1 : 64);
    if (!rc && c->mixed) rc = mg_alloc(c, c->mg.lvf, sizeof(float), 64);
    if (rc) return rc;
  }
  return 0;
}

void poisson_free(nhw_ctx *c) {
  Krylov &k = c->kr;
  double *ps[] = {k.coef, k.rhs, k.x, k.r, k.rh, k.p, k.v, k.s, k.t, k.ph, k.sh, k.lu_c, k.lu_d, k.part, k.sc};
  for (double *p : ps)
    if (p) (void)hipFree(p);
  float *pf[] = {k.fe, k.fr, k.frh, k.fp, k.fv, k.fs, k.ft, k.fph, k.fsh};
  for (float *p : pf)
    if (p) (void)hipFree(p);
  if (k.flag) (void)hipFree(k.flag);
  k = Krylov{};
  mg_free(c->mg.lv);
  mg_free(c->mg.lvf);
}

int launch_p_from_x(nhw_ctx *c, const double *x);

static KV make_kv(nhw_ctx *c) {
  Krylov &kr = c->kr;
  const Geo &g = c->g;
  KV k;
  k.n = kr.n; k.M = g.M; k.MN = (long)g.M * g.N; k.K = g.K;
  k.coef = kr.coef; k.b = kr.rhs;
  k.x = kr.x + kr.pad; k.r = kr.r; k.rh = kr.rh; k.p = kr.p; k.v = kr.v; k.s = kr.s; k.t = kr.t;
  k.ph = kr.ph ? kr.ph + kr.pad : nullptr; k.sh = kr.sh ? kr.sh + kr.pad : nullptr;
  k.lu_c = kr.lu_c; k.lu_d = kr.lu_d; k.part = kr.part; k.sc = kr.sc; k.flag = kr.flag;
  k.itmax = c->cfg.itmax;
  k.tol2 = c->cfg.tol * c->cfg.tol;
  k.itol2 = k.tol2;
  k.inner = 0;
  return k;
}

// FP32 view for the mixed form's inner solve (x is the correction e)
static KVT<float> make_kvf(nhw_ctx *c) {
  Krylov &kr = c->kr;
  const Geo &g = c->g;
  KVT<float> k;
  k.n = kr.n; k.M = g.M; k.MN = (long)g.M * g.N; k.K = g.K;
  k.coef = nullptr; k.b = nullptr;
  k.x = kr.fe; k.r = kr.fr; k.rh = kr.frh; k.p = kr.fp; k.v = kr.fv; k.s = kr.fs; k.t = kr.ft;
  k.ph = kr.fph; k.sh = kr.fsh;
  k.lu_c = nullptr; k.lu_d = nullptr; k.part = kr.part; k.sc = kr.sc; k.flag = kr.flag;
  k.itmax = c->cfg.itmax;
  k.tol2 = c->cfg.tol * c->cfg.tol;
  k.itol2 = kInnerTol * kInnerTol;
  k.inner = 1;
  return k;
}

// end of a solve: report scalars, P from x, statistics.  With `true_res`
// the true residual is reduced from the partials of the last launch (else
// the last k_fin0 already holds it in S_TRUE).
static int finish_solve(nhw_ctx *c, const KV &k, int done, int iters, bool true_res, int nred) {
  Krylov &kr = c->kr;
  hipStream_t st = c->stream;
  int *hflag = (int *)c->pinned;
  if (true_res) {
    Timed _t(c, KID_KREDUCE);
    hipLaunchKernelGGL(k_fin_true, dim3(1), dim3(kBS), 0, st, k, nred);
  }
  double *hsc = (double *)((char *)c->pinned + 64);
  if (hip_check(hipMemcpyAsync(hsc, kr.sc, sizeof(double) * S_N, hipMemcpyDeviceToHost, st), "poisson scalars"))
    return NHW_EHIP;
  if (launch_p_from_x(c, k.x)) return NHW_EHIP;
  if (hip_check(hipStreamSynchronize(st), "poisson sync")) return NHW_EHIP;
  ktimer_flush(c);
  const double bb = hsc[S_BB];
  const double relres = bb > 0.0 ? sqrt(hsc[S_TRUE] / bb) : 0.0;
  c->stats.last_iters = iters;
  c->stats.last_relres = relres;
  c->stats.total_iters += iters;
  c->stats.solves += 1;
  if (iters > c->stats.max_iters) c->stats.max_iters = iters;
  if (done == 3) c->stats.not_converged += 1;
  if (done == 2) {
    set_error("poisson_solver: BiCGStab breakdown (code " + std::to_string(hflag[2]) + ") after " +
              std::to_string(iters) + " iterations");
    return NHW_ESOLVER;
  }
  return 0;
}

// reads the solve's flags after a chunk of queued iterations
static int poll_flags(nhw_ctx *c, int *done, int *iters) {
  Krylov &kr = c->kr;
  int *hflag = (int *)c->pinned;
  if (hip_check(hipGetLastError(), "poisson iteration")) return NHW_EHIP;
  if (hip_check(hipMemcpyAsync(hflag, kr.flag, sizeof(int) * 3, hipMemcpyDeviceToHost, c->stream), "poisson flag"))
    return NHW_EHIP;
  if (hip_check(hipStreamSynchronize(c->stream), "poisson sync")) return NHW_EHIP;
  ktimer_flush(c);
  *done = hflag[0];
  *iters = hflag[1];
  return 0;
}

// right-side of the pressure system (generate_coef_rhs without the matrix)
static void launch_rhs(nhw_ctx *c) {
  const Geo &g = c->g;
  const unsigned nb = (unsigned)((c->kr.n + kBS - 1) / kBS);
  Timed _t(c, KID_RHS);
  if (c->mixed)
    hipLaunchKernelGGL((k_assemble<false, float>), dim3(nb), dim3(kBS), 0, c->stream, g, view<float>(c->d),
                       c->kr.coef, c->kr.rhs);
  else
    hipLaunchKernelGGL((k_assemble<false, double>), dim3(nb), dim3(kBS), 0, c->stream, g, c->d, c->kr.coef,
                       c->kr.rhs);
}

// blocks of the column-walker launches (one thread per column, grid-stride):
// no more than the columns need, so that every block has work and the
// XCD-contiguous block order spreads the columns over all XCDs
static int col_blocks(const nhw_ctx *c) {
  return (int)std::min<long>(c->kr.nblk, ((long)c->g.M * c->g.N + kBS - 1) / kBS);
}

// BiCGStab iterations with the multigrid preconditioner on hierarchy `lvs`,
// queued in chunks sized by the previous solve's count; every kernel of an
// iteration returns at once after the device-side flag is set.
template <class V>
static int bicgstab_mg(nhw_ctx *c, const KVT<V> &k, const std::vector<MGLevel> &lvs, int *done, int *iters) {
  hipStream_t st = c->stream;
  const int nblk = c->kr.nblk, ncb = col_blocks(c);
  const int *stop = k.flag;
  const double n = (double)k.n;
  int chunk = std::max(1, std::min(c->mg.last_iters, 64));
  while (true) {
    for (int q = 0; q < chunk; ++q) {
      { Timed _t(c, KID_KVEC, sizeof(V) * 4 * n); hipLaunchKernelGGL(k_pvec<V>, dim3(nblk), dim3(kBS), 0, st, k); }
      mg_apply<V>(c, lvs, k.p, k.ph, stop);
      mf_apply<V>(c, lvs, k.ph, k.v, k.rh, 1, k.part, stop, ncb, KID_MF_V);
      { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin1<V>, dim3(1), dim3(kBS), 0, st, k, ncb); }
      { Timed _t(c, KID_KVEC, sizeof(V) * 3 * n); hipLaunchKernelGGL(k_svec<V>, dim3(nblk), dim3(kBS), 0, st, k); }
      mg_apply<V>(c, lvs, k.s, k.sh, stop);
      mf_apply<V>(c, lvs, k.sh, k.t, k.s, 2, k.part, stop, ncb, KID_MF_T);
      { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin2<V>, dim3(1), dim3(kBS), 0, st, k, ncb); }
      { Timed _t(c, KID_KUPDATE, sizeof(V) * 8 * n); hipLaunchKernelGGL(k_update<V>, dim3(nblk), dim3(kBS), 0, st, k); }
      { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin3<V>, dim3(1), dim3(kBS), 0, st, k, nblk); }
    }
    const int rc = poll_flags(c, done, iters);
    if (rc) return rc;
    if (*done) break;
    chunk = 2;
  }
  c->mg.last_iters = std::max(1, *iters);
  return 0;
}

// FP64 solve: multigrid-preconditioned BiCGStab on the matrix-free operator
static int poisson_solve_mg(nhw_ctx *c) {
  Krylov &kr = c->kr;
  const Geo &g = c->g;
  hipStream_t st = c->stream;
  const long n = kr.n;
  const unsigned nb = (unsigned)((n + kBS - 1) / kBS);
  launch_rhs(c);
  int rc = mg_setup<double>(c, c->mg.lv);
  if (rc) return rc;
  const KV k = make_kv(c);
  const int nblk = kr.nblk, ncb = col_blocks(c);
  { Timed _t(c, KID_KMISC); hipLaunchKernelGGL(k_x_from_p, dim3(nb), dim3(kBS), 0, st, g, c->d, k.x, zlayout(c)); }
  mf_init<double>(c, c->mg.lv, k.x, k.b, k.r, k.rh, k.p, k.v, (double *)nullptr, k.part, ncb);
  { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin0<double>, dim3(1), dim3(kBS), 0, st, k, ncb); }
  { Timed _t(c, KID_KMISC); hipLaunchKernelGGL(k_zero_if, dim3(nblk), dim3(kBS), 0, st, k.x, n, (const int *)k.flag, 4); }
  int done = 0, iters = 0;
  rc = bicgstab_mg<double>(c, k, c->mg.lv, &done, &iters);
  if (rc) return rc;
  mf_true_resid(c, c->mg.lv, k.x, k.b, k.part, ncb);
  return finish_solve(c, k, done, iters, true, ncb);
}

// Mixed form: FP64 iterative refinement around an FP32 inner solve.  Each
// round forms the FP64 residual r = b - A x of the FP64 solution x (FP64
// operator, FP64 inner products), solves A e = r in FP32 (BiCGStab with the
// FP32 multigrid preconditioner, FP64 dot products) to |r - A e| <= kInnerTol
// |r|, and adds e to x in FP64; it stops when |b - A x| <= TOL |b|, the same
// test as the FP64 solve.
static int poisson_solve_mixed(nhw_ctx *c) {
  Krylov &kr = c->kr;
  const Geo &g = c->g;
  hipStream_t st = c->stream;
  const long n = kr.n;
  const unsigned nb = (unsigned)((n + kBS - 1) / kBS);
  launch_rhs(c);
  int rc = mg_setup<double>(c, c->mg.lv);
  if (!rc) rc = mg_setup<float>(c, c->mg.lvf);
  if (rc) return rc;
  const KV k = make_kv(c);
  const KVT<float> kf = make_kvf(c);
  const int nblk = kr.nblk, ncb = col_blocks(c);
  { Timed _t(c, KID_KMISC); hipLaunchKernelGGL(k_x_from_p, dim3(nb), dim3(kBS), 0, st, g, c->d, k.x, zlayout(c)); }
  int done = 0, inner_total = 0, rounds = 0;
  while (true) {
    // outer residual into the inner vectors (e = 0); S_TRUE = |b - A x|^2
    mf_init<float>(c, c->mg.lv, k.x, k.b, kf.r, kf.rh, kf.p, kf.v, kf.x, k.part, ncb);
    { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin0<float>, dim3(1), dim3(kBS), 0, st, kf, ncb); }
    int it = 0;
    rc = poll_flags(c, &done, &it);
    if (rc) return rc;
    if (done == 4) {  // zero right-hand side
      hipLaunchKernelGGL(k_zero_if, dim3(nblk), dim3(kBS), 0, st, k.x, n, (const int *)k.flag, 4);
      break;
    }
    if (done == 1 || done == 2) break;
    if (rounds >= kMaxRefine || inner_total >= c->cfg.itmax) {
      done = 3;
      break;
    }
    // the k_fin0 above cleared the flags and set the inner reference |r0|
    rc = bicgstab_mg<float>(c, kf, c->mg.lvf, &done, &it);
    if (rc) return rc;
    inner_total += it;
    if (done == 2 && it == 0) break;  // breakdown before any progress
    { Timed _t(c, KID_KVEC, 12.0 * n); hipLaunchKernelGGL(k_add_correction, dim3(nblk), dim3(kBS), 0, st, k.x, (const float *)kf.x, n); }
    ++rounds;
  }
  return finish_solve(c, k, done == 1 || done == 4 ? 1 : done, inner_total, false, kr.nblk);
}

// The Poisson system of the current state as generate_coef_rhs assembles it
// (src/pressure.F:475-840): coef[15*neq] diagonal-major in the order of
// JCoef, rhs[neq], rows i fastest, then j, then k over the interior.
// Device buffers of the call's own; every context can assemble it.
int poisson_system(nhw_ctx *c, double *coef_h, double *rhs_h) {
  const Geo &g = c->g;
  const long n = g.neq();
  double *dc = nullptr, *dr = nullptr;
  if (hipMalloc(&dc, sizeof(double) * 15 * n) != hipSuccess || hipMalloc(&dr, sizeof(double) * n) != hipSuccess) {
    if (dc) (void)hipFree(dc);
    set_error("nhw_generate_coef_rhs: hipMalloc failed");
    return NHW_ENOMEM;
  }
  const unsigned nb = (unsigned)((n + kBS - 1) / kBS);
  if (c->mixed)
    hipLaunchKernelGGL((k_assemble<true, float>), dim3(nb), dim3(kBS), 0, c->stream, g, view<float>(c->d), dc, dr);
  else
    hipLaunchKernelGGL((k_assemble<true, double>), dim3(nb), dim3(kBS), 0, c->stream, g, c->d, dc, dr);
  int rc = hip_check(hipGetLastError(), "nhw_generate_coef_rhs");
  if (!rc)
    rc = hip_check(hipMemcpyAsync(coef_h, dc, sizeof(double) * 15 * n, hipMemcpyDeviceToHost, c->stream),
                   "nhw_generate_coef_rhs") ||
         hip_check(hipMemcpyAsync(rhs_h, dr, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream),
                   "nhw_generate_coef_rhs") ||
         hip_check(hipStreamSynchronize(c->stream), "nhw_generate_coef_rhs");
  (void)hipFree(dc);
  (void)hipFree(dr);
  return rc ? NHW_EHIP : 0;
}

// y = A x with the multigrid solver's matrix-free FP64 operator on the
// current state (host vectors in the row order of poisson_system)
int poisson_apply(nhw_ctx *c, const double *x_h, double *y_h) {
  if (c->mg.lv.empty()) {
    set_error("nhw_poisson_apply: the matrix-free operator belongs to NON_HYDRO contexts of the multigrid "
              "solver (NHW_SOLVER_MG or the mixed form)");
    return NHW_EINVAL;
  }
  const Geo &g = c->g;
  const long n = g.neq(), MN = (long)g.M * g.N;
  // the operator's vectors are in the zebra layout of each sigma plane
  std::vector<double> xz(n), yz(n);
  for (long r = 0; r < n; ++r) {
    const long k = r / MN, q = r - k * MN;
    xz[k * MN + zcol((int)(q % g.M), (int)(q / g.M), g.M)] = x_h[r];
  }
  double *dx = nullptr, *dy = nullptr;
  int *stop = nullptr;
  if (hipMalloc(&dx, sizeof(double) * n) != hipSuccess || hipMalloc(&dy, sizeof(double) * n) != hipSuccess ||
      hipMalloc(&stop, sizeof(int)) != hipSuccess) {
    for (void *p : {(void *)dx, (void *)dy})
      if (p) (void)hipFree(p);
    set_error("nhw_poisson_apply: hipMalloc failed");
    return NHW_ENOMEM;
  }
  int rc = mg_setup<double>(c, c->mg.lv);
  if (!rc)
    rc = hip_check(hipMemsetAsync(stop, 0, sizeof(int), c->stream), "nhw_poisson_apply") ||
         hip_check(hipMemcpyAsync(dx, xz.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream),
                   "nhw_poisson_apply");
  if (!rc) {
    mf_apply<double>(c, c->mg.lv, dx, dy, nullptr, 0, nullptr, stop, c->kr.nblk, KID_MF_V);
    rc = hip_check(hipGetLastError(), "nhw_poisson_apply") ||
         hip_check(hipMemcpyAsync(yz.data(), dy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream),
                   "nhw_poisson_apply") ||
         hip_check(hipStreamSynchronize(c->stream), "nhw_poisson_apply");
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  (void)hipFree(stop);
  if (rc) return NHW_EHIP;
  for (long r = 0; r < n; ++r) {
    const long k = r / MN, q = r - k * MN;
    y_h[r] = yz[k * MN + zcol((int)(q % g.M), (int)(q / g.M), g.M)];
  }
  return 0;
}

// poisson_solver, src/pressure.F:63-245: assemble, solve, P with ghosts.
int poisson_solve(nhw_ctx *c) {
  if (c->mixed) return poisson_solve_mixed(c);
  if (c->cfg.solver != NHW_SOLVER_LINE) return poisson_solve_mg(c);
  Krylov &kr = c->kr;
  const Geo &g = c->g;
  hipStream_t st = c->stream;
  const long n = kr.n;
  const unsigned nb = (unsigned)((n + kBS - 1) / kBS);
  { Timed _t(c, KID_ASSEMBLE); hipLaunchKernelGGL((k_assemble<true, double>), dim3(nb), dim3(kBS), 0, st, g, c->d, kr.coef, kr.rhs); }
  const KV k = make_kv(c);
  const Off15 o = stencil(g);
  const unsigned ncol = (unsigned)((k.MN + kBS - 1) / kBS);
  const int nblk = kr.nblk;
  { Timed _t(c, KID_KMISC); hipLaunchKernelGGL(k_x_from_p, dim3(nb), dim3(kBS), 0, st, g, c->d, k.x, zlayout(c)); }
  { Timed _t(c, KID_LINE_FACTOR); hipLaunchKernelGGL(k_line_factor, dim3(ncol), dim3(kBS), 0, st, k); }
  { Timed _t(c, KID_RESID); hipLaunchKernelGGL(k_resid0, dim3(nblk), dim3(kBS), 0, st, k, o); }
  { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin0<double>, dim3(1), dim3(kBS), 0, st, k, nblk); }
  { Timed _t(c, KID_KMISC); hipLaunchKernelGGL(k_zero_if, dim3(nblk), dim3(kBS), 0, st, k.x, n, (const int *)k.flag, 4); }
  int done = 0, iters = 0, chunk = 4;
  while (true) {
    for (int q = 0; q < chunk; ++q) {
      { Timed _t(c, KID_PREC_P); hipLaunchKernelGGL(k_prec_p, dim3(ncol), dim3(kBS), 0, st, k); }
      { Timed _t(c, KID_SPMV_V); hipLaunchKernelGGL(k_spmv_v, dim3(nblk), dim3(kBS), 0, st, k, o); }
      { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin1<double>, dim3(1), dim3(kBS), 0, st, k, nblk); }
      { Timed _t(c, KID_PREC_S); hipLaunchKernelGGL(k_prec_s, dim3(ncol), dim3(kBS), 0, st, k); }
      { Timed _t(c, KID_SPMV_T); hipLaunchKernelGGL(k_spmv_t, dim3(nblk), dim3(kBS), 0, st, k, o); }
      { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin2<double>, dim3(1), dim3(kBS), 0, st, k, nblk); }
      { Timed _t(c, KID_KUPDATE); hipLaunchKernelGGL(k_update<double>, dim3(nblk), dim3(kBS), 0, st, k); }
      { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin3<double>, dim3(1), dim3(kBS), 0, st, k, nblk); }
    }
    const int rc = poll_flags(c, &done, &iters);
    if (rc) return rc;
    if (done) break;
    chunk = chunk < 32 ? chunk * 2 : 32;
  }
  { Timed _t(c, KID_RESID); hipLaunchKernelGGL(k_true_res, dim3(nblk), dim3(kBS), 0, st, k, o); }
  return finish_solve(c, k, done, iters, true, nblk);
}


// ----------------------------------------------------------------------------
// batched Poisson solve: every member's BiCGStab advances in lockstep (one
// launch per step for all members; each member keeps its own scalars and
// flags, and a converged member's blocks return at once), with the batched
// multigrid V-cycle running each level over the members that have it.
// ----------------------------------------------------------------------------
enum { BV_R = 0, BV_RH, BV_P, BV_V, BV_S, BV_T, BV_PH, BV_SH, BV_E };

template <class R>
__device__ __forceinline__ R *bvec(const BScn<R> &S, int id) {
  switch (id) {
    case BV_R: return S.r;
    case BV_RH: return S.rh;
    case BV_P: return S.pp;
    case BV_V: return S.v;
    case BV_S: return S.s;
    case BV_T: return S.t;
    case BV_PH: return S.ph;
    case BV_SH: return S.sh;
    default: return S.e;
  }
}
template <class R>
__device__ __forceinline__ KScal bscal(const BScn<R> &S) {
  return KScal{S.sc, S.flag, S.itmax, S.tol2, S.itol2, S.inner};
}

template <class R>
__global__ __launch_bounds__(256) void kb_rhs(BTab<R> T) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  assemble_body<false, R>(S.g, S.d, nullptr, S.rhs, b.b * blockDim.x + threadIdx.x);
}

template <class R>
__global__ __launch_bounds__(256) void kb_x_from_p(BTab<R> T) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const Geo &g = S.g;
  const long t = b.b * blockDim.x + threadIdx.x;
  if (t >= g.neq()) return;
  const int i = g.Ibeg + (int)(t % g.M);
  const long r = t / g.M;
  const int j = g.Jbeg + (int)(r % g.N), k = g.Kbeg + (int)(r / g.N);
  S.x[(r / g.N) * ((long)g.M * g.N) + zcol(i - g.Ibeg, j - g.Jbeg, g.M)] = S.d.P[g.c3(i, j, k)];
}

// fine-level column fields: FP64 (outer operator) and, when distinct, R
template <class R>
__global__ __launch_bounds__(256) void kb_geom0(BTab<R> T) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const Geo &g = S.g;
  const long t = b.b * blockDim.x + threadIdx.x;
  if (t >= (long)g.M * g.N) return;
  const int i = (int)(t % g.M), j = (int)(t / g.M);
  const long c = g.c2(g.Ibeg + i, g.Jbeg + j);
  const DevT<R> &d = S.d;
  const long h = S.G64.at(i, j);
  const double r = 1.0 / d.D[c];
  const double ax = d.DelxH[c] * r, bx = (d.DelxH[c] + d.DelxEta[c]) * r;
  const double ay = d.DelyH[c] * r, by = (d.DelyH[c] + d.DelyEta[c]) * r;
  ((double *)S.G64.ax)[h] = ax; ((double *)S.G64.bx)[h] = bx; ((double *)S.G64.ay)[h] = ay;
  ((double *)S.G64.by)[h] = by; ((double *)S.G64.rd)[h] = r;
  if ((const void *)S.G[0].ax != (const void *)S.G64.ax) {
    const G2<R> &G = S.G[0];
    ((R *)G.ax)[h] = (R)ax; ((R *)G.bx)[h] = (R)bx; ((R *)G.ay)[h] = (R)ay; ((R *)G.by)[h] = (R)by;
    ((R *)G.rd)[h] = (R)r;
  }
}

template <class R>
__global__ __launch_bounds__(256) void kb_geom_coarsen(BTab<R> T, int l) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const G2<R> &F = S.G[l], &Cq = S.G[l + 1];
  const int Mf = S.L[l].M, Nf = S.L[l].N, Mc = S.L[l + 1].M, Nc = S.L[l + 1].N, fx = S.lfx[l], fy = S.lfy[l];
  const long t = b.b * blockDim.x + threadIdx.x;
  if (t >= (long)Mc * Nc) return;
  const int I = (int)(t % Mc), J = (int)(t / Mc);
  double a[5] = {0, 0, 0, 0, 0};
  int cnt = 0;
  for (int bb = 0; bb < fy; ++bb)
    for (int aa = 0; aa < fx; ++aa) {
      const int i = I * fx + aa, j = J * fy + bb;
      if (i >= Mf || j >= Nf) continue;
      const long h = F.at(i, j);
      a[0] += F.ax[h]; a[1] += F.bx[h]; a[2] += F.ay[h]; a[3] += F.by[h]; a[4] += F.rd[h];
      ++cnt;
    }
  const double w = 1.0 / cnt;
  const long h = Cq.at(I, J);
  ((R *)Cq.ax)[h] = (R)(a[0] * w); ((R *)Cq.bx)[h] = (R)(a[1] * w); ((R *)Cq.ay)[h] = (R)(a[2] * w);
  ((R *)Cq.by)[h] = (R)(a[3] * w); ((R *)Cq.rd)[h] = (R)(a[4] * w);
}

// FP64 residual of the FP64 solution into the Krylov vectors (e = 0 in the
// mixed form); partials (r,r), (b,b).  One thread walks one column.
template <class R>
__global__ __launch_bounds__(256) void kb_mf_init(BTab<R> T, double *part) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const LvlC<double> &L = S.L64;
  const long col = b.b * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (col < L.MN) {
    const int i = (int)(col % L.M), j = (int)(col / L.M);
    const ColGeo<double> g = load_col(S.G64, L, i, j);
    const Nb o = nbrs(L, i, j);
    const double *x = S.x;
    Pl<double> cu = load_pl(x, o);
    Pl<double> nx = L.K > 1 ? load_pl(plane(x, L, 1), o) : cu;
    Pl<double> pv = cu;
    for (int k = 0; k < L.K; ++k) {
      double c[15];
      coefs(L, g, k, c);
      const long q = (long)k * L.MN + o.c;
      const double bv = S.rhs[q];
      const double rv = bv - arow3(c, pv, cu, nx);
      S.r[q] = (R)rv; S.rh[q] = (R)rv; S.pp[q] = R(0); S.v[q] = R(0);
      if (S.inner) S.e[q] = R(0);
      acc[0] += rv * rv;
      acc[1] += bv * bv;
      pv = cu;
      cu = nx;
      if (k + 2 < L.K) nx = load_pl(plane(x, L, k + 2), o);
    }
  }
  block_sum2(acc, part);
}

// |b - A x|^2 of the FP64 solution (solve report)
template <class R>
__global__ __launch_bounds__(256) void kb_mf_true(BTab<R> T, double *part) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const LvlC<double> &L = S.L64;
  const long col = b.b * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (col < L.MN) {
    const int i = (int)(col % L.M), j = (int)(col / L.M);
    const ColGeo<double> g = load_col(S.G64, L, i, j);
    const Nb o = nbrs(L, i, j);
    const double *x = S.x;
    Pl<double> cu = load_pl(x, o);
    Pl<double> nx = L.K > 1 ? load_pl(plane(x, L, 1), o) : cu;
    Pl<double> pv = cu;
    for (int k = 0; k < L.K; ++k) {
      double c[15];
      coefs(L, g, k, c);
      const double rv = S.rhs[(long)k * L.MN + o.c] - arow3(c, pv, cu, nx);
      acc[0] += rv * rv;
      pv = cu;
      cu = nx;
      if (k + 2 < L.K) nx = load_pl(plane(x, L, k + 2), o);
    }
  }
  block_sum2(acc, part);
}

// per-member scalar steps: which 0..3 = fin0..fin3, 4 = true residual
template <class R>
__global__ void kb_fin(BTab<R> T, const int *off_red, const double *part, int which) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const KScal k = bscal(S);
  if (which >= 1 && which <= 3 && S.flag[0]) return;
  const long b0 = off_red[b.u], b1 = off_red[b.u + 1];
  const double s0 = final_sum(part, 0, b0, b1);
  const double s1 = final_sum(part, 1, b0, b1);
  if (threadIdx.x) return;
  switch (which) {
    case 0: fin0_logic(k, s0, s1); break;
    case 1: fin1_logic(k, s0); break;
    case 2: fin2_logic(k, s0, s1); break;
    case 3: fin3_logic(k, s0, s1); break;
    default: S.sc[S_TRUE] = s0; break;
  }
}

template <class R>
__global__ __launch_bounds__(256) void kb_pvec(BTab<R> T) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  if (S.flag[0]) return;
  const long q = b.b * blockDim.x + threadIdx.x;
  if (q >= S.g.neq()) return;
  const R beta = (R)S.sc[S_BETA], omega = (R)S.sc[S_OMEGA];
  S.pp[q] = S.r[q] + beta * (S.pp[q] - omega * S.v[q]);
}

template <class R>
__global__ __launch_bounds__(256) void kb_svec(BTab<R> T) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  if (S.flag[0]) return;
  const long q = b.b * blockDim.x + threadIdx.x;
  if (q >= S.g.neq()) return;
  S.s[q] = S.r[q] - (R)S.sc[S_ALPHA] * S.v[q];
}

template <class R>
__global__ __launch_bounds__(256) void kb_update(BTab<R> T, double *part) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  if (S.flag[0]) return;
  const long q = b.b * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (q < S.g.neq()) {
    const R alpha = (R)S.sc[S_ALPHA], omega = (R)S.sc[S_OMEGA];
    S.kx[q] = S.kx[q] + alpha * S.ph[q] + omega * S.sh[q];
    const R rv = S.s[q] - omega * S.t[q];
    S.r[q] = rv;
    acc[0] = (double)S.rh[q] * (double)rv;
    acc[1] = (double)rv * (double)rv;
  }
  block_reduce<2>(acc, part);
}

// y = A x on the fine level in R, with dot partials (y, d1) [and (y, y)]
template <class R, int NDOT>
__global__ __launch_bounds__(256) void kb_mf_apply(BTab<R> T, int vin, int vout, int vdot, double *part) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  if (S.flag[0]) return;
  const LvlC<R> &L = S.L[0];
  const long col = b.b * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (col < L.MN) {
    const int i = (int)(col % L.M), j = (int)(col / L.M);
    const ColGeo<R> g = load_col(S.G[0], L, i, j);
    const Nb o = nbrs(L, i, j);
    mf_col<R, NDOT>(L, g, o, bvec(S, vin), bvec(S, vout), bvec(S, vdot), acc);
  }
  block_sum2(acc, part);
}

// multigrid: level-l right side and solution (level 0: the Krylov vectors)
template <class R>
__device__ __forceinline__ const R *lvl_b(const BScn<R> &S, int l, int vin) {
  return l == 0 ? bvec(S, vin) : S.lb[l];
}
template <class R>
__device__ __forceinline__ R *lvl_x(const BScn<R> &S, int l, int vout) {
  return l == 0 ? bvec(S, vout) : S.lx[l];
}

template <class R, int KM, bool ZERO>
__global__ __launch_bounds__(256) void kb_zebra(BTab<R> T, int l, int color, int vin, int vout) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  if (S.flag[0]) return;
  const LvlC<R> &L = S.L[l];
  const int half = (L.M + 1) / 2;
  const long t = b.b * blockDim.x + threadIdx.x;
  if (t >= (long)half * L.N) return;
  const int j = (int)(t / half);
  const int i = 2 * (int)(t % half) + ((j + color) & 1);
  if (i >= L.M) return;
  by_rows<KM>(L.K, [&](auto km) {
    zebra_col<R, decltype(km)::value, ZERO>(L, S.G[l], lvl_b(S, l, vin), lvl_x(S, l, vout), i, j);
  });
}

// the coarsest level of each member in one workgroup (coarse_solve)
template <class R, int KM>
__global__ __launch_bounds__(kCoarseBS) void kb_coarse(BTab<R> T, int l, int vin, int vout, int sweeps) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  if (S.flag[0]) return;
  __shared__ R xs[KM * kMgCoarseCols];
  by_rows<KM>(S.L[l].K, [&](auto km) {
    coarse_solve<R, decltype(km)::value>(S.L[l], S.G[l], lvl_b(S, l, vin), lvl_x(S, l, vout), sweeps, xs);
  });
}

// fused residual + 2x2 restriction of level l into the right side of l+1
// (resid22_body: one thread per coarse cell)
template <class R, int KM>
__global__ __launch_bounds__(256) void kb_resid_restrict22(BTab<R> T, int l, int vin, int vout) {
  const BPos bp = bfind(T);
  const BScn<R> &S = T.scn[bp.s];
  if (S.flag[0]) return;
  const LvlC<R> &L = S.L[l];
  by_rows<KM>(L.K, [&](auto km) {
    resid22_body<R, decltype(km)::value>(L, S.G[l], lvl_x(S, l, vout), lvl_b(S, l, vin), S.L[l + 1].M,
                                         S.L[l + 1].N, S.lb[l + 1], bp.b * blockDim.x + threadIdx.x);
  });
}

// the same on lane pairs (resid22_pair_body), with ZSWEEP starting level
// l+1's V-cycle
template <class R, int KM, bool ZSWEEP>
__global__ __launch_bounds__(256) void kb_resid22_pair(BTab<R> T, int l, int vin, int vout) {
  const BPos bp = bfind(T);
  const BScn<R> &S = T.scn[bp.s];
  if (S.flag[0]) return;
  const LvlC<R> &L = S.L[l];
  by_rows<KM>(L.K, [&](auto km) {
    resid22_pair_body<R, decltype(km)::value, ZSWEEP>(L, S.G[l], lvl_x(S, l, vout), lvl_b(S, l, vin), S.L[l + 1],
                                                       S.G[l + 1], S.lb[l + 1], S.lx[l + 1],
                                                       bp.b * blockDim.x + threadIdx.x);
  });
}

// the same for levels not coarsened 2x2: one thread per coarse column
template <class R, int KM>
__global__ __launch_bounds__(256) void kb_resid_restrict(BTab<R> T, int l, int vin, int vout) {
  const BPos bp = bfind(T);
  const BScn<R> &S = T.scn[bp.s];
  if (S.flag[0]) return;
  const LvlC<R> &L = S.L[l];
  const G2<R> &G = S.G[l];
  const R *x = lvl_x(S, l, vout);
  const R *b = lvl_b(S, l, vin);
  const int Mc = S.L[l + 1].M, Nc = S.L[l + 1].N, fx = S.lfx[l], fy = S.lfy[l];
  const long cc = bp.b * blockDim.x + threadIdx.x;
  const long MNc = (long)Mc * Nc;
  if (cc >= MNc) return;
  const int I = (int)(cc % Mc), J = (int)(cc / Mc);
  by_rows<KM>(L.K, [&](auto km) {
    constexpr int KR = decltype(km)::value;
    R s[KR];
#pragma unroll
    for (int k = 0; k < KR; ++k) s[k] = R(0);
    int cnt = 0;
    for (int bb = 0; bb < fy; ++bb)
      for (int a = 0; a < fx; ++a) {
        const int i = I * fx + a, j = J * fy + bb;
        if (i >= L.M || j >= L.N) continue;
        const ColC<R> q = col_consts(G, L, i, j);
        const Nb o = nbrs(L, i, j);
        R bv[KR];
        load_rows<R, KR>(b, o.c, L, bv);
        walk_col<R, KR>(L, q, o, x, [&](int k, const ZRow<R> &z, const Pl<R> &pv, const Pl<R> &cu, const Pl<R> &nx) {
          s[k] += bv[k] - zrow_apply(z, pv, cu, nx);
        });
        ++cnt;
      }
    R *bc = S.lb[l + 1];
    const long zc = zcol(I, J, Mc);
#pragma unroll
    for (int k = 0; k < KR; ++k)
      if (k < L.K) bc[zc + (long)k * MNc] = s[k] / (R)cnt;
  });
}

// xf += P xc on the even (zebra colour 0) columns of level l (see
// k_prolong_add)
template <class R>
__global__ __launch_bounds__(256) void kb_prolong(BTab<R> T, int l, int vout) {
  const BPos bp = bfind(T);
  const BScn<R> &S = T.scn[bp.s];
  if (S.flag[0]) return;
  const LvlC<R> &L = S.L[l];
  prolong_col<R>(L.M, L.N, S.L[l + 1].M, S.L[l + 1].N, L.K, S.lfx[l], S.lfy[l], S.lx[l + 1], lvl_x(S, l, vout),
                 bp.b * blockDim.x + threadIdx.x);
}

template <class R>
__global__ __launch_bounds__(256) void kb_add_correction(BTab<R> T) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const long q = b.b * blockDim.x + threadIdx.x;
  if (q < S.g.neq()) S.x[q] += (double)S.e[q];
}

template <class R>
__global__ __launch_bounds__(256) void kb_zero_if4(BTab<R> T) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  if (S.flag[0] != 4) return;
  const long q = b.b * blockDim.x + threadIdx.x;
  if (q < S.g.neq()) S.x[q] = 0.0;
}

// P from the solution vector, with its x/y ghost cells
template <class R>
__global__ __launch_bounds__(256) void kb_p_from_x(BTab<R> T) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const Geo &g = S.g;
  const long t = b.b * blockDim.x + threadIdx.x;
  if (t >= g.neq()) return;
  const int i = g.Ibeg + (int)(t % g.M);
  const long r = t / g.M;
  const int j = g.Jbeg + (int)(r % g.N), k = g.Kbeg + (int)(r / g.N);
  const double v = S.x[(r / g.N) * ((long)g.M * g.N) + zcol(i - g.Ibeg, j - g.Jbeg, g.M)];
  double *P = S.d.P;
  P[g.c3(i, j, k)] = v;
  if (i <= g.Ibeg + 1) P[g.c3(2 * g.Ibeg - 1 - i, j, k)] = v;
  if (i >= g.Iend - 1) P[g.c3(2 * g.Iend + 1 - i, j, k)] = v;
  if (j <= g.Jbeg + 1) P[g.c3(i, 2 * g.Jbeg - 1 - j, k)] = v;
  if (j >= g.Jend - 1) P[g.c3(i, 2 * g.Jend + 1 - j, k)] = v;
}

template <class R>
__global__ void kb_gather_sc(BTab<R> T, double *out, int ns) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  if (threadIdx.x == 0) {
    out[b.s] = S.sc[S_TRUE];
    out[ns + b.s] = S.sc[S_BB];
  }
}

template <class R>
static BTab<R> pbtab(Batch *B, const BShape &S) {
  return BTab<R>{(const BScn<R> *)B->tab, S.off, S.ids, S.nsub};
}

// launch of a batched kernel instantiated at R (the batch's storage type)
#define PBL(kid, bytes, kern, shape, ...)                                                             \
  do {                                                                                                 \
    const BShape &S_ = (shape);                                                                        \
    if (S_.nblk > 0) {                                                                                 \
      Timed _t(B->lead, kid, bytes);                                                                   \
      hipLaunchKernelGGL(HIP_KERNEL_NAME(kern), dim3(S_.nblk), dim3(S_.bs), 0, B->lead->stream,        \
                         pbtab<R>(B, S_), ##__VA_ARGS__);                                              \
    }                                                                                                  \
  } while (0)

static bool has_shape(Batch *B, const std::string &k) { return B->shapes.count(k) > 0; }

template <class R, int KM>
static void bzebra(Batch *B, const std::string &key, int l, int color, bool zero, int vin, int vout) {
  if (!has_shape(B, key)) return;
  const int kid = KID_MG_SMOOTH;  // one kernel at every level above the coarsest
  if (zero) PBL(kid, 0.0, (kb_zebra<R, KM, true>), bshape(B, key), l, color, vin, vout);
  else PBL(kid, 0.0, (kb_zebra<R, KM, false>), bshape(B, key), l, color, vin, vout);
}

// batched V(1,1) cycle on level l: members whose coarsest level is l sweep
// there; the others go down
template <class R, int KM>
static void bvcycle(Batch *B, int l, int vin, int vout) {
  const std::string co = "k:co" + std::to_string(l), zb = "k:zb" + std::to_string(l);
  if (has_shape(B, co))
    PBL(KID_MG_COARSE, 0.0, (kb_coarse<R, KM>), bshape(B, co), l, vin, vout, B->lead->mg.coarse_sweeps);
  if (!has_shape(B, zb)) return;
  const std::string L = std::to_string(l);
  // the zero-guess half sweep of the members whose restriction from level
  // l-1 has not done it
  bzebra<R, KM>(B, "k:zz" + L, l, 0, true, vin, vout);
  bzebra<R, KM>(B, zb, l, 1, false, vin, vout);
  const std::string rr = "k:rr" + L, rg = "k:rg" + L, rp = "k:rp" + L, rz = "k:rz" + L;
  if (has_shape(B, rr)) PBL(KID_MG_RESID, 0.0, (kb_resid_restrict22<R, KM>), bshape(B, rr), l, vin, vout);
  if (has_shape(B, rp)) PBL(KID_MG_RESID, 0.0, (kb_resid22_pair<R, KM, false>), bshape(B, rp), l, vin, vout);
  if (has_shape(B, rz)) PBL(KID_MG_RESID, 0.0, (kb_resid22_pair<R, KM, true>), bshape(B, rz), l, vin, vout);
  if (has_shape(B, rg)) PBL(KID_MG_RESID, 0.0, (kb_resid_restrict<R, KM>), bshape(B, rg), l, vin, vout);
  bvcycle<R, KM>(B, l + 1, 0, 0);
  PBL(KID_MG_PROLONG, 0.0, kb_prolong<R>, bshape(B, "k:pr" + std::to_string(l)), l, vout);
  bzebra<R, KM>(B, zb, l, 1, false, vin, vout);
  bzebra<R, KM>(B, zb, l, 0, false, vin, vout);
}

// reads every member's Krylov flag; returns how many are still running
static int bpoll(Batch *B, int *iters_max) {
  if (blaunch_gather_flags(B, 1)) return -1;
  const int ns = (int)B->m.size();
  if (hip_check(hipMemcpyAsync(B->gflag_h, B->gflag, sizeof(int) * 2 * ns, hipMemcpyDeviceToHost,
                               B->lead->stream), "batch poll") ||
      hip_check(hipStreamSynchronize(B->lead->stream), "batch poll"))
    return -1;
  ktimer_flush(B->lead);
  int running = 0, it = 0;
  for (int u = 0; u < ns; ++u) {
    if (B->gflag_h[u] == 0) ++running;
    it = std::max(it, B->gflag_h[ns + u]);
  }
  *iters_max = it;
  return running;
}

// members whose gathered flag[0] is 0 (still iterating), among `set`
static std::vector<int> running_of(Batch *B, const std::vector<int> &set) {
  std::vector<int> out;
  for (int u : set)
    if (B->gflag_h[u] == 0) out.push_back(u);
  return out;
}

// BiCGStab iterations over the members of `round` that still iterate; after
// every poll the launch tables shrink to the members not yet converged
template <class R, int KM>
static int bbicgstab(Batch *B, const std::vector<int> &round, int *iters_out) {
  double *part = B->part;
  int chunk = std::max(1, std::min(B->last_iters, 64)), iters = 0;
  std::vector<int> act = round;
  while (true) {
    for (int q = 0; q < chunk; ++q) {
      const BShape &rows = bshape(B, "k:c3/256"), &cols = bshape(B, "k:c2/256"), &one = bshape(B, "k:one");
      PBL(KID_KVEC, 0.0, kb_pvec<R>, rows);
      bvcycle<R, KM>(B, 0, BV_P, BV_PH);
      PBL(KID_MF_V, 0.0, (kb_mf_apply<R, 1>), cols, BV_PH, BV_V, BV_RH, part);
      PBL(KID_KREDUCE, 0.0, kb_fin<R>, one, cols.off, (const double *)part, 1);
      PBL(KID_KVEC, 0.0, kb_svec<R>, rows);
      bvcycle<R, KM>(B, 0, BV_S, BV_SH);
      PBL(KID_MF_T, 0.0, (kb_mf_apply<R, 2>), cols, BV_SH, BV_T, BV_S, part);
      PBL(KID_KREDUCE, 0.0, kb_fin<R>, one, cols.off, (const double *)part, 2);
      PBL(KID_KUPDATE, 0.0, kb_update<R>, rows, part);
      PBL(KID_KREDUCE, 0.0, kb_fin<R>, one, rows.off, (const double *)part, 3);
    }
    const int running = bpoll(B, &iters);
    if (running < 0) return NHW_EHIP;
    std::vector<int> now = running_of(B, act);
    if (now.empty()) break;
    if (now.size() < act.size()) {
      act = now;
      const int rc = batch_solver_shapes(B, {{"k:", act}, {"r:", round}});
      if (rc) return rc;
    }
    chunk = 2;
  }
  B->last_iters = std::max(1, iters);
  *iters_out = iters;
  return 0;
}

template <class R, int KM>
static int batch_poisson_t(Batch *B) {
  double *part = B->part;
  const int ns = (int)B->m.size();
  std::vector<int> all(ns), its(ns, 0);
  for (int u = 0; u < ns; ++u) all[u] = u;
  int rc = batch_solver_shapes(B, {{"k:", all}, {"r:", all}});
  if (rc) return rc;
  PBL(KID_RHS, 0.0, kb_rhs<R>, bshape(B, "c3/256"));
  PBL(KID_MG_SETUP, 0.0, kb_geom0<R>, bshape(B, "c2/256"));
  for (int l = 0; l + 1 < kBMaxLev; ++l) {
    const std::string key = "gc" + std::to_string(l);
    if (!has_shape(B, key)) break;
    PBL(KID_MG_SETUP, 0.0, kb_geom_coarsen<R>, bshape(B, key), l);
  }
  PBL(KID_KMISC, 0.0, kb_x_from_p<R>, bshape(B, "c3/256"));
  std::vector<int> round = all;
  int total_iters = 0;
  for (int r = 0; r < (B->mixed ? kMaxRefine : 1) && !round.empty(); ++r) {
    {
      const BShape &rows = bshape(B, "r:c3/256"), &cols = bshape(B, "r:c2/256"), &one = bshape(B, "r:one");
      PBL(KID_MF_RESID, 0.0, kb_mf_init<R>, cols, part);
      PBL(KID_KREDUCE, 0.0, kb_fin<R>, one, cols.off, (const double *)part, 0);
      PBL(KID_KMISC, 0.0, kb_zero_if4<R>, rows);
    }
    int it = 0;
    std::vector<int> act = round;
    if (B->mixed) {  // members whose outer residual already meets TOL are done
      if (bpoll(B, &it) < 0) return NHW_EHIP;
      act = running_of(B, round);
      if (act.empty()) break;
      if (act.size() < round.size() && (rc = batch_solver_shapes(B, {{"k:", act}, {"r:", act}}))) return rc;
    }
    if ((rc = bbicgstab<R, KM>(B, act, &it))) return rc;
    total_iters += it;
    for (int u : act) its[u] += B->gflag_h[ns + u];
    if (B->mixed) {
      PBL(KID_KVEC, 0.0, kb_add_correction<R>, bshape(B, "r:c3/256"));
      round = act;
      if ((rc = batch_solver_shapes(B, {{"k:", round}, {"r:", round}}))) return rc;
    }
  }
  const BShape &rows = bshape(B, "c3/256"), &cols = bshape(B, "c2/256"), &one = bshape(B, "one");
  PBL(KID_MF_RESID, 0.0, kb_mf_true<R>, cols, part);
  PBL(KID_KREDUCE, 0.0, kb_fin<R>, one, cols.off, (const double *)part, 4);
  PBL(KID_PROJECTION, 0.0, kb_p_from_x<R>, rows);
  PBL(KID_KMISC, 0.0, kb_gather_sc<R>, one, B->gsc, ns);
  int dummy = 0;
  if (bpoll(B, &dummy) < 0) return NHW_EHIP;  // also waits for the gather
  if (hip_check(hipMemcpyAsync(B->gsc_h, B->gsc, sizeof(double) * 2 * ns, hipMemcpyDeviceToHost, B->lead->stream),
                "batch stats") ||
      hip_check(hipStreamSynchronize(B->lead->stream), "batch stats"))
    return NHW_EHIP;
  for (int u = 0; u < ns; ++u) {
    nhw_stats &st = B->m[u]->stats;
    const int fl = B->gflag_h[u];
    const double bb = B->gsc_h[ns + u];
    st.last_iters = its[u];
    st.last_relres = bb > 0.0 ? sqrt(B->gsc_h[u] / bb) : 0.0;
    st.total_iters += its[u];
    st.solves += 1;
    if (its[u] > st.max_iters) st.max_iters = its[u];
    // FP64 form: ITMAX reached; mixed form: refinement ended above TOL
    if (B->mixed ? st.last_relres > B->m[u]->cfg.tol : fl == 3) st.not_converged += 1;
    if (fl == 2 && !B->mixed) {
      set_error("poisson_solver: BiCGStab breakdown in batch member " + std::to_string(u));
      return NHW_ESOLVER;
    }
  }
  (void)total_iters;
  return 0;
}

int batch_poisson(Batch *B) {
  if (B->mixed) {
    if (B->kmax <= 4) return batch_poisson_t<float, 4>(B);
    if (B->kmax <= 8) return batch_poisson_t<float, 8>(B);
    if (B->kmax <= 12) return batch_poisson_t<float, 12>(B);
    return batch_poisson_t<float, 16>(B);
  }
  if (B->kmax <= 4) return batch_poisson_t<double, 4>(B);
  if (B->kmax <= 8) return batch_poisson_t<double, 8>(B);
  if (B->kmax <= 12) return batch_poisson_t<double, 12>(B);
  return batch_poisson_t<double, 16>(B);
}

}  // namespace nhw
