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
#include "nhw_impl.h"

namespace nhw {

constexpr int kRedBlocks = 8192;
constexpr int kBS = 256;
constexpr double kInnerTol = 1e-4;  // relative residual reduction of one FP32 inner solve
constexpr int kMaxRefine = 20;      // refinement rounds of the mixed form

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
__global__ __launch_bounds__(256) void k_assemble(Geo g, DevT<R> d, double *__restrict__ coef,
                                                  double *__restrict__ rhs) {
  const long n = g.neq();
  long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
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
  rhs[r] = -((ufp - ufm) / (2.0 * dx) + (U[q] - U[q - sk3]) / hs * dxs + (vfp - vfm) / (2.0 * dy) +
             (V[q] - V[q - sk3]) / hs * dys + (W[q] - W[q - sk3]) / hs * dzs - 0.0) *
           kRho0 / dt;
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
};
using KV = KVT<double>;
enum { S_RHO = 0, S_ALPHA, S_OMEGA, S_BETA, S_BB, S_RR, S_RHV, S_TS, S_TT, S_TRUE, S_REF, S_N };

__device__ __forceinline__ double spmv_row(const double *__restrict__ coef, const double *__restrict__ x,
                                           long n, long r, const Off15 &o) {
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 15; ++q) acc += coef[(long)q * n + r] * x[r + o.o[q]];
  return acc;
}

// deterministic block reduction of up to 3 values into part[3*blockIdx+m]
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double *part) {
  __shared__ double sm[NV][kBS];
#pragma unroll
  for (int m = 0; m < NV; ++m) sm[m][threadIdx.x] = v[m];
  __syncthreads();
  for (int s = kBS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int m = 0; m < NV; ++m) sm[m][threadIdx.x] += sm[m][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int m = 0; m < NV; ++m) part[3 * blockIdx.x + m] = sm[m][0];
}

__device__ __forceinline__ double final_sum(const double *part, int m, int nblk) {
  __shared__ double sm[kBS];
  double a = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kBS) a += part[3 * b + m];
  sm[threadIdx.x] = a;
  __syncthreads();
  for (int s = kBS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  double out = sm[0];
  __syncthreads();
  return out;
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
  const double rr = final_sum(k.part, 0, nblk);
  const double bb = final_sum(k.part, 1, nblk);
  if (threadIdx.x == 0) {
    k.sc[S_RR] = rr; k.sc[S_BB] = bb; k.sc[S_RHO] = rr; k.sc[S_TRUE] = rr;
    k.sc[S_REF] = k.inner ? rr : bb;
    k.sc[S_ALPHA] = 1.0; k.sc[S_OMEGA] = 1.0; k.sc[S_BETA] = 0.0;
    k.flag[1] = 0; k.flag[2] = 0;
    if (bb == 0.0) k.flag[0] = 4;                       // zero right-hand side: x = 0
    else if (!(rr == rr)) { k.flag[0] = 2; k.flag[2] = 1; }
    else if (rr <= k.tol2 * bb) k.flag[0] = 1;          // initial guess already converged
    else k.flag[0] = 0;
  }
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
  const double rhv = final_sum(k.part, 0, nblk);
  if (threadIdx.x == 0) {
    k.sc[S_RHV] = rhv;
    if (rhv == 0.0 || !(rhv == rhv)) { k.flag[0] = 2; k.flag[2] = 2; return; }
    k.sc[S_ALPHA] = k.sc[S_RHO] / rhv;
  }
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
  const double ts = final_sum(k.part, 0, nblk);
  const double tt = final_sum(k.part, 1, nblk);
  if (threadIdx.x == 0) {
    k.sc[S_TS] = ts; k.sc[S_TT] = tt;
    // tt == 0 means s == 0: the half step already solves the system
    k.sc[S_OMEGA] = tt == 0.0 ? 0.0 : ts / tt;
    if (!(tt == tt)) { k.flag[0] = 2; k.flag[2] = 3; }
  }
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
  const double rhr = final_sum(k.part, 0, nblk);
  const double rr = final_sum(k.part, 1, nblk);
  if (threadIdx.x == 0) {
    const int it = k.flag[1] + 1;
    k.flag[1] = it;
    k.sc[S_RR] = rr;
    const double omega = k.sc[S_OMEGA];
    if (!(rr == rr)) { k.flag[0] = 2; k.flag[2] = 4; return; }
    if (rr <= (k.inner ? k.itol2 : k.tol2) * k.sc[S_REF]) { k.flag[0] = 1; return; }
    if (it >= k.itmax) { k.flag[0] = 3; return; }
    if (rhr == 0.0 || omega == 0.0) { k.flag[0] = 2; k.flag[2] = 5; return; }
    k.sc[S_BETA] = (rhr / k.sc[S_RHO]) * (k.sc[S_ALPHA] / omega);
    k.sc[S_RHO] = rhr;
  }
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
  const double rr = final_sum(k.part, 0, nblk);
  if (threadIdx.x == 0) k.sc[S_TRUE] = rr;
}

// initial guess x = P (interior), src/pressure.F:105-113; zero when the
// right-hand side vanishes
__global__ void k_x_from_p(Geo g, Dev d, double *__restrict__ x) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.neq()) return;
  const int i = g.Ibeg + (int)(t % g.M);
  const long r = t / g.M;
  const int j = g.Jbeg + (int)(r % g.N), k = g.Kbeg + (int)(r / g.N);
  x[t] = d.P[g.c3(i, j, k)];
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
static int finish_solve(nhw_ctx *c, const KV &k, int done, int iters, bool true_res) {
  Krylov &kr = c->kr;
  hipStream_t st = c->stream;
  int *hflag = (int *)c->pinned;
  if (true_res) {
    Timed _t(c, KID_KREDUCE);
    hipLaunchKernelGGL(k_fin_true, dim3(1), dim3(kBS), 0, st, k, kr.nblk);
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

// BiCGStab iterations with the multigrid preconditioner on hierarchy `lvs`,
// queued in chunks sized by the previous solve's count; every kernel of an
// iteration returns at once after the device-side flag is set.
template <class V>
static int bicgstab_mg(nhw_ctx *c, const KVT<V> &k, const std::vector<MGLevel> &lvs, int *done, int *iters) {
  hipStream_t st = c->stream;
  const int nblk = c->kr.nblk;
  const int *stop = k.flag;
  const double n = (double)k.n;
  int chunk = std::max(1, std::min(c->mg.last_iters, 64));
  while (true) {
    for (int q = 0; q < chunk; ++q) {
      { Timed _t(c, KID_KVEC, sizeof(V) * 4 * n); hipLaunchKernelGGL(k_pvec<V>, dim3(nblk), dim3(kBS), 0, st, k); }
      mg_apply<V>(c, lvs, k.p, k.ph, stop);
      mf_apply<V>(c, lvs, k.ph, k.v, k.rh, 1, k.part, stop, nblk, KID_MF_V);
      { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin1<V>, dim3(1), dim3(kBS), 0, st, k, nblk); }
      { Timed _t(c, KID_KVEC, sizeof(V) * 3 * n); hipLaunchKernelGGL(k_svec<V>, dim3(nblk), dim3(kBS), 0, st, k); }
      mg_apply<V>(c, lvs, k.s, k.sh, stop);
      mf_apply<V>(c, lvs, k.sh, k.t, k.s, 2, k.part, stop, nblk, KID_MF_T);
      { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin2<V>, dim3(1), dim3(kBS), 0, st, k, nblk); }
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
  const int nblk = kr.nblk;
  { Timed _t(c, KID_KMISC); hipLaunchKernelGGL(k_x_from_p, dim3(nb), dim3(kBS), 0, st, g, c->d, k.x); }
  mf_init<double>(c, c->mg.lv, k.x, k.b, k.r, k.rh, k.p, k.v, (double *)nullptr, k.part, nblk);
  { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin0<double>, dim3(1), dim3(kBS), 0, st, k, nblk); }
  { Timed _t(c, KID_KMISC); hipLaunchKernelGGL(k_zero_if, dim3(nblk), dim3(kBS), 0, st, k.x, n, (const int *)k.flag, 4); }
  int done = 0, iters = 0;
  rc = bicgstab_mg<double>(c, k, c->mg.lv, &done, &iters);
  if (rc) return rc;
  mf_true_resid(c, c->mg.lv, k.x, k.b, k.part, nblk);
  return finish_solve(c, k, done, iters, true);
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
  const int nblk = kr.nblk;
  { Timed _t(c, KID_KMISC); hipLaunchKernelGGL(k_x_from_p, dim3(nb), dim3(kBS), 0, st, g, c->d, k.x); }
  int done = 0, inner_total = 0, rounds = 0;
  while (true) {
    // outer residual into the inner vectors (e = 0); S_TRUE = |b - A x|^2
    mf_init<float>(c, c->mg.lv, k.x, k.b, kf.r, kf.rh, kf.p, kf.v, kf.x, k.part, nblk);
    { Timed _t(c, KID_KREDUCE); hipLaunchKernelGGL(k_fin0<float>, dim3(1), dim3(kBS), 0, st, kf, nblk); }
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
  return finish_solve(c, k, done == 1 || done == 4 ? 1 : done, inner_total, false);
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
  { Timed _t(c, KID_KMISC); hipLaunchKernelGGL(k_x_from_p, dim3(nb), dim3(kBS), 0, st, g, c->d, k.x); }
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
  return finish_solve(c, k, done, iters, true);
}

}  // namespace nhw
