// Matrix-free sigma-coordinate Poisson operator and a geometric multigrid
// preconditioner for it.
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
// linear operator, used as the right preconditioner of BiCGStab.
#include <algorithm>
#include <vector>

#include "nhw_impl.h"

namespace nhw {

constexpr int kBSm = 256;

struct LvlC {
  int M, N, K;
  long MN;
  double cxy, idx2, idy2;
  double sig[kMgKMax], r6[kMgKMax], r13[kMgKMax], rx[kMgKMax], ry[kMgKMax];
};

struct G2 {
  const double *ax, *bx, *ay, *by, *rd;
  int ld;  // (M+2): halo layout, interior (i,j) at (i+1) + ld*(j+1)
  __device__ long at(int i, int j) const { return (long)(i + 1) + (long)ld * (j + 1); }
};

struct ColGeo {
  double axc, bxc, ayc, byc, rd, axm, bxm, axp, bxp, aym, bym, ayp, byp;
  bool w, e, s, n;
};

__device__ __forceinline__ ColGeo load_col(const G2 &G, const LvlC &L, int i, int j) {
  ColGeo g;
  const long c = G.at(i, j);
  g.axc = G.ax[c]; g.bxc = G.bx[c]; g.ayc = G.ay[c]; g.byc = G.by[c]; g.rd = G.rd[c];
  g.axm = G.ax[c - 1]; g.bxm = G.bx[c - 1]; g.axp = G.ax[c + 1]; g.bxp = G.bx[c + 1];
  g.aym = G.ay[c - G.ld]; g.bym = G.by[c - G.ld]; g.ayp = G.ay[c + G.ld]; g.byp = G.by[c + G.ld];
  g.w = i == 0; g.e = i == L.M - 1; g.s = j == 0; g.n = j == L.N - 1;
  return g;
}

// the 15 coefficients of row (i,j,k) in JCoef order (c[0] = centre), folds
// applied in the reference's order (west, east, south, north, bottom, top)
__device__ __forceinline__ void coefs(const LvlC &L, const ColGeo &g, int k, double c[15]) {
  const double sg = L.sig[k];
  const double dxs = g.axc - sg * g.bxc, dxm = g.axm - sg * g.bxm, dxp = g.axp - sg * g.bxp;
  const double dys = g.ayc - sg * g.byc, dym = g.aym - sg * g.bym, dyp = g.ayp - sg * g.byp;
  const double A1 = dxs * dxs + dys * dys + g.rd * g.rd;
  c[0] = L.cxy + A1 * (L.r6[k] + L.r13[k]);
  c[1] = -L.idx2; c[8] = -L.idx2;
  c[2] = -L.idy2; c[9] = -L.idy2;
  c[3] = (dym + dys) * L.ry[k]; c[14] = -c[3];
  c[4] = (dxm + dxs) * L.rx[k]; c[13] = -c[4];
  c[5] = -A1 * L.r6[k]; c[12] = -A1 * L.r13[k];
  c[6] = -(dxp + dxs) * L.rx[k]; c[11] = -c[6];
  c[7] = -(dyp + dys) * L.ry[k]; c[10] = -c[7];
  if (g.w) { c[0] += c[8]; c[5] += c[4]; c[12] += c[13]; c[8] = 0.0; c[4] = 0.0; c[13] = 0.0; }
  if (g.e) { c[0] += c[1]; c[5] += c[6]; c[12] += c[11]; c[1] = 0.0; c[6] = 0.0; c[11] = 0.0; }
  if (g.s) { c[0] += c[9]; c[5] += c[3]; c[12] += c[14]; c[9] = 0.0; c[3] = 0.0; c[14] = 0.0; }
  if (g.n) { c[0] += c[2]; c[5] += c[7]; c[12] += c[10]; c[2] = 0.0; c[7] = 0.0; c[10] = 0.0; }
  if (k == 0) {
    c[5] += c[12]; c[6] += c[11]; c[7] += c[10]; c[4] += c[13]; c[3] += c[14];
    c[12] = 0.0; c[11] = 0.0; c[10] = 0.0; c[13] = 0.0; c[14] = 0.0;
  }
  if (k == L.K - 1) { c[3] = 0.0; c[4] = 0.0; c[5] = 0.0; c[6] = 0.0; c[7] = 0.0; }
}

// neighbour offsets of one column, clamped so every read is in bounds; a
// clamped neighbour always carries a folded (zero) coefficient
struct Nb {
  long xm, xp, ym, yp;
};
__device__ __forceinline__ Nb nbrs(const LvlC &L, int i, int j) {
  Nb o;
  o.xm = i > 0 ? -1 : 0;
  o.xp = i < L.M - 1 ? 1 : 0;
  o.ym = j > 0 ? -(long)L.M : 0;
  o.yp = j < L.N - 1 ? (long)L.M : 0;
  return o;
}

// sum of the 12 couplings to other columns for row k (x at base = column)
__device__ __forceinline__ double offcol(const double *__restrict__ x, long base, long MN, int k, int K,
                                         const Nb &o, const double c[15]) {
  const long q = base + (long)k * MN;
  const long up = k < K - 1 ? MN : 0, dn = k > 0 ? -MN : 0;
  return c[1] * x[q + o.xp] + c[2] * x[q + o.yp] + c[3] * x[q + o.ym + up] + c[4] * x[q + o.xm + up] +
         c[6] * x[q + o.xp + up] + c[7] * x[q + o.yp + up] + c[8] * x[q + o.xm] + c[9] * x[q + o.ym] +
         c[10] * x[q + o.yp + dn] + c[11] * x[q + o.xp + dn] + c[13] * x[q + o.xm + dn] +
         c[14] * x[q + o.ym + dn];
}

// ---------------------------------------------------------------------------
// geometry set-up
// ---------------------------------------------------------------------------
__global__ void k_geom0(Geo g, Dev d, double *ax, double *bx, double *ay, double *by, double *rd, int ld) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.M * g.N) return;
  const int i = (int)(t % g.M), j = (int)(t / g.M);
  const long c = g.c2(g.Ibeg + i, g.Jbeg + j);
  const long h = (long)(i + 1) + (long)ld * (j + 1);
  const double r = 1.0 / d.D[c];
  ax[h] = d.DelxH[c] * r;
  bx[h] = (d.DelxH[c] + d.DelxEta[c]) * r;
  ay[h] = d.DelyH[c] * r;
  by[h] = (d.DelyH[c] + d.DelyEta[c]) * r;
  rd[h] = r;
}

// block average of the five column fields onto the next level
__global__ void k_geom_coarsen(int Mf, int Nf, int ldf, int Mc, int Nc, int ldc, int fx, int fy,
                               const double *axf, const double *bxf, const double *ayf, const double *byf,
                               const double *rdf, double *axc, double *bxc, double *ayc, double *byc,
                               double *rdc) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)Mc * Nc) return;
  const int I = (int)(t % Mc), J = (int)(t / Mc);
  double s[5] = {0, 0, 0, 0, 0};
  int cnt = 0;
  for (int b = 0; b < fy; ++b)
    for (int a = 0; a < fx; ++a) {
      const int i = I * fx + a, j = J * fy + b;
      if (i >= Mf || j >= Nf) continue;
      const long h = (long)(i + 1) + (long)ldf * (j + 1);
      s[0] += axf[h]; s[1] += bxf[h]; s[2] += ayf[h]; s[3] += byf[h]; s[4] += rdf[h];
      ++cnt;
    }
  const double w = 1.0 / cnt;
  const long h = (long)(I + 1) + (long)ldc * (J + 1);
  axc[h] = s[0] * w; bxc[h] = s[1] * w; ayc[h] = s[2] * w; byc[h] = s[3] * w; rdc[h] = s[4] * w;
}

// ---------------------------------------------------------------------------
// operator kernels (one thread per column, k loop inside)
// ---------------------------------------------------------------------------

// deterministic fixed-shape block sum of two values into part[3*block + m]
__device__ __forceinline__ void block_sum2(const double (&acc)[2], double *part) {
  __shared__ double sm[2][kBSm];
  sm[0][threadIdx.x] = acc[0];
  sm[1][threadIdx.x] = acc[1];
  __syncthreads();
  for (int s = kBSm / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sm[0][threadIdx.x] += sm[0][threadIdx.x + s];
      sm[1][threadIdx.x] += sm[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = sm[0][0];
    part[3 * blockIdx.x + 1] = sm[1][0];
  }
}

// y = A x; optional dot partials: (y, d1) and (y, y) or (d1, y)
template <int NDOT>
__global__ __launch_bounds__(kBSm) void k_mf_apply(LvlC L, G2 G, const double *__restrict__ x,
                                                   double *__restrict__ y, const double *__restrict__ d1,
                                                   double *part, const int *stop) {
  if (stop && stop[0]) return;
  double acc[2] = {0.0, 0.0};
  for (long col = (long)blockIdx.x * kBSm + threadIdx.x; col < L.MN; col += (long)gridDim.x * kBSm) {
    const int i = (int)(col % L.M), j = (int)(col / L.M);
    const ColGeo g = load_col(G, L, i, j);
    const Nb o = nbrs(L, i, j);
    for (int k = 0; k < L.K; ++k) {
      double c[15];
      coefs(L, g, k, c);
      const long q = col + (long)k * L.MN;
      const long up = k < L.K - 1 ? L.MN : 0, dn = k > 0 ? -L.MN : 0;
      const double v = c[0] * x[q] + c[5] * x[q + up] + c[12] * x[q + dn] + offcol(x, col, L.MN, k, L.K, o, c);
      y[q] = v;
      if (NDOT >= 1) acc[0] += v * d1[q];
      if (NDOT >= 2) acc[1] += v * v;
    }
  }
  if (NDOT > 0) block_sum2(acc, part);
}

// r = b - A x  (all columns)
__global__ __launch_bounds__(kBSm) void k_mf_resid(LvlC L, G2 G, const double *__restrict__ x,
                                                   const double *__restrict__ b, double *__restrict__ r,
                                                   const int *stop) {
  if (stop[0]) return;
  long col = (long)blockIdx.x * kBSm + threadIdx.x;
  if (col >= L.MN) return;
  const int i = (int)(col % L.M), j = (int)(col / L.M);
  const ColGeo g = load_col(G, L, i, j);
  const Nb o = nbrs(L, i, j);
  for (int k = 0; k < L.K; ++k) {
    double c[15];
    coefs(L, g, k, c);
    const long q = col + (long)k * L.MN;
    const long up = k < L.K - 1 ? L.MN : 0, dn = k > 0 ? -L.MN : 0;
    r[q] = b[q] - (c[0] * x[q] + c[5] * x[q + up] + c[12] * x[q + dn] + offcol(x, col, L.MN, k, L.K, o, c));
  }
}

// zebra z-line Gauss-Seidel half sweep over the columns of one colour; with
// ZERO the other colour is taken as zero (first sweep from a zero guess)
template <int KM, bool ZERO>
__global__ __launch_bounds__(kBSm) void k_zebra(LvlC L, G2 G, const double *__restrict__ b,
                                                double *__restrict__ x, int color, const int *stop) {
  if (stop[0]) return;
  const int half = (L.M + 1) / 2;
  long t = (long)blockIdx.x * kBSm + threadIdx.x;
  if (t >= (long)half * L.N) return;
  const int j = (int)(t / half);
  const int i = 2 * (int)(t % half) + ((j + color) & 1);
  if (i >= L.M) return;
  const long col = (long)i + (long)L.M * j;
  const ColGeo g = load_col(G, L, i, j);
  const Nb o = nbrs(L, i, j);
  double sub[KM], dia[KM], sup[KM], rhs[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    if (k < L.K) {
      double c[15];
      coefs(L, g, k, c);
      double r = b[col + (long)k * L.MN];
      if (!ZERO) r -= offcol(x, col, L.MN, k, L.K, o, c);
      sub[k] = c[12];
      dia[k] = c[0];
      sup[k] = c[5];
      rhs[k] = r;
    }
  }
  // Thomas: forward elimination then back substitution
  double cp = sup[0] / dia[0];
  rhs[0] = rhs[0] / dia[0];
  sup[0] = cp;
#pragma unroll
  for (int k = 1; k < KM; ++k) {
    if (k < L.K) {
      const double den = dia[k] - sub[k] * sup[k - 1];
      const double inv = 1.0 / den;
      sup[k] = sup[k] * inv;
      rhs[k] = (rhs[k] - sub[k] * rhs[k - 1]) * inv;
    }
  }
  double xn = 0.0;
#pragma unroll
  for (int k = KM - 1; k >= 0; --k) {
    if (k < L.K) {
      xn = k == L.K - 1 ? rhs[k] : rhs[k] - sup[k] * xn;
      x[col + (long)k * L.MN] = xn;
    }
  }
}

// start of a solve: r = b - A x, rh = r, p = v = 0; partials (r,r), (b,b)
__global__ __launch_bounds__(kBSm) void k_mf_init(LvlC L, G2 G, const double *__restrict__ x,
                                                  const double *__restrict__ b, double *__restrict__ r,
                                                  double *__restrict__ rh, double *__restrict__ p,
                                                  double *__restrict__ v, double *part) {
  double acc[2] = {0.0, 0.0};
  for (long col = (long)blockIdx.x * kBSm + threadIdx.x; col < L.MN; col += (long)gridDim.x * kBSm) {
    const int i = (int)(col % L.M), j = (int)(col / L.M);
    const ColGeo g = load_col(G, L, i, j);
    const Nb o = nbrs(L, i, j);
    for (int k = 0; k < L.K; ++k) {
      double c[15];
      coefs(L, g, k, c);
      const long q = col + (long)k * L.MN;
      const long up = k < L.K - 1 ? L.MN : 0, dn = k > 0 ? -L.MN : 0;
      const double bv = b[q];
      const double rv =
          bv - (c[0] * x[q] + c[5] * x[q + up] + c[12] * x[q + dn] + offcol(x, col, L.MN, k, L.K, o, c));
      r[q] = rv; rh[q] = rv; p[q] = 0.0; v[q] = 0.0;
      acc[0] += rv * rv;
      acc[1] += bv * bv;
    }
  }
  block_sum2(acc, part);
}

// |b - A x|^2 partials for the solve report
__global__ __launch_bounds__(kBSm) void k_mf_true(LvlC L, G2 G, const double *__restrict__ x,
                                                  const double *__restrict__ b, double *part) {
  double acc[2] = {0.0, 0.0};
  for (long col = (long)blockIdx.x * kBSm + threadIdx.x; col < L.MN; col += (long)gridDim.x * kBSm) {
    const int i = (int)(col % L.M), j = (int)(col / L.M);
    const ColGeo g = load_col(G, L, i, j);
    const Nb o = nbrs(L, i, j);
    for (int k = 0; k < L.K; ++k) {
      double c[15];
      coefs(L, g, k, c);
      const long q = col + (long)k * L.MN;
      const long up = k < L.K - 1 ? L.MN : 0, dn = k > 0 ? -L.MN : 0;
      const double rv =
          b[q] - (c[0] * x[q] + c[5] * x[q + up] + c[12] * x[q + dn] + offcol(x, col, L.MN, k, L.K, o, c));
      acc[0] += rv * rv;
    }
  }
  block_sum2(acc, part);
}

// block-average restriction of the fine residual onto the coarse right side
__global__ void k_restrict(int Mf, int Nf, int Mc, int Nc, int K, int fx, int fy, const double *__restrict__ rf,
                           double *__restrict__ bc, const int *stop) {
  if (stop[0]) return;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long MNc = (long)Mc * Nc, MNf = (long)Mf * Nf;
  if (t >= MNc * K) return;
  const int k = (int)(t / MNc);
  const long cc = t % MNc;
  const int I = (int)(cc % Mc), J = (int)(cc / Mc);
  double s = 0.0;
  int cnt = 0;
  for (int b = 0; b < fy; ++b)
    for (int a = 0; a < fx; ++a) {
      const int i = I * fx + a, j = J * fy + b;
      if (i >= Mf || j >= Nf) continue;
      s += rf[(long)i + (long)Mf * j + MNf * k];
      ++cnt;
    }
  bc[t] = s / cnt;
}

// bilinear (cell-centred) prolongation of the coarse correction, added
__global__ void k_prolong_add(int Mf, int Nf, int Mc, int Nc, int K, int fx, int fy,
                              const double *__restrict__ xc, double *__restrict__ xf, const int *stop) {
  if (stop[0]) return;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long MNc = (long)Mc * Nc, MNf = (long)Mf * Nf;
  if (t >= MNf * K) return;
  const int k = (int)(t / MNf);
  const long cf = t % MNf;
  const int i = (int)(cf % Mf), j = (int)(cf / Mf);
  int I0 = i / fx, I1 = I0;
  double wx0 = 1.0, wx1 = 0.0;
  if (fx == 2) {
    I1 = (i & 1) ? I0 + 1 : I0 - 1;
    if (I1 < 0 || I1 >= Mc) I1 = I0;
    wx0 = 0.75; wx1 = 0.25;
  }
  int J0 = j / fy, J1 = J0;
  double wy0 = 1.0, wy1 = 0.0;
  if (fy == 2) {
    J1 = (j & 1) ? J0 + 1 : J0 - 1;
    if (J1 < 0 || J1 >= Nc) J1 = J0;
    wy0 = 0.75; wy1 = 0.25;
  }
  const double *xk = xc + MNc * k;
  const double v = wy0 * (wx0 * xk[I0 + (long)Mc * J0] + wx1 * xk[I1 + (long)Mc * J0]) +
                   wy1 * (wx0 * xk[I0 + (long)Mc * J1] + wx1 * xk[I1 + (long)Mc * J1]);
  xf[t] += v;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline unsigned nb(long n) { return (unsigned)((n + kBSm - 1) / kBSm); }

static LvlC make_lvl(const nhw_ctx *c, const MGLevel &lv) {
  LvlC L{};
  const double dx = lv.dx, dy = lv.dy;
  L.M = lv.M; L.N = lv.N; L.K = c->g.K; L.MN = (long)lv.M * lv.N;
  L.cxy = 2. / (dx * dx) + 2. / (dy * dy);
  L.idx2 = 1. / (dx * dx);
  L.idy2 = 1. / (dy * dy);
  const std::vector<double> &dsig = c->mg.dsig_h, &sig = c->mg.sig_h;  // 1-based
  for (int kk = 0; kk < L.K; ++kk) {
    const int k = c->g.Kbeg + kk;
    const double s = dsig[k] + dsig[k - 1];
    L.sig[kk] = sig[k];
    L.r6[kk] = 1. / (0.5 * s * dsig[k]);
    L.r13[kk] = 1. / (0.5 * s * dsig[k - 1]);
    L.rx[kk] = 1. / (2. * dx * s);
    L.ry[kk] = 1. / (2. * dy * s);
  }
  return L;
}

static G2 g2(const MGLevel &lv) { return G2{lv.ax, lv.bx, lv.ay, lv.by, lv.rd, lv.M + 2}; }

int mg_alloc(nhw_ctx *c) {
  MG &mg = c->mg;
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
    const bool fine = mg.lv.empty();
    bool ok = true;
    for (double **p : {&lv.ax, &lv.bx, &lv.ay, &lv.by, &lv.rd})
      ok = ok && hipMalloc(p, sizeof(double) * n2h) == hipSuccess &&
           hipMemset(*p, 0, sizeof(double) * n2h) == hipSuccess;
    for (double **p : {&lv.x, &lv.b, &lv.r}) {
      if (fine && p != &lv.r) continue;  // level 0 solves into the caller's vectors
      ok = ok && hipMalloc(p, sizeof(double) * n) == hipSuccess && hipMemset(*p, 0, sizeof(double) * n) == hipSuccess;
    }
    mg.lv.push_back(lv);
    if (!ok) {
      set_error("mg_alloc: hipMalloc failed");
      return NHW_ENOMEM;
    }
    const int fx = M >= 4 ? 2 : 1, fy = N >= 4 ? 2 : 1;
    if ((fx == 1 && fy == 1) || (long)M * N <= 64) break;
    mg.lv.back().fx = fx;
    mg.lv.back().fy = fy;
    M = (M + fx - 1) / fx;
    N = (N + fy - 1) / fy;
    dx *= fx;
    dy *= fy;
  }
  return 0;
}

void mg_free(nhw_ctx *c) {
  for (MGLevel &lv : c->mg.lv)
    for (double *p : {lv.ax, lv.bx, lv.ay, lv.by, lv.rd, lv.x, lv.b, lv.r})
      if (p) (void)hipFree(p);
  c->mg.lv.clear();
}

// column fields of every level from the current D, DelxH/Eta, DelyH/Eta
int mg_setup(nhw_ctx *c) {
  MG &mg = c->mg;
  const Geo &g = c->g;
  hipStream_t st = c->stream;
  if (mg.dsig_h.size() != (size_t)g.Kloc + 1) {
    set_error("poisson_solver: vertical grid not set (nhw_set_grid)");
    return NHW_EINVAL;
  }
  MGLevel &l0 = mg.lv[0];
  {
    Timed t(c, KID_MG_SETUP);
    hipLaunchKernelGGL(k_geom0, dim3(nb((long)g.M * g.N)), dim3(kBSm), 0, st, g, c->d, l0.ax, l0.bx, l0.ay, l0.by,
                       l0.rd, l0.M + 2);
  }
  for (size_t l = 0; l + 1 < mg.lv.size(); ++l) {
    const MGLevel &f = mg.lv[l];
    MGLevel &q = mg.lv[l + 1];
    Timed t(c, KID_MG_SETUP, 8.0 * 5 * ((double)f.M * f.N + (double)q.M * q.N));
    hipLaunchKernelGGL(k_geom_coarsen, dim3(nb((long)q.M * q.N)), dim3(kBSm), 0, st, f.M, f.N, f.M + 2, q.M, q.N,
                       q.M + 2, f.fx, f.fy, f.ax, f.bx, f.ay, f.by, f.rd, q.ax, q.bx, q.ay, q.by, q.rd);
  }
  return hip_check(hipGetLastError(), "mg_setup");
}

// one zebra half sweep; algorithmic bytes: b and x of the swept columns, x
// of the other colour (unless zero), five column fields
template <int KM>
static void zebra_k(nhw_ctx *c, const LvlC &L, const G2 &G, const double *b, double *x, int color, bool zero,
                    const int *stop) {
  const long nth = (long)((L.M + 1) / 2) * L.N;
  const double n = (double)L.MN * L.K;
  Timed t(c, KID_MG_SMOOTH, 8.0 * ((zero ? 1.0 : 1.5) * n + 5.0 * L.MN));
  if (zero)
    hipLaunchKernelGGL((k_zebra<KM, true>), dim3(nb(nth)), dim3(kBSm), 0, c->stream, L, G, b, x, color, stop);
  else
    hipLaunchKernelGGL((k_zebra<KM, false>), dim3(nb(nth)), dim3(kBSm), 0, c->stream, L, G, b, x, color, stop);
}

static void smooth(nhw_ctx *c, const LvlC &L, const G2 &G, const double *b, double *x, int color, bool zero,
                   const int *stop) {
  if (L.K <= 4) zebra_k<4>(c, L, G, b, x, color, zero, stop);
  else if (L.K <= 8) zebra_k<8>(c, L, G, b, x, color, zero, stop);
  else zebra_k<16>(c, L, G, b, x, color, zero, stop);
}

// x = V(b): one V(1,1) cycle from a zero guess on level l
static void vcycle(nhw_ctx *c, size_t l, const double *b, double *x, const int *stop) {
  MG &mg = c->mg;
  const MGLevel &lv = mg.lv[l];
  const LvlC L = make_lvl(c, lv);
  const G2 G = g2(lv);
  hipStream_t st = c->stream;
  smooth(c, L, G, b, x, 0, true, stop);
  smooth(c, L, G, b, x, 1, false, stop);
  if (l + 1 == mg.lv.size()) {
    for (int s = 1; s < mg.coarse_sweeps; ++s) {
      smooth(c, L, G, b, x, 0, false, stop);
      smooth(c, L, G, b, x, 1, false, stop);
    }
    return;
  }
  const double n = (double)L.MN * L.K;
  {
    Timed t(c, KID_MG_RESID, 8.0 * (3.0 * n + 5.0 * L.MN));
    hipLaunchKernelGGL(k_mf_resid, dim3(nb(L.MN)), dim3(kBSm), 0, st, L, G, (const double *)x, b, lv.r, stop);
  }
  const MGLevel &q = mg.lv[l + 1];
  const double nc = (double)q.M * q.N * L.K;
  {
    Timed t(c, KID_MG_RESTRICT, 8.0 * (n + nc));
    hipLaunchKernelGGL(k_restrict, dim3(nb((long)q.M * q.N * L.K)), dim3(kBSm), 0, st, lv.M, lv.N, q.M, q.N, L.K,
                       lv.fx, lv.fy, (const double *)lv.r, q.b, stop);
  }
  vcycle(c, l + 1, q.b, q.x, stop);
  {
    Timed t(c, KID_MG_PROLONG, 8.0 * (2.0 * n + nc));
    hipLaunchKernelGGL(k_prolong_add, dim3(nb(L.MN * L.K)), dim3(kBSm), 0, st, lv.M, lv.N, q.M, q.N, L.K, lv.fx,
                       lv.fy, (const double *)q.x, x, stop);
  }
  smooth(c, L, G, b, x, 1, false, stop);
  smooth(c, L, G, b, x, 0, false, stop);
}

void mg_apply(nhw_ctx *c, const double *b, double *x, const int *stop) { vcycle(c, 0, b, x, stop); }

// y = A x on the fine level, with ndot dot partials (y,d1) [and (y,y)]
void mf_apply(nhw_ctx *c, const double *x, double *y, const double *d1, int ndot, double *part, const int *stop,
              int nblk, int kid) {
  const MGLevel &l0 = c->mg.lv[0];
  const LvlC L = make_lvl(c, l0);
  const G2 G = g2(l0);
  Timed t(c, kid);
  if (ndot == 0)
    hipLaunchKernelGGL(k_mf_apply<0>, dim3(nblk), dim3(kBSm), 0, c->stream, L, G, x, y, d1, part, stop);
  else if (ndot == 1)
    hipLaunchKernelGGL(k_mf_apply<1>, dim3(nblk), dim3(kBSm), 0, c->stream, L, G, x, y, d1, part, stop);
  else
    hipLaunchKernelGGL(k_mf_apply<2>, dim3(nblk), dim3(kBSm), 0, c->stream, L, G, x, y, d1, part, stop);
}

void mf_init(nhw_ctx *c, const double *x, const double *b, double *r, double *rh, double *p, double *v,
             double *part, int nblk) {
  const MGLevel &l0 = c->mg.lv[0];
  const LvlC L = make_lvl(c, l0);
  Timed t(c, KID_MF_RESID);
  hipLaunchKernelGGL(k_mf_init, dim3(nblk), dim3(kBSm), 0, c->stream, L, g2(l0), x, b, r, rh, p, v, part);
}

void mf_true_resid(nhw_ctx *c, const double *x, const double *b, double *part, int nblk) {
  const MGLevel &l0 = c->mg.lv[0];
  const LvlC L = make_lvl(c, l0);
  Timed t(c, KID_MF_RESID);
  hipLaunchKernelGGL(k_mf_true, dim3(nblk), dim3(kBSm), 0, c->stream, L, g2(l0), x, b, part);
}

}  // namespace nhw
