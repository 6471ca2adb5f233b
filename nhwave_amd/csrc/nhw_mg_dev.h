// Device side of the matrix-free sigma-coordinate Poisson operator and its
// geometric multigrid preconditioner (see nhw_mg.hip for the method), shared
// by the single-scenario solver and the batched one.
#pragma once
#include "nhw_impl.h"

namespace nhw {

constexpr int kBSm = 256;
constexpr long kMgCoarseCols = 1024;  // coarsest level: at most this many columns

// per-level constants, in the vector type
template <class V>
struct LvlC {
  int M, N, K;
  long MN;
  V cxy, idx2, idy2;
  V sig[kMgKMax], r6[kMgKMax], r13[kMgKMax], rx[kMgKMax], ry[kMgKMax];
};

template <class V>
struct G2 {
  const V *ax, *bx, *ay, *by, *rd;
  int ld;  // (M+2): halo layout, interior (i,j) at (i+1) + ld*(j+1)
  __device__ long at(int i, int j) const { return (long)(i + 1) + (long)ld * (j + 1); }
};

template <class V>
struct ColGeo {
  V axc, bxc, ayc, byc, rd, axm, bxm, axp, bxp, aym, bym, ayp, byp;
  bool w, e, s, n;
};

template <class V>
__device__ __forceinline__ ColGeo<V> load_col(const G2<V> &G, const LvlC<V> &L, int i, int j) {
  ColGeo<V> g;
  const long c = G.at(i, j);
  g.axc = G.ax[c]; g.bxc = G.bx[c]; g.ayc = G.ay[c]; g.byc = G.by[c]; g.rd = G.rd[c];
  g.axm = G.ax[c - 1]; g.bxm = G.bx[c - 1]; g.axp = G.ax[c + 1]; g.bxp = G.bx[c + 1];
  g.aym = G.ay[c - G.ld]; g.bym = G.by[c - G.ld]; g.ayp = G.ay[c + G.ld]; g.byp = G.by[c + G.ld];
  g.w = i == 0; g.e = i == L.M - 1; g.s = j == 0; g.n = j == L.N - 1;
  return g;
}

// the 15 coefficients of row (i,j,k) in JCoef order (c[0] = centre), folds
// applied in the reference's order (west, east, south, north, bottom, top),
// src/pressure.F:540-555, 649-814
template <class V>
__device__ __forceinline__ void coefs(const LvlC<V> &L, const ColGeo<V> &g, int k, V c[15]) {
  const V sg = L.sig[k];
  const V dxs = g.axc - sg * g.bxc, dxm = g.axm - sg * g.bxm, dxp = g.axp - sg * g.bxp;
  const V dys = g.ayc - sg * g.byc, dym = g.aym - sg * g.bym, dyp = g.ayp - sg * g.byp;
  const V A1 = dxs * dxs + dys * dys + g.rd * g.rd;
  const V z = V(0);
  c[0] = L.cxy + A1 * (L.r6[k] + L.r13[k]);
  c[1] = -L.idx2; c[8] = -L.idx2;
  c[2] = -L.idy2; c[9] = -L.idy2;
  c[3] = (dym + dys) * L.ry[k]; c[14] = -c[3];
  c[4] = (dxm + dxs) * L.rx[k]; c[13] = -c[4];
  c[5] = -A1 * L.r6[k]; c[12] = -A1 * L.r13[k];
  c[6] = -(dxp + dxs) * L.rx[k]; c[11] = -c[6];
  c[7] = -(dyp + dys) * L.ry[k]; c[10] = -c[7];
  if (g.w) { c[0] += c[8]; c[5] += c[4]; c[12] += c[13]; c[8] = z; c[4] = z; c[13] = z; }
  if (g.e) { c[0] += c[1]; c[5] += c[6]; c[12] += c[11]; c[1] = z; c[6] = z; c[11] = z; }
  if (g.s) { c[0] += c[9]; c[5] += c[3]; c[12] += c[14]; c[9] = z; c[3] = z; c[14] = z; }
  if (g.n) { c[0] += c[2]; c[5] += c[7]; c[12] += c[10]; c[2] = z; c[7] = z; c[10] = z; }
  if (k == 0) {
    c[5] += c[12]; c[6] += c[11]; c[7] += c[10]; c[4] += c[13]; c[3] += c[14];
    c[12] = z; c[11] = z; c[10] = z; c[13] = z; c[14] = z;
  }
  if (k == L.K - 1) { c[3] = z; c[4] = z; c[5] = z; c[6] = z; c[7] = z; }
}

// neighbour offsets of one column, clamped so every read is in bounds; a
// clamped neighbour always carries a folded (zero) coefficient
struct Nb {
  long xm, xp, ym, yp;
};
template <class V>
__device__ __forceinline__ Nb nbrs(const LvlC<V> &L, int i, int j) {
  Nb o;
  o.xm = i > 0 ? -1 : 0;
  o.xp = i < L.M - 1 ? 1 : 0;
  o.ym = j > 0 ? -(long)L.M : 0;
  o.yp = j < L.N - 1 ? (long)L.M : 0;
  return o;
}

// sum of the 12 couplings to other columns for row k (x at base = column)
template <class V>
__device__ __forceinline__ V offcol(const V *__restrict__ x, long base, long MN, int k, int K, const Nb &o,
                                    const V c[15]) {
  const long q = base + (long)k * MN;
  const long up = k < K - 1 ? MN : 0, dn = k > 0 ? -MN : 0;
  return c[1] * x[q + o.xp] + c[2] * x[q + o.yp] + c[3] * x[q + o.ym + up] + c[4] * x[q + o.xm + up] +
         c[6] * x[q + o.xp + up] + c[7] * x[q + o.yp + up] + c[8] * x[q + o.xm] + c[9] * x[q + o.ym] +
         c[10] * x[q + o.yp + dn] + c[11] * x[q + o.xp + dn] + c[13] * x[q + o.xm + dn] +
         c[14] * x[q + o.ym + dn];
}

// (A x) at row k of column col
template <class V>
__device__ __forceinline__ V arow(const V *__restrict__ x, long col, const LvlC<V> &L, int k, const Nb &o,
                                  const V c[15]) {
  const long q = col + (long)k * L.MN;
  const long up = k < L.K - 1 ? L.MN : 0, dn = k > 0 ? -L.MN : 0;
  return c[0] * x[q] + c[5] * x[q + up] + c[12] * x[q + dn] + offcol<V>(x, col, L.MN, k, L.K, o, c);
}

// the five values of one sigma plane a row's stencil reads: the column and
// its four x-y neighbours (clamped offsets)
template <class V>
struct Pl {
  V c, xm, xp, ym, yp;
};
template <class V>
__device__ __forceinline__ Pl<V> load_pl(const V *__restrict__ x, long q, const Nb &o) {
  return Pl<V>{x[q], x[q + o.xm], x[q + o.xp], x[q + o.ym], x[q + o.yp]};
}
// (A x) at a row from the planes below (pv), at (cu) and above (nx) it; at
// the bottom and top rows the missing plane carries zero coefficients
template <class V>
__device__ __forceinline__ V arow3(const V c[15], const Pl<V> &pv, const Pl<V> &cu, const Pl<V> &nx) {
  return c[0] * cu.c + c[5] * nx.c + c[12] * pv.c + c[1] * cu.xp + c[2] * cu.yp + c[8] * cu.xm + c[9] * cu.ym +
         c[3] * nx.ym + c[4] * nx.xm + c[6] * nx.xp + c[7] * nx.yp + c[10] * pv.yp + c[11] * pv.xp +
         c[13] * pv.xm + c[14] * pv.ym;
}

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

// ---------------------------------------------------------------------------
// geometry set-up (always from the FP64 state)
// ---------------------------------------------------------------------------
template <class V>
__global__ void k_geom0(Geo g, Dev d, V *ax, V *bx, V *ay, V *by, V *rd, int ld) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.M * g.N) return;
  const int i = (int)(t % g.M), j = (int)(t / g.M);
  const long c = g.c2(g.Ibeg + i, g.Jbeg + j);
  const long h = (long)(i + 1) + (long)ld * (j + 1);
  const double r = 1.0 / d.D[c];
  ax[h] = (V)(d.DelxH[c] * r);
  bx[h] = (V)((d.DelxH[c] + d.DelxEta[c]) * r);
  ay[h] = (V)(d.DelyH[c] * r);
  by[h] = (V)((d.DelyH[c] + d.DelyEta[c]) * r);
  rd[h] = (V)r;
}

// block average of the five column fields onto the next level
template <class V>
__global__ void k_geom_coarsen(int Mf, int Nf, int ldf, int Mc, int Nc, int ldc, int fx, int fy, const V *axf,
                               const V *bxf, const V *ayf, const V *byf, const V *rdf, V *axc, V *bxc, V *ayc,
                               V *byc, V *rdc) {
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
  axc[h] = (V)(s[0] * w); bxc[h] = (V)(s[1] * w); ayc[h] = (V)(s[2] * w); byc[h] = (V)(s[3] * w);
  rdc[h] = (V)(s[4] * w);
}

// ---------------------------------------------------------------------------
// operator kernels (one thread per column, k loop inside)
// ---------------------------------------------------------------------------

// y = A x; optional dot partials (y, d1) [and (y, y)], accumulated in FP64.
// One thread walks one column bottom to top, keeping the three sigma planes
// its rows read in registers (each x value is loaded once per column pass).
template <class V, int NDOT>
__global__ __launch_bounds__(kBSm) void k_mf_apply(LvlC<V> L, G2<V> G, const V *__restrict__ x, V *__restrict__ y,
                                                   const V *__restrict__ d1, double *part, const int *stop) {
  if (stop[0]) return;
  double acc[2] = {0.0, 0.0};
  for (long col = (long)blockIdx.x * kBSm + threadIdx.x; col < L.MN; col += (long)gridDim.x * kBSm) {
    const int i = (int)(col % L.M), j = (int)(col / L.M);
    const ColGeo<V> g = load_col(G, L, i, j);
    const Nb o = nbrs(L, i, j);
    Pl<V> cu = load_pl(x, col, o);
    Pl<V> nx = L.K > 1 ? load_pl(x, col + L.MN, o) : cu;
    Pl<V> pv = cu;
    for (int k = 0; k < L.K; ++k) {
      V c[15];
      coefs(L, g, k, c);
      const long q = col + (long)k * L.MN;
      const V v = arow3(c, pv, cu, nx);
      y[q] = v;
      if (NDOT >= 1) acc[0] += (double)v * (double)d1[q];
      if (NDOT >= 2) acc[1] += (double)v * (double)v;
      pv = cu;
      cu = nx;
      if (k + 2 < L.K) nx = load_pl(x, q + 2 * L.MN, o);
    }
  }
  if (NDOT > 0) block_sum2(acc, part);
}

// start of a (sub)solve, in the FP64 operator: res = b - A x with x and b
// FP64; r = rh = res in the Krylov type R, p = v = 0, and e = 0 when given
// (the inner correction of the mixed form); partials (res,res), (b,b)
template <class R>
__global__ __launch_bounds__(kBSm) void k_mf_init(LvlC<double> L, G2<double> G, const double *__restrict__ x,
                                                  const double *__restrict__ b, R *__restrict__ r,
                                                  R *__restrict__ rh, R *__restrict__ p, R *__restrict__ v,
                                                  R *__restrict__ e, double *part) {
  double acc[2] = {0.0, 0.0};
  for (long col = (long)blockIdx.x * kBSm + threadIdx.x; col < L.MN; col += (long)gridDim.x * kBSm) {
    const int i = (int)(col % L.M), j = (int)(col / L.M);
    const ColGeo<double> g = load_col(G, L, i, j);
    const Nb o = nbrs(L, i, j);
    Pl<double> cu = load_pl(x, col, o);
    Pl<double> nx = L.K > 1 ? load_pl(x, col + L.MN, o) : cu;
    Pl<double> pv = cu;
    for (int k = 0; k < L.K; ++k) {
      double c[15];
      coefs(L, g, k, c);
      const long q = col + (long)k * L.MN;
      const double bv = b[q];
      const double rv = bv - arow3(c, pv, cu, nx);
      r[q] = (R)rv; rh[q] = (R)rv; p[q] = R(0); v[q] = R(0);
      if (e) e[q] = R(0);
      acc[0] += rv * rv;
      acc[1] += bv * bv;
      pv = cu;
      cu = nx;
      if (k + 2 < L.K) nx = load_pl(x, q + 2 * L.MN, o);
    }
  }
  block_sum2(acc, part);
}

// |b - A x|^2 partials for the solve report (FP64 operator)
static __global__ __launch_bounds__(kBSm) void k_mf_true(LvlC<double> L, G2<double> G, const double *__restrict__ x,
                                                  const double *__restrict__ b, double *part) {
  double acc[2] = {0.0, 0.0};
  for (long col = (long)blockIdx.x * kBSm + threadIdx.x; col < L.MN; col += (long)gridDim.x * kBSm) {
    const int i = (int)(col % L.M), j = (int)(col / L.M);
    const ColGeo<double> g = load_col(G, L, i, j);
    const Nb o = nbrs(L, i, j);
    Pl<double> cu = load_pl(x, col, o);
    Pl<double> nx = L.K > 1 ? load_pl(x, col + L.MN, o) : cu;
    Pl<double> pv = cu;
    for (int k = 0; k < L.K; ++k) {
      double c[15];
      coefs(L, g, k, c);
      const long q = col + (long)k * L.MN;
      const double rv = b[q] - arow3(c, pv, cu, nx);
      acc[0] += rv * rv;
      pv = cu;
      cu = nx;
      if (k + 2 < L.K) nx = load_pl(x, q + 2 * L.MN, o);
    }
  }
  block_sum2(acc, part);
}

// z-line Gauss-Seidel update of column (i,j): solves its vertical tridiagonal
// (c13, c1, c6) with the couplings to the other columns moved to the right
// side; with ZERO the other columns are taken as zero
template <class V, int KM, bool ZERO>
__device__ __forceinline__ void zebra_col(const LvlC<V> &L, const G2<V> &G, const V *__restrict__ b,
                                          V *__restrict__ x, int i, int j) {
  const long col = (long)i + (long)L.M * j;
  const ColGeo<V> g = load_col(G, L, i, j);
  const Nb o = nbrs(L, i, j);
  // Thomas forward elimination as the rows are formed (only the modified
  // super-diagonal and right side are kept), then back substitution
  V cp[KM], dp[KM];
  V cprev = V(0), dprev = V(0);
  // neighbour planes below / at / above row k (the column's own values are
  // not read: they are the unknowns)
  Pl<V> cu{}, nx{}, pv{};
  if (!ZERO) {
    cu = load_pl(x, col, o);
    nx = L.K > 1 ? load_pl(x, col + L.MN, o) : cu;
    pv = cu;
  }
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    if (k < L.K) {
      V c[15];
      coefs(L, g, k, c);
      const long q = col + (long)k * L.MN;
      V r = b[q];
      if (!ZERO) {
        r -= c[1] * cu.xp + c[2] * cu.yp + c[8] * cu.xm + c[9] * cu.ym + c[3] * nx.ym + c[4] * nx.xm +
             c[6] * nx.xp + c[7] * nx.yp + c[10] * pv.yp + c[11] * pv.xp + c[13] * pv.xm + c[14] * pv.ym;
        pv = cu;
        cu = nx;
        if (k + 2 < L.K) nx = load_pl(x, q + 2 * L.MN, o);
      }
      // c[12] (sub-diagonal) is zero on the bottom row
      const V inv = V(1) / (c[0] - c[12] * cprev);
      cprev = c[5] * inv;
      dprev = (r - c[12] * dprev) * inv;
      cp[k] = cprev;
      dp[k] = dprev;
    }
  }
  V xn = V(0);
#pragma unroll
  for (int k = KM - 1; k >= 0; --k) {
    if (k < L.K) {
      xn = k == L.K - 1 ? dp[k] : dp[k] - cp[k] * xn;
      x[col + (long)k * L.MN] = xn;
    }
  }
}

// zebra half sweep over the columns of one colour ((i+j) parity); columns of
// one colour share no stencil entry, so they update in parallel
template <class V, int KM, bool ZERO>
__global__ __launch_bounds__(kBSm) void k_zebra(LvlC<V> L, G2<V> G, const V *__restrict__ b, V *__restrict__ x,
                                                int color, const int *stop) {
  if (stop[0]) return;
  const int half = (L.M + 1) / 2;
  long t = (long)blockIdx.x * kBSm + threadIdx.x;
  if (t >= (long)half * L.N) return;
  const int j = (int)(t / half);
  const int i = 2 * (int)(t % half) + ((j + color) & 1);
  if (i >= L.M) return;
  zebra_col<V, KM, ZERO>(L, G, b, x, i, j);
}

// residual of a level restricted to the next coarser level's right side in
// one pass: one thread per coarse column (I,J) walks its fx*fy fine columns
// bottom to top, forms b - A x there and stores the block means
// (block-average restriction) for every sigma level
template <class V, int KM>
__global__ __launch_bounds__(kBSm) void k_resid_restrict(LvlC<V> L, G2<V> G, const V *__restrict__ x,
                                                         const V *__restrict__ b, int Mc, int Nc, int fx, int fy,
                                                         V *__restrict__ bc, const int *stop) {
  if (stop[0]) return;
  long cc = (long)blockIdx.x * kBSm + threadIdx.x;
  const long MNc = (long)Mc * Nc;
  if (cc >= MNc) return;
  const int I = (int)(cc % Mc), J = (int)(cc / Mc);
  V s[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) s[k] = V(0);
  int cnt = 0;
  for (int bb = 0; bb < fy; ++bb)
    for (int a = 0; a < fx; ++a) {
      const int i = I * fx + a, j = J * fy + bb;
      if (i >= L.M || j >= L.N) continue;
      const long col = (long)i + (long)L.M * j;
      const ColGeo<V> g = load_col(G, L, i, j);
      const Nb o = nbrs(L, i, j);
      Pl<V> cu = load_pl(x, col, o);
      Pl<V> nx = L.K > 1 ? load_pl(x, col + L.MN, o) : cu;
      Pl<V> pv = cu;
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        if (k < L.K) {
          V c[15];
          coefs(L, g, k, c);
          const long q = col + (long)k * L.MN;
          s[k] += b[q] - arow3(c, pv, cu, nx);
          pv = cu;
          cu = nx;
          if (k + 2 < L.K) nx = load_pl(x, q + 2 * L.MN, o);
        }
      }
      ++cnt;
    }
  const V w = V(1) / (V)cnt;
#pragma unroll
  for (int k = 0; k < KM; ++k)
    if (k < L.K) bc[cc + (long)k * MNc] = s[k] * w;
}

// The same for 2x2 coarsening with one thread per fine column: a 64 x 4
// tile of fine columns per block walks its columns, and the 2x2 block means
// are formed through LDS, one sigma level at a time.
constexpr int kRRx = 64, kRRy = 4;
template <class V, int KM>
__global__ __launch_bounds__(kRRx *kRRy) void k_resid_restrict22(LvlC<V> L, G2<V> G, const V *__restrict__ x,
                                                                 const V *__restrict__ b, int Mc, int Nc,
                                                                 V *__restrict__ bc, const int *stop) {
  if (stop[0]) return;
  __shared__ V sm[kRRy][kRRx];
  const int tx = threadIdx.x % kRRx, ty = threadIdx.x / kRRx;
  const int i = blockIdx.x * kRRx + tx, j = blockIdx.y * kRRy + ty;
  const bool valid = i < L.M && j < L.N;
  V r[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) r[k] = V(0);
  if (valid) {
    const long col = (long)i + (long)L.M * j;
    const ColGeo<V> g = load_col(G, L, i, j);
    const Nb o = nbrs(L, i, j);
    Pl<V> cu = load_pl(x, col, o);
    Pl<V> nx = L.K > 1 ? load_pl(x, col + L.MN, o) : cu;
    Pl<V> pv = cu;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      if (k < L.K) {
        V c[15];
        coefs(L, g, k, c);
        const long q = col + (long)k * L.MN;
        r[k] = b[q] - arow3(c, pv, cu, nx);
        pv = cu;
        cu = nx;
        if (k + 2 < L.K) nx = load_pl(x, q + 2 * L.MN, o);
      }
    }
  }
  // the coarse cell of an even (tx, ty) thread and its fine-cell count
  const int I = i >> 1, J = j >> 1;
  const bool owner = ((tx | ty) & 1) == 0 && valid;
  int cnt = 0;
  if (owner) cnt = 1 + (i + 1 < L.M) + (j + 1 < L.N) + ((i + 1 < L.M) && (j + 1 < L.N));
  const long MNc = (long)Mc * Nc;
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    if (k < L.K) {
      sm[ty][tx] = r[k];
      __syncthreads();
      if (owner)
        bc[(long)I + (long)Mc * J + (long)k * MNc] =
            (sm[ty][tx] + sm[ty][tx + 1] + sm[ty + 1][tx] + sm[ty + 1][tx + 1]) / (V)cnt;
      __syncthreads();
    }
  }
}

// bilinear (cell-centred) prolongation of the coarse correction, added
template <class V>
__global__ void k_prolong_add(int Mf, int Nf, int Mc, int Nc, int K, int fx, int fy, const V *__restrict__ xc,
                              V *__restrict__ xf, const int *stop) {
  if (stop[0]) return;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long MNc = (long)Mc * Nc, MNf = (long)Mf * Nf;
  if (t >= MNf * K) return;
  const int k = (int)(t / MNf);
  const long cf = t % MNf;
  const int i = (int)(cf % Mf), j = (int)(cf / Mf);
  int I0 = i / fx, I1 = I0;
  V wx0 = V(1), wx1 = V(0);
  if (fx == 2) {
    I1 = (i & 1) ? I0 + 1 : I0 - 1;
    if (I1 < 0 || I1 >= Mc) I1 = I0;
    wx0 = V(0.75); wx1 = V(0.25);
  }
  int J0 = j / fy, J1 = J0;
  V wy0 = V(1), wy1 = V(0);
  if (fy == 2) {
    J1 = (j & 1) ? J0 + 1 : J0 - 1;
    if (J1 < 0 || J1 >= Nc) J1 = J0;
    wy0 = V(0.75); wy1 = V(0.25);
  }
  const V *xk = xc + MNc * k;
  const V v = wy0 * (wx0 * xk[I0 + (long)Mc * J0] + wx1 * xk[I1 + (long)Mc * J0]) +
              wy1 * (wx0 * xk[I0 + (long)Mc * J1] + wx1 * xk[I1 + (long)Mc * J1]);
  xf[t] += v;
}

}  // namespace nhw
