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
__global__ __launch_bounds__(kBSm) void k_mf_true(LvlC<double> L, G2<double> G, const double *__restrict__ x,
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
  V sub[KM], dia[KM], sup[KM], rhs[KM];
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
      sub[k] = c[12];
      dia[k] = c[0];
      sup[k] = c[5];
      rhs[k] = r;
    }
  }
  // Thomas: forward elimination then back substitution
  rhs[0] = rhs[0] / dia[0];
  sup[0] = sup[0] / dia[0];
#pragma unroll
  for (int k = 1; k < KM; ++k) {
    if (k < L.K) {
      const V inv = V(1) / (dia[k] - sub[k] * sup[k - 1]);
      sup[k] = sup[k] * inv;
      rhs[k] = (rhs[k] - sub[k] * rhs[k - 1]) * inv;
    }
  }
  V xn = V(0);
#pragma unroll
  for (int k = KM - 1; k >= 0; --k) {
    if (k < L.K) {
      xn = k == L.K - 1 ? rhs[k] : rhs[k] - sup[k] * xn;
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
  const std::vector<double> &dsig = c->mg.dsig_h, &sig = c->mg.sig_h;  // 1-based
  for (int kk = 0; kk < L.K; ++kk) {
    const int k = c->g.Kbeg + kk;
    const double s = dsig[k] + dsig[k - 1];
    L.sig[kk] = (V)sig[k];
    L.r6[kk] = (V)(1. / (0.5 * s * dsig[k]));
    L.r13[kk] = (V)(1. / (0.5 * s * dsig[k - 1]));
    L.rx[kk] = (V)(1. / (2. * dx * s));
    L.ry[kk] = (V)(1. / (2. * dy * s));
  }
  return L;
}

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
    if (maxlev > 1 && !fine)  // level 0 solves into the caller's vectors
      for (double **p : {&lv.x, &lv.b})
        ok = ok && hipMalloc(p, esize * n) == hipSuccess && hipMemset(*p, 0, esize * n) == hipSuccess;
    lvs.push_back(lv);
    if (!ok) {
      set_error("mg_alloc: hipMalloc failed");
      return NHW_ENOMEM;
    }
    const int fx = M >= 4 ? 2 : 1, fy = N >= 4 ? 2 : 1;
    if ((int)lvs.size() >= maxlev || (fx == 1 && fy == 1) || (long)M * N <= kMgCoarseCols) break;
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
  for (MGLevel &lv : lvs)
    for (double *p : {lv.ax, lv.bx, lv.ay, lv.by, lv.rd, lv.x, lv.b, lv.r})
      if (p) (void)hipFree(p);
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
  else zebra_k<V, 16>(c, L, G, b, x, color, zero, stop, kid);
}

// x = V(b): one V(1,1) cycle from a zero guess on level l; the coarsest
// level (at most kMgCoarseCols columns) gets coarse_sweeps zebra sweeps
template <class V>
static void vcycle(nhw_ctx *c, const std::vector<MGLevel> &lvs, size_t l, const V *b, V *x, const int *stop) {
  const MGLevel &lv = lvs[l];
  const LvlC<V> L = make_lvl<V>(c, lv);
  const G2<V> G = g2<V>(lv);
  hipStream_t st = c->stream;
  const double n = (double)L.MN * L.K;
  if (l + 1 == lvs.size()) {
    const int kid = l == 0 ? KID_MG_SMOOTH : KID_MG_COARSE;
    for (int sw = 0; sw < c->mg.coarse_sweeps; ++sw) {
      smooth(c, L, G, b, x, 0, sw == 0, stop, kid);
      smooth(c, L, G, b, x, 1, false, stop, kid);
    }
    return;
  }
  const int kid = l == 0 ? KID_MG_SMOOTH : KID_MG_SMOOTH_C;
  smooth(c, L, G, b, x, 0, true, stop, kid);
  smooth(c, L, G, b, x, 1, false, stop, kid);
  const MGLevel &q = lvs[l + 1];
  const double nc = (double)q.M * q.N * L.K;
  {
    Timed t(c, KID_MG_RESID, sizeof(V) * (2.0 * n + nc + 5.0 * L.MN));
    const dim3 grid(nb((long)q.M * q.N));
    if (lv.fx == 2 && lv.fy == 2) {
      const dim3 g22((lv.M + kRRx - 1) / kRRx, (lv.N + kRRy - 1) / kRRy);
      if (L.K <= 4)
        hipLaunchKernelGGL((k_resid_restrict22<V, 4>), g22, dim3(kRRx * kRRy), 0, st, L, G, (const V *)x, b, q.M,
                           q.N, (V *)q.b, stop);
      else if (L.K <= 8)
        hipLaunchKernelGGL((k_resid_restrict22<V, 8>), g22, dim3(kRRx * kRRy), 0, st, L, G, (const V *)x, b, q.M,
                           q.N, (V *)q.b, stop);
      else
        hipLaunchKernelGGL((k_resid_restrict22<V, 16>), g22, dim3(kRRx * kRRy), 0, st, L, G, (const V *)x, b, q.M,
                           q.N, (V *)q.b, stop);
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
  vcycle<V>(c, lvs, l + 1, (const V *)q.b, (V *)q.x, stop);
  {
    Timed t(c, KID_MG_PROLONG, sizeof(V) * (2.0 * n + nc));
    hipLaunchKernelGGL(k_prolong_add<V>, dim3(nb(L.MN * L.K)), dim3(kBSm), 0, st, lv.M, lv.N, q.M, q.N, L.K,
                       lv.fx, lv.fy, (const V *)q.x, x, stop);
  }
  smooth(c, L, G, b, x, 1, false, stop, kid);
  smooth(c, L, G, b, x, 0, false, stop, kid);
}

template <class V>
void mg_apply(nhw_ctx *c, const std::vector<MGLevel> &lvs, const V *b, V *x, const int *stop) {
  vcycle<V>(c, lvs, 0, b, x, stop);
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
