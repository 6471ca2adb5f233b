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

// element idx of an array at a uniform base, addressed through a 32-bit
// byte offset so the load takes the scalar-base + vector-offset form (one
// VGPR per address, not two).  Every sigma plane and every 2-D field stays
// below 4 GiB.
template <class V>
__device__ __forceinline__ V ldo(const V *__restrict__ base, unsigned idx) {
  return *(const V *)((const char *)base + (unsigned)(idx * (unsigned)sizeof(V)));
}
template <class V>
__device__ __forceinline__ void sto(V *__restrict__ base, unsigned idx, V v) {
  *(V *)((char *)base + (unsigned)(idx * (unsigned)sizeof(V))) = v;
}

template <class V>
__device__ __forceinline__ ColGeo<V> load_col(const G2<V> &G, const LvlC<V> &L, int i, int j) {
  ColGeo<V> g;
  const unsigned c = (unsigned)(i + 1) + (unsigned)G.ld * (unsigned)(j + 1), ld = (unsigned)G.ld;
  g.axc = ldo(G.ax, c); g.bxc = ldo(G.bx, c); g.ayc = ldo(G.ay, c); g.byc = ldo(G.by, c); g.rd = ldo(G.rd, c);
  g.axm = ldo(G.ax, c - 1); g.bxm = ldo(G.bx, c - 1); g.axp = ldo(G.ax, c + 1); g.bxp = ldo(G.bx, c + 1);
  g.aym = ldo(G.ay, c - ld); g.bym = ldo(G.by, c - ld); g.ayp = ldo(G.ay, c + ld); g.byp = ldo(G.by, c + ld);
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
  // folds as selects (branch-free, the same sums in the same order)
  const bool w = g.w, e = g.e, s = g.s, n = g.n;
  c[0] = w ? c[0] + c[8] : c[0]; c[5] = w ? c[5] + c[4] : c[5]; c[12] = w ? c[12] + c[13] : c[12];
  c[8] = w ? z : c[8]; c[4] = w ? z : c[4]; c[13] = w ? z : c[13];
  c[0] = e ? c[0] + c[1] : c[0]; c[5] = e ? c[5] + c[6] : c[5]; c[12] = e ? c[12] + c[11] : c[12];
  c[1] = e ? z : c[1]; c[6] = e ? z : c[6]; c[11] = e ? z : c[11];
  c[0] = s ? c[0] + c[9] : c[0]; c[5] = s ? c[5] + c[3] : c[5]; c[12] = s ? c[12] + c[14] : c[12];
  c[9] = s ? z : c[9]; c[3] = s ? z : c[3]; c[14] = s ? z : c[14];
  c[0] = n ? c[0] + c[2] : c[0]; c[5] = n ? c[5] + c[7] : c[5]; c[12] = n ? c[12] + c[10] : c[12];
  c[2] = n ? z : c[2]; c[7] = n ? z : c[7]; c[10] = n ? z : c[10];
  if (k == 0) {
    c[5] += c[12]; c[6] += c[11]; c[7] += c[10]; c[4] += c[13]; c[3] += c[14];
    c[12] = z; c[11] = z; c[10] = z; c[13] = z; c[14] = z;
  }
  const bool top = k == L.K - 1;
  c[3] = top ? z : c[3]; c[4] = top ? z : c[4]; c[5] = top ? z : c[5]; c[6] = top ? z : c[6];
  c[7] = top ? z : c[7];
}

// plane offsets (zebra layout) of a column and of its four x-y neighbours,
// clamped so every read is in bounds (a clamped neighbour always carries a
// folded, zero coefficient)
struct Nb {
  unsigned c, xm, xp, ym, yp;
};
template <class V>
__device__ __forceinline__ Nb nbrs(const LvlC<V> &L, int i, int j) {
  Nb o;
  o.c = (unsigned)zcol(i, j, L.M);
  o.xm = i > 0 ? (unsigned)zcol(i - 1, j, L.M) : o.c;
  o.xp = i < L.M - 1 ? (unsigned)zcol(i + 1, j, L.M) : o.c;
  o.ym = j > 0 ? (unsigned)zcol(i, j - 1, L.M) : o.c;
  o.yp = j < L.N - 1 ? (unsigned)zcol(i, j + 1, L.M) : o.c;
  return o;
}

// sigma plane k of a level vector (a uniform pointer)
template <class V, class P>
__device__ __forceinline__ P *plane(P *x, const LvlC<V> &L, int k) {
  return x + (long)k * L.MN;
}

// the five values of one sigma plane a row's stencil reads: the column and
// its four x-y neighbours (clamped offsets)
template <class V>
struct Pl {
  V c, xm, xp, ym, yp;
};
template <class V>
__device__ __forceinline__ Pl<V> load_pl(const V *__restrict__ xk, const Nb &o) {
  return Pl<V>{ldo(xk, o.c), ldo(xk, o.xm), ldo(xk, o.xp), ldo(xk, o.ym), ldo(xk, o.yp)};
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

// rows of y = A x for one column (runtime row loop), with the dot partials
template <class V, int NDOT>
__device__ __forceinline__ void mf_col(const LvlC<V> &L, const ColGeo<V> &g, const Nb &o, const V *__restrict__ x,
                                       V *__restrict__ y, const V *__restrict__ d1, double (&acc)[2]) {
  Pl<V> cu = load_pl(x, o);
  Pl<V> nx = L.K > 1 ? load_pl(plane(x, L, 1), o) : cu;
  Pl<V> pv = cu;
  for (int k = 0; k < L.K; ++k) {
    V c[15];
    coefs(L, g, k, c);
    const V v = arow3(c, pv, cu, nx);
    sto(plane(y, L, k), o.c, v);
    if (NDOT >= 1) acc[0] += (double)v * (double)ldo(plane(d1, L, k), o.c);
    if (NDOT >= 2) acc[1] += (double)v * (double)v;
    pv = cu;
    cu = nx;
    if (k + 2 < L.K) nx = load_pl(plane(x, L, k + 2), o);
  }
}

// y = A x; optional dot partials (y, d1) [and (y, y)], accumulated in FP64.
// One thread walks one column bottom to top, keeping the three sigma planes
// its rows read in registers (each x value is loaded once per column pass).
// A runtime row loop: few registers, so many waves hide the latency.
template <class V, int NDOT>
__global__ __launch_bounds__(kBSm) void k_mf_apply(LvlC<V> L, G2<V> G, const V *__restrict__ x, V *__restrict__ y,
                                                   const V *__restrict__ d1, double *part, const int *stop) {
  if (stop[0]) return;
  double acc[2] = {0.0, 0.0};
  for (long col = (long)blockIdx.x * kBSm + threadIdx.x; col < L.MN; col += (long)gridDim.x * kBSm) {
    const int i = (int)(col % L.M), j = (int)(col / L.M);
    const ColGeo<V> g = load_col(G, L, i, j);
    const Nb o = nbrs(L, i, j);
    mf_col<V, NDOT>(L, g, o, x, y, d1, acc);
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
    Pl<double> cu = load_pl(x, o);
    Pl<double> nx = L.K > 1 ? load_pl(plane(x, L, 1), o) : cu;
    Pl<double> pv = cu;
    for (int k = 0; k < L.K; ++k) {
      double c[15];
      coefs(L, g, k, c);
      const long q = (long)k * L.MN + o.c;
      const double bv = b[q];
      const double rv = bv - arow3(c, pv, cu, nx);
      r[q] = (R)rv; rh[q] = (R)rv; p[q] = R(0); v[q] = R(0);
      if (e) e[q] = R(0);
      acc[0] += rv * rv;
      acc[1] += bv * bv;
      pv = cu;
      cu = nx;
      if (k + 2 < L.K) nx = load_pl(plane(x, L, k + 2), o);
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
    Pl<double> cu = load_pl(x, o);
    Pl<double> nx = L.K > 1 ? load_pl(plane(x, L, 1), o) : cu;
    Pl<double> pv = cu;
    for (int k = 0; k < L.K; ++k) {
      double c[15];
      coefs(L, g, k, c);
      const double rv = ldo(plane(b, L, k), o.c) - arow3(c, pv, cu, nx);
      acc[0] += rv * rv;
      pv = cu;
      cu = nx;
      if (k + 2 < L.K) nx = load_pl(plane(x, L, k + 2), o);
    }
  }
  block_sum2(acc, part);
}

// A x at rows 0..KM-1 of column col, bottom to top: f(k, c, pv, cu, nx) gets
// row k's coefficients and the planes below / at / above it.  Plane reads go
// through offsets clamped to the top row, so every load is in bounds and
// none sits behind a branch (the loads of later rows issue while earlier
// rows compute); rows k >= K re-read the top plane and callers discard them.
template <class V, int KM, class F>
__device__ __forceinline__ void walk_col(const LvlC<V> &L, const ColGeo<V> &g, const Nb &o,
                                         const V *__restrict__ x, F &&f) {
  const int K1 = L.K - 1;
  Pl<V> pv = load_pl(x, o);
  Pl<V> cu = pv;
  Pl<V> nx = load_pl(plane(x, L, min(1, K1)), o);
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const Pl<V> nn = load_pl(plane(x, L, min(k + 2, K1)), o);
    V c[15];
    coefs(L, g, k, c);
    f(k, c, pv, cu, nx);
    pv = cu;
    cu = nx;
    nx = nn;
    // keep the loads of later planes in later iterations: two planes ahead
    // cover the latency, and hoisting all of them costs the occupancy
    __builtin_amdgcn_sched_barrier(0);
  }
}

// rows 0..KM-1 of a column vector at byte offset off (clamped like walk_col)
template <class V, int KM>
__device__ __forceinline__ void load_rows(const V *__restrict__ v, unsigned off, const LvlC<V> &L, V (&out)[KM]) {
  const int K1 = L.K - 1;
#pragma unroll
  for (int k = 0; k < KM; ++k) out[k] = ldo(plane(v, L, min(k, K1)), off);
}

// one step of the Thomas forward elimination of a column's tridiagonal
// (sub c[12], diagonal c[0], super c[5]; c[12] is zero on the bottom row)
template <class V>
__device__ __forceinline__ void thomas_fwd(const V (&c)[15], V r, V &cp, V &dp) {
  const V inv = V(1) / (c[0] - c[12] * cp);
  cp = c[5] * inv;
  dp = (r - c[12] * dp) * inv;
}

// z-line Gauss-Seidel update of column (i,j): solves its vertical tridiagonal
// with the couplings to the other columns moved to the right side; with ZERO
// the other columns are taken as zero.  Forward elimination runs as the rows
// are formed (only the modified super-diagonal and right side are kept), then
// back substitution stores the column.
template <class V, int KM, bool ZERO>
__device__ __forceinline__ void zebra_col(const LvlC<V> &L, const G2<V> &G, const V *__restrict__ b,
                                          V *__restrict__ x, int i, int j) {
  const long col = (long)i + (long)L.M * j;
  const ColGeo<V> g = load_col(G, L, i, j);
  const Nb o = nbrs(L, i, j);
  V r[KM], cp[KM];
  load_rows<V, KM>(b, o.c, L, r);
  V cprev = V(0), dprev = V(0);
  if (ZERO) {
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      V c[15];
      coefs(L, g, k, c);
      thomas_fwd(c, r[k], cprev, dprev);
      cp[k] = cprev;
      r[k] = dprev;
    }
  } else {
    // the column's own values are the unknowns: only its neighbours are read
    walk_col<V, KM>(L, g, o, x, [&](int k, const V (&c)[15], const Pl<V> &pv, const Pl<V> &cu, const Pl<V> &nx) {
      const V rk = r[k] - (c[1] * cu.xp + c[2] * cu.yp + c[8] * cu.xm + c[9] * cu.ym + c[3] * nx.ym + c[4] * nx.xm +
                           c[6] * nx.xp + c[7] * nx.yp + c[10] * pv.yp + c[11] * pv.xp + c[13] * pv.xm +
                           c[14] * pv.ym);
      thomas_fwd(c, rk, cprev, dprev);
      cp[k] = cprev;
      r[k] = dprev;
    });
  }
  const int K1 = L.K - 1;
  V xn = V(0);
#pragma unroll
  for (int k = KM - 1; k >= 0; --k) {
    xn = k >= K1 ? r[k] : r[k] - cp[k] * xn;
    if (k < L.K) sto(plane(x, L, k), o.c, xn);
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
      V bv[KM];
      load_rows<V, KM>(b, o.c, L, bv);
      walk_col<V, KM>(L, g, o, x,
                      [&](int k, const V (&c)[15], const Pl<V> &pv, const Pl<V> &cu, const Pl<V> &nx) {
                        s[k] += bv[k] - arow3(c, pv, cu, nx);
                      });
      ++cnt;
    }
  const V w = V(1) / (V)cnt;
  const long zc = zcol(I, J, Mc);
#pragma unroll
  for (int k = 0; k < KM; ++k)
    if (k < L.K) bc[zc + (long)k * MNc] = s[k] * w;
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
    V bv[KM];
    load_rows<V, KM>(b, o.c, L, bv);
    walk_col<V, KM>(L, g, o, x, [&](int k, const V (&c)[15], const Pl<V> &pv, const Pl<V> &cu, const Pl<V> &nx) {
      r[k] = bv[k] - arow3(c, pv, cu, nx);
    });
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
        bc[zcol(I, J, Mc) + (long)k * MNc] =
            (sm[ty][tx] + sm[ty][tx + 1] + sm[ty + 1][tx] + sm[ty + 1][tx + 1]) / (V)cnt;
      __syncthreads();
    }
  }
}

// bilinear (cell-centred) prolongation of the coarse correction at fine
// column (i,j) of plane k (zebra layout on both levels)
template <class V>
__device__ __forceinline__ V prolong_at(const V *__restrict__ xck, int Mc, int Nc, int fx, int fy, int i, int j) {
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
  return wy0 * (wx0 * xck[zcol(I0, J0, Mc)] + wx1 * xck[zcol(I1, J0, Mc)]) +
         wy1 * (wx0 * xck[zcol(I0, J1, Mc)] + wx1 * xck[zcol(I1, J1, Mc)]);
}

// xf += P xc, one thread per fine row in storage order
template <class V>
__global__ void k_prolong_add(int Mf, int Nf, int Mc, int Nc, int K, int fx, int fy, const V *__restrict__ xc,
                              V *__restrict__ xf, const int *stop) {
  if (stop[0]) return;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long MNc = (long)Mc * Nc, MNf = (long)Mf * Nf;
  if (t >= MNf * K) return;
  const int k = (int)(t / MNf);
  int i, j;
  zinv(t - (long)k * MNf, Mf, i, j);
  xf[t] += prolong_at(xc + MNc * k, Mc, Nc, fx, fy, i, j);
}

}  // namespace nhw
