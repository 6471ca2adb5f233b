// Device side of the matrix-free sigma-coordinate Poisson operator and its
// geometric multigrid preconditioner (see nhw_mg.hip for the method), shared
// by the single-scenario solver and the batched one.
#pragma once
#include <type_traits>

#include "nhw_impl.h"

namespace nhw {

constexpr int kBSm = 256;

// calls f(std::integral_constant<int, KM>) with the smallest row bucket KM in
// {4, 8, ..., KMX} that holds K rows: kernels shared by members of different
// K unroll only as far as each member needs (K is uniform per workgroup)
template <int KMX, class F>
__device__ __forceinline__ void by_rows(int K, F &&f) {
  if constexpr (KMX > 4) {
    if (K <= KMX - 4) {
      by_rows<KMX - 4>(K, f);
      return;
    }
  }
  f(std::integral_constant<int, KMX>());
}

// per-sigma-level constants of one multigrid level, in the vector type: a
// small device table read with scalar loads where a row needs them (kernel
// arguments would hold all of them in scalar registers at once)
template <class V>
struct LvlK {
  V sig, r6, r13, rx, ry;
};
// per-level constants, in the vector type
template <class V>
struct LvlC {
  int M, N, K;
  long MN;
  V cxy, idx2, idy2;
  const LvlK<V> *kc;  // [K]
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
  const LvlK<V> kk = L.kc[k];
  const V sg = kk.sig;
  const V dxs = g.axc - sg * g.bxc, dxm = g.axm - sg * g.bxm, dxp = g.axp - sg * g.bxp;
  const V dys = g.ayc - sg * g.byc, dym = g.aym - sg * g.bym, dyp = g.ayp - sg * g.byp;
  const V A1 = dxs * dxs + dys * dys + g.rd * g.rd;
  const V z = V(0);
  c[0] = L.cxy + A1 * (kk.r6 + kk.r13);
  c[1] = -L.idx2; c[8] = -L.idx2;
  c[2] = -L.idy2; c[9] = -L.idy2;
  c[3] = (dym + dys) * kk.ry; c[14] = -c[3];
  c[4] = (dxm + dxs) * kk.rx; c[13] = -c[4];
  c[5] = -A1 * kk.r6; c[12] = -A1 * kk.r13;
  c[6] = -(dxp + dxs) * kk.rx; c[11] = -c[6];
  c[7] = -(dyp + dys) * kk.ry; c[10] = -c[7];
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

// Per-column constants of the smoother's and the restriction's rows: the
// operator of coefs(), re-associated for fewer operations (neighbour sums
// formed once per column, the x/y folds of the reference turned into
// per-column terms, fused multiply-adds).  Rounding differs from coefs() in
// the last bits, which the multigrid preconditioner tolerates; the Krylov
// operator itself (mf_col) keeps coefs().
template <class V>
struct ColC {
  V axc, bxc, ayc, byc, rd2, c0b;
  V Axm, Bxm, Axp, Bxp, Aym, Bym, Ayp, Byp;  // neighbour + column sums, 0 on a folded side
  V FAx, FBx, FAy, FBy;                      // the folded sides' sums, moved onto the column
  V hxm, hxp, hym, hyp;                      // same-plane couplings, 0 on a folded side
};
template <class V>
__device__ __forceinline__ ColC<V> col_consts(const G2<V> &G, const LvlC<V> &L, int i, int j) {
  const ColGeo<V> g = load_col(G, L, i, j);
  ColC<V> q;
  q.axc = g.axc; q.bxc = g.bxc; q.ayc = g.ayc; q.byc = g.byc; q.rd2 = g.rd * g.rd;
  const V o = V(1), z = V(0);
  const V kxm = g.w ? z : o, kxp = g.e ? z : o, kym = g.s ? z : o, kyp = g.n ? z : o;
  const V Axm = g.axm + g.axc, Bxm = g.bxm + g.bxc, Axp = g.axp + g.axc, Bxp = g.bxp + g.bxc;
  const V Aym = g.aym + g.ayc, Bym = g.bym + g.byc, Ayp = g.ayp + g.ayc, Byp = g.byp + g.byc;
  q.Axm = kxm * Axm; q.Bxm = kxm * Bxm; q.Axp = kxp * Axp; q.Bxp = kxp * Bxp;
  q.Aym = kym * Aym; q.Bym = kym * Bym; q.Ayp = kyp * Ayp; q.Byp = kyp * Byp;
  // west / east fold c4 / c6 onto c5 (and their negatives onto c12), south /
  // north fold c3 / c7; c4 = (Axm - sg Bxm) rx, c6 = -(Axp - sg Bxp) rx,
  // c3 = (Aym - sg Bym) ry, c7 = -(Ayp - sg Byp) ry
  q.FAx = (o - kxm) * Axm - (o - kxp) * Axp; q.FBx = (o - kxm) * Bxm - (o - kxp) * Bxp;
  q.FAy = (o - kym) * Aym - (o - kyp) * Ayp; q.FBy = (o - kym) * Bym - (o - kyp) * Byp;
  q.c0b = L.cxy - ((o - kxm) + (o - kxp)) * L.idx2 - ((o - kym) + (o - kyp)) * L.idy2;
  q.hxm = -kxm * L.idx2; q.hxp = -kxp * L.idx2; q.hym = -kym * L.idy2; q.hyp = -kyp * L.idy2;
  return q;
}

// row k from per-column constants: centre c0, couplings to the rows above
// (c5) and below (c12) in the column, and (OFF) the sum s of the couplings
// to the other columns from the planes below / at / above the row
template <class V>
struct ZRow {
  V c0, c5, c12, s;
};
template <class V, bool OFF>
__device__ __forceinline__ ZRow<V> zrow(const LvlC<V> &L, const ColC<V> &q, int k, const Pl<V> &pv,
                                        const Pl<V> &cu, const Pl<V> &nx) {
  const LvlK<V> e = L.kc[k];
  const V sg = e.sig, rxk = e.rx, ryk = e.ry, r6 = e.r6, r13 = e.r13;
  const V dxs = fma(-sg, q.bxc, q.axc), dys = fma(-sg, q.byc, q.ayc);
  const V A1 = fma(dxs, dxs, fma(dys, dys, q.rd2));
  const V F = fma(rxk, fma(-sg, q.FBx, q.FAx), ryk * fma(-sg, q.FBy, q.FAy));
  const bool top = k == L.K - 1;
  ZRow<V> z;
  z.c0 = fma(A1, r6 + r13, q.c0b);
  V c5 = fma(-A1, r6, F), c12 = fma(-A1, r13, -F);
  if (k == 0) {  // bottom: the ghost row below mirrors the row above
    c5 += c12;
    c12 = V(0);
  }
  z.c5 = top ? V(0) : c5;  // top: zero pressure above
  z.c12 = c12;
  z.s = V(0);
  if (OFF) {
    V s = fma(q.hxp, cu.xp, fma(q.hxm, cu.xm, fma(q.hyp, cu.yp, q.hym * cu.ym)));
    if (k > 0) {  // the diagonal couplings cancel on the bottom row
      const V tm = top ? V(0) : V(1);
      const V c4 = rxk * fma(-sg, q.Bxm, q.Axm), c6 = -rxk * fma(-sg, q.Bxp, q.Axp);
      const V c3 = ryk * fma(-sg, q.Bym, q.Aym), c7 = -ryk * fma(-sg, q.Byp, q.Ayp);
      s = fma(c4, fma(tm, nx.xm, -pv.xm), s);
      s = fma(c6, fma(tm, nx.xp, -pv.xp), s);
      s = fma(c3, fma(tm, nx.ym, -pv.ym), s);
      s = fma(c7, fma(tm, nx.yp, -pv.yp), s);
    }
    z.s = s;
  }
  return z;
}

// (A x) at a row from its zrow and planes
template <class V>
__device__ __forceinline__ V zrow_apply(const ZRow<V> &z, const Pl<V> &pv, const Pl<V> &cu, const Pl<V> &nx) {
  return fma(z.c0, cu.c, fma(z.c5, nx.c, fma(z.c12, pv.c, z.s)));
}

// reciprocal for the Thomas pivots: the hardware estimate refined by Newton
// steps (to about an ulp; the pivots are positive and far from the range
// ends)
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float rcp_nr(float d) {
  float r = __builtin_amdgcn_rcpf(d);
  return fmaf(r, fmaf(-d, r, 1.0f), r);
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
#pragma unroll 1
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
  for (long col = (long)xcd_block() * kBSm + threadIdx.x; col < L.MN; col += (long)gridDim.x * kBSm) {
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
  for (long col = (long)xcd_block() * kBSm + threadIdx.x; col < L.MN; col += (long)gridDim.x * kBSm) {
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
  for (long col = (long)xcd_block() * kBSm + threadIdx.x; col < L.MN; col += (long)gridDim.x * kBSm) {
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

// rows 0..KM-1 of a column, bottom to top: f(k, z, pv, cu, nx) gets row k
// (zrow) and the planes below / at / above it.  Plane reads go
// through offsets clamped to the top row, so every load is in bounds and
// none sits behind a branch (the loads of later rows issue while earlier
// rows compute); rows k >= K re-read the top plane and callers discard them.
template <class V, int KM, class F>
__device__ __forceinline__ void walk_col(const LvlC<V> &L, const ColC<V> &q, const Nb &o,
                                         const V *__restrict__ x, F &&f) {
  const int K1 = L.K - 1;
  Pl<V> pv = load_pl(x, o);
  Pl<V> cu = pv;
  Pl<V> nx = load_pl(plane(x, L, min(1, K1)), o);
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const Pl<V> nn = load_pl(plane(x, L, min(k + 2, K1)), o);
    f(k, zrow<V, true>(L, q, k, pv, cu, nx), pv, cu, nx);
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
// (sub c12, diagonal c0, super c5; c12 is zero on the bottom row)
template <class V>
__device__ __forceinline__ void thomas_fwd(const ZRow<V> &z, V r, V &cp, V &dp) {
  const V inv = rcp_nr(fma(-z.c12, cp, z.c0));
  cp = z.c5 * inv;
  dp = fma(-z.c12, dp, r) * inv;
}

// z-line Gauss-Seidel update of column (i,j): solves its vertical tridiagonal
// with the couplings to the other columns moved to the right side; with ZERO
// the other columns are taken as zero.  Forward elimination runs as the rows
// are formed (only the modified super-diagonal and right side are kept), then
// back substitution stores the column.
// back substitution of a column's eliminated tridiagonal into x at plane
// offset off
template <class V, int KM>
__device__ __forceinline__ void col_back(const LvlC<V> &L, const V (&cp)[KM], const V (&r)[KM], V *__restrict__ x,
                                         unsigned off) {
  const int K1 = L.K - 1;
  V xn = V(0);
#pragma unroll
  for (int k = KM - 1; k >= 0; --k) {
    xn = k >= K1 ? r[k] : fma(-cp[k], xn, r[k]);
    if (k < L.K) sto(plane(x, L, k), off, xn);
  }
}

// the zero-guess solve of a column (the other columns taken as zero) with
// its constants q and right side r in registers; r is overwritten, the
// solution stored at plane offset off
template <class V, int KM>
__device__ __forceinline__ void zero_solve(const LvlC<V> &L, const ColC<V> &q, V (&r)[KM], V *__restrict__ x,
                                           unsigned off) {
  V cp[KM];
  V cprev = V(0), dprev = V(0);
  const Pl<V> none{};
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    thomas_fwd(zrow<V, false>(L, q, k, none, none, none), r[k], cprev, dprev);
    cp[k] = cprev;
    r[k] = dprev;
  }
  col_back<V, KM>(L, cp, r, x, off);
}

// the Gauss-Seidel solve of a column against its neighbours' current values
// in x; r is overwritten
template <class V, int KM>
__device__ __forceinline__ void line_solve(const LvlC<V> &L, const ColC<V> &q, const Nb &o, V (&r)[KM],
                                           V *__restrict__ x) {
  V cp[KM];
  V cprev = V(0), dprev = V(0);
  // the column's own values are the unknowns: only its neighbours are read
  walk_col<V, KM>(L, q, o, x, [&](int k, const ZRow<V> &z, const Pl<V> &, const Pl<V> &, const Pl<V> &) {
    thomas_fwd(z, r[k] - z.s, cprev, dprev);
    cp[k] = cprev;
    r[k] = dprev;
  });
  col_back<V, KM>(L, cp, r, x, o.c);
}

// the zero-guess solve of column (i,j) with its right side r in registers
template <class V, int KM>
__device__ __forceinline__ void zero_col_solve(const LvlC<V> &L, const G2<V> &G, int i, int j, V (&r)[KM],
                                               V *__restrict__ x) {
  zero_solve<V, KM>(L, col_consts(G, L, i, j), r, x, (unsigned)zcol(i, j, L.M));
}

template <class V, int KM, bool ZERO>
__device__ __forceinline__ void zebra_col(const LvlC<V> &L, const G2<V> &G, const V *__restrict__ b,
                                          V *__restrict__ x, int i, int j) {
  const Nb o = nbrs(L, i, j);
  V r[KM];
  load_rows<V, KM>(b, o.c, L, r);
  if (ZERO) zero_solve<V, KM>(L, col_consts(G, L, i, j), r, x, o.c);
  else line_solve<V, KM>(L, col_consts(G, L, i, j), o, r, x);
}

// zebra half sweep over the columns of one colour ((i+j) parity); columns of
// one colour share no stencil entry, so they update in parallel
template <class V, int KM, bool ZERO>
__global__ __launch_bounds__(kBSm) void k_zebra(LvlC<V> L, G2<V> G, const V *__restrict__ b, V *__restrict__ x,
                                                int color, const int *stop) {
  if (stop[0]) return;
  const int half = (L.M + 1) / 2;
  long t = (long)xcd_block() * kBSm + threadIdx.x;
  if (t >= (long)half * L.N) return;
  const int j = (int)(t / half);
  const int i = 2 * (int)(t % half) + ((j + color) & 1);
  if (i >= L.M) return;
  zebra_col<V, KM, ZERO>(L, G, b, x, i, j);
}

// The coarsest level (at most kMgCoarseCols columns) solved approximately
// inside one workgroup: `sweeps` zebra pairs from a zero guess with x kept in
// LDS between half sweeps, so the whole solve is one launch.  The arithmetic
// is that of the zebra kernels: the result equals `sweeps` pairs of k_zebra
// launches.
// Each thread owns at most one column of each colour, whose offsets,
// constants and right side it forms once, before the sweeps.
constexpr int kCoarseBS = 512;
template <class V, int KM>
__device__ __forceinline__ void coarse_solve(const LvlC<V> &L, const G2<V> &G, const V *__restrict__ b,
                                             V *__restrict__ x, int sweeps, V *xs /* LDS, KM * kMgCoarseCols */) {
  const int n = (int)L.MN * L.K;
  // clamped neighbours read the column's own (not yet written) value with a
  // zero coefficient: start from zeros so that product is zero
  for (int q = threadIdx.x; q < n; q += kCoarseBS) xs[q] = V(0);
  // thread t takes the t-th column of each colour (at most kCoarseBS of
  // each, as the level has at most 2 kCoarseBS columns): a pair of rows
  // holds M columns of each colour
  bool on[2];
  Nb o[2];
  ColC<V> q[2];
  V rb[2][KM];
  const int t = threadIdx.x;
#pragma unroll
  for (int colour = 0; colour < 2; ++colour) {
    const int P = t / L.M, rem = t - P * L.M;
    const int h0 = (L.M + 1) >> 1, h = colour == 0 ? h0 : L.M - h0;  // in row 2P
    int j = rem < h ? 2 * P : 2 * P + 1, p = rem < h ? rem : rem - h;
    int i = 2 * p + ((j + colour) & 1);
    on[colour] = t < (int)(L.MN + 1) / 2 && j < L.N && i < L.M;
    if (!on[colour]) i = j = 0;  // formed but never used
    o[colour] = nbrs(L, i, j);
    q[colour] = col_consts(G, L, i, j);
    load_rows<V, KM>(b, o[colour].c, L, rb[colour]);
  }
  __syncthreads();
  for (int sw = 0; sw < sweeps; ++sw) {
#pragma unroll
    for (int colour = 0; colour < 2; ++colour) {
      if (on[colour]) {
        V r[KM];
#pragma unroll
        for (int k = 0; k < KM; ++k) r[k] = rb[colour][k];
        if (sw == 0 && colour == 0) zero_solve<V, KM>(L, q[colour], r, xs, o[colour].c);
        else line_solve<V, KM>(L, q[colour], o[colour], r, xs);
      }
      __syncthreads();
    }
  }
  for (int r = threadIdx.x; r < n; r += kCoarseBS) x[r] = xs[r];
}

template <class V, int KM>
__global__ __launch_bounds__(kCoarseBS) void k_coarse(LvlC<V> L, G2<V> G, const V *__restrict__ b,
                                                     V *__restrict__ x, int sweeps, const int *stop) {
  if (stop[0]) return;
  __shared__ V xs[KM * kMgCoarseCols];
  coarse_solve<V, KM>(L, G, b, x, sweeps, xs);
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
  long cc = (long)xcd_block() * kBSm + threadIdx.x;
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
      const ColC<V> q = col_consts(G, L, i, j);
      const Nb o = nbrs(L, i, j);
      V bv[KM];
      load_rows<V, KM>(b, o.c, L, bv);
      walk_col<V, KM>(L, q, o, x, [&](int k, const ZRow<V> &z, const Pl<V> &pv, const Pl<V> &cu, const Pl<V> &nx) {
        s[k] += bv[k] - zrow_apply(z, pv, cu, nx);
      });
      ++cnt;
    }
  const V w = V(1) / (V)cnt;
  const long zc = zcol(I, J, Mc);
#pragma unroll
  for (int k = 0; k < KM; ++k)
    if (k < L.K) bc[zc + (long)k * MNc] = s[k] * w;
}

// The same for 2x2 coarsening right after the pre-smoothing, whose last
// half sweep solved the odd columns ((i+j) odd) exactly: their residual is
// zero, so only the two even columns (2I,2J) and (2I+1,2J+1) of coarse cell
// (I,J) are formed.  One thread per coarse cell; the even columns of rows 2J
// and 2J+1 sit at plane offsets 2J*M + I and (2J+1)*M + I in the zebra
// layout, so neighbouring threads read neighbouring values.
template <class V, int KM>
__device__ __forceinline__ void resid22_body(const LvlC<V> &L, const G2<V> &G, const V *__restrict__ x,
                                             const V *__restrict__ b, int Mc, int Nc, V *__restrict__ bc, long t) {
  if (t >= (long)Mc * Nc) return;
  const int I = (int)(t % Mc), J = (int)(t / Mc);
  const int i0 = 2 * I, j0 = 2 * J;
  const bool ix = i0 + 1 < L.M, jy = j0 + 1 < L.N;
  const int cnt = 1 + ix + jy + (ix && jy);  // fine cells of the block
  V s[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) s[k] = V(0);
  for (int e = 0; e < (ix && jy ? 2 : 1); ++e) {
    const int i = i0 + e, j = j0 + e;
    const ColC<V> q = col_consts(G, L, i, j);
    const Nb o = nbrs(L, i, j);
    V bv[KM];
    load_rows<V, KM>(b, o.c, L, bv);
    walk_col<V, KM>(L, q, o, x, [&](int k, const ZRow<V> &z, const Pl<V> &pv, const Pl<V> &cu, const Pl<V> &nx) {
      s[k] += bv[k] - zrow_apply(z, pv, cu, nx);
    });
  }
  const long MNc = (long)Mc * Nc, zc = zcol(I, J, Mc);
#pragma unroll
  for (int k = 0; k < KM; ++k)
    if (k < L.K) bc[zc + (long)k * MNc] = s[k] / (V)cnt;
}

template <class V, int KM>
__global__ __launch_bounds__(kBSm) void k_resid_restrict22(LvlC<V> L, G2<V> G, const V *__restrict__ x,
                                                           const V *__restrict__ b, int Mc, int Nc,
                                                           V *__restrict__ bc, const int *stop) {
  if (stop[0]) return;
  resid22_body<V, KM>(L, G, x, b, Mc, Nc, bc, (long)xcd_block() * kBSm + threadIdx.x);
}

// resid22_body with the two fine columns of a coarse cell on neighbouring
// lanes (lane pair t, t^1), whose partial sums meet through a shuffle: the
// sum is formed as (0 + r0) + r1, the order of the single-lane loop, so the
// coarse right side is bit-identical.  With ZSWEEP the lane pair then
// starts the coarse level's V-cycle: the zero-guess half sweep over its
// colour-0 columns needs nothing but each column's own right side, which is
// in registers here, so those columns are solved (zero_col_solve) and the
// coarse level skips its first half sweep.
template <class V, int KM, bool ZSWEEP>
__device__ __forceinline__ void resid22_pair_body(const LvlC<V> &L, const G2<V> &G, const V *__restrict__ x,
                                                  const V *__restrict__ b, const LvlC<V> &Lc, const G2<V> &Gc,
                                                  V *__restrict__ bc, V *__restrict__ xc, long t) {
  const long cc = t >> 1;
  const int e = (int)(t & 1);
  if (cc >= Lc.MN) return;  // both lanes of a pair leave together
  const int I = (int)(cc % Lc.M), J = (int)(cc / Lc.M);
  const int i0 = 2 * I, j0 = 2 * J;
  const bool ix = i0 + 1 < L.M, jy = j0 + 1 < L.N;
  const int cnt = 1 + ix + jy + (ix && jy);  // fine cells of the block
  const bool live = e == 0 || (ix && jy);    // a lane without a column walks (i0, j0) and adds zero
  const int i = live ? i0 + e : i0, j = live ? j0 + e : j0;
  V a[KM];
  {
    const ColC<V> q = col_consts(G, L, i, j);
    const Nb o = nbrs(L, i, j);
    V bv[KM];
    load_rows<V, KM>(b, o.c, L, bv);
    walk_col<V, KM>(L, q, o, x, [&](int k, const ZRow<V> &z, const Pl<V> &pv, const Pl<V> &cu, const Pl<V> &nx) {
      const V r = bv[k] - zrow_apply(z, pv, cu, nx);
      a[k] = e == 0 ? V(0) + r : (live ? r : V(0));
    });
  }
  V s[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) s[k] = (a[k] + __shfl_xor(a[k], 1, 64)) / (V)cnt;
  if (e != 0) return;
  const unsigned zc = (unsigned)zcol(I, J, Lc.M);
#pragma unroll
  for (int k = 0; k < KM; ++k)
    if (k < L.K) sto(plane(bc, Lc, k), zc, s[k]);
  if (ZSWEEP && ((I + J) & 1) == 0) zero_col_solve<V, KM>(Lc, Gc, I, J, s, xc);
}

template <class V, int KM, bool ZSWEEP>
__global__ __launch_bounds__(kBSm) void k_resid22_pair(LvlC<V> L, G2<V> G, const V *__restrict__ x,
                                                       const V *__restrict__ b, LvlC<V> Lc, G2<V> Gc,
                                                       V *__restrict__ bc, V *__restrict__ xc, const int *stop) {
  if (stop[0]) return;
  resid22_pair_body<V, KM, ZSWEEP>(L, G, x, b, Lc, Gc, bc, xc, (long)xcd_block() * kBSm + threadIdx.x);
}

// bilinear (cell-centred) prolongation weights and coarse plane offsets
// (zebra layout) of fine column (i,j)
template <class V>
struct Pro {
  unsigned c00, c10, c01, c11;
  V wx0, wx1, wy0, wy1;
};
template <class V>
__device__ __forceinline__ Pro<V> pro_at(int Mc, int Nc, int fx, int fy, int i, int j) {
  Pro<V> w;
  int I0 = fx == 2 ? i >> 1 : i, I1 = I0;
  w.wx0 = V(1); w.wx1 = V(0);
  if (fx == 2) {
    I1 = (i & 1) ? I0 + 1 : I0 - 1;
    if (I1 < 0 || I1 >= Mc) I1 = I0;
    w.wx0 = V(0.75); w.wx1 = V(0.25);
  }
  int J0 = fy == 2 ? j >> 1 : j, J1 = J0;
  w.wy0 = V(1); w.wy1 = V(0);
  if (fy == 2) {
    J1 = (j & 1) ? J0 + 1 : J0 - 1;
    if (J1 < 0 || J1 >= Nc) J1 = J0;
    w.wy0 = V(0.75); w.wy1 = V(0.25);
  }
  w.c00 = (unsigned)zcol(I0, J0, Mc); w.c10 = (unsigned)zcol(I1, J0, Mc);
  w.c01 = (unsigned)zcol(I0, J1, Mc); w.c11 = (unsigned)zcol(I1, J1, Mc);
  return w;
}

// xf += P xc on the fine columns with (i+j) even.  The post-smoothing that
// follows sweeps the odd columns first, which reads only even neighbours,
// and then overwrites the even ones: the odd columns' corrections are never
// read, so they are not formed.  One thread per even column (the first half
// of each zebra row), walking its sigma levels.
template <class V>
__device__ __forceinline__ void prolong_col(int Mf, int Nf, int Mc, int Nc, int K, int fx, int fy,
                                            const V *__restrict__ xc, V *__restrict__ xf, long t) {
  const int half = (Mf + 1) / 2;
  if (t >= (long)half * Nf) return;
  const int p = (int)(t % half), j = (int)(t / half);
  if (p >= (Mf + 1 - (j & 1)) >> 1) return;
  const int i = 2 * p + (j & 1);
  const Pro<V> w = pro_at<V>(Mc, Nc, fx, fy, i, j);
  const long MNc = (long)Mc * Nc, MNf = (long)Mf * Nf;
  const unsigned f = (unsigned)((long)j * Mf + p);
  for (int k = 0; k < K; ++k) {
    const V *xk = xc + MNc * k;
    V *yk = xf + MNf * k;
    const V v = w.wy0 * (w.wx0 * ldo(xk, w.c00) + w.wx1 * ldo(xk, w.c10)) +
                w.wy1 * (w.wx0 * ldo(xk, w.c01) + w.wx1 * ldo(xk, w.c11));
    sto(yk, f, ldo((const V *)yk, f) + v);
  }
}

template <class V>
__global__ void k_prolong_add(int Mf, int Nf, int Mc, int Nc, int K, int fx, int fy, const V *__restrict__ xc,
                              V *__restrict__ xf, const int *stop) {
  if (stop[0]) return;
  prolong_col<V>(Mf, Nf, Mc, Nc, K, fx, fy, xc, xf, (long)xcd_block() * blockDim.x + threadIdx.x);
}

}  // namespace nhw
