// Explicit part of the NHWAVE stage on gfx950: TVD reconstruction + HLL
// fluxes (x, y and sigma sweeps), the mass/momentum stage update, wet-dry
// masking, dt, velocity boundary conditions and the sigma velocity Omega.
//
// Arithmetic follows the reference expression by expression (left-to-right,
// binary32-rounded literals widened to binary64) and the library is built
// with -ffp-contract=off, so FP64 results are bit-identical to the serial
// reference on the same inputs.  Work that the reference spreads over many
// whole-array passes (15 slope arrays, ~40 face-state arrays, flux_bc) is
// fused: each face thread reconstructs its two states from the cell fields,
// applies the wet-dry overrides, wave speeds, HLL and flux_bc in registers,
// and writes only the four final face fluxes.
#include <utility>
#include <vector>

#include "nhw_batch.h"
#include "nhw_impl.h"

namespace nhw {

// ----------------------------------------------------------------------------
// device helpers
// ----------------------------------------------------------------------------

// LIMITER (van Leer), src/nhwave.F:2889-2906
template <class T>
__device__ __forceinline__ T limiter(T a, T b) {
  return (a * fabs(b) + fabs(a) * b) / (fabs(a) + fabs(b));
}

// One cell of delxFun_3D/delyFun_3D (src/nhwave.F:2097-2311): limited slope
// of q at a cell from its two neighbours, zero in dry cells.
template <class T>
__device__ __forceinline__ T tvd_slope(T qm, T q0, T qp, T h, int mask) {
  if (mask == 0) return T(0);
  T t1 = (qp - q0) / h;
  T t2 = (q0 - qm) / h;
  if ((fabs(t1) + fabs(t2)) < T(kSmall)) return T(0);
  return limiter(t1, t2);
}

// One cell of delzFun_3D (src/nhwave.F:2314-2351), interior k.
template <class T>
__device__ __forceinline__ T z_slope(T qm, T q0, T qp, T hp, T hm) {
  T t1 = (qp - q0) / hp;
  T t2 = (q0 - qm) / hm;
  if ((fabs(t1) + fabs(t2)) < T(kSmall)) return T(0);
  return limiter(t1, t2);
}

// HLL, src/nhwave.F:2741-2776
__device__ __forceinline__ double hll(double sl, double sr, double fl, double fr, double ul, double ur) {
  if (sl >= 0.0) return fl;
  if (sr <= 0.0) return fr;
  double f = sr * fl - sl * fr + sl * sr * (ur - ul);
  if (fabs(sr - sl) < kSmall) return f / kSmall;
  return f / (sr - sl);
}

// hydrostatic pressure term 0.5*Grav*(eta*eta+2*eta*h) (src/nhwave.F:1793-1794)
__device__ __forceinline__ double pres(double e, double h) { return 0.5 * kGrav * (e * e + 2.0 * e * h); }

__device__ __forceinline__ bool is_wall(int bc) { return bc == 1 || bc == 2 || bc == 5; }


__device__ __forceinline__ double dmaxd(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double dmind(double a, double b) { return b < a ? b : a; }

// wave speeds at one face, src/nhwave.F:2813-2886.  Returns false when both
// sides are dry (the reference leaves the speeds of such faces untouched and
// flux_bc overwrites every flux there).
__device__ __forceinline__ bool wave_speeds(int ml, int mr, double DL, double DR, double uL, double uR,
                                            double &sL, double &sR) {
  if (ml == 1 && mr == 1) {
    double pl = sqrt(kGrav * fabs(DL));
    double pr = sqrt(kGrav * fabs(DR));
    double ps = 0.5 * (pl + pr) + 0.25 * (uL - uR);
    double us = 0.5 * (uL + uR) + pl - pr;
    sL = dmind(uL - pl, us - ps);
    sR = dmaxd(uR + pr, us + ps);
    return true;
  } else if (ml == 0 && mr == 1) {
    double pr = sqrt(kGrav * fabs(DR));
    sL = uR - 2.0 * pr;
    sR = uR + pr;
    return true;
  } else if (ml == 1 && mr == 0) {
    double pl = sqrt(kGrav * fabs(DL));
    sL = uL - pl;
    sR = uL + 2.0 * pl;
    return true;
  }
  return false;
}

inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ----------------------------------------------------------------------------
// fluxes: x-faces.  delxyzFun + construction (x part) + wave_speed +
// fluxes_at_faces_HLL + flux_bc, src/nhwave.F:1686-2906.
//
// The limited slope of a cell (tvd_slope) is shared by the face west of it
// (right state) and the face east of it (left state), and its second
// difference quotient t2 is the first one, t1, of the cell before.  So each
// lane of a wavefront owns one cell of a row: it forms t1 of its cell (one
// division), takes t2 from the lane below, forms the slope (the limiter's
// division) and both face states, and takes the left state of its face from
// the lane below -- two divisions per cell and field where a thread per face
// makes six.  The same expressions are evaluated as in tvd_slope, so the
// states are bit-identical.  Lanes 0 and 1 of a wavefront only supply the
// cells west of its first face: a wavefront covers kXFaces faces of one row.
// Each lane walks the sigma levels of its face column, so the 2-D part of the
// face (Eta, total depth, wet-dry overrides, Hfx0, wave-speed roots) is
// formed once per column.
// ----------------------------------------------------------------------------
// limited slope of a cell from its difference quotients t1 (forward) and t2
// (backward): tvd_slope after its two divisions
template <class T>
__device__ __forceinline__ T slope_of(T t1, T t2, int mask) {
  if (mask == 0) return T(0);
  if ((fabs(t1) + fabs(t2)) < T(kSmall)) return T(0);
  return limiter(t1, t2);
}

// the value of the lane below (lane 0 gets its own)
template <class T>
__device__ __forceinline__ T from_below(T v) {
  return __shfl_up(v, 1u, 64);
}

// wave_speeds with the roots sqrt(Grav |D|) of both sides formed beforehand
__device__ __forceinline__ bool wave_speeds_r(int ml, int mr, double pl, double pr, double uL, double uR,
                                              double &sL, double &sR) {
  if (ml == 1 && mr == 1) {
    double ps = 0.5 * (pl + pr) + 0.25 * (uL - uR);
    double us = 0.5 * (uL + uR) + pl - pr;
    sL = dmind(uL - pl, us - ps);
    sR = dmaxd(uR + pr, us + ps);
    return true;
  } else if (ml == 0 && mr == 1) {
    sL = uR - 2.0 * pr;
    sR = uR + pr;
    return true;
  } else if (ml == 1 && mr == 0) {
    sL = uL - pl;
    sR = uL + 2.0 * pl;
    return true;
  }
  return false;
}

template <class R>
__device__ __forceinline__ void flux_x_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  const int wpr = (g.M + 1 + kXFaces - 1) / kXFaces;  // wavefronts per row
  const long w = t >> 6;
  if (w >= (long)wpr * g.N) return;  // whole wavefronts leave together
  const int lane = (int)(t & 63);
  const int j = g.Jbeg + (int)(w / wpr);
  const int i = g.Ibeg - 2 + (int)(w % wpr) * kXFaces + lane;  // the lane's cell, and its west face
  const int ic = min(i, g.Iend + 1);                           // loads read cells ic and ic + 1
  const int im = max(ic - 1, 1);
  const bool face = lane >= 2 && i <= g.Iend1;
  const double dx = g.dx, hdx = 0.5 * dx, MinDep = p.min_dep;
  const long c20 = g.c2(ic, j), c2m1 = g.c2(im, j);
  const int ml = d.Mask[c2m1], mr = d.Mask[c20], mrr = d.Mask[c20 + 1];

  // reconstruction: the lane's cell gives the right state of its face and
  // the left state of the next one
  const double em1 = d.Eta[c2m1], e0 = d.Eta[c20];
  double etaL, etaR;
  {
    const double t1 = (d.Eta[c20 + 1] - e0) / dx;
    const double s = slope_of(t1, from_below(t1), mr);
    etaR = e0 - hdx * s;
    etaL = from_below(e0 + hdx * s);
  }
  const double hf = d.Hfx[g.x2(ic, j)];
  double DL = etaL + hf, DR = etaR + hf;
  DL = dmaxd(DL, MinDep);
  DR = dmaxd(DR, MinDep);

  // wet-dry front overrides, src/nhwave.F:1745-1788 (cell i-1 owns the left
  // state of face i, cell i the right one; only interior cells override)
  const bool leftOv = (i > g.Ibeg) && (ml == 0 || mr == 0);
  const bool rightOv = (i <= g.Iend) && (mr == 0 || (mrr != 0 && ml == 0));
  double h0 = d.Hfx0[g.x2(ic, j)];
  const bool mutL = (i > g.Ibeg) && ml == 1 && mr == 0;              // case 1 of cell i-1
  const bool mutR = (i <= g.Iend) && mr == 1 && mrr != 0 && ml == 0; // case 2 of cell i
  if (mutL) h0 = d.Hc[c2m1];
  else if (mutR) h0 = d.Hc[c20];
  if (face && (mutL || mutR)) d.Hfx0[g.x2(i, j)] = h0;
  if (leftOv) { etaL = em1; DL = d.D[c2m1]; }
  if (rightOv) { etaR = e0; DR = d.D[c20]; }
  const double PL = pres(etaL, h0), PR = pres(etaR, h0);
  const double pl = sqrt(kGrav * fabs(DL)), pr = sqrt(kGrav * fabs(DR));

  const R dxR = (R)dx, hdxR = (R)hdx;
  for (int k = g.Kbeg; k <= g.Kend; ++k) {
    const long c0 = g.c3(ic, j, k);
    auto recon = [&](const R *F, R &qL, R &qR) {
      const R c = F[c0];
      const R t1 = (F[c0 + 1] - c) / dxR;
      const R s = slope_of<R>(t1, from_below(t1), mr);
      qR = c - hdxR * s;
      qL = from_below(c + hdxR * s);
    };
    R uL, uR, vL, vR, wL, wR, duL, duR, dvL, dvR, dwL, dwR;
    recon(d.U, uL, uR);
    recon(d.V, vL, vR);
    recon(d.W, wL, wR);
    recon(d.DU, duL, duR);
    recon(d.DV, dvL, dvR);
    recon(d.DW, dwL, dwR);
    if (leftOv) { uL = 0.0; vL = 0.0; wL = 0.0; duL = 0.0; dvL = 0.0; dwL = 0.0; }
    if (rightOv) { uR = 0.0; vR = 0.0; wR = 0.0; duR = 0.0; dvR = 0.0; dwR = 0.0; }

    double E = 0.0, F = 0.0, G = 0.0, H = 0.0, sL, sR;
    if (wave_speeds_r(ml, mr, pl, pr, uL, uR, sL, sR)) {
      E = hll(sL, sR, duL, duR, DL, DR);
      F = hll(sL, sR, duL * uL + PL, duR * uR + PR, duL, duR);
      G = hll(sL, sR, DL * uL * vL, DR * uR * vR, dvL, dvR);
      H = hll(sL, sR, DL * uL * wL, DR * uR * wR, dwL, dwR);
    }
    // flux_bc, src/nhwave.F:2354-2552 (inflow at the west, :2381-2386)
    if (i == g.Ibeg && is_wall(p.bc_x0)) { E = 0.0; F = PR; G = 0.0; H = 0.0; }
    if (i == g.Ibeg && p.bc_x0 == 3) {
      const long b = inflow_idx(g, j, k);
      const double din = d.Din0[j - 1], ein = d.Ein0[j - 1];
      const double uin = d.Uin0[b], vin = d.Vin0[b], win = d.Win0[b];
      E = din * uin;
      F = din * uin * uin + pres(ein, h0);
      G = din * uin * vin;
      H = din * uin * win;
    }
    if (i == g.Iend1 && is_wall(p.bc_xn)) { E = 0.0; F = PL; G = 0.0; H = 0.0; }
    if (i == g.Iend1 && p.bc_xn == 3) {  // FLUX_LR's open east boundary, :2411-2416
      const long b = inflow_idx(g, j, k);
      const double din = d.Din_Xn[j - 1], ein = d.Ein_Xn[j - 1];
      const double uin = d.Uin_Xn[b], vin = d.Vin_Xn[b], win = d.Win_Xn[b];
      E = din * uin;
      F = din * uin * uin + pres(ein, h0);
      G = din * uin * vin;
      H = din * uin * win;
    }
    if (ml == 0) { E = 0.0; G = 0.0; H = 0.0; F = (i - 1 == g.Iend) ? 0.0 : PR * (double)mr; }
    if (mr == 0) { E = 0.0; G = 0.0; H = 0.0; F = (i == g.Ibeg) ? 0.0 : PL * (double)ml; }
    if (face) {
      const long f = g.x3(i, j, k);
      d.Ex[f] = E; d.Fx[f] = F; d.Gx[f] = G; d.Hx[f] = H;
    }
  }
}

// fluxes: y-faces, the y half of the same routines.  A thread takes one
// (i, k) and walks kYFaces faces of increasing j: the cell entering at each
// step gives the right state of its south face and the left state of its
// north face, its backward difference quotient is the previous cell's
// forward one, and every value is loaded once.  Threads run along i, so each
// step's loads are contiguous rows.
template <class R>
__device__ __forceinline__ void flux_y_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  if (t >= flux_y_threads(g)) return;
  const int i = g.Ibeg + (int)(t % g.M);
  const long r = t / g.M;
  const int k = g.Kbeg + (int)(r % g.K);
  const int jf0 = g.Jbeg + (int)(r / g.K) * kYFaces;
  const int jf1 = min(jf0 + kYFaces - 1, g.Jend1);
  const double dy = g.dy, hdy = 0.5 * dy, MinDep = p.min_dep;
  const R dyR = (R)dy, hdyR = (R)hdy;
  const long sj = g.Mloc;
  const R *const Fl[6] = {d.U, d.V, d.W, d.DU, d.DV, d.DW};
  // per field: the current cell's value, its backward difference quotient
  // (the previous cell's forward one) and the left state of its south face
  R q[6], t2[6], qL[6];
  double eq, et2, eL;
  // prime with cells jf0-2 and jf0-1: the left state of face jf0
  long c2 = g.c2(i, jf0 - 2), c3 = g.c3(i, jf0 - 2, k);
  {
    const double en = d.Eta[c2 + sj];
    et2 = (en - d.Eta[c2]) / dy;
    eq = en;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const R vn = Fl[n][c3 + sj];
      t2[n] = (vn - Fl[n][c3]) / dyR;
      q[n] = vn;
    }
    c2 += sj;
    c3 += sj;
    const int m = d.Mask[c2];
    const double en2 = d.Eta[c2 + sj];
    const double t1 = (en2 - eq) / dy;
    eL = eq + hdy * slope_of(t1, et2, m);
    et2 = t1;
    eq = en2;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const R vn = Fl[n][c3 + sj];
      const R t1r = (vn - q[n]) / dyR;
      qL[n] = q[n] + hdyR * slope_of<R>(t1r, t2[n], m);
      t2[n] = t1r;
      q[n] = vn;
    }
    c2 += sj;
    c3 += sj;
  }
  int ml = d.Mask[c2 - sj], mr = d.Mask[c2];
  double em1 = d.Eta[c2 - sj];
  // row j+1 of the walk, loaded one step ahead (the loads of a step then
  // overlap the arithmetic of the one before)
  int mnx = d.Mask[c2 + sj];
  double enx = d.Eta[c2 + sj];
  R fnx[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) fnx[n] = Fl[n][c3 + sj];
  double hfnx = d.Hfy[g.y2(i, jf0)], h0nx = d.Hfy0[g.y2(i, jf0)];
  for (int j = jf0; j <= jf1; ++j, c2 += sj, c3 += sj) {
    const long a2 = j + 2 <= g.Nloc ? sj : 0;  // row j+2, clamped to the last row
    const int mnn = d.Mask[c2 + sj + a2];
    const double enn = d.Eta[c2 + sj + a2];
    R fnn[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) fnn[n] = Fl[n][c3 + sj + a2];
    const int jn = min(j + 1, g.Jend1);
    const double hfnn = d.Hfy[g.y2(i, jn)], h0nn = d.Hfy0[g.y2(i, jn)];
    const int mrr = mnx;
    // the cell j: right state of face j, left state of face j+1
    const double e0 = eq;
    double etaL = eL, etaR;
    {
      const double en = enx;
      const double t1 = (en - e0) / dy;
      const double s = slope_of(t1, et2, mr);
      etaR = e0 - hdy * s;
      eL = e0 + hdy * s;
      et2 = t1;
      eq = en;
    }
    R vL_[6], vR_[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const R c = q[n];
      const R vn = fnx[n];
      const R t1 = (vn - c) / dyR;
      const R s = slope_of<R>(t1, t2[n], mr);
      vL_[n] = qL[n];
      vR_[n] = c - hdyR * s;
      qL[n] = c + hdyR * s;
      t2[n] = t1;
      q[n] = vn;
      fnx[n] = fnn[n];
    }
    mnx = mnn;
    enx = enn;
    const double hf = hfnx, h0f = h0nx;
    hfnx = hfnn;
    h0nx = h0nn;
    R uL = vL_[0], uR = vR_[0], vL = vL_[1], vR = vR_[1], wL = vL_[2], wR = vR_[2];
    R duL = vL_[3], duR = vR_[3], dvL = vL_[4], dvR = vR_[4], dwL = vL_[5], dwR = vR_[5];
    double DL = etaL + hf, DR = etaR + hf;
    DL = dmaxd(DL, MinDep);
    DR = dmaxd(DR, MinDep);
    const bool leftOv = (j > g.Jbeg) && (ml == 0 || mr == 0);
    const bool rightOv = (j <= g.Jend) && (mr == 0 || (mrr != 0 && ml == 0));
    double h0 = h0f;
    const bool mutL = (j > g.Jbeg) && ml == 1 && mr == 0;
    const bool mutR = (j <= g.Jend) && mr == 1 && mrr != 0 && ml == 0;
    if (mutL) h0 = d.Hc[c2 - sj];
    else if (mutR) h0 = d.Hc[c2];
    if ((mutL || mutR) && k == g.Kbeg) d.Hfy0[g.y2(i, j)] = h0;
    if (leftOv) {
      etaL = em1; DL = d.D[c2 - sj];
      uL = 0.0; vL = 0.0; wL = 0.0; duL = 0.0; dvL = 0.0; dwL = 0.0;
    }
    if (rightOv) {
      etaR = e0; DR = d.D[c2];
      uR = 0.0; vR = 0.0; wR = 0.0; duR = 0.0; dvR = 0.0; dwR = 0.0;
    }
    const double PL = pres(etaL, h0), PR = pres(etaR, h0);
    double E = 0.0, F = 0.0, G = 0.0, H = 0.0, sL, sR;
    if (wave_speeds(ml, mr, DL, DR, vL, vR, sL, sR)) {
      E = hll(sL, sR, dvL, dvR, DL, DR);
      F = hll(sL, sR, DL * uL * vL, DR * uR * vR, duL, duR);
      G = hll(sL, sR, dvL * vL + PL, dvR * vR + PR, dvL, dvR);
      H = hll(sL, sR, DL * vL * wL, DR * vR * wR, dwL, dwR);
    }
    if (j == g.Jbeg && is_wall(p.bc_y0)) { E = 0.0; F = 0.0; G = PR; H = 0.0; }
    if (j == g.Jend1 && is_wall(p.bc_yn)) { E = 0.0; F = 0.0; G = PL; H = 0.0; }
    if (ml == 0) { E = 0.0; F = 0.0; H = 0.0; G = (j - 1 == g.Jend) ? 0.0 : PR * (double)mr; }
    if (mr == 0) { E = 0.0; F = 0.0; H = 0.0; G = (j == g.Jbeg) ? 0.0 : PL * (double)ml; }
    const long f = g.y3(i, j, k);
    d.Ey[f] = E; d.Fy[f] = F; d.Gy[f] = G; d.Hy[f] = H;
    ml = mr;
    mr = mrr;
    em1 = e0;
  }
}

// fluxes: sigma faces (k in Kbeg..Kend1) of interior columns; vertical upwind
// fluxes, src/nhwave.F:2719-2730, zero at bottom and surface (:2475-2485).
// One thread per column walks the faces upward: the backward difference
// quotient of a cell (z_slope's t2) is the forward one (t1) of the cell
// below, over the same level spacing, and the left (lower) state of a face
// is formed with the cell below's slope one step earlier; each value is
// loaded once, with the arithmetic of z_slope.
template <class R>
__device__ __forceinline__ void flux_z_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long sk = (long)g.Mloc * g.Nloc;
  const long q0 = g.c3(i, j, g.Kbeg);
  d.Fz[q0] = 0.0; d.Gz[q0] = 0.0; d.Hz[q0] = 0.0;
  const long qt = q0 + (long)g.K * sk;  // the surface face Kend1
  d.Fz[qt] = 0.0; d.Gz[qt] = 0.0; d.Hz[qt] = 0.0;
  if (g.K < 2) return;
  const double *sigc = d.sigc - 1, *dsig = d.dsig - 1;
  const R *const Fl[3] = {d.U, d.V, d.W};
  R *const Out[3] = {d.Fz, d.Gz, d.Hz};
  // per field: the forward difference quotient of the cell below the face,
  // the face's left state and the value of the cell above it
  R t1b[3], zl[3], ca[3];
  {
    const int k = g.Kbeg;  // the bottom cell: its slope gives the left state of face Kbeg+1
    const R hm = (R)(sigc[k] - sigc[k - 1]), hp = (R)(sigc[k + 1] - sigc[k]);
    const R hd = (R)(0.5 * dsig[k]);
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const R a = Fl[n][q0 - sk], b = Fl[n][q0], c = Fl[n][q0 + sk];
      const R t2 = (b - a) / hm, t1 = (c - b) / hp;
      const R s = (fabs(t1) + fabs(t2)) < R(kSmall) ? R(0) : limiter(t1, t2);
      zl[n] = b + hd * s;
      t1b[n] = t1;
      ca[n] = c;
    }
  }
  for (int k = g.Kbeg + 1; k <= g.Kend; ++k) {
    const long q = q0 + (long)(k - g.Kbeg) * sk;
    const R om = d.Omega[q];
    const R hp = (R)(sigc[k + 1] - sigc[k]);
    const R hd0 = (R)(0.5 * dsig[k]);
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const R c = ca[n], e = Fl[n][q + sk];
      const R t2 = t1b[n], t1 = (e - c) / hp;
      const R s = (fabs(t1) + fabs(t2)) < R(kSmall) ? R(0) : limiter(t1, t2);
      const R zr = c - hd0 * s;
      Out[n][q] = R(0.5) * (om * (zl[n] + zr) - fabs(om) * (zr - zl[n]));
      zl[n] = c + hd0 * s;
      t1b[n] = t1;
      ca[n] = e;
    }
  }
}

// source_terms (bed slope), src/source_terms.F:446-453
template <class R>
__device__ __forceinline__ void source_terms_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c = g.c2(i, j);
  const double m = (double)d.Mask[c];
  d.SourceX[c] = kGrav * 0.5 * (d.EtaxL[g.x2(i + 1, j)] + d.EtaxR[g.x2(i, j)]) *
                 (d.Hfx0[g.x2(i + 1, j)] - d.Hfx0[g.x2(i, j)]) / g.dx * m;
  d.SourceY[c] = kGrav * 0.5 * (d.EtayL[g.y2(i, j + 1)] + d.EtayR[g.y2(i, j)]) *
                 (d.Hfy0[g.y2(i, j + 1)] - d.Hfy0[g.y2(i, j)]) / g.dy * m;
}

// mirror image of a ghost index for wall boundaries (wl_bc / phi_2D_coll)
__device__ __forceinline__ int mir(int i, int beg, int end) {
  return i < beg ? 2 * beg - 1 - i : (i > end ? 2 * end + 1 - i : i);
}

// eval_duvw explicit part, src/nhwave.F:611-634, 637-688, 841-845,
// 1004-1008, 1163-1182: one thread per interior column.  Mass update, D floor,
// wall ghosts of D (wl_bc, :5119-5262), Eta = D - Hc, momentum update.
template <class R>
__device__ __forceinline__ void eval_explicit_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  const double A = p.alpha[arg - 1], B = p.beta[arg - 1];
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c = g.c2(i, j);
  const double dt = d.scal[0], dx = g.dx, dy = g.dy;
  const double *dsig = d.dsig - 1;
  const int m = d.Mask[c];
  double Dn = d.D[c];
  double R1 = 0.0;
  if (m != 0) {
    for (int k = g.Kbeg; k <= g.Kend; ++k)
      R1 = R1 - 1.0 / dx * (d.Ex[g.x3(i + 1, j, k)] - d.Ex[g.x3(i, j, k)]) * dsig[k] -
           1.0 / dy * (d.Ey[g.y3(i, j + 1, k)] - d.Ey[g.y3(i, j, k)]) * dsig[k];
    R1 = R1 + Dn * (d.SourceC ? d.SourceC[c] : 0.0);  // + D*SourceC (internal wavemaker; else 0)
    Dn = A * d.D0[c] + B * (Dn + dt * R1);
  }
  d.R1[c] = R1;
  Dn = dmaxd(Dn, p.min_dep);
  // D and Eta at the cell and at every ghost cell that mirrors it (walls:
  // the same D; west inflow: 2 Din_X0(j) - D, wl_bc :5144-5153).  Ghost
  // rows and columns whose mirror source lies outside the interior (the
  // second ghost row of a one-row grid) are left as they are.
  int xs[3], ys[3], nx = 0, ny = 0;
  double xv[3];
  xs[nx] = i; xv[nx++] = Dn;
  // FLUX_LR pins the first column and its ghosts to Din_X0 (:5145-5148);
  // its open east boundary radiates Eta from the previous steps (Orlanski,
  // :5182-5193): D = Hc + Cg Eta0(i-1) + (1-Cg) Eta0(i) at each ghost
  const bool flr = p.wave.kind == NHW_WAVE_FLUX_LR;
  if (flr && i == g.Ibeg) xv[0] = Dn = d.Din0[j - 1];
  const double xin = flr ? d.Din0[j - 1] : p.bc_x0 == 3 ? 2.0 * d.Din0[j - 1] - Dn : Dn;
  if (i <= g.Ibeg + 1) { xs[nx] = 2 * g.Ibeg - 1 - i; xv[nx++] = xin; }
  if (i >= g.Iend - 1) {
    const int ge = 2 * g.Iend + 1 - i;
    double xe = Dn;
    if (p.bc_xn == 3) {
      const long a = g.c2(ge - 1, j), b = g.c2(ge - 2, j), e = g.c2(ge, j);
      double cg = -(d.Eta0[a] - d.Eta00[a] + kSmall) / (d.Eta00[a] - d.Eta00[b] + kSmall);
      cg = fmax(fmin(cg, 1.0), 0.0);
      xe = d.Hc[e] + (cg * d.Eta0[a] + (1.0 - cg) * d.Eta0[e]);
    }
    xs[nx] = ge; xv[nx++] = xe;
  }
  ys[ny++] = j;
  if (j <= g.Jbeg + 1) ys[ny++] = 2 * g.Jbeg - 1 - j;
  if (j >= g.Jend - 1) ys[ny++] = 2 * g.Jend + 1 - j;
  for (int a = 0; a < nx; ++a)
    for (int b = 0; b < ny; ++b) {
      const long cc = g.c2(xs[a], ys[b]);
      d.D[cc] = xv[a];
      d.Eta[cc] = xv[a] - d.Hc[cc];
    }
  // momentum (the viscous branch solves it implicitly in k_visc_solve)
  if (p.viscous) return;
  const double fcor = p.fcor, SX = d.SourceX[c], SY = d.SourceY[c];
  for (int k = g.Kbeg; k <= g.Kend; ++k) {
    const long q = g.c3(i, j, k);
    if (m == 0) {
      d.DU[q] = 0.0; d.DV[q] = 0.0; d.DW[q] = 0.0;
      continue;
    }
    const double du = d.DU[q], dv = d.DV[q], dw = d.DW[q];
    const double rdz = 1.0 / dsig[k];
    const long fx0 = g.x3(i, j, k), fy0 = g.y3(i, j, k);
    const long fz0 = q, fz1 = q + (long)g.Mloc * g.Nloc;
    double R2 = -1.0 / dx * (d.Fx[fx0 + 1] - d.Fx[fx0]) - 1.0 / dy * (d.Fy[fy0 + g.Mloc] - d.Fy[fy0]) -
                rdz * (d.Fz[fz1] - d.Fz[fz0]) + fcor * dv + 0.0 + SX + 0.0 + 0.0 + 0.0;
    double R3 = -1.0 / dx * (d.Gx[fx0 + 1] - d.Gx[fx0]) - 1.0 / dy * (d.Gy[fy0 + g.Mloc] - d.Gy[fy0]) -
                rdz * (d.Gz[fz1] - d.Gz[fz0]) - fcor * du + 0.0 + SY + 0.0 + 0.0 + 0.0;
    double R4 = -1.0 / dx * (d.Hx[fx0 + 1] - d.Hx[fx0]) - 1.0 / dy * (d.Hy[fy0 + g.Mloc] - d.Hy[fy0]) -
                rdz * (d.Hz[fz1] - d.Hz[fz0]) + 0.0 + 0.0;
    d.DU[q] = du + dt * R2;
    d.DV[q] = dv + dt * R3;
    d.DW[q] = dw + dt * R4;
  }
}

// DelxEta/DelyEta at interior cells, src/nhwave.F:630-631
template <class R>
__device__ __forceinline__ void eta_slopes_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c = g.c2(i, j);
  const int m = d.Mask[c];
  d.DelxEta[c] = tvd_slope(d.Eta[c - 1], d.Eta[c], d.Eta[c + 1], g.dx, m);
  d.DelyEta[c] = tvd_slope(d.Eta[c - g.Mloc], d.Eta[c], d.Eta[c + g.Mloc], g.dy, m);
}

// ----------------------------------------------------------------------------
// get_UVW + vel_bc, src/nhwave.F:1651-1683, 4416-5066 (lateral BC 1/2,
// Bc_Z0 = 1).  Pass A: U = DU/D, lateral wall ghosts, and the surface/bottom
// z-face states with the kinematic BCs (KBC_surface/KBC_bottom).  Pass B:
// limited x/y slopes of those face states and the vertical ghost cells.
// ----------------------------------------------------------------------------
// vel_bc's wall friction boundary (BC 5, :4486-4501, :4583-4598,
// :4676-4691, :4737-4752): Cdrag = Cd0 (Ibot = 1) or 1/(1/Kappa log(30 h/2/
// Zob))**2, and Phi = min(h Cdrag |u_t| / (Cmu + CmuHt), 2) from the two
// velocity components a, b along the wall at the first interior cell c (Cmu
// = Visc; CmuHt the uniform device scalar, or the k-epsilon array)
template <class R>
__device__ __forceinline__ double wall_phi(const Par &p, const DevT<R> &d, double h, double a, double b, long c) {
  double cdrag = p.cd0;
  if (p.ibot != 1) {
    const double t = 1.0 / kKappa * log(30.0 * h / 2.0 / p.zob);
    cdrag = 1.0 / (t * t);
  }
  const double cmuht = p.turb == 10 ? d.CmuVt[c] : d.scal[6];
  const double phi = h * cdrag * sqrt(a * a + b * b) / (p.visc + cmuht);
  return phi < 2.0 ? phi : 2.0;
}

template <class R>
__device__ __forceinline__ void uvw_a_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  const int divide = arg;
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c = g.c2(i, j);
  const double Dd = d.D[c];
  const long sk = (long)g.Mloc * g.Nloc;
  const long q0 = g.c3(i, j, g.Kbeg);
  if (divide)
    for (int k = g.Kbeg; k <= g.Kend; ++k) {
      const long q = q0 + (k - g.Kbeg) * sk;
      d.U[q] = d.DU[q] / Dd;
      d.V[q] = d.DV[q] / Dd;
      d.W[q] = d.DW[q] / Dd;
    }
  // the velocity of interior cell q of column cq as this pass leaves it (a
  // neighbour's own thread may be writing it: formed again, the same way)
  auto vel = [&](const R *F, const R *DF, long q, long cq) -> double {
    return divide ? (double)(R)((double)DF[q] / d.D[cq]) : (double)F[q];
  };
  // wall friction ghosts: the tangential components scaled by 1 - Phi, the
  // normal one mirrored; their D U take D(n, j) (x sides) or D(i, n) (y
  // sides), n = 1, 2 the ghost's distance from the wall, as the reference
  // writes them
  auto wall5 = [&](long gq, long s, double ph, double dg, bool xside) {
    d.U[gq] = xside ? -d.U[s] : (R)((1.0 - ph) * (double)d.U[s]);
    d.V[gq] = xside ? (R)((1.0 - ph) * (double)d.V[s]) : -d.V[s];
    d.W[gq] = (R)((1.0 - ph) * (double)d.W[s]);
    d.DU[gq] = (R)(dg * (double)d.U[gq]);
    d.DV[gq] = (R)(dg * (double)d.V[gq]);
    d.DW[gq] = (R)(dg * (double)d.W[gq]);
  };
  // lateral ghosts, :4440-4761 (inflow at the west, :4468-4476)
  auto xghost = [&](int gi, int bc) {
    if (bc == 3 && gi > g.Iend) return;  // FLUX_LR's open east boundary: k_east_ghosts
    if (bc == 5) {
      const int ib = gi < g.Ibeg ? g.Ibeg : g.Iend, n = gi < g.Ibeg ? g.Ibeg - gi : gi - g.Iend;
      const long cb = g.c2(ib, j);
      const double dg = d.D[g.c2(n, j)];
      for (int k = g.Kbeg; k <= g.Kend; ++k) {
        const long b = g.c3(ib, j, k);
        const double ph = wall_phi(p, d, g.dx, vel(d.V, d.DV, b, cb), vel(d.W, d.DW, b, cb), b);
        wall5(g.c3(gi, j, k), q0 + (k - g.Kbeg) * sk, ph, dg, true);
      }
      return;
    }
    const bool ns = bc == 2;
    for (int k = g.Kbeg; k <= g.Kend; ++k) {
      const long s = q0 + (k - g.Kbeg) * sk, gq = g.c3(gi, j, k);
      if (bc == 3) {
        const long b = inflow_idx(g, j, k);
        const double din = d.Din0[j - 1], uin = d.Uin0[b], vin = d.Vin0[b], win = d.Win0[b];
        d.U[gq] = (R)(2.0 * uin - (double)d.U[s]);
        d.V[gq] = (R)(2.0 * vin - (double)d.V[s]);
        d.W[gq] = (R)(2.0 * win - (double)d.W[s]);
        d.DU[gq] = (R)(2.0 * din * uin - (double)d.DU[s]);
        d.DV[gq] = (R)(2.0 * din * vin - (double)d.DV[s]);
        d.DW[gq] = (R)(2.0 * din * win - (double)d.DW[s]);
        continue;
      }
      d.U[gq] = -d.U[s];
      d.V[gq] = ns ? -d.V[s] : d.V[s];
      d.W[gq] = ns ? -d.W[s] : d.W[s];
      d.DU[gq] = -d.DU[s];
      d.DV[gq] = ns ? -d.DV[s] : d.DV[s];
      d.DW[gq] = ns ? -d.DW[s] : d.DW[s];
    }
  };
  auto yghost = [&](int gj, int bc) {
    if (bc == 5) {
      const int jb = gj < g.Jbeg ? g.Jbeg : g.Jend, n = gj < g.Jbeg ? g.Jbeg - gj : gj - g.Jend;
      const long cb = g.c2(i, jb);
      const double dg = d.D[g.c2(i, n)];
      for (int k = g.Kbeg; k <= g.Kend; ++k) {
        const long b = g.c3(i, jb, k);
        const double ph = wall_phi(p, d, g.dy, vel(d.U, d.DU, b, cb), vel(d.W, d.DW, b, cb), b);
        wall5(g.c3(i, gj, k), q0 + (k - g.Kbeg) * sk, ph, dg, false);
      }
      return;
    }
    const bool ns = bc == 2;
    for (int k = g.Kbeg; k <= g.Kend; ++k) {
      const long s = q0 + (k - g.Kbeg) * sk, gq = g.c3(i, gj, k);
      d.U[gq] = ns ? -d.U[s] : d.U[s];
      d.V[gq] = -d.V[s];
      d.W[gq] = ns ? -d.W[s] : d.W[s];
      d.DU[gq] = ns ? -d.DU[s] : d.DU[s];
      d.DV[gq] = -d.DV[s];
      d.DW[gq] = ns ? -d.DW[s] : d.DW[s];
    }
  };
  if (i <= g.Ibeg + 1) xghost(2 * g.Ibeg - 1 - i, p.bc_x0);
  if (i >= g.Iend - 1) xghost(2 * g.Iend + 1 - i, p.bc_xn);
  if (j <= g.Jbeg + 1) yghost(2 * g.Jbeg - 1 - j, p.bc_y0);
  if (j >= g.Jend - 1) yghost(2 * g.Jend + 1 - j, p.bc_yn);
  // surface/bottom face states, :4778-4785 (delzFun_3D + construct_3D_z +
  // KBC); the vertical ghosts read here are those of the previous vel_bc
  const double *sigc = d.sigc - 1, *dsig = d.dsig - 1;
  const int Kb = g.Kbeg, Ke = g.Kend;
  const long qb = q0, qe = q0 + (long)(Ke - Kb) * sk;
  const double hpt = sigc[Ke + 1] - sigc[Ke], hmt = sigc[Ke] - sigc[Ke - 1];
  const double hpb = sigc[Kb + 1] - sigc[Kb], hmb = sigc[Kb] - sigc[Kb - 1];
  auto top = [&](const R *F) {
    double a = F[qe - sk], b = F[qe], e = F[qe + sk];
    return b + 0.5 * dsig[Ke] * z_slope(a, b, e, hpt, hmt);
  };
  auto bot = [&](const R *F) {
    double a = F[qb - sk], b = F[qb], e = F[qb + sk];
    return b - 0.5 * dsig[Kb] * z_slope(a, b, e, hpb, hmb);
  };
  const double uzl = top(d.U), vzl = top(d.V);
  // KBC_bottom (:5104-5112): a no-slip bottom (Bc_Z0 = 2) zeroes the
  // horizontal bottom-face states
  const double uzr = p.bc_z0 == 2 ? 0.0 : bot(d.U), vzr = p.bc_z0 == 2 ? 0.0 : bot(d.V);
  const double dt = d.scal[0];
  const double wzl = (d.Eta[c] - d.Eta0[c]) / dt + d.DelxEta[c] * uzl + d.DelyEta[c] * vzl;
  const double deltH = 0.0;  // DeltH: fixed bathymetry
  const double wzr = -deltH - d.DelxH[c] * uzr - d.DelyH[c] * vzr;
  d.ztU[c] = uzl; d.ztV[c] = vzl; d.ztW[c] = wzl;
  d.zbU[c] = uzr; d.zbV[c] = vzr; d.zbW[c] = wzr;
}

template <class R>
__device__ __forceinline__ void uvw_b_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c = g.c2(i, j), sj = g.Mloc;
  const int m = d.Mask[c];
  const double dx = g.dx, dy = g.dy;
  // delxFun1_3D / delyFun1_3D at k = Kend1, :4786-4800
  const double dxuzl = tvd_slope(d.ztU[c - 1], d.ztU[c], d.ztU[c + 1], dx, m);
  const double dxvzl = tvd_slope(d.ztV[c - 1], d.ztV[c], d.ztV[c + 1], dx, m);
  const double dxwzl = tvd_slope(d.ztW[c - 1], d.ztW[c], d.ztW[c + 1], dx, m);
  const double dyuzl = tvd_slope(d.ztU[c - sj], d.ztU[c], d.ztU[c + sj], dy, m);
  const double dyvzl = tvd_slope(d.ztV[c - sj], d.ztV[c], d.ztV[c + sj], dy, m);
  const double dywzl = tvd_slope(d.ztW[c - sj], d.ztW[c], d.ztW[c + sj], dy, m);
  const double dxe = d.DelxEta[c], dye = d.DelyEta[c], Dd = d.D[c];
  const double Atop = sqrt(1.0 + dxe * dxe + dye * dye);
  const double fw1 = 1.0 / Atop * ((1.0 - dxe * dxe) * d.Wsx[c] - dxe * dye * d.Wsy[c]);
  const double fw2 = 1.0 / Atop * ((1.0 - dye * dye) * d.Wsy[c] - dxe * dye * d.Wsx[c]);
  const double *sigc = d.sigc - 1;
  const double uzl = d.ztU[c], vzl = d.ztV[c], wzl = d.ztW[c], wzr = d.zbW[c];
  const long sk = (long)g.Mloc * g.Nloc;
  const int Kb = g.Kbeg, Ke = g.Kend;
  // Cmu = Visc, CmuR = 0; CmuVt of the bottom row and of the two levels around
  // the surface: per cell with the k-epsilon model, else uniform (zero on the
  // inviscid path)
  const double visc = p.visc;
  const bool kef = p.turb == 10;
  const double cvu = kef ? 0.0 : d.scal[7];  // the uniform CmuVt of constant viscosities
  const double cv = kef ? d.CmuVt[g.c3(i, j, Kb)] : cvu;
  const double cmu6 = 0.5 * (((((visc + 0.0) + (kef ? d.CmuVt[g.c3(i, j, Ke)] : cvu)) + visc) + 0.0) +
                             (kef ? d.CmuVt[g.c3(i, j, Ke + 1)] : cvu));
  // bottom stress, :4800-4858 (Bc_Z0 = 2 no-slip, 5 friction law), with the
  // limited x/y slopes of the bottom-face states
  const double dxh = d.DelxH[c], dyh = d.DelyH[c];
  double fb1 = 0.0, fb2 = 0.0, Abot = 1.0;
  double dxuzr = 0.0, dxvzr = 0.0, dxwzr = 0.0, dyuzr = 0.0, dyvzr = 0.0, dywzr = 0.0;
  if (p.bc_z0 != 1) {
    dxuzr = tvd_slope(d.zbU[c - 1], d.zbU[c], d.zbU[c + 1], dx, m);
    dxvzr = tvd_slope(d.zbV[c - 1], d.zbV[c], d.zbV[c + 1], dx, m);
    dxwzr = tvd_slope(d.zbW[c - 1], d.zbW[c], d.zbW[c + 1], dx, m);
    dyuzr = tvd_slope(d.zbU[c - sj], d.zbU[c], d.zbU[c + sj], dy, m);
    dyvzr = tvd_slope(d.zbV[c - sj], d.zbV[c], d.zbV[c + sj], dy, m);
    dywzr = tvd_slope(d.zbW[c - sj], d.zbW[c], d.zbW[c + sj], dy, m);
    Abot = sqrt(1.0 + dxh * dxh + dyh * dyh);
    const double Bbot = sqrt(1.0 + dxh * dxh);
    const long b0 = g.c3(i, j, Kb);
    const double du0 = d.DU[b0], dv0 = d.DV[b0], dw0 = d.DW[b0];
    const double dup = (du0 - dxh * dw0) / Bbot;
    // `-DelxH*DelyH*DU* +(1.0+DelxH**2)*DV -DelyH*DW`, unary plus on the next primary
    const double dvp = (-(dxh * dyh * du0 * (1.0 + dxh * dxh) * dv0) - dyh * dw0) / Abot / Bbot;
    const double dumag = sqrt(dup * dup + dvp * dvp);
    const double Dz1 = 0.5 * Dd * (d.dsig - 1)[Kb] / Abot;
    double cdrag = p.cd0;
    if (p.ibot != 1) {
      const double tk = kKappa / log(Dz1 / (p.zob / 30.0));
      cdrag = tk * tk;
    }
    cdrag = dmind(cdrag, ((visc + 0.0) + cv) / Dz1 / (dumag + kSmall) * Dd);
    double tb1, tb2;
    if (p.bc_z0 == 2) {
      tb1 = dup / Dd / Dz1 * ((visc + 0.0) + cv);
      tb2 = dvp / Dd / Dz1 * ((visc + 0.0) + cv);
    } else {
      tb1 = cdrag * dup * dumag / (Dd * Dd);
      tb2 = cdrag * dvp * dumag / (Dd * Dd);
    }
    fb1 = Bbot * tb1;
    fb2 = (dxh * dyh * tb1 + Abot * tb2) / Bbot;
  }
  const double uzr = d.zbU[c], vzr = d.zbV[c];
  for (int k = 1; k <= 2; ++k) {
    const long gb = g.c3(i, j, Kb - k), sb = g.c3(i, j, Kb + k - 1);
    const double dwb = 2.0 * Dd * wzr - d.DW[sb];
    double dub = d.DU[sb], dvb = d.DV[sb];
    if (p.bc_z0 != 1) {
      const double dsb = sigc[Kb + k - 1] - sigc[Kb - k];
      const double dwsb = d.DW[sb], dusb = dub, dvsb = dvb;
      double t1 = fb1 * (Dd * Dd / Abot / (((kSmall + visc) + 0.0) + cv));
      double t2 = dxh * (dwsb - Dd * wzr) / (dsb / 2.0);
      double t3 = Dd * Dd / (Abot * Abot) *
                  (-2.0 * dxh * dxuzr - (1.0 - dxh * dxh) * dxwzr - dyh * (dyuzr + dxvzr - dxh * dywzr));
      dub = dusb - dsb * (t1 + t2 + t3);
      double old = 2.0 * Dd * uzr - dusb;
      if (old > 0.) dub = dmind(3.0 * old, dmaxd(-1.0 * old, dub));
      else if (old < 0.) dub = dmaxd(3.0 * old, dmind(-1.0 * old, dub));
      if (Dd < kMinDepBC) dub = dusb;
      t1 = fb2 * (Dd * Dd / Abot / (((kSmall + visc) + 0.0) + cv));
      t2 = dyh * (dwsb - Dd * wzr) / (dsb / 2.0);
      t3 = Dd * Dd / (Abot * Abot) *
           (-2.0 * dyh * dyvzr - (1.0 - dyh * dyh) * dywzr - dxh * (dxvzr + dyuzr - dyh * dxwzr));
      dvb = dvsb - dsb * (t1 + t2 + t3);
      old = 2.0 * Dd * vzr - dvsb;
      if (old > 0.) dvb = dmind(3.0 * old, dmaxd(-1.0 * old, dvb));
      else if (old < 0.) dvb = dmaxd(3.0 * old, dmind(-1.0 * old, dvb));
      if (Dd < kMinDepBC) dvb = dvsb;
    }
    d.DW[gb] = dwb; d.DU[gb] = dub; d.DV[gb] = dvb;

    const long gt = g.c3(i, j, Ke + k), st = g.c3(i, j, Ke - k + 1);
    const double dsg = sigc[Ke + k] - sigc[Ke - k + 1];
    const double dwst = d.DW[st], dust = d.DU[st], dvst = d.DV[st];
    const double dwt = 2.0 * Dd * wzl - dwst;
    double t1 = fw1 * (Dd * Dd / Atop / (kSmall + cmu6));
    double t2 = dxe * (Dd * wzl - dwst) / (dsg / 2.0);
    double t3 = Dd * Dd / (Atop * Atop) *
                (+2.0 * dxe * dxuzl - (1.0 - dxe * dxe) * dxwzl + dye * (dyuzl + dxvzl + dxe * dywzl));
    double dut = dust + dsg * (t1 - t2 + t3);
    double old = 2.0 * Dd * uzl - dust;
    if (old > 0.) dut = dmind(3.0 * old, dmaxd(-1.0 * old, dut));
    else if (old < 0.) dut = dmaxd(3.0 * old, dmind(-1.0 * old, dut));
    if (Dd < kMinDepBC) dut = dust;
    t1 = fw2 * (Dd * Dd / Atop / (kSmall + cmu6));
    t2 = dye * (Dd * wzl - dwst) / (dsg / 2.0);
    t3 = Dd * Dd / (Atop * Atop) *
         (+2.0 * dye * dyvzl - (1.0 - dye * dye) * dywzl + dxe * (dxvzl + dyuzl + dye * dxwzl));
    double dvt = dvst + dsg * (t1 - t2 + t3);
    old = 2.0 * Dd * vzl - dvst;
    if (old > 0.) dvt = dmind(3.0 * old, dmaxd(-1.0 * old, dvt));
    else if (old < 0.) dvt = dmaxd(3.0 * old, dmind(-1.0 * old, dvt));
    if (Dd < kMinDepBC) dvt = dvst;
    d.DW[gt] = dwt; d.DU[gt] = dut; d.DV[gt] = dvt;
    d.U[gb] = dub / Dd; d.V[gb] = dvb / Dd; d.W[gb] = dwb / Dd;
    d.U[gt] = dut / Dd; d.V[gt] = dvt / Dd; d.W[gt] = dwt / Dd;
  }
  (void)sk;
}

// projection_corrector, src/nhwave.F:487-536 (DelxSc/DelySc of
// sigma_transform, :6505-6541, formed in place)
template <class R>
__device__ __forceinline__ void projection_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  if (t >= (long)g.M * g.N * g.K) return;
  const int i = g.Ibeg + (int)(t % g.M);
  const long r = t / g.M;
  const int j = g.Jbeg + (int)(r % g.N), k = g.Kbeg + (int)(r / g.N);
  const long c = g.c2(i, j);
  if (d.Mask[c] == 0) return;
  const double dx = g.dx, dy = g.dy, dt = d.scal[0];
  const double *dsig = d.dsig - 1, *sigc = d.sigc - 1;
  const long sk = (long)g.Mloc * g.Nloc, sj = g.Mloc;
  const long q = g.c3(i, j, k);
  const double *P = d.P;
  const double dxp = 0.5 * ((P[q + 1] - P[q - 1]) / (2. * dx) + (P[q + sk + 1] - P[q + sk - 1]) / (2. * dx));
  const double dyp = 0.5 * ((P[q + sj] - P[q - sj]) / (2. * dy) + (P[q + sk + sj] - P[q + sk - sj]) / (2. * dy));
  const double dzp = (P[q + sk] - P[q]) / dsig[k];
  const double Dd = d.D[c];
  const double dxsc = (1.0 - sigc[k]) / Dd * d.DelxH[c] - sigc[k] / Dd * d.DelxEta[c];
  const double dysc = (1.0 - sigc[k]) / Dd * d.DelyH[c] - sigc[k] / Dd * d.DelyEta[c];
  d.DU[q] = d.DU[q] - Dd * dt / kRho0 * (dxp + dzp * dxsc);
  d.DV[q] = d.DV[q] - Dd * dt / kRho0 * (dyp + dzp * dysc);
  d.DW[q] = d.DW[q] - dt / kRho0 * dzp;
}

// SSP-RK blend + dry zeroing + Froude cap, src/nhwave.F:1227-1261; one
// thread per interior column.
template <class R>
__device__ __forceinline__ void rk_blend_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  const double A = p.alpha[arg - 1], B = p.beta[arg - 1];
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c = g.c2(i, j);
  const int m = d.Mask[c];
  const double Dd = d.D[c];
  const double fu = sqrt(kGrav * Dd) * p.froude_cap * Dd;
  for (int k = g.Kbeg; k <= g.Kend; ++k) {
    const long q = g.c3(i, j, k);
    double du = A * d.DU0[q] + B * d.DU[q];
    double dv = A * d.DV0[q] + B * d.DV[q];
    double dw = A * d.DW0[q] + B * d.DW[q];
    if (m == 0) { du = 0.0; dv = 0.0; dw = 0.0; }
    if (p.froude_on && m > 0) {
      const double duu = sqrt(du * du + dv * dv + dw * dw);
      if (duu > fu) {
        const double ang = atan2(dv, du);
        du = fu * cos(ang);
        dv = fu * sin(ang);
        dw = 0.0;
      }
    }
    d.DU[q] = du; d.DV[q] = dv; d.DW[q] = dw;
  }
}

// ----------------------------------------------------------------------------
// get_Omega, src/nhwave.F:1411-1648.  Faces: reconstruction of Eta, U (V) and
// DU (DV) with the reduced wet-dry overrides, wave speeds, HLL mass flux and
// wall BCs; also stores EtaxL/R (EtayL/R) that source_terms reads next stage.
// ----------------------------------------------------------------------------
// Thread layouts as in the flux kernels: the x-face kernel shares each
// cell's slope across lanes and walks the sigma levels of a face column, the
// y-face kernel walks kYFaces faces of increasing j per (i, k).  The
// arithmetic is the reference's, in FP64 in both forms (the mixed form's
// FP32 U, DU, V, DV are widened on load).
template <class R>
__device__ __forceinline__ void omega_x_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  const int wpr = (g.M + 1 + kXFaces - 1) / kXFaces;
  const long w = t >> 6;
  if (w >= (long)wpr * g.N) return;  // whole wavefronts leave together
  const int lane = (int)(t & 63);
  const int j = g.Jbeg + (int)(w / wpr);
  const int i = g.Ibeg - 2 + (int)(w % wpr) * kXFaces + lane;
  const int ic = min(i, g.Iend + 1), im = max(ic - 1, 1);
  const bool face = lane >= 2 && i <= g.Iend1;
  const double dx = g.dx, hdx = 0.5 * dx, MinDep = p.min_dep;
  const long c20 = g.c2(ic, j), c2m1 = g.c2(im, j);
  const int ml = d.Mask[c2m1], mr = d.Mask[c20], mrr = d.Mask[c20 + 1];
  const double em1 = d.Eta[c2m1], e0 = d.Eta[c20];
  double etaL, etaR;
  {
    const double t1 = (d.Eta[c20 + 1] - e0) / dx;
    const double s = slope_of(t1, from_below(t1), mr);
    etaR = e0 - hdx * s;
    etaL = from_below(e0 + hdx * s);
  }
  const double hf = d.Hfx[g.x2(ic, j)];
  double DL = dmaxd(etaL + hf, MinDep), DR = dmaxd(etaR + hf, MinDep);
  const bool leftOv = (i > g.Ibeg) && (ml == 0 || mr == 0);
  const bool rightOv = (i <= g.Iend) && (mr == 0 || (mrr != 0 && ml == 0));
  if (leftOv) { etaL = em1; DL = d.D[c2m1]; }
  if (rightOv) { etaR = e0; DR = d.D[c20]; }
  const double pl = sqrt(kGrav * fabs(DL)), pr = sqrt(kGrav * fabs(DR));
  if (face) {
    d.EtaxL[g.x2(i, j)] = etaL;
    d.EtaxR[g.x2(i, j)] = etaR;
  }
  for (int k = g.Kbeg; k <= g.Kend; ++k) {
    const long c0 = g.c3(ic, j, k);
    auto recon = [&](const R *F, double &qL, double &qR) {
      const double c = F[c0];
      const double t1 = ((double)F[c0 + 1] - c) / dx;
      const double s = slope_of(t1, from_below(t1), mr);
      qR = c - hdx * s;
      qL = from_below(c + hdx * s);
    };
    double uL, uR, duL, duR;
    recon(d.U, uL, uR);
    recon(d.DU, duL, duR);
    if (leftOv) duL = 0.0;
    if (rightOv) duR = 0.0;
    double E = 0.0, sL, sR;
    if (wave_speeds_r(ml, mr, pl, pr, uL, uR, sL, sR)) E = hll(sL, sR, duL, duR, DL, DR);
    if (i == g.Ibeg && is_wall(p.bc_x0)) E = 0.0;
    if (i == g.Ibeg && p.bc_x0 == 3) E = d.Din0[j - 1] * d.Uin0[inflow_idx(g, j, k)];  // :1564-1565
    if (i == g.Iend1 && is_wall(p.bc_xn)) E = 0.0;
    if (i == g.Iend1 && p.bc_xn == 3) E = d.Din_Xn[j - 1] * d.Uin_Xn[inflow_idx(g, j, k)];  // :1580-1581
    if (face) d.Ex[g.x3(i, j, k)] = E;
  }
}

template <class R>
__device__ __forceinline__ void omega_y_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  if (t >= flux_y_threads(g)) return;
  const int i = g.Ibeg + (int)(t % g.M);
  const long r = t / g.M;
  const int k = g.Kbeg + (int)(r % g.K);
  const int jf0 = g.Jbeg + (int)(r / g.K) * kYFaces;
  const int jf1 = min(jf0 + kYFaces - 1, g.Jend1);
  const double dy = g.dy, hdy = 0.5 * dy, MinDep = p.min_dep;
  const long sj = g.Mloc;
  const R *const Fl[2] = {d.V, d.DV};
  // per field: the current cell's value, its backward difference quotient
  // and the left state of its south face (Eta first)
  double q[3], t2[3], qL[3];
  long c2 = g.c2(i, jf0 - 2), c3 = g.c3(i, jf0 - 2, k);
  auto val = [&](int n, long off) -> double { return n == 0 ? d.Eta[c2 + off] : (double)Fl[n - 1][c3 + off]; };
  {
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const double vn = val(n, sj);
      t2[n] = (vn - val(n, 0)) / dy;
      q[n] = vn;
    }
    c2 += sj;
    c3 += sj;
    const int m = d.Mask[c2];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const double vn = val(n, sj);
      const double t1 = (vn - q[n]) / dy;
      qL[n] = q[n] + hdy * slope_of(t1, t2[n], m);
      t2[n] = t1;
      q[n] = vn;
    }
    c2 += sj;
    c3 += sj;
  }
  int ml = d.Mask[c2 - sj], mr = d.Mask[c2];
  double em1 = d.Eta[c2 - sj];
  // row j+1 of the walk and the face's depth, loaded one step ahead
  int mnx = d.Mask[c2 + sj];
  double fnx[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) fnx[n] = val(n, sj);
  double hfnx = d.Hfy[g.y2(i, jf0)];
  for (int j = jf0; j <= jf1; ++j, c2 += sj, c3 += sj) {
    const long a2 = j + 2 <= g.Nloc ? sj : 0;  // row j+2, clamped to the last row
    const int mnn = d.Mask[c2 + sj + a2];
    double fnn[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) fnn[n] = val(n, sj + a2);
    const double hfnn = d.Hfy[g.y2(i, min(j + 1, g.Jend1))];
    const int mrr = mnx;
    const double e0 = q[0];
    double sL_[3], sR_[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const double c = q[n];
      const double t1 = (fnx[n] - c) / dy;
      const double s = slope_of(t1, t2[n], mr);
      sL_[n] = qL[n];
      sR_[n] = c - hdy * s;
      qL[n] = c + hdy * s;
      t2[n] = t1;
      q[n] = fnx[n];
      fnx[n] = fnn[n];
    }
    mnx = mnn;
    const double hf = hfnx;
    hfnx = hfnn;
    double etaL = sL_[0], etaR = sR_[0], vL = sL_[1], vR = sR_[1], dvL = sL_[2], dvR = sR_[2];
    double DL = dmaxd(etaL + hf, MinDep), DR = dmaxd(etaR + hf, MinDep);
    const bool leftOv = (j > g.Jbeg) && (ml == 0 || mr == 0);
    const bool rightOv = (j <= g.Jend) && (mr == 0 || (mrr != 0 && ml == 0));
    if (leftOv) { etaL = em1; DL = d.D[c2 - sj]; dvL = 0.0; }
    if (rightOv) { etaR = e0; DR = d.D[c2]; dvR = 0.0; }
    double E = 0.0, sL, sR;
    if (wave_speeds(ml, mr, DL, DR, vL, vR, sL, sR)) E = hll(sL, sR, dvL, dvR, DL, DR);
    if (j == g.Jbeg && is_wall(p.bc_y0)) E = 0.0;
    if (j == g.Jend1 && is_wall(p.bc_yn)) E = 0.0;
    d.Ey[g.y3(i, j, k)] = E;
    if (k == g.Kbeg) {
      d.EtayL[g.y2(i, j)] = etaL;
      d.EtayR[g.y2(i, j)] = etaR;
    }
    ml = mr;
    mr = mrr;
    em1 = e0;
  }
}

// Omega column sweep, src/nhwave.F:1619-1628
template <class R>
__device__ __forceinline__ void omega_col_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c = g.c2(i, j);
  const int m = d.Mask[c];
  const double R1 = d.R1[c], dx = g.dx, dy = g.dy;
  const double *dsig = d.dsig - 1;
  double om = 0.0;
  for (int k = g.Kbeg + 1; k <= g.Kend1; ++k) {
    const long fx = g.x3(i, j, k - 1), fy = g.y3(i, j, k - 1);
    om = om - dsig[k - 1] * (R1 + (d.Ex[fx + 1] - d.Ex[fx]) / dx + (d.Ey[fy + g.Mloc] - d.Ey[fy]) / dy);
    if (m == 0) om = 0.0;
    d.Omega[g.c3(i, j, k)] = om;
  }
}

// ----------------------------------------------------------------------------
// per-step routines
// ----------------------------------------------------------------------------

// estimate_dt, src/nhwave.F:3694-3874: block minima ...
template <class R>
__global__ __launch_bounds__(256) void k_dt_partial(Geo g, DevT<R> d) {
  __shared__ double sm[256];
  const long n = (long)g.M * g.N * g.K;
  double v = kLarge;
  const double *dsig = d.dsig - 1;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int i = g.Ibeg + (int)(t % g.M);
    const long r = t / g.M;
    const int j = g.Jbeg + (int)(r % g.N), k = g.Kbeg + (int)(r / g.N);
    const double Dd = d.D[g.c2(i, j)];
    const long q = g.c3(i, j, k);
    const double sq = sqrt(kGrav * Dd);
    double t1 = fabs(d.U[q]) + sq;
    t1 = dmaxd(t1, kSmall);
    double w = g.dx / t1;
    if (w < v) v = w;
    t1 = fabs(d.V[q]) + sq;
    t1 = dmaxd(t1, kSmall);
    w = g.dy / t1;
    if (w < v) v = w;
    t1 = dmaxd(fabs(d.W[q]), kSmall);
    w = dsig[k] * Dd / t1;
    if (w < v) v = w;
  }
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = dmind(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) d.red[blockIdx.x] = sm[0];
}

// explicit viscous stress limit, src/nhwave.F:3812-3832 (1.e-16 is a
// REAL(4) literal): the minimum over cells of dx**2/(|CmuHt|+1.e-16) is the
// value at the largest |CmuHt| (rounded division is monotonic); that is
// uniform with constant viscosities (scal[6]), gathered by the k-epsilon
// pass into scal[5] otherwise
__device__ __forceinline__ double dt_viscous_limit(const Geo &g, const Par &p, const double *scal) {
  if (!p.viscous) return kLarge;
  const double ch = p.turb == 10 ? scal[5] : scal[6];
  double t2 = kLarge;
  double t1 = g.dx * g.dx / (fabs(ch) + kSmall);
  if (t1 < t2) t2 = t1;
  t1 = g.dy * g.dy / (fabs(ch) + kSmall);
  if (t1 < t2) t2 = t1;
  return p.visc_number * t2;
}

// ... and the step decision
__global__ void k_dt_final(Geo g, Par p, Dev d, int nblk) {
  __shared__ double sm[256];
  double v = kLarge;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) v = dmind(v, d.red[b]);
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = dmind(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double dt_old = d.scal[0];
    const double dt_growth = kDtGrowth * dt_old;
    const double dt_courant = p.cfl * sm[0];
    const double dt_viscous = dt_viscous_limit(g, p, d.scal);
    double dt = dmind(dmind(dmind(dt_growth, dt_courant), dt_viscous), p.dt_max);
    d.scal[2] = dt_old;
    d.scal[0] = dt;
    d.scal[3] = dt_courant;
    if (dt < p.dt_min) {
      d.iscal[1] |= 1;
    } else {
      d.scal[1] = d.scal[1] + dt;
      d.iscal[0] += 1;
    }
  }
}

// update_mask, src/nhwave.F:5403-5456.  Pass A decides the new mask from
// the old one (flooding reads neighbour Mask/Eta/D; a neighbour that dries
// this step fails `D-MinDep > 1e-6` before and after drying, so reading the
// pre-drying values is exact); pass B applies drying to Eta/D and forms
// Mask9 from the new mask.
template <class R>
__device__ __forceinline__ void mask_a_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  if (t >= (long)g.Mloc * g.Nloc) return;
  const int i = 1 + (int)(t % g.Mloc), j = 1 + (int)(t / g.Mloc);
  const long c = t;
  const int *M = d.Mask;
  int m = M[c];
  const int ms = d.Mask_Struct[c];
  if (i >= g.Ibeg && i <= g.Iend && j >= g.Jbeg && j <= g.Jend && ms != 0) {
    const double MinDep = p.min_dep;
    const double *Eta = d.Eta, *D = d.D;
    if (m == 0) {
      const long nb[4] = {c - 1, c + 1, c - g.Mloc, c + g.Mloc};
      for (int q = 0; q < 4; ++q)
        if (M[nb[q]] == 1 && Eta[nb[q]] > Eta[c] && D[nb[q]] - MinDep > kWetDry) m = 1;
    } else if (fabs(D[c] - MinDep) <= kWetDry) {
      m = 0;
    }
  }
  d.MaskNew[c] = m * ms;
}

template <class R>
__device__ __forceinline__ void mask_b_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c = g.c2(i, j), sj = g.Mloc;
  if (d.Mask_Struct[c] != 0 && d.Mask[c] != 0 && fabs(d.D[c] - p.min_dep) <= kWetDry) {
    const double e = p.min_dep - d.Hc[c];
    d.Eta[c] = e;
    d.D[c] = e + d.Hc[c];
  }
  const int *M = d.MaskNew;
  d.Mask9[c] = M[c] * M[c - 1] * M[c + 1] * M[c + sj + 1] * M[c + sj] * M[c + sj - 1] * M[c - sj + 1] *
               M[c - sj] * M[c - sj - 1];
}

// update_wind (Iws = 1), src/nhwave.F:5333-5400
template <class R>
__device__ __forceinline__ void wind_body(const Geo &g, const Par &p, const DevT<R> &d, long t, int arg) {
  const double wu = p.wind_u, wv = p.wind_v;
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c = g.c2(i, j);
  // constant expressions of default reals fold in binary32
  const double cmin = (double)(1.e-3f * (0.61f + 0.063f * 6.0f));
  const double cmax = (double)(1.e-3f * (0.61f + 0.063f * 50.0f));
  const double wds = sqrt(wu * wu + wv * wv);
  double cds = (double)1.e-3f * ((double)0.61f + (double)0.063f * wds);
  cds = dmind(dmaxd(cds, cmin), cmax);
  d.Wsx[c] = (double)0.001293f * cds * wu * wds;
  d.Wsy[c] = (double)0.001293f * cds * wv * wds;
}

// initial, src/initialize.F:3118-3153: mask, dry Eta, over all cells
__global__ void k_init_mask(Geo g, Par p, Dev d) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.Mloc * g.Nloc) return;
  int m = 1;
  if ((d.Eta[t] + d.Hc[t]) <= p.min_dep) {
    m = 0;
    d.Eta[t] = p.min_dep - d.Hc[t];
  }
  d.Mask[t] = m * d.Mask_Struct[t];
}

// phi_2D_coll(Eta), Eta0 = Eta (all cells)
__global__ void k_init_eta_ghost(Geo g, Dev d) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.Mloc * g.Nloc) return;
  const int i = 1 + (int)(t % g.Mloc), j = 1 + (int)(t / g.Mloc);
  const int si = mir(i, g.Ibeg, g.Iend), sj = mir(j, g.Jbeg, g.Jend);
  const double e = d.Eta[g.c2(si, sj)];
  // ghosts only; the interior value is its own source
  if (si != i || sj != j) d.Eta[t] = e;
}

__global__ void k_init_eta0_d(Geo g, Par p, Dev d) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.Mloc * g.Nloc) return;
  const int i = 1 + (int)(t % g.Mloc), j = 1 + (int)(t / g.Mloc);
  d.Eta0[t] = d.Eta[t];
  d.D[t] = dmaxd(d.Hc[t] + d.Eta[t], p.min_dep);
  if (i >= g.Ibeg && i <= g.Iend && j >= g.Jbeg && j <= g.Jend) {
    const int *M = d.Mask;
    const long c = t, sj = g.Mloc;
    d.Mask9[c] = M[c] * M[c - 1] * M[c + 1] * M[c + sj + 1] * M[c + sj] * M[c + sj - 1] * M[c - sj + 1] *
                 M[c - sj] * M[c - sj - 1];
  }
}

// DU = D*U*Mask over every cell, src/initialize.F:3162-3170
template <class R>
__global__ void k_init_du(Geo g, DevT<R> d) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n2 = (long)g.Mloc * g.Nloc;
  if (t >= n2 * g.Kloc) return;
  const long c = t % n2;
  const double Dd = d.D[c], m = (double)d.Mask[c];
  d.DU[t] = Dd * d.U[t] * m;
  d.DV[t] = Dd * d.V[t] * m;
  d.DW[t] = Dd * d.W[t] * m;
}

// P from the interior solution vector, with its x/y ghost cells
// (src/pressure.F:133-141, 185-239)
__global__ void k_p_from_x(Geo g, const double *__restrict__ x, Dev d, int zl) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.M * g.N * g.K) return;
  const int i = g.Ibeg + (int)(t % g.M);
  const long r = t / g.M;
  const int j = g.Jbeg + (int)(r % g.N), k = g.Kbeg + (int)(r / g.N);
  const double v = x[zl ? (r / g.N) * ((long)g.M * g.N) + zcol(i - g.Ibeg, j - g.Jbeg, g.M) : t];
  d.P[g.c3(i, j, k)] = v;
  if (i <= g.Ibeg + 1) d.P[g.c3(2 * g.Ibeg - 1 - i, j, k)] = v;
  if (i >= g.Iend - 1) d.P[g.c3(2 * g.Iend + 1 - i, j, k)] = v;
  if (j <= g.Jbeg + 1) d.P[g.c3(i, 2 * g.Jbeg - 1 - j, k)] = v;
  if (j >= g.Jend - 1) d.P[g.c3(i, 2 * g.Jend + 1 - j, k)] = v;
}

// ----------------------------------------------------------------------------
// kernels: one per body, for a single scenario (k_*) and for a batch (kb_*)
// ----------------------------------------------------------------------------
// XCD = 1: the single-scenario kernel takes its blocks in XCD-aware order, so
// the neighbouring rows and planes its stencil reads are mostly worked by the
// same XCD and found in its L2.  Measured per kernel at 1024x1024x10: it pays
// for projection (176 -> 109 us) and the Omega column sweep (65 -> 55 us) and
// costs the sigma-face fluxes (190 -> 225 us); the others are unchanged.
#define NHW_KERNELS(NAME, XCD)                                                                  \
  template <class R>                                                                            \
  __global__ __launch_bounds__(256) void k_##NAME(Geo g, Par p, DevT<R> d, int arg) {          \
    const unsigned blk = XCD ? xcd_block() : blockIdx.x;                                        \
    NAME##_body<R>(g, p, d, (long)blk * blockDim.x + threadIdx.x, arg);                         \
  }                                                                                             \
  template <class R>                                                                            \
  __global__ __launch_bounds__(256) void kb_##NAME(BTab<R> T, int arg) {                       \
    const BPos b = bfind(T);                                                                    \
    const BScn<R> &S = T.scn[b.s];                                                              \
    NAME##_body<R>(S.g, S.p, S.d, b.b * blockDim.x + threadIdx.x, arg);                         \
  }
NHW_KERNELS(flux_x, 0)
NHW_KERNELS(flux_y, 0)
NHW_KERNELS(flux_z, 0)
NHW_KERNELS(source_terms, 0)
NHW_KERNELS(eval_explicit, 0)
NHW_KERNELS(eta_slopes, 0)
NHW_KERNELS(uvw_a, 0)
NHW_KERNELS(uvw_b, 0)
NHW_KERNELS(projection, 1)
NHW_KERNELS(rk_blend, 0)
NHW_KERNELS(omega_x, 0)
NHW_KERNELS(omega_y, 0)
NHW_KERNELS(omega_col, 1)
NHW_KERNELS(mask_a, 0)
NHW_KERNELS(mask_b, 0)
NHW_KERNELS(wind, 0)
#undef NHW_KERNELS

// ----------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------
#define LAUNCH(kid, kern, n, ...)                                                         \
  do {                                                                                    \
    long _n = (n);                                                                        \
    if (_n > 0) {                                                                         \
      Timed _t(c, kid);                                                                   \
      hipLaunchKernelGGL(kern, dim3(nblocks(_n, 256)), dim3(256), 0, c->stream, __VA_ARGS__); \
    }                                                                                     \
  } while (0)

// launch of a kernel templated on the 3-D storage type, with `dv` the
// matching view of the context's device state
#define LAUNCH_R(kid, kern, n, ...)                    \
  do {                                                 \
    if (c->mixed) {                                    \
      const DevT<float> dv = view<float>(c->d);        \
      LAUNCH(kid, kern<float>, n, __VA_ARGS__);        \
    } else {                                           \
      const Dev &dv = c->d;                            \
      LAUNCH(kid, kern<double>, n, __VA_ARGS__);       \
    }                                                  \
  } while (0)

static int post(const char *what) { return hip_check(hipGetLastError(), what); }

int launch_initial(nhw_ctx *c) {
  const Geo &g = c->g;
  LAUNCH(KID_MASK, k_init_mask, g.n2(), g, c->p, c->d);
  LAUNCH(KID_MASK, k_init_eta_ghost, g.n2(), g, c->d);
  LAUNCH(KID_MASK, k_init_eta0_d, g.n2(), g, c->p, c->d);
  LAUNCH_R(KID_UVW_A, k_uvw_a, (long)g.M * g.N, g, c->p, dv, 0);
  if (int rc = launch_east_ghosts(c)) return rc;
  LAUNCH_R(KID_UVW_B, k_uvw_b, (long)g.M * g.N, g, c->p, dv, 0);
  LAUNCH_R(KID_RK_BLEND, k_init_du, g.n3(), g, dv);
  return post("initial");
}

int launch_estimate_dt(nhw_ctx *c) {
  const Geo &g = c->g;
  const int nblk = 1024;
  {
    Timed t(c, KID_DT);
    if (c->mixed)
      hipLaunchKernelGGL(k_dt_partial<float>, dim3(nblk), dim3(256), 0, c->stream, g, view<float>(c->d));
    else
      hipLaunchKernelGGL(k_dt_partial<double>, dim3(nblk), dim3(256), 0, c->stream, g, c->d);
  }
  {
    Timed t(c, KID_KREDUCE);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(256), 0, c->stream, c->g, c->p, c->d, nblk);
  }
  return post("estimate_dt");
}

int launch_update_mask(nhw_ctx *c) {
  const Geo &g = c->g;
  LAUNCH_R(KID_MASK, k_mask_a, g.n2(), g, c->p, dv, 0);
  LAUNCH_R(KID_MASK, k_mask_b, (long)g.M * g.N, g, c->p, dv, 0);
  int rc = post("update_mask");
  std::swap(c->d.Mask, c->d.MaskNew);
  return rc;
}

int launch_update_wind(nhw_ctx *c) {
  const Geo &g = c->g;
  LAUNCH_R(KID_SOURCE, k_wind, (long)g.M * g.N, g, c->p, dv, 0);
  return post("update_wind");
}

// FLUX_LR's open east boundary in vel_bc (:4547-4573), after k_uvw_a has
// set U, V, W in the interior: one thread per interior (j, k) row, the ghost
// columns Iend+1 then Iend+2 in the reference's order (the second reads the
// first's new U).  Cg = -(F0(i-1) - F00(i-1) + 1e-16)/(F00(i-1) - F00(i-2) +
// 1e-16) clipped to [0, 1], from the previous steps' velocities.
__global__ __launch_bounds__(256) void k_east_ghosts(Geo g, Dev d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.N * g.K) return;
  const int j = g.Jbeg + (int)(t % g.N), k = g.Kbeg + (int)(t / g.N);
  const long b = inflow_idx(g, j, k);
  const double din = d.Din_Xn[j - 1];
  auto clip = [](double cg) { return fmax(fmin(cg, 1.0), 0.0); };
  for (int i = 1; i <= 2; ++i) {
    const long gq = g.c3(g.Iend + i, j, k), a = g.c3(g.Iend + i - 1, j, k), a2 = g.c3(g.Iend + i - 2, j, k);
    const long s = g.c3(g.Iend - i + 1, j, k);
    double cg = clip(-(d.U0[a] - d.U00[a] + kSmall) / (d.U00[a] - d.U00[a2] + kSmall));
    d.U[gq] = i == 1 ? cg * d.U[a] + d.Uin_Xni0[b] - cg * d.Uin_Xni[b] : cg * d.U[a] + (1.0 - cg) * d.U[gq];
    cg = clip(-(d.V0[a] - d.V00[a] + kSmall) / (d.V00[a] - d.V00[a2] + kSmall));
    d.V[gq] = cg * d.V0[a] + (1.0 - cg) * d.V0[gq];
    cg = clip(-(d.W0[a] - d.W00[a] + kSmall) / (d.W00[a] - d.W00[a2] + kSmall));
    d.W[gq] = cg * d.W0[a] + (1.0 - cg) * d.W0[gq];
    d.DU[gq] = 2.0 * din * d.Uin_Xn[b] - d.DU[s];
    d.DV[gq] = 2.0 * din * d.Vin_Xn[b] - d.DV[s];
    d.DW[gq] = 2.0 * din * d.Win_Xn[b] - d.DW[s];
  }
}

// update_vars' history of FLUX_LR (:5470-5478): Eta00 = Eta0, U00 = U0, ...
// before Eta0 = Eta (the copies that follow on the same stream), U0 = U
__global__ __launch_bounds__(256) void k_flux_lr_history(Geo g, Dev d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < g.n2()) d.Eta00[t] = d.Eta0[t];
  if (t < g.n3()) {
    d.U00[t] = d.U0[t];
    d.V00[t] = d.V0[t];
    d.W00[t] = d.W0[t];
    d.U0[t] = d.U[t];
    d.V0[t] = d.V[t];
    d.W0[t] = d.W[t];
  }
}

int launch_east_ghosts(nhw_ctx *c) {
  if (c->p.wave.kind != NHW_WAVE_FLUX_LR) return 0;
  const Geo &g = c->g;
  LAUNCH(KID_UVW_B, k_east_ghosts, (long)g.N * g.K, g, c->d);
  return post("vel_bc east");
}

int launch_flux_lr_history(nhw_ctx *c) {
  if (c->p.wave.kind != NHW_WAVE_FLUX_LR) return 0;
  const Geo &g = c->g;
  LAUNCH(KID_KMISC, k_flux_lr_history, g.n3(), g, c->d);
  return post("update_vars history");
}

int launch_update_vars(nhw_ctx *c) {
  const Geo &g = c->g;
  hipStream_t s = c->stream;
  const size_t b3 = (c->mixed ? sizeof(float) : sizeof(double)) * g.n3();
  int rc = launch_flux_lr_history(c);
  if (rc) return rc;
  rc |= hip_check(hipMemcpyAsync(c->d.D0, c->d.D, sizeof(double) * g.n2(), hipMemcpyDeviceToDevice, s), "update_vars");
  rc |= hip_check(hipMemcpyAsync(c->d.Eta0, c->d.Eta, sizeof(double) * g.n2(), hipMemcpyDeviceToDevice, s), "update_vars");
  rc |= hip_check(hipMemcpyAsync(c->d.DU0, c->d.DU, b3, hipMemcpyDeviceToDevice, s), "update_vars");
  rc |= hip_check(hipMemcpyAsync(c->d.DV0, c->d.DV, b3, hipMemcpyDeviceToDevice, s), "update_vars");
  rc |= hip_check(hipMemcpyAsync(c->d.DW0, c->d.DW, b3, hipMemcpyDeviceToDevice, s), "update_vars");
  return rc ? NHW_EHIP : 0;
}

int launch_source_terms(nhw_ctx *c) {
  const Geo &g = c->g;
  int rc = launch_int_source(c);
  if (rc) return rc;
  LAUNCH_R(KID_SOURCE, k_source_terms, (long)g.M * g.N, g, c->p, dv, 0);
  return post("source_terms");
}

int launch_fluxes(nhw_ctx *c) {
  const Geo &g = c->g;
  LAUNCH_R(KID_FLUX_X, k_flux_x, flux_x_threads(g), g, c->p, dv, 0);
  LAUNCH_R(KID_FLUX_Y, k_flux_y, flux_y_threads(g), g, c->p, dv, 0);
  LAUNCH_R(KID_FLUX_Z, k_flux_z, (long)g.M * g.N, g, c->p, dv, 0);
  return post("fluxes");
}

int launch_eval_explicit(nhw_ctx *c, int istep) {
  const Geo &g = c->g;
  LAUNCH_R(KID_EXPLICIT, k_eval_explicit, (long)g.M * g.N, g, c->p, dv, istep);
  LAUNCH_R(KID_SOURCE, k_eta_slopes, (long)g.M * g.N, g, c->p, dv, 0);
  return post("eval_duvw");
}

int launch_get_uvw(nhw_ctx *c) {
  const Geo &g = c->g;
  LAUNCH_R(KID_UVW_A, k_uvw_a, (long)g.M * g.N, g, c->p, dv, 1);
  if (int rc = launch_east_ghosts(c)) return rc;
  LAUNCH_R(KID_UVW_B, k_uvw_b, (long)g.M * g.N, g, c->p, dv, 0);
  return post("get_UVW");
}

int launch_p_from_x(nhw_ctx *c, const double *x) {
  const Geo &g = c->g;
  LAUNCH(KID_PROJECTION, k_p_from_x, g.neq(), g, x, c->d, zlayout(c));
  return post("poisson P");
}

int launch_projection(nhw_ctx *c) {
  const Geo &g = c->g;
  LAUNCH_R(KID_PROJECTION, k_projection, g.neq(), g, c->p, dv, 0);
  return post("projection_corrector");
}

int launch_rk_blend(nhw_ctx *c, int istep) {
  const Geo &g = c->g;
  LAUNCH_R(KID_RK_BLEND, k_rk_blend, (long)g.M * g.N, g, c->p, dv, istep);
  return post("rk blend");
}

int launch_get_omega(nhw_ctx *c) {
  const Geo &g = c->g;
  LAUNCH_R(KID_OMEGA_X, k_omega_x, flux_x_threads(g), g, c->p, dv, 0);
  LAUNCH_R(KID_OMEGA_Y, k_omega_y, flux_y_threads(g), g, c->p, dv, 0);
  LAUNCH_R(KID_OMEGA_COL, k_omega_col, (long)g.M * g.N, g, c->p, dv, 0);
  return post("get_Omega");
}


// ----------------------------------------------------------------------------
// batched form: the same bodies over every member of a batch per launch
// ----------------------------------------------------------------------------
template <class R>
static BTab<R> btab(Batch *B, const BShape &S) {
  return BTab<R>{(const BScn<R> *)B->tab, S.off, S.ids, S.nsub};
}

// sum over the members of a kernel's algorithmic bytes per launch
static double bbytes(Batch *B, int kid) {
  double b = 0.0;
  for (nhw_ctx *c : B->m) b += kernel_bytes(c, kid);
  return b;
}

#define BLAUNCH(kid, kern, shape, ...)                                                                   \
  do {                                                                                                   \
    const BShape &S_ = (shape);                                                                          \
    if (S_.nblk > 0) {                                                                                   \
      Timed _t(B->lead, kid, bbytes(B, kid));                                                            \
      if (B->mixed)                                                                                      \
        hipLaunchKernelGGL(kern<float>, dim3(S_.nblk), dim3(S_.bs), 0, B->lead->stream, btab<float>(B, S_), \
                           __VA_ARGS__);                                                                 \
      else                                                                                               \
        hipLaunchKernelGGL(kern<double>, dim3(S_.nblk), dim3(S_.bs), 0, B->lead->stream,                 \
                           btab<double>(B, S_), __VA_ARGS__);                                            \
    }                                                                                                    \
  } while (0)

static const BShape &tuned(Batch *B, const char *key, int kind) {
  return bshape(B, std::string(key) + "/" + std::to_string(B->bs[kind]));
}

// estimate_dt per member: block minima over the member's interior rows ...
template <class R>
__global__ __launch_bounds__(256) void kb_dt_partial(BTab<R> T, double *part) {
  __shared__ double sm[256];
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const Geo &g = S.g;
  const long t = b.b * blockDim.x + threadIdx.x;
  double v = kLarge;
  if (t < (long)g.M * g.N * g.K) {
    const DevT<R> &d = S.d;
    const double *dsig = d.dsig - 1;
    const int i = g.Ibeg + (int)(t % g.M);
    const long r = t / g.M;
    const int j = g.Jbeg + (int)(r % g.N), k = g.Kbeg + (int)(r / g.N);
    const double Dd = d.D[g.c2(i, j)];
    const long q = g.c3(i, j, k);
    const double sq = sqrt(kGrav * Dd);
    double t1 = fabs(d.U[q]) + sq;
    t1 = dmaxd(t1, kSmall);
    double w = g.dx / t1;
    if (w < v) v = w;
    t1 = fabs(d.V[q]) + sq;
    t1 = dmaxd(t1, kSmall);
    w = g.dy / t1;
    if (w < v) v = w;
    t1 = dmaxd(fabs(d.W[q]), kSmall);
    w = dsig[k] * Dd / t1;
    if (w < v) v = w;
  }
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = dmind(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[3 * blockIdx.x] = sm[0];
}

// ... and the step decision, one block per member over its block range
template <class R>
__global__ void kb_dt_final(BTab<R> T, const int *off_rows, const double *part) {
  __shared__ double sm[256];
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  double v = kLarge;
  for (long q = off_rows[b.u] + threadIdx.x; q < off_rows[b.u + 1]; q += blockDim.x) v = dmind(v, part[3 * q]);
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = dmind(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const Par &p = S.p;
    const DevT<R> &d = S.d;
    const double dt_old = d.scal[0];
    const double dt_growth = kDtGrowth * dt_old;
    const double dt_courant = p.cfl * sm[0];
    const double dt_viscous = dt_viscous_limit(S.g, p, d.scal);
    double dt = dmind(dmind(dmind(dt_growth, dt_courant), dt_viscous), p.dt_max);
    d.scal[2] = dt_old;
    d.scal[0] = dt;
    d.scal[3] = dt_courant;
    if (dt < p.dt_min) {
      d.iscal[1] |= 1;
    } else {
      d.scal[1] = d.scal[1] + dt;
      d.iscal[0] += 1;
    }
  }
}

// update_vars over all cells (ghosts included): D0, Eta0 and DU0..DW0
template <class R>
__global__ __launch_bounds__(256) void kb_update_vars(BTab<R> T, int unused) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const long t = b.b * blockDim.x + threadIdx.x;
  const DevT<R> &d = S.d;
  if (t < S.g.n3()) {
    d.DU0[t] = d.DU[t];
    d.DV0[t] = d.DV[t];
    d.DW0[t] = d.DW[t];
  }
  if (t < S.g.n2()) {
    d.D0[t] = d.D[t];
    d.Eta0[t] = d.Eta[t];
  }
}

// the new mask becomes the current one (the single form swaps pointers)
template <class R>
__global__ __launch_bounds__(256) void kb_mask_commit(BTab<R> T, int unused) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  const long t = b.b * blockDim.x + threadIdx.x;
  if (t < S.g.n2()) S.d.Mask[t] = S.d.MaskNew[t];
}

// per-member flags into one array: which 0 -> dt error bits, 1 -> Krylov
// state (flag[0]) and iteration count (flag[1])
template <class R>
__global__ void kb_gather(BTab<R> T, int which, int *out, int ns) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  if (threadIdx.x == 0) {
    if (which == 0) {
      out[b.s] = S.d.iscal[1];
    } else {
      out[b.s] = S.flag[0];
      out[ns + b.s] = S.flag[1];
    }
  }
}

static int bpost(const char *what) { return hip_check(hipGetLastError(), what); }

int blaunch_estimate_dt(Batch *B) {
  const BShape &rows = bshape(B, "c3/256"), &one = bshape(B, "one");
  BLAUNCH(KID_DT, kb_dt_partial, rows, B->part);
  BLAUNCH(KID_KREDUCE, kb_dt_final, one, rows.off, (const double *)B->part);
  return bpost("batch estimate_dt");
}

int blaunch_update_mask(Batch *B) {
  BLAUNCH(KID_MASK, kb_mask_a, bshape(B, "a2/256"), 0);
  BLAUNCH(KID_MASK, kb_mask_b, bshape(B, "c2/256"), 0);
  BLAUNCH(KID_MASK, kb_mask_commit, bshape(B, "a2/256"), 0);
  return bpost("batch update_mask");
}

int blaunch_update_wind(Batch *B) {
  BLAUNCH(KID_SOURCE, kb_wind, bshape(B, "c2/256"), 0);
  return bpost("batch update_wind");
}

int blaunch_update_vars(Batch *B) {
  BLAUNCH(KID_KMISC, kb_update_vars, bshape(B, "a3/256"), 0);
  return bpost("batch update_vars");
}

int blaunch_source_terms(Batch *B) {
  BLAUNCH(KID_SOURCE, kb_source_terms, bshape(B, "c2/256"), 0);
  return bpost("batch source_terms");
}

int blaunch_fluxes(Batch *B) {
  BLAUNCH(KID_FLUX_X, kb_flux_x, tuned(B, "fx", BT_FLUX_X), 0);
  BLAUNCH(KID_FLUX_Y, kb_flux_y, tuned(B, "fy", BT_FLUX_Y), 0);
  BLAUNCH(KID_FLUX_Z, kb_flux_z, tuned(B, "fz", BT_FLUX_Z), 0);
  return bpost("batch fluxes");
}

int blaunch_eval_explicit(Batch *B, int istep) {
  BLAUNCH(KID_EXPLICIT, kb_eval_explicit, bshape(B, "c2/256"), istep);
  BLAUNCH(KID_SOURCE, kb_eta_slopes, bshape(B, "c2/256"), 0);
  return bpost("batch eval_duvw");
}

int blaunch_get_uvw(Batch *B) {
  BLAUNCH(KID_UVW_A, kb_uvw_a, tuned(B, "c2", BT_UVW_A), 1);
  BLAUNCH(KID_UVW_B, kb_uvw_b, tuned(B, "c2", BT_UVW_B), 0);
  return bpost("batch get_UVW");
}

int blaunch_projection(Batch *B) {
  BLAUNCH(KID_PROJECTION, kb_projection, bshape(B, "c3/256"), 0);
  return bpost("batch projection_corrector");
}

int blaunch_rk_blend(Batch *B, int istep) {
  BLAUNCH(KID_RK_BLEND, kb_rk_blend, bshape(B, "c2/256"), istep);
  return bpost("batch rk blend");
}

int blaunch_get_omega(Batch *B) {
  BLAUNCH(KID_OMEGA_X, kb_omega_x, tuned(B, "fx", BT_OMEGA_X), 0);
  BLAUNCH(KID_OMEGA_Y, kb_omega_y, tuned(B, "fy", BT_OMEGA_Y), 0);
  BLAUNCH(KID_OMEGA_COL, kb_omega_col, bshape(B, "c2/256"), 0);
  return bpost("batch get_Omega");
}

int blaunch_gather_flags(Batch *B, int which) {
  BLAUNCH(KID_KMISC, kb_gather, bshape(B, "one"), which, B->gflag, (int)B->m.size());
  return bpost("batch flags");
}

// the arrays the timed sweeps write that the next step reads before
// rewriting them: front-mutated face depths, face surface values, and the
// velocity/momentum fields with their ghosts (U = DU/D is not the uploaded U
// to the last bit)
struct TuneSave {
  std::vector<std::pair<void *, size_t>> what;
  void *buf = nullptr;
};
static int tune_save(Batch *B, TuneSave &ts) {
  const size_t r = B->mixed ? sizeof(float) : sizeof(double);
  size_t total = 0;
  for (nhw_ctx *c : B->m) {
    const Geo &g = c->g;
    Dev &d = c->d;
    for (double *p : {d.Hfx0, d.EtaxL, d.EtaxR}) ts.what.push_back({p, sizeof(double) * g.nx2()});
    for (double *p : {d.Hfy0, d.EtayL, d.EtayR}) ts.what.push_back({p, sizeof(double) * g.ny2()});
    for (double *p : {d.U, d.V, d.W, d.DU, d.DV, d.DW}) ts.what.push_back({p, r * g.n3()});
  }
  for (auto &w : ts.what) total += w.second;
  if (hip_check(hipMalloc(&ts.buf, total), "autotune save")) return NHW_ENOMEM;
  size_t at = 0;
  for (auto &w : ts.what) {
    if (hip_check(hipMemcpyAsync((char *)ts.buf + at, w.first, w.second, hipMemcpyDeviceToDevice,
                                 B->lead->stream), "autotune save"))
      return NHW_EHIP;
    at += w.second;
  }
  return 0;
}
static int tune_restore(Batch *B, TuneSave &ts) {
  size_t at = 0;
  int rc = 0;
  for (auto &w : ts.what) {
    rc |= hip_check(hipMemcpyAsync(w.first, (char *)ts.buf + at, w.second, hipMemcpyDeviceToDevice,
                                   B->lead->stream), "autotune restore");
    at += w.second;
  }
  rc |= hip_check(hipStreamSynchronize(B->lead->stream), "autotune restore");
  (void)hipFree(ts.buf);
  return rc ? NHW_EHIP : 0;
}

// Threads per block of the flux, Omega and get_UVW sweeps, picked by timing
// each candidate on the batch's own state, which is saved before and put
// back after (the sweeps rewrite front face depths, face values and U = DU/D)
int batch_autotune(Batch *B) {
  static const int cand[3] = {64, 128, 256};
  TuneSave ts;
  if (int rc0 = tune_save(B, ts)) return rc0;
  hipEvent_t e0, e1;
  if (hip_check(hipEventCreate(&e0), "autotune") || hip_check(hipEventCreate(&e1), "autotune")) return NHW_EHIP;
  hipStream_t st = B->lead->stream;
  const bool was_on = B->lead->kt.on;
  B->lead->kt.on = false;
  int rc = 0;
  for (int kind = 0; kind < BT_COUNT && !rc; ++kind) {
    double best = 1e30;
    int best_bs = 256;
    for (int ci = 0; ci < 3; ++ci) {
      B->bs[kind] = cand[ci];
      float ms = 0.f;
      for (int rep = 0; rep < 4; ++rep) {  // first repetition warms up
        (void)hipEventRecord(e0, st);
        switch (kind) {
          case BT_FLUX_X: BLAUNCH(KID_FLUX_X, kb_flux_x, tuned(B, "fx", kind), 0); break;
          case BT_FLUX_Y: BLAUNCH(KID_FLUX_Y, kb_flux_y, tuned(B, "fy", kind), 0); break;
          case BT_FLUX_Z: BLAUNCH(KID_FLUX_Z, kb_flux_z, tuned(B, "fz", kind), 0); break;
          case BT_OMEGA_X: BLAUNCH(KID_OMEGA_X, kb_omega_x, tuned(B, "fx", kind), 0); break;
          case BT_OMEGA_Y: BLAUNCH(KID_OMEGA_Y, kb_omega_y, tuned(B, "fy", kind), 0); break;
          case BT_UVW_A: BLAUNCH(KID_UVW_A, kb_uvw_a, tuned(B, "c2", kind), 1); break;
          default: BLAUNCH(KID_UVW_B, kb_uvw_b, tuned(B, "c2", kind), 0); break;
        }
        (void)hipEventRecord(e1, st);
        if ((rc = hip_check(hipEventSynchronize(e1), "autotune"))) break;
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        if (rep > 0) ms += t;
      }
      if (rc) break;
      B->tune_ms[kind][ci] = ms / 3.0;
      if (ms < best) {
        best = ms;
        best_bs = cand[ci];
      }
    }
    B->bs[kind] = best_bs;
  }
  B->lead->kt.on = was_on;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const int rc2 = tune_restore(B, ts);
  return rc ? rc : rc2;
}

}  // namespace nhw
