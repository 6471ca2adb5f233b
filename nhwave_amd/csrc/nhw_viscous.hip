// Viscous branch of eval_duvw with constant viscosities (VISCOUS_FLOW = T,
// IVTURB = IHTURB = 1), single scenario, FP64:
//   * k_diffusion: the stress terms of diffusion (src/nhwave.F:6785-7141),
//     one thread per interior cell, at the start of eval_duvw;
//   * k_visc_sigma: sigma_transform's DelxSc/DelySc (cell) and DelxSl/DelySl
//     (sigma level) after the mass update (src/nhwave.F:634, 6505-6541); the
//     next stage's diffusion reads them;
//   * k_visc_solve: the momentum right-hand sides with the stress terms and
//     the implicit vertical solves of DU, DV, DW per wet column
//     (src/nhwave.F:637-688, 693-1144, trig :13182-13224), one thread per
//     column.
// The viscosities are uniform: Cmu = Visc everywhere (src/initialize.F:2666),
// CmuR = 0, and CmuHt/CmuVt hold the turbulence seed Cmut_min
// (src/initialize.F:3172-3196) until the first eval_turb and Chs/Cvs after it
// (src/nhwave.F:6650-6685).  They enter as scalars inside the reference's
// own expressions, which round the same as the uniform arrays do.
#include "nhw_impl.h"

namespace nhw {

namespace {

// sigma-derivative weights of level k on the non-uniform grid and the
// Lagrange weights at k +- 1/2 (src/nhwave.F:6813-6827)
struct SigW {
  double L1top, L2top, L1bot, L2bot, dsigck, dsigck1, ac, bc, gc;
};
__device__ __forceinline__ SigW sig_weights(const double *dsig, int k) {
  SigW w;
  w.L1top = dsig[k + 1] / (dsig[k] + dsig[k + 1]);
  w.L2top = dsig[k] / (dsig[k] + dsig[k + 1]);
  w.L1bot = dsig[k] / (dsig[k] + dsig[k - 1]);
  w.L2bot = dsig[k - 1] / (dsig[k] + dsig[k - 1]);
  w.dsigck = (dsig[k] + dsig[k + 1]) / 2.0;
  w.dsigck1 = (dsig[k - 1] + dsig[k]) / 2.0;
  w.ac = -w.dsigck / (w.dsigck + w.dsigck1) / w.dsigck1;
  w.bc = (w.dsigck - w.dsigck1) / (w.dsigck * w.dsigck1);
  w.gc = w.dsigck1 / (w.dsigck + w.dsigck1) / w.dsigck;
  return w;
}

// a 3-D cell field seen from cell (i,j,k): f(di,dj,dk)
struct Nb {
  const double *p;
  long c, sj, sk;
  __device__ double operator()(int di, int dj, int dk) const { return p[c + di + dj * sj + dk * sk]; }
};

}  // namespace

// Diffxx, Diffxy, Diffyx, Diffyy, Diffzx, Diffzy at interior cell t;
// Diffxz/Diffyz feed no momentum equation and are not formed.
__global__ __launch_bounds__(256) void k_diffusion(Geo g, Par p, Dev d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long mn = (long)g.M * g.N;
  if (t >= mn * g.K) return;
  // i fastest, then k, then j: the k-1 and k+1 planes of a row are read by
  // neighbouring blocks and the j-1, j+1 rows a few dozen blocks apart, so
  // both are still in L2 (row-plane-slowest order fetched 3.3x the bytes)
  const int i = g.Ibeg + (int)(t % g.M), k = g.Kbeg + (int)((t / g.M) % g.K), j = g.Jbeg + (int)(t / ((long)g.M * g.K));
  const long c = g.c3(i, j, k), n3 = g.n3();
  double *out = d.Diff;
  if (d.Mask[g.c2(i, j)] == 0) {
    for (int q = 0; q < 6; ++q) out[q * n3 + c] = 0.0;
    return;
  }
  const double dx = g.dx, dy = g.dy, S = p.schmidt, cm = p.visc, ch = p.cmu_ht;
  const double *dsig = d.dsig - 1;
  const SigW w = sig_weights(dsig, k);
  const double ac = w.ac, bc = w.bc, gc = w.gc, dsigck = w.dsigck, dsigck1 = w.dsigck1;
  const long sj = g.Mloc, sk = (long)g.Mloc * g.Nloc;
  const Nb DU{d.DU, c, sj, sk}, DV{d.DV, c, sj, sk}, DW{d.DW, c, sj, sk};
  const Nb SX{d.DelxSc, c, sj, sk}, SY{d.DelySc, c, sj, sk};
  const Nb LX{d.DelxSl, c, sj, sk}, LY{d.DelySl, c, sj, sk};  // z-face arrays share the cell strides
  const double Dc = d.D[g.c2(i, j)];
  const double nuH_top = (w.L1top * cm + w.L2top * cm) + (w.L1top * ch + w.L2top * ch) / S;
  const double nuH_bot = (w.L1bot * cm + w.L2bot * cm) + (w.L1bot * ch + w.L2bot * ch) / S;
  const double hf = 0.5 * (cm + cm) + 0.5 * (ch + ch) / S;  // average over two cells
  const double hc = cm + ch / S;                             // one cell
  const double lxd = nuH_top * LX(0, 0, 1) - nuH_bot * LX(0, 0, 0);
  const double lxs = nuH_top * LX(0, 0, 1) + nuH_bot * LX(0, 0, 0);
  const double lyd = nuH_top * LY(0, 0, 1) - nuH_bot * LY(0, 0, 0);
  const double lys = nuH_top * LY(0, 0, 1) + nuH_bot * LY(0, 0, 0);
  // sigma-sigma coupling weights of levels k-1, k, k+1 (:6911-6928 and alike)
  const double wxm = lxd * ac / dsig[k] + lxs / dsigck1 / (dsigck + dsigck1);
  const double wx0 = lxd * bc / dsig[k] - lxs / dsigck1 / dsigck;
  const double wxp = lxd * gc / dsig[k] + lxs / dsigck / (dsigck + dsigck1);
  const double wym = lyd * ac / dsig[k] + lys / dsigck1 / (dsigck + dsigck1);
  const double wy0 = lyd * bc / dsig[k] - lys / dsigck1 / dsigck;
  const double wyp = lyd * gc / dsig[k] + lys / dsigck / (dsigck + dsigck1);

  // d/dx[2 nu d/dx DU] + d/dx[2 nu d/dsig(DU DelxSc)] + d/dsig[2 nu DelxSc d/dx DU]
  const double xx =
      (2.0 * hf * (DU(1, 0, 0) - DU(0, 0, 0)) / dx - 2.0 * hf * (DU(0, 0, 0) - DU(-1, 0, 0)) / dx) / dx +
      (2.0 * hc * (ac * DU(1, 0, -1) * SX(1, 0, -1) + bc * DU(1, 0, 0) * SX(1, 0, 0) + gc * DU(1, 0, 1) * SX(1, 0, 1)) -
       2.0 * hc * (ac * DU(-1, 0, -1) * SX(-1, 0, -1) + bc * DU(-1, 0, 0) * SX(-1, 0, 0) + gc * DU(-1, 0, 1) * SX(-1, 0, 1))) /
          2.0 / dx +
      (ac * 2.0 * hc * SX(0, 0, -1) * (DU(1, 0, -1) - DU(-1, 0, -1)) / 2.0 / dx +
       bc * 2.0 * hc * SX(0, 0, 0) * (DU(1, 0, 0) - DU(-1, 0, 0)) / 2.0 / dx +
       gc * 2.0 * hc * SX(0, 0, 1) * (DU(1, 0, 1) - DU(-1, 0, 1)) / 2.0 / dx);
  // cross-term helper: DelxSc DV + DelySc DU at a neighbour level
  auto mix = [&](int di, int dj, int dk) { return SX(di, dj, dk) * DV(di, dj, dk) + SY(di, dj, dk) * DU(di, dj, dk); };
  const double xy =
      (hf * (DU(0, 1, 0) - DU(0, 0, 0)) / dy - hf * (DU(0, 0, 0) - DU(0, -1, 0)) / dy) / dy +
      (hc * ((DV(1, 1, 0) - DV(-1, 1, 0)) / 2.0 / dx + ac * mix(0, 1, -1) + bc * mix(0, 1, 0) + gc * mix(0, 1, 1)) -
       hc * ((DV(1, -1, 0) - DV(-1, -1, 0)) / 2.0 / dx + ac * mix(0, -1, -1) + bc * mix(0, -1, 0) + gc * mix(0, -1, 1))) /
          2.0 / dy +
      (ac * hc * SY(0, 0, -1) * ((DU(0, 1, -1) - DU(0, -1, -1)) / 2.0 / dy + (DV(1, 0, -1) - DV(-1, 0, -1)) / 2.0 / dx) +
       bc * hc * SY(0, 0, 0) * ((DU(0, 1, 0) - DU(0, -1, 0)) / 2.0 / dy + (DV(1, 0, 0) - DV(-1, 0, 0)) / 2.0 / dx) +
       gc * hc * SY(0, 0, 1) * ((DU(0, 1, 1) - DU(0, -1, 1)) / 2.0 / dy + (DV(1, 0, 1) - DV(-1, 0, 1)) / 2.0 / dx)) +
      (wym * SX(0, 0, -1) * DV(0, 0, -1) + wy0 * SX(0, 0, 0) * DV(0, 0, 0) + wyp * SX(0, 0, 1) * DV(0, 0, 1));
  const double yx =
      (hf * (DV(1, 0, 0) - DV(0, 0, 0)) / dx - hf * (DV(0, 0, 0) - DV(-1, 0, 0)) / dx) / dx +
      (hc * ((DU(1, 1, 0) - DU(1, -1, 0)) / 2.0 / dy + ac * mix(1, 0, -1) + bc * mix(1, 0, 0) + gc * mix(1, 0, 1)) -
       hc * ((DU(-1, 1, 0) - DU(-1, -1, 0)) / 2.0 / dy + ac * mix(-1, 0, -1) + bc * mix(-1, 0, 0) + gc * mix(-1, 0, 1))) /
          2.0 / dx +
      (ac * hc * SX(0, 0, -1) * ((DU(0, 1, -1) - DU(0, -1, -1)) / 2.0 / dy + (DV(1, 0, -1) - DV(-1, 0, -1)) / 2.0 / dx) +
       bc * hc * SX(0, 0, 0) * ((DU(0, 1, 0) - DU(0, -1, 0)) / 2.0 / dy + (DV(1, 0, 0) - DV(-1, 0, 0)) / 2.0 / dx) +
       gc * hc * SX(0, 0, 1) * ((DU(0, 1, 1) - DU(0, -1, 1)) / 2.0 / dy + (DV(1, 0, 1) - DV(-1, 0, 1)) / 2.0 / dx)) +
      (wxm * SY(0, 0, -1) * DU(0, 0, -1) + wx0 * SY(0, 0, 0) * DU(0, 0, 0) + wxp * SY(0, 0, 1) * DU(0, 0, 1));
  const double yy =
      (2.0 * hf * (DV(0, 1, 0) - DV(0, 0, 0)) / dy - 2.0 * hf * (DV(0, 0, 0) - DV(0, -1, 0)) / dy) / dy +
      (2.0 * hc * (ac * DV(0, 1, -1) * SY(0, 1, -1) + bc * DV(0, 1, 0) * SY(0, 1, 0) + gc * DV(0, 1, 1) * SY(0, 1, 1)) -
       2.0 * hc * (ac * DV(0, -1, -1) * SY(0, -1, -1) + bc * DV(0, -1, 0) * SY(0, -1, 0) + gc * DV(0, -1, 1) * SY(0, -1, 1))) /
          2.0 / dy +
      (ac * 2.0 * hc * SY(0, 0, -1) * (DV(0, 1, -1) - DV(0, -1, -1)) / 2.0 / dy +
       bc * 2.0 * hc * SY(0, 0, 0) * (DV(0, 1, 0) - DV(0, -1, 0)) / 2.0 / dy +
       gc * 2.0 * hc * SY(0, 0, 1) * (DV(0, 1, 1) - DV(0, -1, 1)) / 2.0 / dy);
  const double zx =
      (hf * (DW(1, 0, 0) - DW(0, 0, 0)) / dx - hf * (DW(0, 0, 0) - DW(-1, 0, 0)) / dx) / dx +
      (hc * (ac * (DW(1, 0, -1) * SX(1, 0, -1) + DU(1, 0, -1) / Dc) + bc * (DW(1, 0, 0) * SX(1, 0, 0) + DU(1, 0, 0) / Dc) +
             gc * (DW(1, 0, 1) * SX(1, 0, 1) + DU(1, 0, 1) / Dc)) -
       hc * (ac * (DW(-1, 0, -1) * SX(-1, 0, -1) + DU(-1, 0, -1) / Dc) + bc * (DW(-1, 0, 0) * SX(-1, 0, 0) + DU(-1, 0, 0) / Dc) +
             gc * (DW(-1, 0, 1) * SX(-1, 0, 1) + DU(-1, 0, 1) / Dc))) /
          2.0 / dx +
      (ac * hc * SX(0, 0, -1) * (DW(1, 0, -1) - DW(-1, 0, -1)) / 2.0 / dx +
       bc * hc * SX(0, 0, 0) * (DW(1, 0, 0) - DW(-1, 0, 0)) / 2.0 / dx +
       gc * hc * SX(0, 0, 1) * (DW(1, 0, 1) - DW(-1, 0, 1)) / 2.0 / dx) +
      (wxm * DU(0, 0, -1) / Dc + wx0 * DU(0, 0, 0) / Dc + wxp * DU(0, 0, 1) / Dc);
  const double zy =
      (hf * (DW(0, 1, 0) - DW(0, 0, 0)) / dy - hf * (DW(0, 0, 0) - DW(0, -1, 0)) / dy) / dy +
      (hc * (ac * (DW(0, 1, -1) * SY(0, 1, -1) + DV(0, 1, -1) / Dc) + bc * (DW(0, 1, 0) * SY(0, 1, 0) + DV(0, 1, 0) / Dc) +
             gc * (DW(0, 1, 1) * SY(0, 1, 1) + DV(0, 1, 1) / Dc)) -
       hc * (ac * (DW(0, -1, -1) * SY(0, -1, -1) + DV(0, -1, -1) / Dc) + bc * (DW(0, -1, 0) * SY(0, -1, 0) + DV(0, -1, 0) / Dc) +
             gc * (DW(0, -1, 1) * SY(0, -1, 1) + DV(0, -1, 1) / Dc))) /
          2.0 / dy +
      (ac * hc * SY(0, 0, -1) * (DW(0, 1, -1) - DW(0, -1, -1)) / 2.0 / dy +
       bc * hc * SY(0, 0, 0) * (DW(0, 1, 0) - DW(0, -1, 0)) / 2.0 / dy +
       gc * hc * SY(0, 0, 1) * (DW(0, 1, 1) - DW(0, -1, 1)) / 2.0 / dy) +
      (wym * DV(0, 0, -1) / Dc + wy0 * DV(0, 0, 0) / Dc + wyp * DV(0, 0, 1) / Dc);
  out[c] = xx;
  out[n3 + c] = xy;
  out[2 * n3 + c] = yx;
  out[3 * n3 + c] = yy;
  out[4 * n3 + c] = zx;
  out[5 * n3 + c] = zy;
}

// one limited slope of delxFun_2D/delyFun_2D (src/nhwave.F:1890-1960)
__device__ __forceinline__ double eta_slope(double qm, double q0, double qp, double h, int mask) {
  if (mask == 0) return 0.0;
  const double t1 = (qp - q0) / h, t2 = (q0 - qm) / h;
  if ((fabs(t1) + fabs(t2)) < kSmall) return 0.0;
  return (t1 * fabs(t2) + fabs(t1) * t2) / (fabs(t1) + fabs(t2));
}

// sigma_transform over every column (ghosts included), src/nhwave.F:6505-6541.
// DelxEta/DelyEta at the ghost cells are those of delxFun_2D/delyFun_2D over
// the whole array (limited slopes at 2..Mloc-1, one-sided at the two ends);
// the interior ones, already written by the mass-update pass, are recomputed
// to the same values.
__global__ __launch_bounds__(256) void k_visc_sigma(Geo g, Dev d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.n2()) return;
  const int i = 1 + (int)(t % g.Mloc), j = 1 + (int)(t / g.Mloc);
  const double *E = d.Eta;
  const int m = d.Mask[t];
  const long sj = g.Mloc;
  const double ex = i == 1 ? (E[t + 1] - E[t]) / g.dx
                  : i == g.Mloc ? (E[t] - E[t - 1]) / g.dx : eta_slope(E[t - 1], E[t], E[t + 1], g.dx, m);
  const double ey = j == 1 ? (E[t + sj] - E[t]) / g.dy
                  : j == g.Nloc ? (E[t] - E[t - sj]) / g.dy : eta_slope(E[t - sj], E[t], E[t + sj], g.dy, m);
  d.DelxEta[t] = ex;
  d.DelyEta[t] = ey;
  const double Dc = d.D[t], hx = d.DelxH[t], hy = d.DelyH[t];
  const double *sigc = d.sigc - 1, *sig = d.sig - 1;
  const long n2 = g.n2();
  for (int k = 1; k <= g.Kloc; ++k) {
    d.DelxSc[(k - 1) * n2 + t] = (1.0 - sigc[k]) / Dc * hx - sigc[k] / Dc * ex;
    d.DelySc[(k - 1) * n2 + t] = (1.0 - sigc[k]) / Dc * hy - sigc[k] / Dc * ey;
  }
  for (int k = 1; k <= g.Kloc1; ++k) {
    d.DelxSl[(k - 1) * n2 + t] = (1. - sig[k]) / Dc * hx - sig[k] / Dc * ex;
    d.DelySl[(k - 1) * n2 + t] = (1. - sig[k]) / Dc * hy - sig[k] / Dc * ey;
  }
}

// Momentum right-hand sides and the implicit vertical solves (src/nhwave.F:
// 637-688, 693-1144), one thread per interior column, 64 columns per block.
// The three tridiagonal systems run through trig's Thomas sweeps
// (:13182-13224) together, level by level: the forward elimination writes
// each level's eliminated right side over DU/DV/DW (the old value at a level
// is read before it is overwritten, and nothing reads it later) and keeps the
// eliminated diagonals in LDS (DU and DV share theirs: same coefficients);
// the back substitution then overwrites them with the solution.  Dry columns
// get DU = DV = DW = 0 (:1172-1182).  Runs after the mass update,
// DelxEta/DelyEta and sigma_transform of this stage.
__global__ __launch_bounds__(64) void k_visc_solve(Geo g, Par p, Dev d) {
  extern __shared__ double lds[];  // [2][K][64]: eliminated diagonals of DU/DV, of DW
  const int lane = threadIdx.x;
  const long t = (long)blockIdx.x * blockDim.x + lane;
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c2 = g.c2(i, j), n3 = g.n3(), sk = g.n2();
  if (d.Mask[c2] == 0) {
    for (int k = g.Kbeg; k <= g.Kend; ++k) {
      const long q = g.c3(i, j, k);
      d.DU[q] = 0.0; d.DV[q] = 0.0; d.DW[q] = 0.0;
    }
    return;
  }
  double *bpu = lds, *bpw = lds + (long)g.K * 64;
  const double dt = d.scal[0], dx = g.dx, dy = g.dy, fcor = p.fcor;
  const double *dsig = d.dsig - 1, *Df = d.Diff;
  const double SX = d.SourceX[c2], SY = d.SourceY[c2], wsx = d.Wsx[c2], wsy = d.Wsy[c2];
  const double Dc = d.D[c2], S = p.schmidt, cm = p.visc, cv = p.cmu_vt;
  const double nu = 0.5 * (cm + cm + 0.0 + 0.0) + 0.5 * (cv + cv) / S;  // face viscosity
  double bu = 0.0, bw = 0.0, ru = 0.0, rv = 0.0, rw = 0.0, cprev = 0.0;
  for (int k = g.Kbeg; k <= g.Kend; ++k) {
    const int q = k - g.Kbeg;
    const long c = g.c3(i, j, k), fx0 = g.x3(i, j, k), fy0 = g.y3(i, j, k), fz1 = c + sk;
    const double du = d.DU[c], dv = d.DV[c], dw = d.DW[c];
    const double rdz = 1.0 / dsig[k];
    const double r2 = -1.0 / dx * (d.Fx[fx0 + 1] - d.Fx[fx0]) - 1.0 / dy * (d.Fy[fy0 + g.Mloc] - d.Fy[fy0]) -
                      rdz * (d.Fz[fz1] - d.Fz[c]) + fcor * dv + 0.0 + SX + Df[c] + Df[n3 + c] + 0.0;
    const double r3 = -1.0 / dx * (d.Gx[fx0 + 1] - d.Gx[fx0]) - 1.0 / dy * (d.Gy[fy0 + g.Mloc] - d.Gy[fy0]) -
                      rdz * (d.Gz[fz1] - d.Gz[c]) - fcor * du + 0.0 + SY + Df[2 * n3 + c] + Df[3 * n3 + c] + 0.0;
    const double r4 = -1.0 / dx * (d.Hx[fx0 + 1] - d.Hx[fx0]) - 1.0 / dy * (d.Hy[fy0 + g.Mloc] - d.Hy[fy0]) -
                      rdz * (d.Hz[fz1] - d.Hz[c]) + Df[4 * n3 + c] + Df[5 * n3 + c];
    const double a = k == g.Kbeg ? 0.0 : -dt / (Dc * Dc) * nu / (0.5 * dsig[k] * (dsig[k] + dsig[k - 1]));
    const double cc = k == g.Kend ? 0.0 : -dt / (Dc * Dc) * nu / (0.5 * dsig[k] * (dsig[k] + dsig[k + 1]));
    // rows of DU and DV (wind stress on the surface row) and of DW (kinematic
    // W on the bottom and surface rows)
    // no-slip bottom row (Bc_Z0 = 2), :718-722
    double b = k == g.Kbeg && p.bc_z0 == 2
                   ? 1.0 - cc + dt / (Dc * Dc) * (cm + 0.0 + cv / S) / (0.5 * dsig[k] * dsig[k])
                   : 1.0 - a - cc;
    double xu, xv;
    if (k == g.Kbeg && p.bc_z0 == 5) {  // friction-law bottom row, :773-785
      const double Dz1 = 0.5 * Dc * dsig[g.Kbeg];
      double cdrag = p.cd0;
      if (p.ibot != 1) {
        const double tk = 1. / kKappa * log(30.0 * Dz1 / p.zob);
        cdrag = 1. / (tk * tk);
      }
      const double u = d.U[c], v = d.V[c], sp = sqrt(u * u + v * v);
      xu = du + dt * r2 - dt * (cdrag * sp * u) / dsig[k];
      xv = dv + dt * r3 - dt * (cdrag * sp * v) / dsig[k];
    } else {
      xu = k == g.Kend ? du + dt * r2 + dt * wsx / dsig[k] : du + dt * r2;
      xv = k == g.Kend ? dv + dt * r3 + dt * wsy / dsig[k] : dv + dt * r3;
    }
    double b3, xw;
    if (k == g.Kbeg) {
      const double y = 0.25 * dsig[k] * (dsig[k] + dsig[k - 1]);
      b3 = 1.0 - cc + dt / (Dc * Dc) * nu / y;
      // -DeltH - U DelxH - V DelyH with a fixed bed (DeltH = 0)
      const double Wbot = -0.0 - d.U[c] * d.DelxH[c2] - d.V[c] * d.DelyH[c2];
      xw = dw + dt * r4 + dt / Dc * nu / y * Wbot;
    } else if (k == g.Kend) {
      const double y = 0.25 * dsig[k] * (dsig[k] + dsig[k + 1]);
      b3 = 1.0 - a + dt / (Dc * Dc) * nu / y;
      const double Wtop = (d.Eta[c2] - d.Eta0[c2]) / dt + d.U[c] * d.DelxEta[c2] + d.V[c] * d.DelyEta[c2];
      xw = dw + dt * r4 + dt / Dc * nu / y * Wtop;
    } else {
      b3 = 1.0 - a - cc;
      xw = dw + dt * r4;
    }
    if (q > 0) {  // forward elimination
      const double coeff = a / bu;
      b = b - coeff * cprev;
      xu = xu - coeff * ru;
      xv = xv - coeff * rv;
      const double coeffw = a / bw;
      b3 = b3 - coeffw * cprev;
      xw = xw - coeffw * rw;
    }
    bu = b; bw = b3; ru = xu; rv = xv; rw = xw; cprev = cc;
    bpu[q * 64 + lane] = b;
    bpw[q * 64 + lane] = b3;
    d.DU[c] = xu; d.DV[c] = xv; d.DW[c] = xw;
  }
  // back substitution
  double xu = ru / bu, xv = rv / bu, xw = rw / bw;
  {
    const long c = g.c3(i, j, g.Kend);
    d.DU[c] = xu; d.DV[c] = xv; d.DW[c] = xw;
  }
  for (int k = g.Kend - 1; k >= g.Kbeg; --k) {
    const int q = k - g.Kbeg;
    const long c = g.c3(i, j, k);
    const double cc = -dt / (Dc * Dc) * nu / (0.5 * dsig[k] * (dsig[k] + dsig[k + 1]));
    const double b = bpu[q * 64 + lane], b3 = bpw[q * 64 + lane];
    xu = (d.DU[c] - cc * xu) / b;
    xv = (d.DV[c] - cc * xv) / b;
    xw = (d.DW[c] - cc * xw) / b3;
    d.DU[c] = xu; d.DV[c] = xv; d.DW[c] = xw;
  }
}

int viscous_setup(nhw_ctx *c) {
  if (!c->cfg.viscous) return 0;
  const Geo &g = c->g;
  Dev &d = c->d;
  const long n3 = g.n3(), nz = g.nz3();
  struct A { double **p; long n; } al[] = {{&d.DelxSc, n3}, {&d.DelySc, n3}, {&d.DelxSl, nz}, {&d.DelySl, nz},
                                           {&d.Diff, 6 * n3}};
  for (auto &x : al) {
    if (hipMalloc(x.p, sizeof(double) * x.n) != hipSuccess ||
        hipMemset(*x.p, 0, sizeof(double) * x.n) != hipSuccess) {
      set_error("nhw_create: device allocation failed (viscous arrays)");
      return NHW_ENOMEM;
    }
  }
  return 0;
}

void viscous_free(nhw_ctx *c) {
  Dev &d = c->d;
  for (double **p : {&d.DelxSc, &d.DelySc, &d.DelxSl, &d.DelySl, &d.Diff}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

int launch_diffusion(nhw_ctx *c) {
  if (!c->cfg.viscous) return 0;
  const Geo &g = c->g;
  const long n = (long)g.M * g.N * g.K;
  Timed tm(c, KID_DIFFUSION);
  hipLaunchKernelGGL(k_diffusion, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, g, c->p, c->d);
  return hip_check(hipGetLastError(), "k_diffusion");
}

int launch_visc_solve(nhw_ctx *c) {
  if (!c->cfg.viscous) return 0;
  const Geo &g = c->g;
  {
    Timed tm(c, KID_VISC_SIGMA);
    hipLaunchKernelGGL(k_visc_sigma, dim3((unsigned)((g.n2() + 255) / 256)), dim3(256), 0, c->stream, g, c->d);
    if (hip_check(hipGetLastError(), "k_visc_sigma")) return NHW_EHIP;
  }
  const long n = (long)g.M * g.N;
  const size_t lds = sizeof(double) * 2 * (size_t)g.K * 64;
  Timed tm(c, KID_VISC_SOLVE);
  hipLaunchKernelGGL(k_visc_solve, dim3((unsigned)((n + 63) / 64)), dim3(64), lds, c->stream, g, c->p, c->d);
  return hip_check(hipGetLastError(), "k_visc_solve");
}

// eval_turb with IVturb = IHturb = 1 (src/nhwave.F:6650-6685): from the first
// call on, CmuVt = Cvs and CmuHt = Chs.  Kernel arguments carry them.
void eval_turb(nhw_ctx *c) {
  if (!c->cfg.viscous) return;
  c->p.cmu_vt = c->cfg.cvs;
  c->p.cmu_ht = c->cfg.chs;
}

// the turbulence seed, src/initialize.F:3172-3196 (REAL(4) literals)
void turb_seed(nhw_ctx *c) {
  if (!c->cfg.viscous) return;
  const double tke_min = (double)1.e-9f, eps_min = (double)1.e-9f;
  const double cmut = (c->cfg.rng ? (double)8.5e-2f : (double)9.0e-2f) * (tke_min * tke_min) / eps_min;
  c->p.cmu_ht = cmut;
  c->p.cmu_vt = cmut;
}

}  // namespace nhw
