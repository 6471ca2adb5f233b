// Viscous branch of eval_duvw (VISCOUS_FLOW = T), single scenario, FP64, with
// constant viscosities (IVTURB = IHTURB = 1) or the k-epsilon closure
// (IVTURB = 10, IHTURB >= 10):
//   * k_diffusion: the stress terms of diffusion (src/nhwave.F:6785-7141),
//     one thread per interior cell, at the start of eval_duvw;
//   * k_visc_sigma: sigma_transform's DelxSc/DelySc (cell) and DelxSl/DelySl
//     (sigma level) after the mass update (src/nhwave.F:634, 6505-6541); the
//     next stage's diffusion reads them;
//   * k_visc_solve: the momentum right-hand sides with the stress terms and
//     the implicit vertical solves of DU, DV, DW per wet column
//     (src/nhwave.F:637-688, 693-1144, trig :13182-13224), one thread per
//     column;
//   * k_ke_column, k_ke_final: eval_turb's k-epsilon model (below).
// Cmu = Visc everywhere (src/initialize.F:2666) and CmuR = 0.  With constant
// viscosities CmuHt/CmuVt hold the turbulence seed Cmut_min
// (src/initialize.F:3172-3196) until the first eval_turb and Chs/Cvs after it
// (src/nhwave.F:6650-6685); they enter as scalars inside the reference's own
// expressions, which round the same as the uniform arrays do.  With the
// k-epsilon model they are per-cell arrays (CmuHt = CmuVt).
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

// a 3-D cell field seen from cell (i,j,k): f(di,dj,dk), read as FP64 (T:
// the field's storage type)
template <class T>
struct NbT {
  const T *p;
  long c, sj, sk;
  __device__ double operator()(int di, int dj, int dk) const { return (double)p[c + di + dj * sj + dk * sk]; }
};
using Nb = NbT<double>;

// a turbulent viscosity seen from cell (i,j,k): the per-cell array of the
// k-epsilon model (F) or the uniform value of constant viscosities
template <bool F>
struct Visc {
  const double *p;
  double u;
  long c, sj, sk;
  __device__ double operator()(int di, int dj, int dk) const { return F ? p[c + di + dj * sj + dk * sk] : u; }
};

}  // namespace

// Diffxx, Diffxy, Diffyx, Diffyy, Diffzx, Diffzy at interior cell t;
// Diffxz/Diffyz feed no momentum equation and are not formed.  F: CmuHt
// per cell (k-epsilon model), else the uniform value.  R: the storage type
// of the 3-D momenta (float in the mixed form; the arithmetic is FP64).
template <class R, bool F>
__global__ __launch_bounds__(256) void k_diffusion(Geo g, Par p, DevT<R> d) {
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
  const double dx = g.dx, dy = g.dy, S = p.schmidt, cm = p.visc;
  const double *dsig = d.dsig - 1;
  const SigW w = sig_weights(dsig, k);
  const double ac = w.ac, bc = w.bc, gc = w.gc, dsigck = w.dsigck, dsigck1 = w.dsigck1;
  const long sj = g.Mloc, sk = (long)g.Mloc * g.Nloc;
  const NbT<R> DU{d.DU, c, sj, sk}, DV{d.DV, c, sj, sk}, DW{d.DW, c, sj, sk};
  const Nb SX{d.DelxSc, c, sj, sk}, SY{d.DelySc, c, sj, sk};
  const Nb LX{d.DelxSl, c, sj, sk}, LY{d.DelySl, c, sj, sk};  // z-face arrays share the cell strides
  const Visc<F> CH{d.CmuHt, F ? 0.0 : d.scal[6], c, sj, sk};
  const double Dc = d.D[g.c2(i, j)];
  const double nuH_top = (w.L1top * cm + w.L2top * cm) + (w.L1top * CH(0, 0, 0) + w.L2top * CH(0, 0, 1)) / S;
  const double nuH_bot = (w.L1bot * cm + w.L2bot * cm) + (w.L1bot * CH(0, 0, -1) + w.L2bot * CH(0, 0, 0)) / S;
  // Cmu + CmuHt/Schmidt averaged over the two cells of the faces i+-1/2,
  // j+-1/2, and at single cells (east, west, north, south, below, here, above)
  const double hfe = 0.5 * (cm + cm) + 0.5 * (CH(1, 0, 0) + CH(0, 0, 0)) / S;
  const double hfw = 0.5 * (cm + cm) + 0.5 * (CH(0, 0, 0) + CH(-1, 0, 0)) / S;
  const double hfn = 0.5 * (cm + cm) + 0.5 * (CH(0, 1, 0) + CH(0, 0, 0)) / S;
  const double hfs = 0.5 * (cm + cm) + 0.5 * (CH(0, 0, 0) + CH(0, -1, 0)) / S;
  const double hE = cm + CH(1, 0, 0) / S, hW = cm + CH(-1, 0, 0) / S, hN = cm + CH(0, 1, 0) / S;
  const double hS = cm + CH(0, -1, 0) / S, hB = cm + CH(0, 0, -1) / S, hC = cm + CH(0, 0, 0) / S;
  const double hT = cm + CH(0, 0, 1) / S;
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
      (2.0 * hfe * (DU(1, 0, 0) - DU(0, 0, 0)) / dx - 2.0 * hfw * (DU(0, 0, 0) - DU(-1, 0, 0)) / dx) / dx +
      (2.0 * hE * (ac * DU(1, 0, -1) * SX(1, 0, -1) + bc * DU(1, 0, 0) * SX(1, 0, 0) + gc * DU(1, 0, 1) * SX(1, 0, 1)) -
       2.0 * hW * (ac * DU(-1, 0, -1) * SX(-1, 0, -1) + bc * DU(-1, 0, 0) * SX(-1, 0, 0) + gc * DU(-1, 0, 1) * SX(-1, 0, 1))) /
          2.0 / dx +
      (ac * 2.0 * hB * SX(0, 0, -1) * (DU(1, 0, -1) - DU(-1, 0, -1)) / 2.0 / dx +
       bc * 2.0 * hC * SX(0, 0, 0) * (DU(1, 0, 0) - DU(-1, 0, 0)) / 2.0 / dx +
       gc * 2.0 * hT * SX(0, 0, 1) * (DU(1, 0, 1) - DU(-1, 0, 1)) / 2.0 / dx);
  // cross-term helper: DelxSc DV + DelySc DU at a neighbour level
  auto mix = [&](int di, int dj, int dk) { return SX(di, dj, dk) * DV(di, dj, dk) + SY(di, dj, dk) * DU(di, dj, dk); };
  const double xy =
      (hfn * (DU(0, 1, 0) - DU(0, 0, 0)) / dy - hfs * (DU(0, 0, 0) - DU(0, -1, 0)) / dy) / dy +
      (hN * ((DV(1, 1, 0) - DV(-1, 1, 0)) / 2.0 / dx + ac * mix(0, 1, -1) + bc * mix(0, 1, 0) + gc * mix(0, 1, 1)) -
       hS * ((DV(1, -1, 0) - DV(-1, -1, 0)) / 2.0 / dx + ac * mix(0, -1, -1) + bc * mix(0, -1, 0) + gc * mix(0, -1, 1))) /
          2.0 / dy +
      (ac * hB * SY(0, 0, -1) * ((DU(0, 1, -1) - DU(0, -1, -1)) / 2.0 / dy + (DV(1, 0, -1) - DV(-1, 0, -1)) / 2.0 / dx) +
       bc * hC * SY(0, 0, 0) * ((DU(0, 1, 0) - DU(0, -1, 0)) / 2.0 / dy + (DV(1, 0, 0) - DV(-1, 0, 0)) / 2.0 / dx) +
       gc * hT * SY(0, 0, 1) * ((DU(0, 1, 1) - DU(0, -1, 1)) / 2.0 / dy + (DV(1, 0, 1) - DV(-1, 0, 1)) / 2.0 / dx)) +
      (wym * SX(0, 0, -1) * DV(0, 0, -1) + wy0 * SX(0, 0, 0) * DV(0, 0, 0) + wyp * SX(0, 0, 1) * DV(0, 0, 1));
  const double yx =
      (hfe * (DV(1, 0, 0) - DV(0, 0, 0)) / dx - hfw * (DV(0, 0, 0) - DV(-1, 0, 0)) / dx) / dx +
      (hE * ((DU(1, 1, 0) - DU(1, -1, 0)) / 2.0 / dy + ac * mix(1, 0, -1) + bc * mix(1, 0, 0) + gc * mix(1, 0, 1)) -
       hW * ((DU(-1, 1, 0) - DU(-1, -1, 0)) / 2.0 / dy + ac * mix(-1, 0, -1) + bc * mix(-1, 0, 0) + gc * mix(-1, 0, 1))) /
          2.0 / dx +
      (ac * hB * SX(0, 0, -1) * ((DU(0, 1, -1) - DU(0, -1, -1)) / 2.0 / dy + (DV(1, 0, -1) - DV(-1, 0, -1)) / 2.0 / dx) +
       bc * hC * SX(0, 0, 0) * ((DU(0, 1, 0) - DU(0, -1, 0)) / 2.0 / dy + (DV(1, 0, 0) - DV(-1, 0, 0)) / 2.0 / dx) +
       gc * hT * SX(0, 0, 1) * ((DU(0, 1, 1) - DU(0, -1, 1)) / 2.0 / dy + (DV(1, 0, 1) - DV(-1, 0, 1)) / 2.0 / dx)) +
      (wxm * SY(0, 0, -1) * DU(0, 0, -1) + wx0 * SY(0, 0, 0) * DU(0, 0, 0) + wxp * SY(0, 0, 1) * DU(0, 0, 1));
  const double yy =
      (2.0 * hfn * (DV(0, 1, 0) - DV(0, 0, 0)) / dy - 2.0 * hfs * (DV(0, 0, 0) - DV(0, -1, 0)) / dy) / dy +
      (2.0 * hN * (ac * DV(0, 1, -1) * SY(0, 1, -1) + bc * DV(0, 1, 0) * SY(0, 1, 0) + gc * DV(0, 1, 1) * SY(0, 1, 1)) -
       2.0 * hS * (ac * DV(0, -1, -1) * SY(0, -1, -1) + bc * DV(0, -1, 0) * SY(0, -1, 0) + gc * DV(0, -1, 1) * SY(0, -1, 1))) /
          2.0 / dy +
      (ac * 2.0 * hB * SY(0, 0, -1) * (DV(0, 1, -1) - DV(0, -1, -1)) / 2.0 / dy +
       bc * 2.0 * hC * SY(0, 0, 0) * (DV(0, 1, 0) - DV(0, -1, 0)) / 2.0 / dy +
       gc * 2.0 * hT * SY(0, 0, 1) * (DV(0, 1, 1) - DV(0, -1, 1)) / 2.0 / dy);
  const double zx =
      (hfe * (DW(1, 0, 0) - DW(0, 0, 0)) / dx - hfw * (DW(0, 0, 0) - DW(-1, 0, 0)) / dx) / dx +
      (hE * (ac * (DW(1, 0, -1) * SX(1, 0, -1) + DU(1, 0, -1) / Dc) + bc * (DW(1, 0, 0) * SX(1, 0, 0) + DU(1, 0, 0) / Dc) +
             gc * (DW(1, 0, 1) * SX(1, 0, 1) + DU(1, 0, 1) / Dc)) -
       hW * (ac * (DW(-1, 0, -1) * SX(-1, 0, -1) + DU(-1, 0, -1) / Dc) + bc * (DW(-1, 0, 0) * SX(-1, 0, 0) + DU(-1, 0, 0) / Dc) +
             gc * (DW(-1, 0, 1) * SX(-1, 0, 1) + DU(-1, 0, 1) / Dc))) /
          2.0 / dx +
      (ac * hB * SX(0, 0, -1) * (DW(1, 0, -1) - DW(-1, 0, -1)) / 2.0 / dx +
       bc * hC * SX(0, 0, 0) * (DW(1, 0, 0) - DW(-1, 0, 0)) / 2.0 / dx +
       gc * hT * SX(0, 0, 1) * (DW(1, 0, 1) - DW(-1, 0, 1)) / 2.0 / dx) +
      (wxm * DU(0, 0, -1) / Dc + wx0 * DU(0, 0, 0) / Dc + wxp * DU(0, 0, 1) / Dc);
  const double zy =
      (hfn * (DW(0, 1, 0) - DW(0, 0, 0)) / dy - hfs * (DW(0, 0, 0) - DW(0, -1, 0)) / dy) / dy +
      (hN * (ac * (DW(0, 1, -1) * SY(0, 1, -1) + DV(0, 1, -1) / Dc) + bc * (DW(0, 1, 0) * SY(0, 1, 0) + DV(0, 1, 0) / Dc) +
             gc * (DW(0, 1, 1) * SY(0, 1, 1) + DV(0, 1, 1) / Dc)) -
       hS * (ac * (DW(0, -1, -1) * SY(0, -1, -1) + DV(0, -1, -1) / Dc) + bc * (DW(0, -1, 0) * SY(0, -1, 0) + DV(0, -1, 0) / Dc) +
             gc * (DW(0, -1, 1) * SY(0, -1, 1) + DV(0, -1, 1) / Dc))) /
          2.0 / dy +
      (ac * hB * SY(0, 0, -1) * (DW(0, 1, -1) - DW(0, -1, -1)) / 2.0 / dy +
       bc * hC * SY(0, 0, 0) * (DW(0, 1, 0) - DW(0, -1, 0)) / 2.0 / dy +
       gc * hT * SY(0, 0, 1) * (DW(0, 1, 1) - DW(0, -1, 1)) / 2.0 / dy) +
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
template <class R, bool F>
__global__ __launch_bounds__(64) void k_visc_solve(Geo g, Par p, DevT<R> d) {
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
  const double Dc = d.D[c2], S = p.schmidt, cm = p.visc;
  // CmuVt of level k of this column (per cell with the k-epsilon model)
  const Visc<F> CV{d.CmuVt, F ? 0.0 : d.scal[7], g.c3(i, j, 0), 0, sk};
  double bu = 0.0, bw = 0.0, ru = 0.0, rv = 0.0, rw = 0.0, cprev = 0.0;
  for (int k = g.Kbeg; k <= g.Kend; ++k) {
    const int q = k - g.Kbeg;
    const long c = g.c3(i, j, k), fx0 = g.x3(i, j, k), fy0 = g.y3(i, j, k), fz1 = c + sk;
    // viscosity at the sigma levels below and above level k
    const double cv = CV(0, 0, k);
    const double nub = 0.5 * (cm + cm + 0.0 + 0.0) + 0.5 * (CV(0, 0, k - 1) + cv) / S;
    const double nut = 0.5 * (cm + cm + 0.0 + 0.0) + 0.5 * (cv + CV(0, 0, k + 1)) / S;
    const double du = d.DU[c], dv = d.DV[c], dw = d.DW[c];
    const double rdz = 1.0 / dsig[k];
    const double r2 = -1.0 / dx * (d.Fx[fx0 + 1] - d.Fx[fx0]) - 1.0 / dy * (d.Fy[fy0 + g.Mloc] - d.Fy[fy0]) -
                      rdz * (d.Fz[fz1] - d.Fz[c]) + fcor * dv + 0.0 + SX + Df[c] + Df[n3 + c] + 0.0;
    const double r3 = -1.0 / dx * (d.Gx[fx0 + 1] - d.Gx[fx0]) - 1.0 / dy * (d.Gy[fy0 + g.Mloc] - d.Gy[fy0]) -
                      rdz * (d.Gz[fz1] - d.Gz[c]) - fcor * du + 0.0 + SY + Df[2 * n3 + c] + Df[3 * n3 + c] + 0.0;
    const double r4 = -1.0 / dx * (d.Hx[fx0 + 1] - d.Hx[fx0]) - 1.0 / dy * (d.Hy[fy0 + g.Mloc] - d.Hy[fy0]) -
                      rdz * (d.Hz[fz1] - d.Hz[c]) + Df[4 * n3 + c] + Df[5 * n3 + c];
    const double a = k == g.Kbeg ? 0.0 : -dt / (Dc * Dc) * nub / (0.5 * dsig[k] * (dsig[k] + dsig[k - 1]));
    const double cc = k == g.Kend ? 0.0 : -dt / (Dc * Dc) * nut / (0.5 * dsig[k] * (dsig[k] + dsig[k + 1]));
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
      b3 = 1.0 - cc + dt / (Dc * Dc) * nub / y;
      // -DeltH - U DelxH - V DelyH with a fixed bed (DeltH = 0)
      const double Wbot = -0.0 - d.U[c] * d.DelxH[c2] - d.V[c] * d.DelyH[c2];
      xw = dw + dt * r4 + dt / Dc * nub / y * Wbot;
    } else if (k == g.Kend) {
      const double y = 0.25 * dsig[k] * (dsig[k] + dsig[k + 1]);
      b3 = 1.0 - a + dt / (Dc * Dc) * nut / y;
      const double Wtop = (d.Eta[c2] - d.Eta0[c2]) / dt + d.U[c] * d.DelxEta[c2] + d.V[c] * d.DelyEta[c2];
      xw = dw + dt * r4 + dt / Dc * nut / y * Wtop;
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
    const double nut = 0.5 * (cm + cm + 0.0 + 0.0) + 0.5 * (CV(0, 0, k) + CV(0, 0, k + 1)) / S;
    const double cc = -dt / (Dc * Dc) * nut / (0.5 * dsig[k] * (dsig[k] + dsig[k + 1]));
    const double b = bpu[q * 64 + lane], b3 = bpw[q * 64 + lane];
    xu = (d.DU[c] - cc * xu) / b;
    xv = (d.DV[c] - cc * xv) / b;
    xw = (d.DW[c] - cc * xw) / b3;
    d.DU[c] = xu; d.DV[c] = xv; d.DW[c] = xw;
  }
}

// ---- k-epsilon model: kepsilon_3D (src/nhwave.F:10030-11114) as the
// reference's serial build executes it.
//
// kepsilon_3D reads its local variables alpha_c, beta_c and gamma_c (declared
// at :10058-10059, never assigned in it) in the interior-row coefficients of
// both tridiagonal systems (:10468-10541, :10669-10742).  The reference's
// build folds those reads to NaN constants, so with Kglob >= 3 trig returns
// NaN for every solved column and the clamp max(DEps, D*Eps_min), compiled as
// (a > b ? a : b), leaves D*Eps_min (and D*Tke_min) wherever D >= Dmin and
// Mask = 1, NaN where a column is solved but not clamped.  The production
// terms, the HLPA advection and the diffusion coefficients enter only those
// systems and leave no trace.  The state the model produces is therefore
// formed here directly: the solves' outcome, the bottom wall function
// (:10766-10836), the free-surface gradient condition of the ghost levels
// (:10838-10862), the lateral ghosts (:10869-10996), and the new Tke, Eps and
// eddy viscosity (:10999-11030), with CmuHt = CmuVt (eval_turb :6773-6778).

constexpr double kTkeMin = (double)1.e-9f;  // Tke_min, src/initialize.F:3178 (REAL(4))
constexpr double kEpsMin = (double)1.e-9f;  // Eps_min, :3179
__host__ __device__ inline double cmut_min(int rng) {  // :3180-3184
  return (rng ? (double)8.5e-2f : (double)9.0e-2f) * (kTkeMin * kTkeMin) / kEpsMin;
}
__host__ __device__ inline double ke_cmiu(int rng) { return rng ? (double)0.085f : (double)0.09f; }  // :10091-10110

// delzFun_3D of one column at level k (src/nhwave.F:2314-2351; no mask)
__device__ __forceinline__ double ke_delz(const double *q, long sk, const double *sigc, int k) {
  const double t1 = (q[sk] - q[0]) / (sigc[k + 1] - sigc[k]);
  const double t2 = (q[0] - q[-sk]) / (sigc[k] - sigc[k - 1]);
  if ((fabs(t1) + fabs(t2)) < kSmall) return 0.0;
  return (t1 * fabs(t2) + fabs(t1) * t2) / (fabs(t1) + fabs(t2));
}

// the lower-side reconstruction of Q at sigma level Kend1 of column (i,j)
// (construct_3D_z, :1998-2023: zero outside the interior columns), floored
// at qmin as kepsilon_3D does (:10406-10414)
__device__ __forceinline__ double ke_top_state(const Geo &g, const double *Q, const double *dsig, const double *sigc,
                                               int i, int j, double qmin) {
  double v = 0.0;
  if (i >= g.Ibeg && i <= g.Iend && j >= g.Jbeg && j <= g.Jend) {
    const long c = g.c3(i, j, g.Kend), sk = g.n2();
    v = Q[c] + 0.5 * dsig[g.Kend] * ke_delz(Q + c, sk, sigc, g.Kend);
  }
  return qmin > v ? qmin : v;
}

// One thread per interior column: the outcome of both solves, the bottom
// wall function and the surface ghosts; then the column's values copied to
// the lateral ghost cells that mirror it.  Writes DTke/DEps only (Tke/Eps,
// which the surface condition reads in the neighbouring columns, are
// replaced by k_ke_final).
template <class R>
__global__ __launch_bounds__(64) void k_ke_column(Geo g, Par p, DevT<R> d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) d.scal[5] = 0.0;  // largest |CmuHt|, gathered by k_ke_final
  if (t >= (long)g.M * g.N) return;
  const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
  const long c2 = g.c2(i, j), sk = g.n2();
  const double *dsig = d.dsig - 1, *sigc = d.sigc - 1;
  const int Kb = g.Kbeg, Ke = g.Kend;
  const double Dd = d.D[c2], Dmin = 2.0 * p.min_dep;  // :10087
  const int m = d.Mask[c2];
  // level k of this column at index cb + k * sk
  const long cb = g.c3(i, j, 1) - sk;
  double *DT = d.DTke, *DE = d.DEps;
  // the two solves (:10421-10558, :10562-10759)
  if (!(Dd < Dmin && m == 0)) {
    const bool clamp = Dd >= Dmin && m == 1;
    const double qnan = __builtin_nan("");
    const double e = clamp ? Dd * kEpsMin : qnan, kk = clamp ? Dd * kTkeMin : qnan;
    for (int k = Kb; k <= Ke; ++k) {
      DE[cb + k * sk] = e;
      DT[cb + k * sk] = kk;
    }
  }
  if (Dd >= Dmin && m == 1) {
    // bottom wall function, :10763-10836
    const long b0 = g.c3(i, j, Kb);
    const double dxh = d.DelxH[c2], dyh = d.DelyH[c2];
    const double Abot = sqrt(1.0 + dxh * dxh + dyh * dyh);
    const double Bbot = sqrt(1.0 + dxh * dxh);
    const double dup = (d.DU[b0] - dxh * d.DW[b0]) / Bbot;
    const double dvp = (-(dxh * dyh * d.DU[b0]) + (1.0 + dxh * dxh) * d.DV[b0] - dyh * d.DW[b0]) / Abot / Bbot;
    const double dumag = sqrt(dup * dup + dvp * dvp);
    double tkeb = kTkeMin, epsb = kEpsMin;
    if (!(dumag / Dd < (double)1.e-6f)) {
      const double Dz1 = 0.5 * Dd * dsig[Kb] / Abot;
      const double Z0 = p.zob / 30.0;
      double cdrag = p.cd0;
      if (p.ibot != 1) {
        const double tk = kKappa / log(Dz1 / Z0);
        cdrag = tk * tk;
      }
      const double lim = ((p.visc + 0.0) + d.CmuVt[b0]) / Dz1 / (dumag + kSmall) * Dd;
      cdrag = lim < cdrag ? lim : cdrag;
      const double ustar = sqrt(cdrag) * dumag / Dd;
      const double a = ustar * ustar / sqrt(ke_cmiu(p.rng_on)), b = ustar * ustar * ustar / (kKappa * Dz1);
      tkeb = kTkeMin > a ? kTkeMin : a;
      epsb = kEpsMin > b ? kEpsMin : b;
    }
    DT[cb + Kb * sk] = Dd * tkeb;
    DE[cb + Kb * sk] = Dd * epsb;
    for (int k = 1; k <= 2; ++k) {
      DT[cb + (Kb - k) * sk] = DT[cb + Kb * sk];
      DE[cb + (Kb - k) * sk] = DE[cb + Kb * sk];
    }
    // free surface: the slopes at level Kend1 of the floored reconstructions
    // of the old Tke and Eps (delxFun1_3D / delyFun1_3D, :10415-10419)
    const double dx = g.dx, dy = g.dy;
    const double kw = ke_top_state(g, d.Tke, dsig, sigc, i - 1, j, kTkeMin);
    const double k0 = ke_top_state(g, d.Tke, dsig, sigc, i, j, kTkeMin);
    const double ke = ke_top_state(g, d.Tke, dsig, sigc, i + 1, j, kTkeMin);
    const double ks = ke_top_state(g, d.Tke, dsig, sigc, i, j - 1, kTkeMin);
    const double kn = ke_top_state(g, d.Tke, dsig, sigc, i, j + 1, kTkeMin);
    const double ew = ke_top_state(g, d.Eps, dsig, sigc, i - 1, j, kEpsMin);
    const double e0 = ke_top_state(g, d.Eps, dsig, sigc, i, j, kEpsMin);
    const double ee = ke_top_state(g, d.Eps, dsig, sigc, i + 1, j, kEpsMin);
    const double es = ke_top_state(g, d.Eps, dsig, sigc, i, j - 1, kEpsMin);
    const double en = ke_top_state(g, d.Eps, dsig, sigc, i, j + 1, kEpsMin);
    const double dxk = eta_slope(kw, k0, ke, dx, m), dyk = eta_slope(ks, k0, kn, dy, m);
    const double dxe_ = eta_slope(ew, e0, ee, dx, m), dye_ = eta_slope(es, e0, en, dy, m);
    const double dxe = d.DelxEta[c2], dye = d.DelyEta[c2];
    for (int k = 1; k <= 2; ++k) {
      const double Atop = sqrt(1.0 + dxe * dxe + dye * dye);
      const double ds = sigc[Ke + k] - sigc[Ke - k + 1];
      DT[cb + (Ke + k) * sk] = DT[cb + (Ke - k + 1) * sk] + ds * (Dd * Dd) / (Atop * Atop) * (dxe * dxk + dye * dyk);
      DE[cb + (Ke + k) * sk] =
          DE[cb + (Ke - k + 1) * sk] + ds * (Dd * Dd) / (Atop * Atop) * (dxe * dxe_ + dye * dye_);
    }
  }
  // lateral ghosts (:10869-10996: west, east, then south, north over every
  // column): a ghost cell takes the value of the interior cell its x and y
  // mirrors lead to.  The columns this one feeds, by x and by y:
  int xs[3], ys[4], nx = 0, ny = 0;
  xs[nx++] = i;
  if (i == g.Ibeg) xs[nx++] = g.Ibeg - 1;
  if (i == g.Ibeg + 1) xs[nx++] = g.Ibeg - 2;
  if (i == g.Iend) xs[nx++] = g.Iend + 1;
  if (i == g.Iend - 1) xs[nx++] = g.Iend + 2;
  ys[ny++] = j;
  if (j == g.Jbeg) ys[ny++] = g.Jbeg - 1;
  if (j == g.Jend) ys[ny++] = g.Jend + 1;
  if (g.N > 1) {
    if (j == g.Jbeg + 1) ys[ny++] = g.Jbeg - 2;
    if (j == g.Jend - 1) ys[ny++] = g.Jend + 2;
  } else {
    // one row: row Jbeg-2 takes row Jend+1 before the north pass rewrites it
    // (its value from the previous call), and row Jend+2 takes row Jbeg-1
    ys[ny++] = g.Jend + 2;
    for (int a = 0; a < nx; ++a)
      for (int k = 1; k <= g.Kloc; ++k) {
        d.DTke[g.c3(xs[a], g.Jbeg - 2, k)] = d.DTke[g.c3(xs[a], g.Jend + 1, k)];
        d.DEps[g.c3(xs[a], g.Jbeg - 2, k)] = d.DEps[g.c3(xs[a], g.Jend + 1, k)];
      }
  }
  for (int a = 0; a < nx; ++a)
    for (int b = 0; b < ny; ++b) {
      if (a == 0 && b == 0) continue;
      for (int k = 1; k <= g.Kloc; ++k) {
        d.DTke[g.c3(xs[a], ys[b], k)] = DT[cb + k * sk];
        d.DEps[g.c3(xs[a], ys[b], k)] = DE[cb + k * sk];
      }
    }
}

// kepsilon_3D's wall function at a wall friction side boundary (BC 5,
// :10874-10892 and its east, south, north twins): from the speed along the
// wall |u_t| (below 1.e-6, a REAL(4) literal: Tke_min, Eps_min) the friction
// velocity sqrt(Cd0)|u_t| (Ibot = 1) or |u_t|/(1/Kappa log(30 h/2/Zob)),
// then DTke = D u*^2/sqrt(cmiu), DEps = D u*^3/(Kappa h/2) at cell c
template <class R>
__device__ __forceinline__ void ke_wall(const Par &p, const DevT<R> &d, double h, double a, double b, long c, long c2) {
  const double umag = sqrt(a * a + b * b);
  double tkeb = kTkeMin, epsb = kEpsMin;
  if (!(umag < (double)1.e-6f)) {
    const double fricu = p.ibot == 1 ? sqrt(p.cd0) * umag : umag / (1.0 / kKappa * log(30.0 * h / 2.0 / p.zob));
    tkeb = fricu * fricu / sqrt(ke_cmiu(p.rng_on));
    epsb = fricu * fricu * fricu / (kKappa * h / 2.0);
  }
  d.DTke[c] = d.D[c2] * tkeb;
  d.DEps[c] = d.D[c2] * epsb;
}

// The lateral ghost passes again where a side is a wall friction boundary,
// after k_ke_column, in the reference's order (:10869-10996).  West then
// east, one thread per interior row j and level k (every level, ghosts
// included): the wall function at the first interior cell, then the two
// ghost cells mirror it.
template <class R>
__global__ __launch_bounds__(64) void k_ke_wall_x(Geo g, Par p, DevT<R> d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.N * g.Kloc) return;
  const int j = g.Jbeg + (int)(t % g.N), k = 1 + (int)(t / g.N);
  auto copy = [&](int to, int from) {
    d.DTke[g.c3(to, j, k)] = d.DTke[g.c3(from, j, k)];
    d.DEps[g.c3(to, j, k)] = d.DEps[g.c3(from, j, k)];
  };
  if (p.bc_x0 == 5) {
    const long b = g.c3(g.Ibeg, j, k);
    ke_wall(p, d, g.dx, d.V[b], d.W[b], b, g.c2(g.Ibeg, j));
  }
  for (int i = 1; i <= 2; ++i) copy(g.Ibeg - i, g.Ibeg + i - 1);
  if (p.bc_xn == 5) {
    const long b = g.c3(g.Iend, j, k);
    ke_wall(p, d, g.dx, d.V[b], d.W[b], b, g.c2(g.Iend, j));
  }
  for (int i = 1; i <= 2; ++i) copy(g.Iend + i, g.Iend - i + 1);
}

// South then north, one thread per column i (every column, ghosts included)
// and level k
template <class R>
__global__ __launch_bounds__(64) void k_ke_wall_y(Geo g, Par p, DevT<R> d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.Mloc * g.Kloc) return;
  const int i = 1 + (int)(t % g.Mloc), k = 1 + (int)(t / g.Mloc);
  auto copy = [&](int to, int from) {
    d.DTke[g.c3(i, to, k)] = d.DTke[g.c3(i, from, k)];
    d.DEps[g.c3(i, to, k)] = d.DEps[g.c3(i, from, k)];
  };
  if (p.bc_y0 == 5) {
    const long b = g.c3(i, g.Jbeg, k);
    ke_wall(p, d, g.dy, d.U[b], d.W[b], b, g.c2(i, g.Jbeg));
  }
  for (int j = 1; j <= 2; ++j) copy(g.Jbeg - j, g.Jbeg + j - 1);
  if (p.bc_yn == 5) {
    const long b = g.c3(i, g.Jend, k);
    ke_wall(p, d, g.dy, d.U[b], d.W[b], b, g.c2(i, g.Jend));
  }
  for (int j = 1; j <= 2; ++j) copy(g.Jend + j, g.Jend - j + 1);
}

// One thread per column, ghosts included: Tke, Eps and the eddy viscosity
// (:10999-11030), the ghost levels of CmuVt (:11031-11038); the largest
// |CmuVt| of the interior cells for estimate_dt's viscous limit.
__global__ __launch_bounds__(256) void k_ke_final(Geo g, Par p, Dev d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double vmax = 0.0;
  if (t < g.n2()) {
    const int i = 1 + (int)(t % g.Mloc), j = 1 + (int)(t / g.Mloc);
    const long sk = g.n2();
    const double Dd = d.D[t], Dmin = 2.0 * p.min_dep, cmiu = ke_cmiu(p.rng_on), cmut = cmut_min(p.rng_on);
    const bool wet = Dd >= Dmin && d.Mask[t] == 1;
    for (int k = 1; k <= g.Kloc; ++k) {
      const long c = t + (long)(k - 1) * sk;
      double cv = cmut;
      if (wet) {
        const double tk = d.DTke[c] / Dd, ep = d.DEps[c] / Dd;
        d.Tke[c] = tk;
        d.Eps[c] = ep;
        if (!(tk < (double)1.e-6f || ep < (double)1.e-6f)) cv = cmiu * (tk * tk) / ep;
      } else {
        d.Tke[c] = kTkeMin;
        d.Eps[c] = kEpsMin;
        d.DTke[c] = Dd * kTkeMin;
        d.DEps[c] = Dd * kEpsMin;
      }
      d.CmuVt[c] = cv;
    }
    const double cb = d.CmuVt[t + (long)(g.Kbeg - 1) * sk], ct = d.CmuVt[t + (long)(g.Kend - 1) * sk];
    for (int k = 1; k <= 2; ++k) {
      d.CmuVt[t + (long)(g.Kbeg - k - 1) * sk] = cb;
      d.CmuVt[t + (long)(g.Kend + k - 1) * sk] = ct;
    }
    // no turbulence in the internal wavemaker's source box (:11055-11071;
    // the west wavemakers' box at the origin holds no cell centre)
    const Wave &w = p.wave;
    if (i >= w.bi0 && i <= w.bi1 && j >= w.bj0 && j <= w.bj1)
      for (int k = 1; k <= g.Kloc; ++k) {
        const long c = t + (long)(k - 1) * sk;
        d.Tke[c] = kTkeMin;
        d.Eps[c] = kEpsMin;
        d.DTke[c] = Dd * kTkeMin;
        d.DEps[c] = Dd * kEpsMin;
        d.CmuVt[c] = cmut;
      }
    if (i >= g.Ibeg && i <= g.Iend && j >= g.Jbeg && j <= g.Jend)
      for (int k = g.Kbeg; k <= g.Kend; ++k) vmax = fmax(vmax, fabs(d.CmuVt[t + (long)(k - 1) * sk]));
  }
  // block maximum, then one atomic per block on the bits of a non-negative
  // double (their integer order is the numeric order)
  __shared__ double sm[256];
  sm[threadIdx.x] = vmax;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    atomicMax(reinterpret_cast<unsigned long long *>(d.scal + 5), (unsigned long long)__double_as_longlong(sm[0]));
}

// the seed of initial (src/initialize.F:3185-3196): Tke, Eps and the eddy
// viscosity everywhere (depth = 0: before D and Mask exist), or DTke = D Tke
// Mask, DEps = D Eps Mask (depth = 1); depth = 2: the eddy viscosity as
// allocated (Zero, :2667-2668), which initial's vel_bc reads before the seed
__global__ __launch_bounds__(256) void k_ke_seed(Geo g, Par p, Dev d, int depth) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.n3()) return;
  const long c2 = t % g.n2();
  if (depth == 2) {
    d.CmuVt[t] = 0.0;
  } else if (depth) {
    d.DTke[t] = d.D[c2] * d.Tke[t] * d.Mask[c2];
    d.DEps[t] = d.D[c2] * d.Eps[t] * d.Mask[c2];
  } else {
    d.Tke[t] = kTkeMin;
    d.Eps[t] = kEpsMin;
    d.CmuVt[t] = cmut_min(p.rng_on);
    if (t == 0) d.scal[5] = cmut_min(p.rng_on);
  }
}

int viscous_setup(nhw_ctx *c) {
  if (!c->cfg.viscous) return 0;
  const Geo &g = c->g;
  Dev &d = c->d;
  const long n3 = g.n3(), nz = g.nz3();
  const bool ke = c->p.turb == 10;
  struct A { double **p; long n; } al[] = {{&d.DelxSc, n3}, {&d.DelySc, n3}, {&d.DelxSl, nz}, {&d.DelySl, nz},
                                           {&d.Diff, 6 * n3}, {&d.CmuVt, ke ? n3 : 0}, {&d.Tke, ke ? n3 : 0},
                                           {&d.Eps, ke ? n3 : 0}, {&d.DTke, ke ? n3 : 0}, {&d.DEps, ke ? n3 : 0}};
  for (auto &x : al) {
    if (!x.n) continue;
    if (hipMalloc(x.p, sizeof(double) * x.n) != hipSuccess ||
        hipMemset(*x.p, 0, sizeof(double) * x.n) != hipSuccess) {
      set_error("nhw_create: device allocation failed (viscous arrays)");
      return NHW_ENOMEM;
    }
  }
  d.CmuHt = d.CmuVt;  // IHturb >= 10: CmuHt = CmuVt (src/nhwave.F:6773-6778)
  return 0;
}

void viscous_free(nhw_ctx *c) {
  Dev &d = c->d;
  for (double **p : {&d.DelxSc, &d.DelySc, &d.DelxSl, &d.DelySl, &d.Diff, &d.CmuVt, &d.Tke, &d.Eps, &d.DTke, &d.DEps}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  d.CmuHt = nullptr;
}

int launch_diffusion(nhw_ctx *c) {
  if (!c->cfg.viscous) return 0;
  const Geo &g = c->g;
  const long n = (long)g.M * g.N * g.K;
  const dim3 grid((unsigned)((n + 255) / 256));
  Timed tm(c, KID_DIFFUSION);
  if (c->p.turb == 10)
    if (c->mixed)
      hipLaunchKernelGGL((k_diffusion<float, true>), grid, dim3(256), 0, c->stream, g, c->p, view<float>(c->d));
    else
      hipLaunchKernelGGL((k_diffusion<double, true>), grid, dim3(256), 0, c->stream, g, c->p, c->d);
  else
    if (c->mixed)
      hipLaunchKernelGGL((k_diffusion<float, false>), grid, dim3(256), 0, c->stream, g, c->p, view<float>(c->d));
    else
      hipLaunchKernelGGL((k_diffusion<double, false>), grid, dim3(256), 0, c->stream, g, c->p, c->d);
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
  const dim3 grid((unsigned)((n + 63) / 64));
  Timed tm(c, KID_VISC_SOLVE);
  if (c->p.turb == 10)
    if (c->mixed)
      hipLaunchKernelGGL((k_visc_solve<float, true>), grid, dim3(64), lds, c->stream, g, c->p, view<float>(c->d));
    else
      hipLaunchKernelGGL((k_visc_solve<double, true>), grid, dim3(64), lds, c->stream, g, c->p, c->d);
  else
    if (c->mixed)
      hipLaunchKernelGGL((k_visc_solve<float, false>), grid, dim3(64), lds, c->stream, g, c->p, view<float>(c->d));
    else
      hipLaunchKernelGGL((k_visc_solve<double, false>), grid, dim3(64), lds, c->stream, g, c->p, c->d);
  return hip_check(hipGetLastError(), "k_visc_solve");
}

// The viscous branch's state that outlives a step, for checkpoints: the
// largest |CmuHt| the k-epsilon model left for the next estimate_dt and the
// uniform CmuHt, CmuVt of constant viscosities (the seed until the first
// eval_turb), the device scalars scal[5..7], and sigma_transform's slopes,
// which the next stage's diffusion reads.  Layout: the three FP64 scalars,
// then DelxSc, DelySc (Mloc,Nloc,Kloc) and DelxSl, DelySl (Mloc,Nloc,Kloc1).
// (The k-epsilon arrays are fields of nhw_fields and travel with them.)
size_t viscous_state_bytes(const nhw_ctx *c) {
  const Geo &g = c->g;
  return sizeof(double) * (3 + 2 * (size_t)g.n3() + 2 * (size_t)g.nz3());
}

int viscous_state(nhw_ctx *c, std::vector<char> &buf, bool save) {
  const Geo &g = c->g;
  const size_t b3 = sizeof(double) * (size_t)g.n3(), bz = sizeof(double) * (size_t)g.nz3();
  if (save) buf.assign(viscous_state_bytes(c), 0);
  else if (buf.size() != viscous_state_bytes(c)) {
    set_error("viscous checkpoint state: size differs from the context's");
    return NHW_EINVAL;
  }
  double *sc = (double *)buf.data();
  char *a = buf.data() + 3 * sizeof(double);
  struct Arr { double *dev; size_t bytes; } arrs[] = {{c->d.DelxSc, b3}, {c->d.DelySc, b3}, {c->d.DelxSl, bz},
                                                       {c->d.DelySl, bz}};
  int rc = save ? hip_check(hipMemcpyAsync(sc, c->d.scal + 5, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                            "viscous checkpoint state")
               : hip_check(hipMemcpyAsync(c->d.scal + 5, sc, 3 * sizeof(double), hipMemcpyHostToDevice, c->stream),
                            "viscous checkpoint state");
  for (const Arr &x : arrs) {
    if (rc) break;
    rc = save ? hip_check(hipMemcpyAsync(a, x.dev, x.bytes, hipMemcpyDeviceToHost, c->stream), "viscous checkpoint state")
              : hip_check(hipMemcpyAsync(x.dev, a, x.bytes, hipMemcpyHostToDevice, c->stream), "viscous checkpoint state");
    a += x.bytes;
  }
  if (!rc) rc = hip_check(hipStreamSynchronize(c->stream), "viscous checkpoint state");
  return rc ? NHW_EHIP : 0;
}

// the uniform CmuHt, CmuVt of constant viscosities (device scalars, so that
// captured launches see the current values)
__global__ void k_set_cmu(Dev d, double ht, double vt) {
  if (threadIdx.x == 0) {
    d.scal[6] = ht;
    d.scal[7] = vt;
  }
}

// eval_turb (src/nhwave.F:6544-6782).  IVturb = IHturb = 1: from the first
// call on, CmuVt = Cvs and CmuHt = Chs, held in device scalars
// (:6650-6685).  k-epsilon: kepsilon_3D as described above, then CmuHt =
// CmuVt (one array).
int eval_turb(nhw_ctx *c) {
  if (!c->cfg.viscous) return 0;
  if (c->p.turb != 10) {
    hipLaunchKernelGGL(k_set_cmu, dim3(1), dim3(64), 0, c->stream, c->d, c->cfg.chs, c->cfg.cvs);
    return hip_check(hipGetLastError(), "k_set_cmu");
  }
  const Geo &g = c->g;
  {
    Timed tm(c, KID_KEPS, 8.0 * (8.0 * g.M * g.N * g.Kloc + 12.0 * g.M * g.N));
    const dim3 cg((unsigned)(((long)g.M * g.N + 63) / 64));
    if (c->mixed)
      hipLaunchKernelGGL(k_ke_column<float>, cg, dim3(64), 0, c->stream, g, c->p, view<float>(c->d));
    else
      hipLaunchKernelGGL(k_ke_column<double>, cg, dim3(64), 0, c->stream, g, c->p, c->d);
    if (hip_check(hipGetLastError(), "k_ke_column")) return NHW_EHIP;
    const Par &p = c->p;
    if (p.bc_x0 == 5 || p.bc_xn == 5 || p.bc_y0 == 5 || p.bc_yn == 5) {
      const dim3 gx((unsigned)(((long)g.N * g.Kloc + 63) / 64)), gy((unsigned)(((long)g.Mloc * g.Kloc + 63) / 64));
      if (c->mixed) {
        hipLaunchKernelGGL(k_ke_wall_x<float>, gx, dim3(64), 0, c->stream, g, p, view<float>(c->d));
        hipLaunchKernelGGL(k_ke_wall_y<float>, gy, dim3(64), 0, c->stream, g, p, view<float>(c->d));
      } else {
        hipLaunchKernelGGL(k_ke_wall_x<double>, gx, dim3(64), 0, c->stream, g, p, c->d);
        hipLaunchKernelGGL(k_ke_wall_y<double>, gy, dim3(64), 0, c->stream, g, p, c->d);
      }
      if (hip_check(hipGetLastError(), "k_ke_wall")) return NHW_EHIP;
    }
  }
  Timed tm(c, KID_KEPS, 8.0 * 7.0 * g.n3() + 12.0 * g.n2());
  hipLaunchKernelGGL(k_ke_final, dim3((unsigned)((g.n2() + 255) / 256)), dim3(256), 0, c->stream, g, c->p, c->d);
  return hip_check(hipGetLastError(), "k_ke_final");
}

// the turbulence seed, src/initialize.F:3172-3196 (REAL(4) literals), before
// initial's vel_bc
int turb_seed(nhw_ctx *c) {
  if (!c->cfg.viscous) return 0;
  const double cmut = cmut_min(c->cfg.rng);
  hipLaunchKernelGGL(k_set_cmu, dim3(1), dim3(64), 0, c->stream, c->d, cmut, cmut);
  if (hip_check(hipGetLastError(), "k_set_cmu")) return NHW_EHIP;
  if (c->p.turb != 10) return 0;
  const Geo &g = c->g;
  hipLaunchKernelGGL(k_ke_seed, dim3((unsigned)((g.n3() + 255) / 256)), dim3(256), 0, c->stream, g, c->p, c->d, 0);
  return hip_check(hipGetLastError(), "k_ke_seed");
}

// initial's vel_bc (src/initialize.F:3155) runs before the seed, on CmuHt =
// CmuVt = Zero (:2667-2668)
int turb_zero(nhw_ctx *c) {
  if (!c->cfg.viscous) return 0;
  hipLaunchKernelGGL(k_set_cmu, dim3(1), dim3(64), 0, c->stream, c->d, 0.0, 0.0);
  if (hip_check(hipGetLastError(), "k_set_cmu")) return NHW_EHIP;
  if (c->p.turb != 10) return 0;
  const Geo &g = c->g;
  hipLaunchKernelGGL(k_ke_seed, dim3((unsigned)((g.n3() + 255) / 256)), dim3(256), 0, c->stream, g, c->p, c->d, 2);
  return hip_check(hipGetLastError(), "k_ke_seed");
}

// DTke = D Tke Mask, DEps = D Eps Mask once initial has formed D and Mask
int turb_seed_depth(nhw_ctx *c) {
  if (!c->cfg.viscous || c->p.turb != 10) return 0;
  const Geo &g = c->g;
  hipLaunchKernelGGL(k_ke_seed, dim3((unsigned)((g.n3() + 255) / 256)), dim3(256), 0, c->stream, g, c->p, c->d, 1);
  return hip_check(hipGetLastError(), "k_ke_seed");
}

}  // namespace nhw
