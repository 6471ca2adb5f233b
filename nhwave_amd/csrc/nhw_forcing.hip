// Boundary forcing on gfx950: the west wavemakers of update_wave_bc
// (solitary, linear, second-order Stokes and cnoidal waves; src/nhwave.F:
// 5514-5562, src/waves.F:92-361, 518-903), the internal irregular
// wavemaker's mass source (source_terms, src/source_terms.F:194-305) and the
// sponge layer (calculate_sponge and sponge_damping, src/nhwave.F:6405-6502).
//
// Every constant of the wavemakers (the dispersion-relation wave number,
// sinh(k h), cos/sin of THETA, the solitary wave's celerity and decay rate,
// the per-row phase lag) and the whole Sponge table are computed once on the
// host, with the host libm and the reference's expression order, so they are
// the reference's values.  The device evaluates only what moves with TIME:
// cos, sin, cosh and sinh of time-dependent arguments, from the device libm,
// which can differ from the host's in the last bit.
//
// x**n with a constant integer n is evaluated as the reference build
// evaluates it: n-1 multiplications from the left, ((x*x)*x)*...
#include <cmath>
#include <vector>

#include "nhw_batch.h"

namespace nhw {

constexpr double kPi = (double)3.141592653f;  // pi, src/mod_global.F:21

// wave_dispersion, src/waves.F:61-80: Newton iteration on the linear
// dispersion relation from Eckart's guess, to a relative step below 1e-8
static double wave_dispersion(double segma, double dep) {
  double wk = (segma * segma) / (kGrav * std::sqrt(std::tanh(segma * segma * dep / kGrav)));
  double err = 100.0;
  while (err >= (double)1.0e-8f) {
    const double th = std::tanh(wk * dep), ch = std::cosh(wk * dep);
    const double fk = segma * segma - kGrav * wk * th;
    const double fkdif = -kGrav * (th + (wk * dep) / (ch * ch));
    const double wn = wk - fk / fkdif;
    err = std::fabs((wk - wn) / wk);
    wk = wn;
  }
  return wk;
}

// The cnoidal wave's elliptic functions (src/waves.F:184-361).  agm: the
// arithmetic-geometric mean from a = 1, b = sqrt(1-m), c = sqrt(m) until
// |c| < 1.0e-15 (a REAL(4) literal) or n > 1000, keeping a(i), c(i)
// (1-based); K(m) = pi/2/a(n) with the REAL(4) pi literal; E(m) = K (1 -
// sum 2**(i-2) c(i)**2); cn(u|m) by the descending Landen recurrence.
struct Agm {
  std::vector<double> a, c;
  int n;
};
static Agm agm(double m) {
  Agm r{std::vector<double>(2, 1.0), std::vector<double>(2, 0.0), 1};
  double a0 = 1.0, b0 = std::sqrt(1.0 - m), c0 = std::sqrt(m);
  r.c[1] = c0;
  while (!(std::fabs(c0) < (double)1.0e-15f || r.n > 1000)) {
    ++r.n;
    const double a1 = (a0 + b0) / 2.0, b1 = std::sqrt(a0 * b0), c1 = (a0 - b0) / 2.0;
    a0 = a1;
    b0 = b1;
    c0 = c1;
    r.a.push_back(a0);
    r.c.push_back(c0);
  }
  return r;
}
constexpr double kPiAgm = (double)3.1415926535897932384626f;
static double ell_k(double m) { return kPiAgm / 2.0 / agm(m).a.back(); }
static double ell_e(double m) {
  const Agm r = agm(m);
  const double ck = kPiAgm / 2.0 / r.a[r.n];
  double sum = 0.0;
  for (int i = 1; i <= r.n; ++i) sum = sum + std::ldexp(1.0, i - 2) * (r.c[i] * r.c[i]);
  return ck * (1.0 - sum);
}
static double ell_cn(double u, double m) {
  const Agm r = agm(m);
  double y = std::ldexp(1.0, r.n - 1) * r.a[r.n] * u;
  for (int i = r.n - 1; i >= 1; --i) y = (y + std::asin(r.c[i + 1] / r.a[i + 1] * std::sin(y))) / 2.0;
  return std::cos(y);
}

// cnoidal, src/waves.F:184-220: Newton iteration on the modulus from
// 0.99999999 until |xa| <= 1.0e-8 (REAL(4)) or 1000 steps; then the wave
// length and trough level.  cnoidal_wave_left_boundary (:101-130) then
// bisects [0, 1] for the surface's zero crossing (|Eta| <= 1.0e-6 within
// 1000 steps, else the reference stops).
static bool cnoidal_setup(double H, double d, double T, Wave &w) {
  double m = 0.99999999;
  for (int n = 1;; ++n) {
    const double ck = ell_k(m), ce = ell_e(m);
    const double d3 = d * d * d, t2 = T * T, ck2 = ck * ck;
    const double xa = m * d + 2.0 * H - m * H - 3.0 * H * ce / ck - 16.0 * d3 * (m * m) * ck2 / 3.0 / kGrav / H / t2;
    if (std::fabs(xa) <= (double)1.0e-8f || n > 1000) break;
    const double xb = d - H + 3.0 * H / 2.0 / m / (1.0 - m) / ck2 * ((1.0 - m) * ck2 + ce * ce - 2.0 * (1.0 - m) * ck * ce) -
                      16.0 * d3 * m * ck / 3.0 / kGrav / (1.0 - m) / H / t2 * ((1.0 - m) * ck + ce);
    m = m - xa / xb;
  }
  const double ck = ell_k(m), ce = ell_e(m);
  w.cm = m;
  w.cck = ck;
  w.cwl = 4.0 * ck * d * std::sqrt(m * d / H / 3.0);
  w.cyt = H / m * (1.0 - m - ce / ck);
  double zup = 1.0, zlow = 0.0, zmid = (zup + zlow) / 2.0;
  for (int nn = 1;; ++nn) {
    const double cn = ell_cn(zmid * 0.5 * ck, m);
    const double z1 = w.cyt + H * (cn * cn);
    if (std::fabs(z1) <= (double)1.0e-6f) break;
    if (nn > 1000) return false;
    if (z1 < 0.0)
      zup = zmid;
    else
      zlow = zmid;
    zmid = (zup + zlow) / 2.0;
  }
  w.cu0 = zmid * 0.5 * ck;
  w.cu1 = 2.0 * ck;
  w.cper = T;
  return true;
}

// cell centres of generate_grid (serial), src/initialize.F:1666-1698: face
// coordinates by repeated addition from 0 at the first interior face
static std::vector<double> centres(int n, int beg, double h) {
  const int nloc = n + 4, nloc1 = nloc + 1;
  std::vector<double> x(nloc1 + 1, 0.0), xc(nloc + 1, 0.0);  // 1-based
  x[beg] = 0.0;
  for (int i = beg + 1; i <= nloc1; ++i) {
    x[i] = x[i - 1] + h;
    xc[i - 1] = x[i - 1] + 0.5 * h;
  }
  for (int i = beg - 1; i >= 1; --i) {
    x[i] = x[i + 1] - h;
    xc[i] = x[i + 1] - 0.5 * h;
  }
  return xc;
}

// calculate_sponge, src/nhwave.F:6405-6459 (Sponge = 1 elsewhere,
// src/initialize.F:2665): a later side overrides an earlier one
static std::vector<double> sponge_table(const nhw_config &cf, const Geo &g) {
  std::vector<double> sp((size_t)g.n2(), 1.0);
  const std::vector<double> xc = centres(g.M, g.Ibeg, g.dx), yc = centres(g.N, g.Jbeg, g.dy);
  const double W = cf.sponge_w, E = cf.sponge_e, S = cf.sponge_s, N = cf.sponge_n;
  const double xe = cf.mglob * cf.dx, yn = cf.nglob * cf.dy;
  auto side = [&](double width, bool xdir, bool low, double edge) {
    if (!(width > 0.0)) return;
    for (int j = 1; j <= g.Nloc; ++j)
      for (int i = 1; i <= g.Mloc; ++i) {
        const double c = xdir ? xc[i] : yc[j];
        const double start = low ? width : edge - width;
        if (low ? !(c <= start) : !(c >= start)) continue;
        const double r = (c - (low ? width : start)) / width;
        sp[(size_t)g.c2(i, j)] = std::sqrt(1.0 - std::min(r * r, 1.0));
      }
  };
  side(W, true, true, 0.0);
  side(E, true, false, xe);
  side(S, false, true, 0.0);
  side(N, false, false, yn);
  return sp;
}

int forcing_setup(nhw_ctx *c) {
  const nhw_config &cf = c->cfg;
  const Geo &g = c->g;
  Wave &w = c->p.wave;
  w = Wave{};
  w.kind = cf.wavemaker;
  w.sponge_on = cf.sponge_on != 0;
  w.amp = cf.amp_wave;
  w.dep = cf.dep_wave;
  // CUR is read with the wavemaker keys only (src/initialize.F:607-613)
  w.cur = cf.wavemaker != NHW_WAVE_NONE ? cf.cur_wave : 0.0;
  w.tramp = cf.tramp;
  const double A = cf.amp_wave, Dp = cf.dep_wave;
  std::vector<double> lag;
  if (w.kind == NHW_WAVE_LEF_SOL) {  // src/waves.F:877-883, 888
    w.cel = std::sqrt(kGrav * Dp * (1.0 + A / Dp));
    w.at = std::sqrt(0.75 * A / (Dp * Dp * Dp));
    w.xs = 4.0 * Dp / std::sqrt(A / Dp);
    w.c2 = std::sqrt(kGrav * Dp);
  } else if (w.kind == NHW_WAVE_LEF_LIN || w.kind == NHW_WAVE_LEF_STK) {
    w.segma = 2.0 * kPi / cf.per_wave;
    w.wk = wave_dispersion(w.segma, Dp);
    w.shd = std::sinh(w.wk * Dp);
    if (w.kind == NHW_WAVE_LEF_LIN) {  // src/waves.F:822-835
      const double wl = 2.0 * kPi / w.wk;
      w.drift = kGrav * (A * A) / (wl / cf.per_wave) / Dp / 8.0;
      const double th = cf.theta_wave * kPi / 180.0;
      w.cth = std::cos(th);
      w.sth = std::sin(th);
      const std::vector<double> yc = centres(g.N, g.Jbeg, g.dy);
      lag.resize(g.Nloc);
      for (int j = 1; j <= g.Nloc; ++j) lag[j - 1] = yc[j] * std::sin(th) * w.wk;
    } else {  // src/waves.F:536-540
      const double sh3 = w.shd * w.shd * w.shd;
      w.sh4 = sh3 * w.shd;
      w.eta2 = (A * A) * w.wk / 16.0 * std::cosh(w.wk * Dp) / sh3 * (2.0 + std::cosh(2.0 * w.wk * Dp));
    }
  } else if (w.kind == NHW_WAVE_LEF_CON) {  // src/waves.F:99-130, 139-140
    if (!cnoidal_setup(A, Dp, cf.per_wave, w)) {
      set_error("nhw_create: cnoidal wave: too many iterations for its zero crossing");
      return NHW_EINVAL;
    }
    w.drift = kGrav * (A * A) / (w.cwl / cf.per_wave) / Dp / 8.0;
    w.c2 = std::sqrt(kGrav * Dp);
    const Agm r = agm(w.cm);
    w.cn_n = r.n;
    w.cysc = std::ldexp(1.0, r.n - 1) * r.a[r.n];
    lag.assign(r.n, 0.0);
    for (int i = 2; i <= r.n; ++i) lag[i - 1] = r.c[i] / r.a[i];
  }
  Dev &d = c->d;
  auto alloc = [&](double **p, long n, const double *host) -> int {
    if (hip_check(hipMalloc(p, sizeof(double) * n), "forcing: hipMalloc")) return NHW_ENOMEM;
    if (host) return hip_check(hipMemcpy(*p, host, sizeof(double) * n, hipMemcpyHostToDevice), "forcing: upload");
    return hip_check(hipMemset(*p, 0, sizeof(double) * n), "forcing: clear");
  };
  int rc = 0;
  // the source box of the internal wavemaker, empty until
  // nhw_set_internal_wavemaker (the west wavemakers leave the box keys at 0:
  // a box at the origin, which no cell centre reaches)
  w.si0 = w.bi0 = 1;
  w.si1 = w.bi1 = 0;
  if (w.kind == NHW_WAVE_INT_IRR && (rc = alloc(&d.SourceC, g.n2(), nullptr))) return rc;
  const long nb = (long)g.Nloc * g.Kloc;
  if (west_wavemaker(w.kind) || w.kind == NHW_WAVE_FLUX_LR) {
    if ((rc = alloc(&d.Ein0, g.Nloc, nullptr)) || (rc = alloc(&d.Din0, g.Nloc, nullptr)) ||
        (rc = alloc(&d.Uin0, nb, nullptr)) || (rc = alloc(&d.Vin0, nb, nullptr)) ||
        (rc = alloc(&d.Win0, nb, nullptr)))
      return rc;
    if (!lag.empty() && (rc = alloc(&d.wtab, (long)lag.size(), lag.data()))) return rc;
  }
  // FLUX_LR's east boundary values and the two previous steps' Eta, U, V, W
  // (zero before the first update_vars, as the reference's allocations)
  if (w.kind == NHW_WAVE_FLUX_LR &&
      ((rc = alloc(&d.Ein_Xn, g.Nloc, nullptr)) || (rc = alloc(&d.Din_Xn, g.Nloc, nullptr)) ||
       (rc = alloc(&d.Uin_Xn, nb, nullptr)) || (rc = alloc(&d.Vin_Xn, nb, nullptr)) ||
       (rc = alloc(&d.Win_Xn, nb, nullptr)) || (rc = alloc(&d.Uin_Xni, nb, nullptr)) ||
       (rc = alloc(&d.Uin_Xni0, nb, nullptr)) || (rc = alloc(&d.Eta00, g.n2(), nullptr)) ||
       (rc = alloc(&d.U0, g.n3(), nullptr)) || (rc = alloc(&d.V0, g.n3(), nullptr)) ||
       (rc = alloc(&d.W0, g.n3(), nullptr)) || (rc = alloc(&d.U00, g.n3(), nullptr)) ||
       (rc = alloc(&d.V00, g.n3(), nullptr)) || (rc = alloc(&d.W00, g.n3(), nullptr))))
    return rc;
  if (w.sponge_on) {
    const std::vector<double> sp = sponge_table(cf, g);
    if ((rc = alloc(&d.Sponge, g.n2(), sp.data()))) return rc;
  }
  return 0;
}

void forcing_free(nhw_ctx *c) {
  Dev &d = c->d;
  for (double **p : {&d.Ein0, &d.Din0, &d.Uin0, &d.Vin0, &d.Win0, &d.wtab, &d.Sponge, &d.SourceC, &d.Ein_Xn,
                     &d.Din_Xn, &d.Uin_Xn, &d.Vin_Xn, &d.Win_Xn, &d.Uin_Xni, &d.Uin_Xni0, &d.Eta00, &d.U0, &d.V0,
                     &d.W0, &d.U00, &d.V00, &d.W00}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

// The internal irregular wavemaker's constants (src/source_terms.F:194-305),
// computed on the host as the reference computes them at every call: per
// component (frequency outer, direction inner) Segma, the wave number of
// the dispersion relation, the celerity, the direction moved to the nearest
// one a periodic y boundary admits (:230-275, as written there: for
// negative angles the smaller candidate keeps its positive sign) and its
// y wave number, the amplitude 2*Wave_Spc2d.  Table: five doubles a
// component -- Segma, the coefficient (2*Celerity for the first frequency,
// Celerity*Amp after), Amp/2, WnumY, Random_Phs -- then dy/2 + yc(j) for
// j = 1..Nloc.
int internal_wavemaker_setup(nhw_ctx *c, const double box[4], int nfreq, int ndir, const double *freq,
                             const double *dire, const double *spc, const double *phs) {
  const Geo &g = c->g;
  Wave &w = c->p.wave;
  if (w.kind != NHW_WAVE_INT_IRR) {
    set_error("nhw_set_internal_wavemaker: the context's WAVEMAKER is not INT_IRR");
    return NHW_EINVAL;
  }
  if (nfreq < 1 || ndir < 1 || !freq || !dire || !spc || !phs) {
    set_error("nhw_set_internal_wavemaker: need NumFreq, NumDir >= 1 and the four arrays");
    return NHW_EINVAL;
  }
  for (int f = 0; f < nfreq; ++f)
    if (!(freq[f] > 0)) {
      set_error("nhw_set_internal_wavemaker: frequencies must be positive");
      return NHW_EINVAL;
    }
  const double dy = g.dy, Dp = c->cfg.dep_wave;
  const int ng = c->cfg.nglob;
  const long n = (long)nfreq * ndir;
  std::vector<double> tab((size_t)(5 * n + g.Nloc));
  for (int f = 1; f <= nfreq; ++f)
    for (int q = 1; q <= ndir; ++q) {
      const double per = 1.0 / freq[f - 1];
      const double segma = 2.0 * kPi / per;
      const double wk = wave_dispersion(segma, Dp);
      const double wl = 2.0 * kPi / wk;
      const double cel = wl / per;
      double angle = dire[q - 1] * kPi / 180.0;
      if (angle != 0.0) {
        const bool pos = angle > 0.0;
        double t3 = 0.0;
        int nk = 0;
        while (pos ? t3 < angle : t3 > angle) {
          ++nk;
          const double t2 = nk * 2.0 * kPi / (ng * dy);
          t3 = t2 >= wk ? (pos ? 0.5 * kPi - kSmall : -0.5 * kPi + kSmall) : (pos ? std::asin(t2 / wk)
                                                                                : -std::asin(t2 / wk));
        }
        const double t4 = std::asin((nk - 1) * 2.0 * kPi / (ng * dy) / wk);
        angle = std::fabs(t4 - angle) < std::fabs(angle - t3) ? t4 : t3;
      } else {
        angle = 0.0;
      }
      const double amp = 2.0 * spc[(long)(f - 1) * ndir + (q - 1)];
      double *e = &tab[(size_t)(5 * ((long)(f - 1) * ndir + (q - 1)))];
      e[0] = segma;
      e[1] = f == 1 ? 2.0 * cel : cel * amp;
      e[2] = amp / 2.0;
      e[3] = wk * std::sin(angle);
      e[4] = phs[(long)(f - 1) * ndir + (q - 1)];
    }
  const std::vector<double> xc = centres(g.M, g.Ibeg, g.dx), yc = centres(g.N, g.Jbeg, g.dy);
  for (int j = 1; j <= g.Nloc; ++j) tab[(size_t)(5 * n + j - 1)] = dy / 2.0 + yc[j];
  // the box over the whole array and over the interior
  auto in_x = [&](int i) { return xc[i] >= box[0] && xc[i] <= box[1]; };
  auto in_y = [&](int j) { return yc[j] >= box[2] && yc[j] <= box[3]; };
  auto range = [](int lo, int hi, auto in, int &a, int &b) {
    a = 1;
    b = 0;
    for (int q = lo; q <= hi; ++q)
      if (in(q)) {
        if (a > b) a = q;
        b = q;
      }
  };
  range(1, g.Mloc, in_x, w.bi0, w.bi1);
  range(1, g.Nloc, in_y, w.bj0, w.bj1);
  range(g.Ibeg, g.Iend, in_x, w.si0, w.si1);
  range(g.Jbeg, g.Jend, in_y, w.sj0, w.sj1);
  if (w.bj0 > w.bj1) w.bi0 = 1, w.bi1 = 0;
  if (w.sj0 > w.sj1) w.si0 = 1, w.si1 = 0;
  Dev &d = c->d;
  if (d.wtab) (void)hipFree(d.wtab);
  d.wtab = nullptr;
  if (hip_check(hipMalloc(&d.wtab, sizeof(double) * tab.size()), "internal wavemaker: hipMalloc")) return NHW_ENOMEM;
  if (hip_check(hipMemcpy(d.wtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice),
                "internal wavemaker: upload"))
    return NHW_EHIP;
  w.irr_n = (int)n;
  w.irr_nd = ndir;
  return 0;
}

// update_wave_bc for the west wavemakers: one thread per row j = 1..Nloc at
// the device TIME (after estimate_dt).  Ein/Din for every row; Uin/Vin/Win
// at k = Kbeg..Kend for every row (solitary, which also zeroes the other
// levels) or for the interior rows (linear, Stokes), as the reference's
// loops run.  The inflow arrays are FP64 in both precisions.
template <class R>
__device__ __forceinline__ void wave_bc_body(const Geo &g, const Par &p, const DevT<R> &d, long t) {
  const int j = 1 + (int)t;
  if (j > g.Nloc) return;
  const Wave &w = p.wave;
  const double T = d.scal[1];
  const double *sigc = d.sigc - 1;
  const double hf = d.Hfx[g.x2(g.Ibeg, j)];
  const double A = w.amp, Dp = w.dep;
  auto bi = [&](int k) { return (long)(k - 1) * g.Nloc + (j - 1); };
  if (w.kind == NHW_WAVE_LEF_SOL) {  // solitary_wave_left_boundary, src/waves.F:869-903
    const double at = w.at;
    const double xc = w.xs - w.cel * T;
    const double ch = cosh(at * xc), sh = sinh(at * xc);
    const double ein = A / (ch * ch);
    const double din = ein + hf;
    d.Ein0[j - 1] = ein;
    d.Din0[j - 1] = din;
    const double ch2 = ch * ch, ch3 = ch2 * ch, ch4 = ch3 * ch, ch5 = ch4 * ch;
    const double d2 = 2.0 * A * (at * at) * (2.0 * ch2 - 3.0) / ch4;
    const double d1 = -2.0 * A * sh * at / ch3;
    const double d3 = -8.0 * A * sh * (at * at * at) * (ch2 - 3.0) / ch5;
    for (int k = 1; k <= g.Kloc; ++k) {
      double u = 0.0, wv = 0.0;
      if (k >= g.Kbeg && k <= g.Kend) {
        const double z = sigc[k] * din;
        u = w.c2 * ein / Dp *
            (1.0 - 0.25 * ein / Dp + Dp / 3.0 * (Dp / ein) * (1.0 - 1.5 * (z * z) / (Dp * Dp)) * d2);
        wv = -w.c2 * z / Dp *
             ((1.0 - 0.5 * ein / Dp) * d1 +
              (double)(1.0f / 3.0f) * (Dp * Dp) * (1.0 - 0.5 * (z * z) / (Dp * Dp)) * d3);
      }
      d.Uin0[bi(k)] = u;
      d.Vin0[bi(k)] = 0.0;
      d.Win0[bi(k)] = wv;
    }
    return;
  }
  if (w.kind == NHW_WAVE_LEF_CON) {  // cnoidal_wave_left_boundary, src/waves.F:132-166
    double y = w.cysc * (w.cu0 + w.cu1 * (-(T / w.cper)));
    for (int i = w.cn_n - 1; i >= 1; --i) y = (y + asin(d.wtab[i] * sin(y))) / 2.0;
    const double xcn = cos(y);
    const double ein = w.cyt + A * (xcn * xcn);
    const double din = ein + hf;
    d.Ein0[j - 1] = ein;
    d.Din0[j - 1] = din;
    if (j < g.Jbeg || j > g.Jend) return;
    const double m = w.cm, ck = w.cck, L = w.cwl, yt = w.cyt;
    const double c2 = xcn * xcn, c4 = c2 * xcn * xcn;
    const double xsn = sqrt(1.0 - c2), xdn = sqrt(1.0 - m * (1.0 - c2));
    const double s2 = xsn * xsn, d2 = xdn * xdn, m2 = m * m, dy = Dp + yt;
    for (int k = g.Kbeg; k <= g.Kend; ++k) {
      const double z = sigc[k] * din;
      d.Uin0[bi(k)] = w.c2 *
                          (-1.25 + 3.0 * dy / 2.0 / Dp - (dy * dy) / 4.0 / (Dp * Dp) +
                           (3.0 * A / 2.0 / Dp - dy * A / 2.0 / (Dp * Dp)) * c2 - (A * A) / 4.0 / (Dp * Dp) * c4 -
                           8.0 * A * (ck * ck) / (L * L) * (Dp / 3.0 - (z * z) / 2.0 / Dp) *
                               (-(m2 * s2 * c2) + c2 * d2 - s2 * d2)) -
                      w.drift * 1.0;
      d.Win0[bi(k)] = w.c2 * z * 2.0 * A * ck / L / Dp *
                      (1.0 + dy / Dp + A / Dp * c2 +
                       32.0 * (ck * ck) / 3.0 / (L * L) * (Dp * Dp - (z * z) / 2.0) * (m2 * s2 - m2 * c2 - d2)) *
                      xsn * xcn * xdn;
      d.Vin0[bi(k)] = 0.0;
    }
    return;
  }
  const double ph = kPi / 2.0 - w.segma * T;
  if (w.kind == NHW_WAVE_LEF_LIN) {  // linear_wave_left_boundary, src/waves.F:808-855
    const double arg = ph + d.wtab[j - 1];
    const double c = cos(arg), sn = sin(arg);
    const double ein = 0.5 * A * c;
    const double din = ein + hf;
    d.Ein0[j - 1] = ein;
    d.Din0[j - 1] = din;
    if (j < g.Jbeg || j > g.Jend) return;
    for (int k = g.Kbeg; k <= g.Kend; ++k) {
      const double z = sigc[k] * din;
      const double uh = 0.5 * A * w.segma * cosh(w.wk * z) / w.shd * c - w.drift * 1.0;
      d.Uin0[bi(k)] = uh * w.cth;
      d.Win0[bi(k)] = 0.5 * A * w.segma * sinh(w.wk * z) / w.shd * sn;
      d.Vin0[bi(k)] = uh * w.sth;
    }
    return;
  }
  // stokes_wave_left_boundary, src/waves.F:518-560
  const double c1 = cos(ph), c2 = cos(2.0 * ph);
  const double ein = 0.5 * A * c1 + w.eta2 * c2;
  const double din = ein + hf;
  d.Ein0[j - 1] = ein;
  d.Din0[j - 1] = din;
  if (j < g.Jbeg || j > g.Jend) return;
  const double s1 = sin(ph), s2 = sin(2.0 * ph);
  for (int k = g.Kbeg; k <= g.Kend; ++k) {
    const double z = sigc[k] * din;
    d.Uin0[bi(k)] = 0.5 * A * w.segma * cosh(w.wk * z) / w.shd * c1 +
                    0.1875 * (A * A) * w.segma * w.wk * cosh(2.0 * w.wk * z) / w.sh4 * c2;
    d.Win0[bi(k)] = 0.5 * A * w.segma * sinh(w.wk * z) / w.shd * s1 +
                    0.1875 * (A * A) * w.segma * w.wk * sinh(2.0 * w.wk * z) / w.sh4 * s2;
    d.Vin0[bi(k)] = 0.0;
  }
}

// sponge_damping, src/nhwave.F:6462-6502: every wet cell of the whole array
// (ghosts included), one thread per column, k innermost in its loop.
template <class R>
__device__ __forceinline__ void sponge_body(const Geo &g, const Par &p, const DevT<R> &d, long c) {
  const long n2 = g.n2();
  if (c >= n2) return;
  if (d.Mask[c] != 1) return;
  const double ramp = p.wave.tramp > 0.0 ? tanh(d.scal[1] / p.wave.tramp) : 1.0;
  const double ucur = ramp * p.wave.cur;
  const double sp = d.Sponge[c];
  const double eta = d.Eta[c] * sp;
  d.Eta[c] = eta;
  const double D = eta + d.Hc[c];
  d.D[c] = D;
  for (int k = 0; k < g.Kloc; ++k) {
    const long q = c + (long)k * n2;
    const R u = (R)(((double)d.U[q] - ucur) * sp + ucur);
    const R v = (R)((double)d.V[q] * sp);
    d.U[q] = u;
    d.V[q] = v;
    d.DU[q] = (R)(D * (double)u);
    d.DV[q] = (R)(D * (double)v);
  }
}

__global__ __launch_bounds__(64) void k_wave_bc(Geo g, Par p, Dev d) {
  wave_bc_body<double>(g, p, d, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <class R>
__global__ __launch_bounds__(256) void k_sponge(Geo g, Par p, DevT<R> d) {
  sponge_body<R>(g, p, d, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// batched form: the members with a wavemaker (a sponge) only, each block on
// its member's own state
template <class R>
__global__ __launch_bounds__(64) void kb_wave_bc(BTab<R> T) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  wave_bc_body<R>(S.g, S.p, S.d, b.b * blockDim.x + threadIdx.x);
}

template <class R>
__global__ __launch_bounds__(256) void kb_sponge(BTab<R> T) {
  const BPos b = bfind(T);
  const BScn<R> &S = T.scn[b.s];
  sponge_body<R>(S.g, S.p, S.d, b.b * blockDim.x + threadIdx.x);
}

// source_terms' internal irregular wavemaker, src/source_terms.F:194-305:
// one thread per interior cell of the source box at the device TIME.  Each
// sums dx*D(i,Jbeg+1) over the box's columns (Source_Area, in the
// reference's order), then the components in the reference's order.  The
// first frequency's terms carry "+ eta_mean", a module variable the
// reference never assigns (zero).  Cells outside the box keep SourceC = 0.
__global__ __launch_bounds__(64) void k_int_source(Geo g, Par p, Dev d) {
  const Wave &w = p.wave;
  const int ni = w.si1 - w.si0 + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ni * (w.sj1 - w.sj0 + 1)) return;
  const int i = w.si0 + (int)(t % ni), j = w.sj0 + (int)(t / ni);
  double area = 0.0;
  for (int q = w.si0; q <= w.si1; ++q) area = area + g.dx * d.D[g.c2(q, g.Jbeg + 1)];
  const double T = d.scal[1], yl = d.wtab[5L * w.irr_n + (j - 1)];
  double sc = 0.0;
  for (int q = 0; q < w.irr_n; ++q) {
    const double *e = d.wtab + 5L * q;
    const double arg = kPi / 2.0 - e[0] * T + yl * e[3] + e[4];
    if (q < w.irr_nd)
      sc = sc + e[1] / area * (e[2] * cos(arg) + 0.0);
    else
      sc = sc + e[1] / area * cos(arg);
  }
  d.SourceC[g.c2(i, j)] = sc;
}

int launch_int_source(nhw_ctx *c) {
  const Wave &w = c->p.wave;
  if (w.kind != NHW_WAVE_INT_IRR) return 0;
  if (!w.irr_n) {
    set_error("source_terms: INT_IRR needs its spectrum (nhw_set_internal_wavemaker) before the first step");
    return NHW_EINVAL;
  }
  const long n = (long)(w.si1 - w.si0 + 1) * (w.sj1 - w.sj0 + 1);
  if (n <= 0) return 0;
  Timed tm(c, KID_WAVE_BC);
  hipLaunchKernelGGL(k_int_source, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, c->stream, c->g, c->p, c->d);
  return hip_check(hipGetLastError(), "k_int_source");
}

// flux_left_right_boundary, src/nhwave.F:5565-5712 (WAVEMAKER = FLUX_LR), one
// workgroup at the device TIME.  West: a still surface (Ein = 0, Din =
// Hfx(Ibeg)) and the logarithmic profile Uin = u*/kappa log(30 z/Zob)
// tanh(TIME/TRAMP) with u* = 0.115 kappa/log(30 Din(Jbeg)/Zob) (0.115, 0.41
// and 1.e-16 are default REALs); its discharge FluxL summed by one lane in
// the reference's order (k outer, j inner).  East: Orlanski radiation of Eta,
// V, W from the two previous steps and the outflow Uin_Xni = FluxL/(Din_Xn dy
// Nglob), the previous Uin_Xni kept in Uin_Xni0.  With `staged`, each
// thread forms its columns' FluxL terms (the same products, left to right)
// into LDS, in the sum's order, and lane 0 adds them from there; otherwise
// lane 0 forms them from global memory.
__global__ __launch_bounds__(256) void k_flux_lr(Geo g, Par p, Dev d, int staged) {
  extern __shared__ double terms[];
  __shared__ double fl;
  const double T = d.scal[1], zob = p.zob, dy = g.dy;
  const double uavg = (double)0.115f, kappa = (double)0.41f;
  const double ramp = p.wave.tramp > 0.0 ? tanh(T / p.wave.tramp) : 1.0;
  const int nb = g.N * g.K;
  auto clip = [](double cg) { return fmax(fmin(cg, 1.0), 0.0); };
  for (int j = 1 + threadIdx.x; j <= g.Nloc; j += blockDim.x) {
    d.Ein0[j - 1] = 0.0;
    d.Din0[j - 1] = d.Hfx[g.x2(g.Ibeg, j)];
    const long a = g.c2(g.Iend, j), b = g.c2(g.Iend - 1, j);
    const double cg = clip(-(d.Eta0[a] - d.Eta00[a] + kSmall) / (d.Eta00[a] - d.Eta00[b] + kSmall));
    const double ein = (cg * d.Eta0[a] + (1.0 - cg) * d.Eta0[g.c2(g.Iend + 1, j)] + d.Eta[a]) / 2.0;
    d.Ein_Xn[j - 1] = ein;
    d.Din_Xn[j - 1] = d.Hfx[g.x2(g.Iend + 1, j)] + ein;
  }
  __syncthreads();
  const double ufric = uavg * kappa / log(30.0 * d.Din0[g.Jbeg - 1] / zob);
  for (int t = threadIdx.x; t < nb; t += blockDim.x) {
    const int j = g.Jbeg + t % g.N, k = g.Kbeg + t / g.N;
    const long q = inflow_idx(g, j, k);
    const double uin = ufric / kappa * log(30.0 * (d.sigc[k - 1] * d.Din0[j - 1]) / zob) * ramp;
    d.Uin0[q] = uin;
    d.Vin0[q] = 0.0;
    d.Win0[q] = 0.0;
    if (staged) terms[t] = d.dsig[k - 1] * d.Din0[j - 1] * dy * uin;  // t: k outer, j inner
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    if (staged)
      for (int t = 0; t < nb; ++t) s = s + terms[t];
    else
      for (int k = g.Kbeg; k <= g.Kend; ++k)
        for (int j = g.Jbeg; j <= g.Jend; ++j)
          s = s + d.dsig[k - 1] * d.Din0[j - 1] * dy * d.Uin0[inflow_idx(g, j, k)];
    fl = s;
  }
  __syncthreads();
  const double fluxl = fl;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) {
    const int j = g.Jbeg + t % g.N, k = g.Kbeg + t / g.N;
    const long q = inflow_idx(g, j, k);
    const long a = g.c3(g.Iend, j, k), b = g.c3(g.Iend - 1, j, k), e = g.c3(g.Iend + 1, j, k);
    const double xi0 = d.Uin_Xni[q];
    const double xi = fluxl / (d.Din_Xn[j - 1] * dy * (double)g.N);
    d.Uin_Xni0[q] = xi0;
    d.Uin_Xni[q] = xi;
    double cg = clip(-(d.U0[a] - d.U00[a] + kSmall) / (d.U00[a] - d.U00[b] + kSmall));
    d.Uin_Xn[q] = (cg * d.U[a] + xi0 - cg * xi + d.U[a]) / 2.0;
    cg = clip(-(d.V0[a] - d.V00[a] + kSmall) / (d.V00[a] - d.V00[b] + kSmall));
    d.Vin_Xn[q] = (cg * d.V0[a] + (1.0 - cg) * d.V0[e] + d.V[a]) / 2.0;
    cg = clip(-(d.W0[a] - d.W00[a] + kSmall) / (d.W00[a] - d.W00[b] + kSmall));
    d.Win_Xn[q] = (cg * d.W0[a] + (1.0 - cg) * d.W0[e] + d.W[a]) / 2.0;
  }
}

size_t flux_lr_state_bytes(const nhw_ctx *c) {
  if (c->p.wave.kind != NHW_WAVE_FLUX_LR) return 0;
  const Geo &g = c->g;
  return sizeof(double) * ((size_t)g.Nloc * g.Kloc + (size_t)g.n2() + 6 * (size_t)g.n3());
}

// FLUX_LR's history, a checkpoint section: Uin_Xni (the next call's
// Uin_Xni0) and Eta00, U0, V0, W0, U00, V00, W00.  The other boundary arrays
// are recomputed from these by the next update_wave_bc before any reads them.
int flux_lr_state(nhw_ctx *c, std::vector<char> &buf, bool save) {
  const size_t n = flux_lr_state_bytes(c);
  if (!n) return 0;
  if (save) buf.assign(n, 0);
  else if (buf.size() != n) {
    set_error("FLUX_LR checkpoint state: size differs from the context's");
    return NHW_EINVAL;
  }
  const Geo &g = c->g;
  const Dev &d = c->d;
  const size_t b3 = sizeof(double) * (size_t)g.n3();
  struct Arr { double *dev; size_t bytes; } arrs[] = {
      {d.Uin_Xni, sizeof(double) * (size_t)g.Nloc * g.Kloc}, {d.Eta00, sizeof(double) * (size_t)g.n2()},
      {d.U0, b3}, {d.V0, b3}, {d.W0, b3}, {d.U00, b3}, {d.V00, b3}, {d.W00, b3}};
  char *a = buf.data();
  int rc = 0;
  for (const Arr &x : arrs) {
    if (rc) break;
    rc = save ? hip_check(hipMemcpyAsync(a, x.dev, x.bytes, hipMemcpyDeviceToHost, c->stream), "FLUX_LR checkpoint state")
              : hip_check(hipMemcpyAsync(x.dev, a, x.bytes, hipMemcpyHostToDevice, c->stream), "FLUX_LR checkpoint state");
    a += x.bytes;
  }
  if (!rc) rc = hip_check(hipStreamSynchronize(c->stream), "FLUX_LR checkpoint state");
  return rc ? NHW_EHIP : 0;
}

int launch_wave_bc(nhw_ctx *c) {
  if (c->p.wave.kind == NHW_WAVE_FLUX_LR) {
    Timed tm(c, KID_WAVE_BC);
    // the FluxL terms staged in LDS while they fit in 64 KiB
    const size_t lds = sizeof(double) * (size_t)c->g.N * c->g.K;
    const int staged = lds <= 65536 - 64;
    hipLaunchKernelGGL(k_flux_lr, dim3(1), dim3(256), staged ? lds : 0, c->stream, c->g, c->p, c->d, staged);
    return hip_check(hipGetLastError(), "k_flux_lr");
  }
  if (!west_wavemaker(c->p.wave.kind)) return 0;
  Timed tm(c, KID_WAVE_BC);
  hipLaunchKernelGGL(k_wave_bc, dim3((c->g.Nloc + 63) / 64), dim3(64), 0, c->stream, c->g, c->p, c->d);
  return hip_check(hipGetLastError(), "k_wave_bc");
}

int launch_sponge(nhw_ctx *c) {
  if (!c->p.wave.sponge_on) return 0;
  Timed tm(c, KID_SPONGE);
  const unsigned nb = (unsigned)((c->g.n2() + 255) / 256);
  if (c->mixed)
    hipLaunchKernelGGL(k_sponge<float>, dim3(nb), dim3(256), 0, c->stream, c->g, c->p, view<float>(c->d));
  else
    hipLaunchKernelGGL(k_sponge<double>, dim3(nb), dim3(256), 0, c->stream, c->g, c->p, c->d);
  return hip_check(hipGetLastError(), "k_sponge");
}

template <class R>
static BTab<R> ftab(Batch *B, const BShape &S) {
  return BTab<R>{(const BScn<R> *)B->tab, S.off, S.ids, S.nsub};
}

// launches over the members that have the forcing ("wb/64": a wavemaker,
// "sp/256": a sponge; absent when no member has one)
static int blaunch_forcing(Batch *B, const char *key, int kid, bool wave) {
  const BShape &S = bshape(B, key);
  if (S.nblk <= 0) return 0;
  double bytes = 0.0;
  for (nhw_ctx *c : B->m)
    if (wave ? west_wavemaker(c->p.wave.kind) : c->p.wave.sponge_on) bytes += kernel_bytes(c, kid);
  Timed tm(B->lead, kid, bytes);
  if (wave) {
    if (B->mixed)
      hipLaunchKernelGGL(kb_wave_bc<float>, dim3(S.nblk), dim3(S.bs), 0, B->lead->stream, ftab<float>(B, S));
    else
      hipLaunchKernelGGL(kb_wave_bc<double>, dim3(S.nblk), dim3(S.bs), 0, B->lead->stream, ftab<double>(B, S));
  } else {
    if (B->mixed)
      hipLaunchKernelGGL(kb_sponge<float>, dim3(S.nblk), dim3(S.bs), 0, B->lead->stream, ftab<float>(B, S));
    else
      hipLaunchKernelGGL(kb_sponge<double>, dim3(S.nblk), dim3(S.bs), 0, B->lead->stream, ftab<double>(B, S));
  }
  return hip_check(hipGetLastError(), wave ? "batch update_wave_bc" : "batch sponge_damping");
}

int blaunch_wave_bc(Batch *B) { return blaunch_forcing(B, "wb/64", KID_WAVE_BC, true); }
int blaunch_sponge(Batch *B) { return blaunch_forcing(B, "sp/256", KID_SPONGE, false); }

}  // namespace nhw

using namespace nhw;

extern "C" {

int nhw_set_internal_wavemaker(nhw_ctx *c, double xs_west, double xs_east, double ys_south, double ys_north,
                               int32_t nfreq, int32_t ndir, const double *freq, const double *dire,
                               const double *spc2d, const double *phs) {
  if (!c) {
    set_error("nhw_set_internal_wavemaker: null context");
    return NHW_EINVAL;
  }
  const double box[4] = {xs_west, xs_east, ys_south, ys_north};
  return internal_wavemaker_setup(c, box, nfreq, ndir, freq, dire, spc2d, phs);
}

int nhw_update_wave_bc(nhw_ctx *c) {
  if (!c) {
    set_error("nhw_update_wave_bc: null context");
    return NHW_EINVAL;
  }
  return launch_wave_bc(c);
}

int nhw_sponge_damping(nhw_ctx *c) {
  if (!c) {
    set_error("nhw_sponge_damping: null context");
    return NHW_EINVAL;
  }
  return launch_sponge(c);
}

}  // extern "C"
