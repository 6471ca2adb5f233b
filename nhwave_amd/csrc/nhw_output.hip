// Output and restart path of a resident context:
//  * nhw_statistics: the reference's `statistics` quantities
//    (src/nhwave.F:3119-3208) as device reductions of the resident state,
//    so a driver needs no download to print them;
//  * nhw_snapshot_begin / nhw_snapshot_wait: a copy of chosen fields taken
//    in stream order and drained to the host on a second stream while the
//    time stepping continues;
//  * nhw_checkpoint_save / nhw_checkpoint_load: the full state (every array
//    of nhw_fields, the vertical grid, TIME, dt, dt_old, RUN_STEP) in a
//    binary file, from which a run continues bit for bit (the role of the
//    reference's hot start, src/nhwave.F:13338-13692, at full precision).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "nhw_impl.h"

namespace nhw {

constexpr int kStatBS = 256;
constexpr int kStatBlocks = 512;  // kStatN partials each, inside the context's 4096-entry reduction buffer
constexpr int kStatN = 7;  // mass, energy, max eta, min eta, max |u|, max |v|, max |w|

// per-block partials in a fixed shape: columns of the interior in a
// grid-stride loop, sigma levels inside (the reference's loops,
// src/nhwave.F:3137-3154; each term in its expression order)
template <class R>
__global__ __launch_bounds__(kStatBS) void k_stat_partial(Geo g, DevT<R> d, double *part) {
  const double *dsig = d.dsig - 1, *sigc = d.sigc - 1;
  const double dx = g.dx, dy = g.dy;
  double a[kStatN] = {0.0, 0.0, -INFINITY, INFINITY, 0.0, 0.0, 0.0};
  const long n2 = (long)g.M * g.N;
  for (long t = (long)blockIdx.x * kStatBS + threadIdx.x; t < n2; t += (long)gridDim.x * kStatBS) {
    const int i = g.Ibeg + (int)(t % g.M), j = g.Jbeg + (int)(t / g.M);
    const long c = g.c2(i, j);
    const double Dd = d.D[c], hc = d.Hc[c], eta = d.Eta[c];
    a[0] += Dd * dx * dy;
    for (int k = g.Kbeg; k <= g.Kend; ++k) {
      const long q = g.c3(i, j, k);
      const double u = (double)d.U[q], v = (double)d.V[q], w = (double)d.W[q];
      const double cm = kRho0 * dsig[k] * Dd * dx * dy;
      a[1] += cm * kGrav * (Dd * sigc[k] - hc) + 0.5 * cm * (u * u + v * v + w * w);
      a[4] = fmax(a[4], fabs(u));
      a[5] = fmax(a[5], fabs(v));
      a[6] = fmax(a[6], fabs(w));
    }
    a[2] = fmax(a[2], eta);
    a[3] = fmin(a[3], eta);
  }
  __shared__ double sm[kStatN][kStatBS];
  for (int m = 0; m < kStatN; ++m) sm[m][threadIdx.x] = a[m];
  __syncthreads();
  for (int s = kStatBS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sm[0][threadIdx.x] += sm[0][threadIdx.x + s];
      sm[1][threadIdx.x] += sm[1][threadIdx.x + s];
      sm[2][threadIdx.x] = fmax(sm[2][threadIdx.x], sm[2][threadIdx.x + s]);
      sm[3][threadIdx.x] = fmin(sm[3][threadIdx.x], sm[3][threadIdx.x + s]);
      for (int m = 4; m < kStatN; ++m) sm[m][threadIdx.x] = fmax(sm[m][threadIdx.x], sm[m][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int m = 0; m < kStatN; ++m) part[kStatN * blockIdx.x + m] = sm[m][0];
}

// the block partials in a fixed order into out[0..6]
__global__ __launch_bounds__(kStatBS) void k_stat_final(const double *part, int nblk, double *out) {
  __shared__ double sm[kStatN][kStatBS];
  double a[kStatN] = {0.0, 0.0, -INFINITY, INFINITY, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += kStatBS) {
    const double *p = part + kStatN * b;
    a[0] += p[0];
    a[1] += p[1];
    a[2] = fmax(a[2], p[2]);
    a[3] = fmin(a[3], p[3]);
    for (int m = 4; m < kStatN; ++m) a[m] = fmax(a[m], p[m]);
  }
  for (int m = 0; m < kStatN; ++m) sm[m][threadIdx.x] = a[m];
  __syncthreads();
  for (int s = kStatBS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sm[0][threadIdx.x] += sm[0][threadIdx.x + s];
      sm[1][threadIdx.x] += sm[1][threadIdx.x + s];
      sm[2][threadIdx.x] = fmax(sm[2][threadIdx.x], sm[2][threadIdx.x + s]);
      sm[3][threadIdx.x] = fmin(sm[3][threadIdx.x], sm[3][threadIdx.x + s]);
      for (int m = 4; m < kStatN; ++m) sm[m][threadIdx.x] = fmax(sm[m][threadIdx.x], sm[m][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int m = 0; m < kStatN; ++m) out[m] = sm[m][0];
}

void snap_free(nhw_ctx *c) {
  Snap &s = c->snap;
  if (s.stream) (void)hipStreamSynchronize(s.stream);
  if (s.dev) (void)hipFree(s.dev);
  if (s.host) (void)hipHostFree(s.host);
  if (s.taken) (void)hipEventDestroy(s.taken);
  if (s.done) (void)hipEventDestroy(s.done);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  s = Snap{};
}

// unpacks a finished snapshot into the caller's arrays
static int snap_finish(nhw_ctx *c) {
  Snap &s = c->snap;
  if (!s.busy) return 0;
  if (hip_check(hipEventSynchronize(s.done), "nhw_snapshot_wait")) return NHW_EHIP;
  for (const SnapItem &it : s.items) {
    const char *src = (const char *)s.host + it.off;
    if (it.widen) {
      const float *f = (const float *)src;
      double *dst = (double *)it.host;
      for (size_t q = 0; q < it.dbytes / sizeof(float); ++q) dst[q] = (double)f[q];
    } else {
      std::memcpy(it.host, src, it.dbytes);
    }
  }
  s.items.clear();
  s.busy = false;
  return 0;
}

// the checkpoint file: a fixed header, the vertical grid, the device
// scalars, then every array of nhw_fields in the header's order as the
// host sees it (FP64 / int32)
struct CkptHeader {
  char magic[8];
  int32_t version, sizeof_config;
  nhw_config cfg;
  int64_t nbytes_arrays;
};
static const char kCkptMagic[8] = {'N', 'H', 'W', 'C', 'K', 'P', 'T', '1'};

// nhw_fields is a sequence of pointers in field_table's order: member q
static void set_slot(nhw_fields &f, size_t q, void *p) {
  std::memcpy((char *)&f + q * sizeof(void *), &p, sizeof(void *));
}

static bool same_shape(const nhw_config &a, const nhw_config &b) {
  return a.mglob == b.mglob && a.nglob == b.nglob && a.kglob == b.kglob && a.precision == b.precision &&
         a.non_hydro == b.non_hydro;
}

// The first physics setting in which a checkpoint's configuration and the
// context's differ (nullptr: none).  A run continued under other physics
// would go on silently from a state made by different equations, so the
// load refuses it.  DT_INI (used only by the first step), ITMAX and the
// Poisson solver choice (both solve to the same TOL) may differ.
static const char *physics_mismatch(const nhw_config &a, const nhw_config &b) {
#define NHW_SAME(f) \
  if (!(a.f == b.f)) return #f;
  NHW_SAME(dx) NHW_SAME(dy) NHW_SAME(cfl) NHW_SAME(min_dep) NHW_SAME(dt_min) NHW_SAME(dt_max)
  NHW_SAME(froude_cap_on) NHW_SAME(visc) NHW_SAME(fcor) NHW_SAME(wind_u) NHW_SAME(wind_v)
  NHW_SAME(time_order) NHW_SAME(hllc) NHW_SAME(bc_x0) NHW_SAME(bc_xn) NHW_SAME(bc_y0) NHW_SAME(bc_yn)
  NHW_SAME(bc_z0) NHW_SAME(tol) NHW_SAME(wavemaker) NHW_SAME(sponge_on) NHW_SAME(viscous)
  if (a.froude_cap_on) {
    NHW_SAME(froude_cap)
  }
  if (a.bc_z0 == 5) {
    NHW_SAME(ibot)
    if (a.ibot == 1) {
      NHW_SAME(cd0)
    } else {
      NHW_SAME(zob)
    }
  }
  if (a.wavemaker != NHW_WAVE_NONE) {
    NHW_SAME(amp_wave) NHW_SAME(per_wave) NHW_SAME(dep_wave) NHW_SAME(theta_wave)
  }
  if (a.sponge_on) {
    NHW_SAME(sponge_w) NHW_SAME(sponge_e) NHW_SAME(sponge_s) NHW_SAME(sponge_n) NHW_SAME(cur_wave)
    NHW_SAME(tramp)
  }
  if (a.viscous) {
    NHW_SAME(rng) NHW_SAME(cvs) NHW_SAME(chs) NHW_SAME(schmidt) NHW_SAME(visc_number)
  }
#undef NHW_SAME
  return nullptr;
}

}  // namespace nhw

using namespace nhw;

extern "C" {

int nhw_statistics(nhw_ctx *c, nhw_stat_values *out) {
  if (!c || !out) {
    set_error("nhw_statistics: null argument");
    return NHW_EINVAL;
  }
  const Geo &g = c->g;
  const int nblk = (int)std::min<long>(kStatBlocks, ((long)g.M * g.N + kStatBS - 1) / kStatBS);
  double *part = c->d.red;  // kStatN * kStatBlocks partials, then the kStatN results
  double *res = c->d.red + kStatN * kStatBlocks;
  if (c->mixed)
    hipLaunchKernelGGL(k_stat_partial<float>, dim3(nblk), dim3(kStatBS), 0, c->stream, g, view<float>(c->d), part);
  else
    hipLaunchKernelGGL(k_stat_partial<double>, dim3(nblk), dim3(kStatBS), 0, c->stream, g, c->d, part);
  hipLaunchKernelGGL(k_stat_final, dim3(1), dim3(kStatBS), 0, c->stream, (const double *)part, nblk, res);
  double *h = (double *)((char *)c->pinned + 2048);
  if (hip_check(hipGetLastError(), "nhw_statistics") ||
      hip_check(hipMemcpyAsync(h, res, sizeof(double) * kStatN, hipMemcpyDeviceToHost, c->stream), "nhw_statistics") ||
      hip_check(hipStreamSynchronize(c->stream), "nhw_statistics"))
    return NHW_EHIP;
  out->mass_volume = h[0];
  out->energy = h[1];
  out->max_eta = h[2];
  out->min_eta = h[3];
  out->max_u = h[4];
  out->max_v = h[5];
  out->max_w = h[6];
  return 0;
}

int nhw_snapshot_begin(nhw_ctx *c, const nhw_fields *host) {
  if (!c || !host) {
    set_error("nhw_snapshot_begin: null argument");
    return NHW_EINVAL;
  }
  Snap &s = c->snap;
  int rc = snap_finish(c);  // a snapshot still in flight completes first
  if (rc) return rc;
  std::vector<SnapItem> items;
  size_t total = 0;
  std::vector<const Xfer *> src;
  const std::vector<Xfer> tab = field_table(c, host);
  if ((rc = check_present(tab.data(), tab.size(), "nhw_snapshot_begin"))) return rc;
  for (const Xfer &x : tab) {
    if (!x.host) continue;
    const bool widen = x.r3 && c->mixed;
    const size_t db = widen ? x.bytes / 2 : x.bytes;
    items.push_back(SnapItem{x.host, total, db, widen});
    src.push_back(&x);
    total += (db + 255) / 256 * 256;
  }
  if (items.empty()) return 0;
  if (!s.stream &&
      (hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "nhw_snapshot_begin") ||
       hip_check(hipEventCreateWithFlags(&s.taken, hipEventDisableTiming), "nhw_snapshot_begin") ||
       hip_check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "nhw_snapshot_begin")))
    return NHW_EHIP;
  if (total > s.cap) {
    if (hip_check(hipStreamSynchronize(c->stream), "nhw_snapshot_begin")) return NHW_EHIP;
    if (s.dev) (void)hipFree(s.dev);
    if (s.host) (void)hipHostFree(s.host);
    s.dev = s.host = nullptr;
    s.cap = 0;
    if (hip_check(hipMalloc(&s.dev, total), "nhw_snapshot_begin staging") ||
        hip_check(hipHostMalloc(&s.host, total, hipHostMallocDefault), "nhw_snapshot_begin staging"))
      return NHW_ENOMEM;
    s.cap = total;
  }
  // device copies in stream order: the state as it is after the work queued so far
  for (size_t q = 0; q < items.size(); ++q)
    if (hip_check(hipMemcpyAsync((char *)s.dev + items[q].off, src[q]->dev, items[q].dbytes,
                                 hipMemcpyDeviceToDevice, c->stream),
                  "nhw_snapshot_begin"))
      return NHW_EHIP;
  if (hip_check(hipEventRecord(s.taken, c->stream), "nhw_snapshot_begin") ||
      hip_check(hipStreamWaitEvent(s.stream, s.taken, 0), "nhw_snapshot_begin") ||
      hip_check(hipMemcpyAsync(s.host, s.dev, total, hipMemcpyDeviceToHost, s.stream), "nhw_snapshot_begin") ||
      hip_check(hipEventRecord(s.done, s.stream), "nhw_snapshot_begin"))
    return NHW_EHIP;
  s.items = std::move(items);
  s.busy = true;
  return 0;
}

int nhw_snapshot_wait(nhw_ctx *c) {
  if (!c) {
    set_error("nhw_snapshot_wait: null context");
    return NHW_EINVAL;
  }
  return snap_finish(c);
}

int nhw_checkpoint_save(nhw_ctx *c, const char *path) {
  if (!c || !path) {
    set_error("nhw_checkpoint_save: null argument");
    return NHW_EINVAL;
  }
  if (!c->grid_set) {
    set_error("nhw_checkpoint_save: vertical grid not set");
    return NHW_EINVAL;
  }
  const Geo &g = c->g;
  // every array of nhw_fields, downloaded in one pass
  nhw_fields f;
  std::memset(&f, 0, sizeof f);
  std::vector<Xfer> tab = field_table(c, &f);
  if (tab.size() != kFieldSlots) {
    set_error("nhw_checkpoint_save: field table out of step with nhw_fields");
    return NHW_EINVAL;
  }
  std::vector<std::vector<char>> bufs(tab.size());
  int64_t nbytes = 0;
  for (size_t q = 0; q < tab.size(); ++q) {
    if (!tab[q].bytes) continue;  // an array this context does not hold
    bufs[q].resize(tab[q].bytes);
    set_slot(f, q, bufs[q].data());
    nbytes += (int64_t)tab[q].bytes;
  }
  int rc = nhw_download(c, &f);
  if (rc) return rc;
  double scal[4];
  int iscal[2];
  if (hip_check(hipMemcpy(scal, c->d.scal, sizeof scal, hipMemcpyDeviceToHost), "nhw_checkpoint_save") ||
      hip_check(hipMemcpy(iscal, c->d.iscal, sizeof iscal, hipMemcpyDeviceToHost), "nhw_checkpoint_save"))
    return NHW_EHIP;
  FILE *fp = std::fopen(path, "wb");
  if (!fp) {
    set_error(std::string("nhw_checkpoint_save: cannot open ") + path);
    return NHW_EINVAL;
  }
  CkptHeader h;
  std::memset(&h, 0, sizeof h);
  std::memcpy(h.magic, kCkptMagic, 8);
  h.version = 1;
  h.sizeof_config = (int32_t)sizeof(nhw_config);
  h.cfg = c->cfg;
  h.nbytes_arrays = nbytes;
  bool ok = std::fwrite(&h, sizeof h, 1, fp) == 1;
  ok = ok && std::fwrite(c->mg.dsig_h.data() + 1, sizeof(double), g.Kloc, fp) == (size_t)g.Kloc;
  ok = ok && std::fwrite(c->mg.sig_h.data() + 1, sizeof(double), g.Kloc1, fp) == (size_t)g.Kloc1;
  std::vector<double> sigc(g.Kloc);
  ok = ok && hipMemcpy(sigc.data(), c->d.sigc, sizeof(double) * g.Kloc, hipMemcpyDeviceToHost) == hipSuccess;
  ok = ok && std::fwrite(sigc.data(), sizeof(double), g.Kloc, fp) == (size_t)g.Kloc;
  ok = ok && std::fwrite(scal, sizeof scal, 1, fp) == 1 && std::fwrite(iscal, sizeof iscal, 1, fp) == 1;
  for (size_t q = 0; ok && q < tab.size(); ++q) ok = std::fwrite(bufs[q].data(), 1, bufs[q].size(), fp) == bufs[q].size();
  if (ok && c->cfg.viscous) {
    std::vector<char> v;
    ok = viscous_state(c, v, true) == 0 && std::fwrite(v.data(), 1, v.size(), fp) == v.size();
  }
  if (ok && flux_lr_state_bytes(c)) {
    std::vector<char> v;
    ok = flux_lr_state(c, v, true) == 0 && std::fwrite(v.data(), 1, v.size(), fp) == v.size();
  }
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) {
    set_error(std::string("nhw_checkpoint_save: write failed: ") + path);
    return NHW_EINVAL;
  }
  return 0;
}

int nhw_checkpoint_load(nhw_ctx *c, const char *path) {
  if (!c || !path) {
    set_error("nhw_checkpoint_load: null argument");
    return NHW_EINVAL;
  }
  const Geo &g = c->g;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) {
    set_error(std::string("nhw_checkpoint_load: cannot open ") + path);
    return NHW_EINVAL;
  }
  auto fail = [&](const char *why) {
    std::fclose(fp);
    set_error(std::string("nhw_checkpoint_load: ") + why + ": " + path);
    return NHW_EINVAL;
  };
  CkptHeader h;
  if (std::fread(&h, sizeof h, 1, fp) != 1) return fail("short header");
  if (std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != 1 || h.sizeof_config != (int32_t)sizeof(nhw_config))
    return fail("not a checkpoint of this library version");
  if (!same_shape(h.cfg, c->cfg)) return fail("grid, precision or NON_HYDRO differ from the context's");
  if (const char *f = physics_mismatch(h.cfg, c->cfg)) {
    std::fclose(fp);
    set_error(std::string("nhw_checkpoint_load: the checkpoint's ") + f +
              " differs from the context's (a run continues only under the physics that made it): " + path);
    return NHW_EINVAL;
  }
  std::vector<double> dsig(g.Kloc), sig(g.Kloc1), sigc(g.Kloc);
  double scal[4];
  int iscal[2];
  if (std::fread(dsig.data(), sizeof(double), g.Kloc, fp) != (size_t)g.Kloc ||
      std::fread(sig.data(), sizeof(double), g.Kloc1, fp) != (size_t)g.Kloc1 ||
      std::fread(sigc.data(), sizeof(double), g.Kloc, fp) != (size_t)g.Kloc ||
      std::fread(scal, sizeof scal, 1, fp) != 1 || std::fread(iscal, sizeof iscal, 1, fp) != 1)
    return fail("short grid or scalars");
  nhw_fields f;
  std::memset(&f, 0, sizeof f);
  std::vector<Xfer> tab = field_table(c, &f);
  std::vector<std::vector<char>> bufs(tab.size());
  int64_t nbytes = 0;
  for (size_t q = 0; q < tab.size(); ++q) nbytes += (int64_t)tab[q].bytes;
  if (nbytes != h.nbytes_arrays) return fail("array sizes differ");
  for (size_t q = 0; q < tab.size(); ++q) {
    if (!tab[q].bytes) continue;  // an array this context does not hold
    bufs[q].resize(tab[q].bytes);
    if (std::fread(bufs[q].data(), 1, bufs[q].size(), fp) != bufs[q].size()) return fail("short arrays");
    set_slot(f, q, bufs[q].data());
  }
  std::vector<char> visc;
  if (c->cfg.viscous) {
    visc.resize(viscous_state_bytes(c));
    if (std::fread(visc.data(), 1, visc.size(), fp) != visc.size()) return fail("short viscous state");
  }
  std::vector<char> flr(flux_lr_state_bytes(c));
  if (!flr.empty() && std::fread(flr.data(), 1, flr.size(), fp) != flr.size()) return fail("short FLUX_LR state");
  std::fclose(fp);
  int rc = nhw_set_grid(c, dsig.data(), sig.data(), sigc.data());
  if (rc) return rc;
  if ((rc = nhw_upload(c, &f))) return rc;
  if (c->cfg.viscous && (rc = viscous_state(c, visc, false))) return rc;
  if (!flr.empty() && (rc = flux_lr_state(c, flr, false))) return rc;
  if (hip_check(hipMemcpy(c->d.scal, scal, sizeof scal, hipMemcpyHostToDevice), "nhw_checkpoint_load") ||
      hip_check(hipMemcpy(c->d.iscal, iscal, sizeof iscal, hipMemcpyHostToDevice), "nhw_checkpoint_load"))
    return NHW_EHIP;
  return 0;
}

}  // extern "C"
