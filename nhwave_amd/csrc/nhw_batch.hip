// Batched (ensemble) form of the time step: many scenarios of uneven grid
// sizes advance together.  Each member is an ordinary context (its own
// resident state, created and filled through the single-scenario C ABI);
// the batch adds a device table of the members and, per kernel, a block
// prefix over the members, so that every launch of the step covers the whole
// ensemble: the same per-cell bodies as the single form, one launch instead
// of one per member, and one BiCGStab/multigrid pass for all Poisson solves.
// The threads per block of the face and column sweeps are picked by an
// autotuner that times the candidates on the batch's own state.
#include <algorithm>
#include <cstring>
#include <iterator>
#include <utility>

#include "nhw_batch.h"
#include "nhw_krylov_dev.h"

struct nhw_batch {
  nhw::Batch b;
};

namespace nhw {

const BShape &bshape(Batch *B, const std::string &key) {
  auto it = B->shapes.find(key);
  if (it == B->shapes.end()) {
    static BShape empty;
    return empty;
  }
  return it->second;
}

void bshape_add(Batch *B, const std::string &key, const std::vector<long> &items, const std::vector<int> &ids,
                int bs, bool blocks_given, int table) {
  BShape S;
  if (bs % 64) bs = (bs + 63) / 64 * 64;  // bfind works on whole waves
  S.bs = bs;
  S.table = table;
  S.nsub = (int)items.size();
  if (S.nsub == 0) {
    B->shapes.erase(key);
    return;
  }
  std::vector<int> &buf = table ? B->kbuf_h : B->ibuf_h;
  S.off_at = (long)buf.size();
  long acc = 0;
  buf.push_back(0);
  for (long it : items) {
    acc += blocks_given ? it : (it + bs - 1) / bs;
    buf.push_back((int)acc);
  }
  S.nblk = acc;
  if (!ids.empty()) {
    S.ids_at = (long)buf.size();
    for (int v : ids) buf.push_back(v);
  }
  B->shapes[key] = S;
}

// the Krylov and multigrid launch shapes over members `mem` (key prefix pre)
static void solver_shapes(Batch *B, const std::vector<int> &mem, const std::string &pre, int table) {
  std::vector<long> c2, c3, one;
  int maxlev = 0;
  for (int u : mem) {
    const Geo &g = B->m[u]->g;
    c2.push_back((long)g.M * g.N);
    c3.push_back(g.neq());
    one.push_back(1);
    maxlev = std::max(maxlev, (int)(B->mixed ? B->m[u]->mg.lvf : B->m[u]->mg.lv).size());
  }
  bshape_add(B, pre + "c3/256", c3, mem, 256, false, table);
  bshape_add(B, pre + "c2/256", c2, mem, 256, false, table);
  bshape_add(B, pre + "one", one, mem, 256, true, table);
  for (int l = 0; l < kBMaxLev; ++l) {
    std::vector<long> zb, zz, co, rr, rp, rz, rg, pr, gc;
    std::vector<int> izb, izz, ico, irr, irp, irz, irg;
    for (int u : mem) {
      const std::vector<MGLevel> &h = B->mixed ? B->m[u]->mg.lvf : B->m[u]->mg.lv;
      const int nl = (int)h.size();
      if (nl <= l) continue;
      const MGLevel &lv = h[l];
      const long half = (long)((lv.M + 1) / 2) * lv.N;
      if (nl == l + 1) {  // coarsest: one workgroup (kb_coarse)
        co.push_back(1);
        ico.push_back(u);
        continue;
      }
      const MGLevel &q = h[l + 1];
      zb.push_back(half);
      pr.push_back((long)((lv.M + 1) / 2) * lv.N);  // even columns only, each walks its levels
      gc.push_back((long)q.M * q.N);
      izb.push_back(u);
      // the zero-guess half sweep, unless the restriction from level l-1
      // did it (lane pairs into a level that is not the coarsest)
      const bool zdone = l >= 1 && mg_pair_restrict(h[l - 1]);
      if (!zdone) {
        zz.push_back(half);
        izz.push_back(u);
      }
      if (mg_pair_restrict(lv)) {  // lane pairs per coarse cell
        (nl > l + 2 ? rz : rp).push_back(2 * (long)q.M * q.N);
        (nl > l + 2 ? irz : irp).push_back(u);
      } else if (lv.fx == 2 && lv.fy == 2) {
        rr.push_back((long)q.M * q.N);  // one thread per coarse cell
        irr.push_back(u);
      } else {
        rg.push_back((long)q.M * q.N);
        irg.push_back(u);
      }
    }
    const std::string L = std::to_string(l);
    bshape_add(B, pre + "co" + L, co, ico, kCoarseBS, true, table);
    bshape_add(B, pre + "zb" + L, zb, izb, 256, false, table);
    bshape_add(B, pre + "zz" + L, zz, izz, 256, false, table);
    bshape_add(B, pre + "pr" + L, pr, izb, 256, false, table);
    bshape_add(B, pre + "gc" + L, gc, izb, 256, false, table);
    bshape_add(B, pre + "rr" + L, rr, irr, 256, false, table);
    bshape_add(B, pre + "rp" + L, rp, irp, 256, false, table);
    bshape_add(B, pre + "rz" + L, rz, irz, 256, false, table);
    bshape_add(B, pre + "rg" + L, rg, irg, 256, false, table);
  }
  (void)maxlev;
}

// device pointers of the shapes of one table after its upload
static void bind_shapes(Batch *B, int table) {
  int *base = table ? B->kbuf : B->ibuf;
  for (auto &kv : B->shapes) {
    BShape &S = kv.second;
    if (S.table != table) continue;
    S.off = base + S.off_at;
    S.ids = S.ids_at >= 0 ? base + S.ids_at : nullptr;
  }
}

int batch_solver_shapes(Batch *B, const std::vector<std::pair<std::string, std::vector<int>>> &sets) {
  // the previous upload may still be reading the host tables: drain the
  // batch's stream before they are rewritten.  The upload goes on that
  // stream, ordered with the kernels around it; nothing here touches the
  // legacy stream, which another batch's graph capture on another host
  // thread forbids.
  if (B->lead && hip_check(hipStreamSynchronize(B->lead->stream), "batch solver tables")) return NHW_EHIP;
  B->kbuf_h.clear();
  for (auto it = B->shapes.begin(); it != B->shapes.end();)
    it = it->second.table == 1 ? B->shapes.erase(it) : std::next(it);
  for (const auto &st : sets) solver_shapes(B, st.second, st.first, 1);
  if ((long)B->kbuf_h.size() > B->kbuf_cap) {
    if (B->kbuf) (void)hipFree(B->kbuf);
    B->kbuf = nullptr;
    const long cap = std::max<long>((long)B->kbuf_h.size(), 2 * B->kbuf_cap);
    if (hip_check(hipMalloc(&B->kbuf, sizeof(int) * cap), "batch solver tables")) return NHW_ENOMEM;
    B->kbuf_cap = cap;
  }
  if (!B->kbuf_h.empty() &&
      hip_check(hipMemcpyAsync(B->kbuf, B->kbuf_h.data(), sizeof(int) * B->kbuf_h.size(), hipMemcpyHostToDevice,
                               B->lead->stream),
                "batch solver tables"))
    return NHW_EHIP;
  bind_shapes(B, 1);
  return 0;
}

template <class R>
static BScn<R> member_record(Batch *B, nhw_ctx *c) {
  BScn<R> S;
  std::memset((void *)&S, 0, sizeof S);
  Krylov &kr = c->kr;
  S.g = c->g;
  S.p = c->p;
  S.d = view<R>(c->d);
  if (c->cfg.non_hydro) {
    S.rhs = kr.rhs;
    S.x = kr.x + kr.pad;
    S.sc = kr.sc;
    S.flag = kr.flag;
    S.itmax = c->cfg.itmax;
    S.tol2 = c->cfg.tol * c->cfg.tol;
    S.inner = B->mixed ? 1 : 0;
    S.itol2 = B->mixed ? kInnerTol * kInnerTol : S.tol2;
    if (B->mixed) {
      S.r = (R *)kr.fr; S.rh = (R *)kr.frh; S.pp = (R *)kr.fp; S.v = (R *)kr.fv; S.s = (R *)kr.fs;
      S.t = (R *)kr.ft; S.ph = (R *)kr.fph; S.sh = (R *)kr.fsh; S.e = (R *)kr.fe; S.kx = S.e;
    } else {
      S.r = (R *)kr.r; S.rh = (R *)kr.rh; S.pp = (R *)kr.p; S.v = (R *)kr.v; S.s = (R *)kr.s;
      S.t = (R *)kr.t; S.ph = (R *)(kr.ph + kr.pad); S.sh = (R *)(kr.sh + kr.pad); S.e = nullptr;
      S.kx = (R *)S.x;
    }
    const std::vector<MGLevel> &h = B->mixed ? c->mg.lvf : c->mg.lv;
    S.nlev = (int)h.size();
    for (int l = 0; l < S.nlev; ++l) {
      S.L[l] = mg_level_consts<R>(c, h[l]);
      S.G[l] = G2<R>{(const R *)h[l].ax, (const R *)h[l].bx, (const R *)h[l].ay, (const R *)h[l].by,
                     (const R *)h[l].rd, h[l].M + 2};
      S.lx[l] = (R *)h[l].x;
      S.lb[l] = (R *)h[l].b;
      S.lfx[l] = h[l].fx;
      S.lfy[l] = h[l].fy;
    }
    const MGLevel &f = c->mg.lv[0];
    S.L64 = mg_level_consts<double>(c, f);
    S.G64 = G2<double>{f.ax, f.bx, f.ay, f.by, f.rd, f.M + 2};
  }
  return S;
}

static int batch_build(Batch *B) {
  const int ns = (int)B->m.size();
  B->shapes.clear();
  B->ibuf_h.clear();
  B->kbuf_h.clear();
  std::vector<long> fx, fy, fz, c2, c3, a2, a3, one;
  for (nhw_ctx *c : B->m) {
    const Geo &g = c->g;
    fx.push_back(flux_x_threads(g));
    fy.push_back(flux_y_threads(g));
    fz.push_back((long)g.M * g.N);
    c2.push_back((long)g.M * g.N);
    c3.push_back(g.neq());
    a2.push_back(g.n2());
    a3.push_back(g.n3());
    one.push_back(1);
  }
  const std::vector<int> all;
  for (int bs : {64, 128, 256}) {
    const std::string s = "/" + std::to_string(bs);
    bshape_add(B, "fx" + s, fx, all, bs);
    bshape_add(B, "fy" + s, fy, all, bs);
    bshape_add(B, "fz" + s, fz, all, bs);
    bshape_add(B, "c2" + s, c2, all, bs);
  }
  bshape_add(B, "c3/256", c3, all, 256);
  bshape_add(B, "a2/256", a2, all, 256);
  bshape_add(B, "a3/256", a3, all, 256);
  bshape_add(B, "one", one, all, 256, true);
  // boundary forcing over the members that have it
  std::vector<long> wb, sp;
  std::vector<int> wbi, spi;
  for (int u = 0; u < ns; ++u) {
    const nhw_ctx *c = B->m[u];
    if (west_wavemaker(c->p.wave.kind)) { wb.push_back(c->g.Nloc); wbi.push_back(u); }
    if (c->p.wave.sponge_on) { sp.push_back(c->g.n2()); spi.push_back(u); }
  }
  bshape_add(B, "wb/64", wb, wbi, 64);
  bshape_add(B, "sp/256", sp, spi, 256);
  if (B->non_hydro) {
    int maxlev = 0;
    for (nhw_ctx *c : B->m) maxlev = std::max(maxlev, (int)(B->mixed ? c->mg.lvf : c->mg.lv).size());
    if (maxlev > kBMaxLev) {
      set_error("nhw_batch_prepare: too many multigrid levels");
      return NHW_EINVAL;
    }
    for (nhw_ctx *c : B->m) {  // kb_coarse holds the coarsest level in one workgroup
      const MGLevel &lc = (B->mixed ? c->mg.lvf : c->mg.lv).back();
      if ((long)lc.M * lc.N > kMgCoarseCols) {
        set_error("nhw_batch_prepare: coarsest multigrid level above the single-workgroup size");
        return NHW_EINVAL;
      }
    }
    std::vector<int> mem(ns);
    for (int u = 0; u < ns; ++u) mem[u] = u;
    solver_shapes(B, mem, "", 0);  // geometry coarsening ("gc") over every member
  }
  // device tables
  if (B->ibuf) (void)hipFree(B->ibuf);
  B->ibuf = nullptr;
  if (hip_check(hipMalloc(&B->ibuf, sizeof(int) * B->ibuf_h.size()), "batch tables") ||
      hip_check(hipMemcpy(B->ibuf, B->ibuf_h.data(), sizeof(int) * B->ibuf_h.size(), hipMemcpyHostToDevice),
                "batch tables"))
    return NHW_EHIP;
  bind_shapes(B, 0);
  long maxblk = 0;
  for (auto &kv : B->shapes) maxblk = std::max(maxblk, kv.second.nblk);
  if (maxblk > B->part_blocks) {
    if (B->part) (void)hipFree(B->part);
    B->part = nullptr;
    if (hip_check(hipMalloc(&B->part, sizeof(double) * 3 * maxblk), "batch partials")) return NHW_ENOMEM;
    B->part_blocks = maxblk;
  }
  // member table
  const size_t rec = B->mixed ? sizeof(BScn<float>) : sizeof(BScn<double>);
  std::vector<char> tab(rec * ns);
  for (int u = 0; u < ns; ++u) {
    if (B->mixed) {
      BScn<float> S = member_record<float>(B, B->m[u]);
      std::memcpy(tab.data() + rec * u, &S, rec);
    } else {
      BScn<double> S = member_record<double>(B, B->m[u]);
      std::memcpy(tab.data() + rec * u, &S, rec);
    }
  }
  if (B->tab) (void)hipFree(B->tab);
  B->tab = nullptr;
  if (hip_check(hipMalloc(&B->tab, tab.size()), "batch table") ||
      hip_check(hipMemcpy(B->tab, tab.data(), tab.size(), hipMemcpyHostToDevice), "batch table"))
    return NHW_EHIP;
  if (ns > B->gcap) {  // gathered flags and scalars: grow with the membership
    if (B->gflag) (void)hipFree(B->gflag);
    if (B->gflag_h) (void)hipHostFree(B->gflag_h);
    if (B->gsc) (void)hipFree(B->gsc);
    if (B->gsc_h) (void)hipHostFree(B->gsc_h);
    B->gflag = nullptr; B->gflag_h = nullptr; B->gsc = nullptr; B->gsc_h = nullptr;
    const int cap = std::max(ns, 2 * B->gcap);
    if (hip_check(hipMalloc(&B->gflag, sizeof(int) * 2 * cap), "batch flags") ||
        hip_check(hipHostMalloc(&B->gflag_h, sizeof(int) * 2 * cap, hipHostMallocDefault), "batch flags") ||
        hip_check(hipMalloc(&B->gsc, sizeof(double) * 2 * cap), "batch flags") ||
        hip_check(hipHostMalloc(&B->gsc_h, sizeof(double) * 2 * cap, hipHostMallocDefault), "batch flags"))
      return NHW_EHIP;
    B->gcap = cap;
  }
  if (B->non_hydro) {
    // size the solver tables for the whole membership now, so that a step
    // never allocates or frees: with every member iterating they are at
    // their largest
    std::vector<int> all_m(ns);
    for (int u = 0; u < ns; ++u) all_m[u] = u;
    const int rc = batch_solver_shapes(B, {{"k:", all_m}, {"r:", all_m}});
    if (rc) return rc;
  }
  return 0;
}

static int batch_check_flags(Batch *B) {
  const int ns = (int)B->m.size();
  int rc = blaunch_gather_flags(B, 0);
  if (rc) return rc;
  if (hip_check(hipMemcpyAsync(B->gflag_h, B->gflag, sizeof(int) * ns, hipMemcpyDeviceToHost, B->lead->stream),
                "batch flags") ||
      hip_check(hipStreamSynchronize(B->lead->stream), "batch flags"))
    return NHW_EHIP;
  ktimer_flush(B->lead);
  for (int u = 0; u < ns; ++u)
    if (B->gflag_h[u] & 1) {
      set_error("estimate_dt: time step too small (dt < DT_MIN) in batch member " + std::to_string(u));
      return NHW_EDT;
    }
  return 0;
}

// The launches of one step fall into segments separated by the Poisson
// solves, whose host-polled iteration counts vary: hydrostatic, one segment
// (the whole step); non-hydrostatic with It_Order stages, segment 0 = the
// per-step calls and stage 1 up to its solve, segment s = the end of stage s
// and stage s+1 up to its solve, segment It_Order = the end of the last stage
// (src/nhwave.F:119-258, 1186-1264).
static int step_head(Batch *B) {
  int rc;
  if ((rc = blaunch_estimate_dt(B))) return rc;
  if ((rc = blaunch_wave_bc(B))) return rc;
  if ((rc = blaunch_update_mask(B))) return rc;
  if ((rc = blaunch_update_wind(B))) return rc;
  return blaunch_update_vars(B);
}
// The viscous branch of a member (diffusion, the implicit momentum solve,
// eval_turb) and the internal wavemaker's mass source run that member's own
// single-scenario launches, issued on the batch stream so that they order
// with the batched launches and are captured into the same graphs.  Their
// kernels read TIME, CmuHt and CmuVt from device scalars, so a captured
// launch sees the current values.  Kernel timers of members other than the
// lead are held off meanwhile (the lead's timers are on only in the eager,
// uncaptured form).
static bool is_viscous(const nhw_ctx *c) { return c->cfg.viscous != 0; }
static bool is_internal(const nhw_ctx *c) { return c->p.wave.kind == NHW_WAVE_INT_IRR; }
static int own_launches(Batch *B, bool (*want)(const nhw_ctx *), int (*launch)(nhw_ctx *)) {
  for (nhw_ctx *c : B->m) {
    if (!want(c)) continue;
    const hipStream_t own = c->stream;
    const bool timed = c->kt.on;
    c->stream = B->lead->stream;
    c->kt.on = timed && c == B->lead;
    const int rc = launch(c);
    c->stream = own;
    c->kt.on = timed;
    if (rc) return rc;
  }
  return 0;
}

static int stage_pre(Batch *B, int st) {
  int rc;
  if ((rc = own_launches(B, is_internal, launch_int_source))) return rc;
  if ((rc = blaunch_source_terms(B))) return rc;
  if ((rc = blaunch_fluxes(B))) return rc;
  if ((rc = own_launches(B, is_viscous, launch_diffusion))) return rc;
  if ((rc = blaunch_eval_explicit(B, st))) return rc;
  if ((rc = own_launches(B, is_viscous, launch_visc_solve))) return rc;
  return B->non_hydro ? blaunch_get_uvw(B) : 0;
}
static int stage_post(Batch *B, int st) {
  int rc;
  if (B->non_hydro && (rc = blaunch_projection(B))) return rc;
  if ((rc = blaunch_rk_blend(B, st))) return rc;
  if ((rc = blaunch_get_uvw(B))) return rc;
  if ((rc = blaunch_get_omega(B))) return rc;
  if ((rc = blaunch_sponge(B))) return rc;
  return own_launches(B, is_viscous, eval_turb);
}
static int nsegments(const Batch *B) { return B->non_hydro ? B->order + 1 : 1; }
static int segment(Batch *B, int s) {
  int rc;
  if (!B->non_hydro) {
    if ((rc = step_head(B))) return rc;
    for (int st = 1; st <= B->order; ++st)
      if ((rc = stage_pre(B, st)) || (rc = stage_post(B, st))) return rc;
    return 0;
  }
  if (s == 0) return (rc = step_head(B)) ? rc : stage_pre(B, 1);
  if ((rc = stage_post(B, s))) return rc;
  return s < B->order ? stage_pre(B, s + 1) : 0;
}

void batch_drop_graphs(Batch *B) {
  for (hipGraphExec_t x : B->gexec) (void)hipGraphExecDestroy(x);
  B->gexec.clear();
}

// records every segment's launches on the batch stream into a graph; the
// kernels take only device pointers and tables fixed until the next prepare
static int capture_segments(Batch *B) {
  hipStream_t st = B->lead->stream;
  for (int s = 0; s < nsegments(B); ++s) {
    hipGraph_t g = nullptr;
    if (hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "batch graph capture")) {
      batch_drop_graphs(B);
      return NHW_EHIP;
    }
    const int rc = segment(B, s);
    const hipError_t e = hipStreamEndCapture(st, &g);
    hipGraphExec_t x = nullptr;
    if (rc || hip_check(e, "batch graph capture") ||
        hip_check(hipGraphInstantiate(&x, g, nullptr, nullptr, 0), "batch graph instantiate")) {
      if (g) (void)hipGraphDestroy(g);
      batch_drop_graphs(B);
      return rc ? rc : NHW_EHIP;
    }
    (void)hipGraphDestroy(g);
    B->gexec.push_back(x);
  }
  return 0;
}

static int batch_step(Batch *B, int nsteps) {
  int rc = 0;
  // per-launch event timing records events between the launches: run eagerly then
  const bool graphs = B->graphs && !B->lead->kt.on;
  if (graphs && B->gexec.empty() && (rc = capture_segments(B))) return rc;
  const int ns = nsegments(B);
  for (int s = 0; s < nsteps; ++s) {
    for (int g = 0; g < ns; ++g) {
      if (graphs)
        rc = hip_check(hipGraphLaunch(B->gexec[g], B->lead->stream), "batch graph launch");
      else
        rc = segment(B, g);
      if (rc) return rc;
      if (B->non_hydro && g < B->order && (rc = batch_poisson(B))) return rc;
    }
    if ((rc = batch_check_flags(B))) return rc;
  }
  return 0;
}

}  // namespace nhw

using namespace nhw;

static int bad_member(const char *msg) {
  set_error(msg);
  return NHW_EINVAL;
}

// members of one batch share precision, TIME_ORDER and NON_HYDRO
static bool compatible(const nhw_config &a, const nhw_config &b) {
  if (a.precision != b.precision || a.time_order != b.time_order || a.non_hydro != b.non_hydro) {
    set_error("nhw_batch: members must share precision, TIME_ORDER and NON_HYDRO");
    return false;
  }
  return true;
}

static const char *tune_names[BT_COUNT] = {"k_flux_x", "k_flux_y", "k_flux_z", "k_omega_x",
                                           "k_omega_y", "k_uvw_a", "k_uvw_b"};

extern "C" {

int nhw_batch_create(const nhw_config *cfgs, int32_t n, nhw_batch **out) {
  if (!out || n < 0 || (n > 0 && !cfgs)) {
    set_error("nhw_batch_create: need n >= 0 configurations and an output pointer");
    return NHW_EINVAL;
  }
  *out = nullptr;
  for (int u = 1; u < n; ++u)
    if (!compatible(cfgs[0], cfgs[u])) return NHW_EINVAL;
  for (int u = 0; u < n; ++u)
    if (cfgs[u].non_hydro && cfgs[u].solver != NHW_SOLVER_MG)
      return bad_member("nhw_batch_create: batched Poisson solves use the multigrid solver");
  nhw_batch *h = new nhw_batch();
  Batch &B = h->b;
  for (int k = 0; k < BT_COUNT; ++k) B.bs[k] = 256;
  for (int u = 0; u < n; ++u) {
    nhw_ctx *c = nullptr;
    const int rc = nhw_batch_add(h, &cfgs[u], &c);
    if (rc) {
      nhw_batch_destroy(h);
      return rc;
    }
  }
  *out = h;
  return 0;
}

int nhw_batch_add(nhw_batch *h, const nhw_config *cfg, nhw_ctx **ctx) {
  if (!h || !cfg) {
    set_error("nhw_batch_add: null argument");
    return NHW_EINVAL;
  }
  Batch &B = h->b;
  if (!B.m.empty() && !compatible(B.m[0]->cfg, *cfg)) return NHW_EINVAL;
  if (cfg->non_hydro && cfg->solver != NHW_SOLVER_MG)
    return bad_member("nhw_batch_add: batched Poisson solves use the multigrid solver");
  if (cfg->wavemaker == NHW_WAVE_FLUX_LR)
    return bad_member("nhw_batch_add: FLUX_LR runs in single contexts only");
  nhw_ctx *c = nullptr;
  const int rc = nhw_create(cfg, &c);
  if (rc) return rc;
  if (B.m.empty()) {
    B.mixed = cfg->precision == NHW_MIXED;
    B.order = cfg->time_order;
    B.non_hydro = cfg->non_hydro;
  }
  B.m.push_back(c);
  B.kmax = std::max(B.kmax, (int)cfg->kglob);
  B.lead = B.m[0];
  B.prepared = false;
  if (ctx) *ctx = c;
  return 0;
}

int nhw_batch_remove(nhw_batch *h, int32_t i) {
  if (!h || i < 0 || i >= (int)h->b.m.size()) {
    set_error("nhw_batch_remove: bad batch or index");
    return NHW_EINVAL;
  }
  Batch &B = h->b;
  if (B.lead && hip_check(hipStreamSynchronize(B.lead->stream), "nhw_batch_remove")) return NHW_EHIP;
  batch_drop_graphs(&B);
  nhw_destroy(B.m[i]);
  B.m.erase(B.m.begin() + i);
  B.lead = B.m.empty() ? nullptr : B.m[0];
  B.kmax = 1;
  for (nhw_ctx *c : B.m) B.kmax = std::max(B.kmax, (int)c->cfg.kglob);
  B.prepared = false;
  return 0;
}

int nhw_batch_graphs(nhw_batch *h, int32_t enable) {
  if (!h) return NHW_EINVAL;
  if (h->b.lead && hip_check(hipStreamSynchronize(h->b.lead->stream), "nhw_batch_graphs")) return NHW_EHIP;
  h->b.graphs = enable != 0;
  if (!h->b.graphs) batch_drop_graphs(&h->b);
  return 0;
}

int nhw_batch_size(nhw_batch *h, int32_t *n) {
  if (!h || !n) return NHW_EINVAL;
  *n = (int32_t)h->b.m.size();
  return 0;
}

int nhw_batch_member(nhw_batch *h, int32_t i, nhw_ctx **ctx) {
  if (!h || !ctx || i < 0 || i >= (int)h->b.m.size()) {
    set_error("nhw_batch_member: bad batch or index");
    return NHW_EINVAL;
  }
  *ctx = h->b.m[i];
  return 0;
}

int nhw_batch_prepare(nhw_batch *h, int32_t autotune) {
  if (!h) {
    set_error("nhw_batch_prepare: null batch");
    return NHW_EINVAL;
  }
  Batch &B = h->b;
  batch_drop_graphs(&B);  // tables, pointers and block sizes may change below
  if (B.m.empty()) {
    B.prepared = true;
    return 0;
  }
  for (nhw_ctx *c : B.m) {
    if (!c->grid_set) {
      set_error("nhw_batch_prepare: every member needs nhw_set_grid first");
      return NHW_EINVAL;
    }
    if (hip_check(hipStreamSynchronize(c->stream), "nhw_batch_prepare")) return NHW_EHIP;
  }
  int rc = batch_build(&B);
  if (rc) return rc;
  if (autotune) {
    rc = batch_autotune(&B);
    if (rc) return rc;
  }
  if (hip_check(hipStreamSynchronize(B.lead->stream), "nhw_batch_prepare")) return NHW_EHIP;
  B.prepared = true;
  return 0;
}

int nhw_batch_step(nhw_batch *h, int32_t nsteps) {
  if (!h || !h->b.prepared) {
    set_error("nhw_batch_step: batch not prepared since its last change (nhw_batch_prepare)");
    return NHW_EINVAL;
  }
  if (h->b.m.empty()) return 0;
  return batch_step(&h->b, nsteps);
}

int nhw_batch_tuning(nhw_batch *h, int32_t kind, const char **name, int32_t *threads, double *ms3) {
  if (!h || kind < 0 || kind >= BT_COUNT) return NHW_EINVAL;
  if (name) *name = tune_names[kind];
  if (threads) *threads = h->b.bs[kind];
  if (ms3)
    for (int q = 0; q < 3; ++q) ms3[q] = h->b.tune_ms[kind][q];
  return 0;
}

int nhw_batch_destroy(nhw_batch *h) {
  if (!h) return 0;
  Batch &B = h->b;
  if (B.lead && B.lead->stream) (void)hipStreamSynchronize(B.lead->stream);
  batch_drop_graphs(&B);
  if (B.tab) (void)hipFree(B.tab);
  if (B.ibuf) (void)hipFree(B.ibuf);
  if (B.kbuf) (void)hipFree(B.kbuf);
  if (B.part) (void)hipFree(B.part);
  if (B.gflag) (void)hipFree(B.gflag);
  if (B.gflag_h) (void)hipHostFree(B.gflag_h);
  if (B.gsc) (void)hipFree(B.gsc);
  if (B.gsc_h) (void)hipHostFree(B.gsc_h);
  for (nhw_ctx *c : B.m) nhw_destroy(c);
  delete h;
  return 0;
}

}  // extern "C"
