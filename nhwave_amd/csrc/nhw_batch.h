// Device-side scenario table of the batched (ensemble) form: many scenarios
// of uneven sizes advance together, each kernel launch covering every
// scenario of the batch.  A launch's blocks are split between scenarios by a
// prefix table `off` (scenario u owns blocks off[u] .. off[u+1]-1); a block
// finds its scenario by binary search and runs the same per-item body as the
// single-scenario kernel, on that scenario's own geometry and state.
#pragma once
#include "nhw_impl.h"
#include "nhw_mg_dev.h"

namespace nhw {

constexpr int kBMaxLev = 16;  // multigrid levels per scenario in a batch

// one scenario as the batched kernels see it (device memory, read-only)
template <class R>
struct BScn {
  Geo g;
  Par p;
  DevT<R> d;
  // Poisson: FP64 right side and solution; Krylov vectors in R (the inner
  // FP32 solve of the mixed form when R is float); kx is the vector the
  // iteration updates (x itself in the FP64 form, the correction e in the
  // mixed form)
  double *rhs, *x;
  R *r, *rh, *pp, *v, *s, *t, *ph, *sh, *e, *kx;
  double *sc;
  int *flag;
  int itmax;
  double tol2, itol2;
  int inner;
  int nlev;  // multigrid levels
  LvlC<R> L[kBMaxLev];
  G2<R> G[kBMaxLev];
  R *lx[kBMaxLev], *lb[kBMaxLev];  // level vectors (levels >= 1)
  int lfx[kBMaxLev], lfy[kBMaxLev];
  // FP64 fine-level operator (the residual of the FP64 solution)
  LvlC<double> L64;
  G2<double> G64;
};

// a launch over a subset of the batch: blocks of subset member u are
// off[u] .. off[u+1]-1; member u is scenario ids[u] (ids == nullptr: u)
template <class R>
struct BTab {
  const BScn<R> *scn;
  const int *off;
  const int *ids;
  int nsub;
};

struct BPos {
  int s;   // scenario index
  int u;   // subset member
  long b;  // block index within the member
};

// the member subset entry u owning this block: the last u with off[u] <= block
// (off is non-decreasing).  Each wave compares 64 entries at once and counts
// them, so a block start costs one table load rather than a chain of
// dependent loads of a binary search.  Blocks are whole waves (their sizes
// are multiples of 64) and every thread calls this before any exits.
template <class R>
__device__ __forceinline__ BPos bfind(const BTab<R> &T) {
  const int blk = blockIdx.x;
  const int lane = threadIdx.x & 63;
  int cnt = 0;
  for (int base = 0; base < T.nsub; base += 64) {
    const int q = base + lane;
    const bool le = q < T.nsub && T.off[q] <= blk;
    const unsigned long long m = __ballot(le);
    cnt += __popcll(m);
    if (m != ~0ull) break;
  }
  const int u = cnt - 1;
  return BPos{T.ids ? T.ids[u] : u, u, (long)(blk - T.off[u])};
}

}  // namespace nhw

// ---------------------------------------------------------------------------
// host side of a batch
// ---------------------------------------------------------------------------
#include <map>
#include <string>
#include <vector>

namespace nhw {

// block prefix of one launch over a subset of the batch
struct BShape {
  int nsub = 0;
  long nblk = 0;
  int bs = 256;               // threads per block
  const int *off = nullptr;   // device, nsub+1 entries
  const int *ids = nullptr;   // device, nsub entries (nullptr: identity)
  long off_at = -1, ids_at = -1;  // positions in the host table
  int table = 0;              // 0: the batch table, 1: the solver's active-member table
};

// kernels whose threads-per-block the batch autotuner picks
enum BTuneKind { BT_FLUX_X = 0, BT_FLUX_Y, BT_FLUX_Z, BT_OMEGA_X, BT_OMEGA_Y, BT_UVW_A, BT_UVW_B, BT_COUNT };

struct Batch {
  std::vector<nhw_ctx *> m;   // members: each owns its state, like a single context
  nhw_ctx *lead = nullptr;    // m[0]: the batch's stream and kernel timers
  bool mixed = false;
  int order = 2, non_hydro = 1, kmax = 1;
  void *tab = nullptr;        // device BScn<float|double>[m.size()]
  std::map<std::string, BShape> shapes;
  std::vector<int> ibuf_h;    // all prefix/id tables, uploaded once into ibuf
  int *ibuf = nullptr;
  std::vector<int> kbuf_h;    // tables of the members a solve still iterates ("k:" and "r:" shapes)
  int *kbuf = nullptr;
  long kbuf_cap = 0;
  double *part = nullptr;     // reduction partials, 3 per block
  long part_blocks = 0;
  int *gflag = nullptr;       // gathered per-member flags (device) ...
  int *gflag_h = nullptr;     // ... and their pinned host mirror
  double *gsc = nullptr;      // gathered per-member solve scalars (device) ...
  double *gsc_h = nullptr;    // ... and their pinned host mirror
  int gcap = 0;               // members the gathered arrays hold
  int bs[BT_COUNT];           // autotuned threads per block
  double tune_ms[BT_COUNT][4];
  bool prepared = false;
  int last_iters = 4;
  // the explicit launches between Poisson solves, captured once per prepare
  // into HIP graphs and replayed every step (empty: not captured yet)
  bool graphs = true;
  std::vector<hipGraphExec_t> gexec;
};

void batch_drop_graphs(Batch *B);  // after any change of tables, members or block sizes

// per-level constants of a multigrid level (nhw_mg.hip)
template <class V>
LvlC<V> mg_level_consts(const nhw_ctx *c, const MGLevel &lv);

// shape registry (nhw_batch.hip); items[u] per subset member u, ids[u] its
// member index (empty ids: all members in order)
const BShape &bshape(Batch *B, const std::string &key);
void bshape_add(Batch *B, const std::string &key, const std::vector<long> &items, const std::vector<int> &ids,
                int bs, bool blocks_given = false, int table = 0);
// rebuilds the solver shapes ("<pre>c3", "<pre>c2", "<pre>one" and the
// multigrid levels) over members `mem` into the active-member table; every
// prefix in `pres` is rebuilt from its own member list
int batch_solver_shapes(Batch *B, const std::vector<std::pair<std::string, std::vector<int>>> &sets);

// batched launchers: explicit part (nhw_explicit.hip), Poisson (nhw_poisson.hip)
int blaunch_estimate_dt(Batch *B);
int blaunch_update_mask(Batch *B);
int blaunch_update_wind(Batch *B);
int blaunch_update_vars(Batch *B);
int blaunch_source_terms(Batch *B);
int blaunch_fluxes(Batch *B);
int blaunch_eval_explicit(Batch *B, int istep);
int blaunch_get_uvw(Batch *B);
int blaunch_projection(Batch *B);
int blaunch_rk_blend(Batch *B, int istep);
int blaunch_get_omega(Batch *B);
int blaunch_wave_bc(Batch *B);   // boundary forcing (nhw_forcing.hip)
int blaunch_sponge(Batch *B);
int blaunch_gather_flags(Batch *B, int which);  // 0: iscal[1] (dt), 1: Krylov flag[0], [1]
int batch_poisson(Batch *B);
// autotuner: time each block-size candidate of the tuned kernels
int batch_autotune(Batch *B);

}  // namespace nhw
