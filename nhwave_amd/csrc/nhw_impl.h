// Internal definitions shared by the HIP translation units of libnhwave_amd.
// Layout: every field keeps the reference's Fortran layout (1-based (i,j,k),
// i fastest, two ghost cells), so host<->device copies are plain memcpys and
// the x direction is the coalesced one.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nhwave_amd.h"

namespace nhw {

// Literals of the reference that are default REAL (binary32) and widened to
// REAL(8) in its -DDOUBLE_PRECISION build (src/mod_global.F:21-32).
constexpr double kGrav = (double)9.81f;        // Grav
constexpr double kSmall = (double)1.0e-16f;    // Small
constexpr double kLarge = 10000000.0;          // Large
constexpr double kRho0 = 1000.0;               // Rho0
constexpr double kDtGrowth = (double)1.05f;    // src/nhwave.F:3715
constexpr double kWetDry = (double)1.e-6f;     // src/nhwave.F:5425-5431
constexpr double kMinDepBC = (double)0.01f;    // src/nhwave.F:4438
constexpr double kKappa = (double)0.41f;        // Kappa, src/mod_global.F:30

// Zebra layout of one sigma plane of the multigrid solver's vectors: row j
// holds its columns with (i+j) even, then those with (i+j) odd, each colour
// in increasing i, so a half sweep over one colour reads and writes
// contiguous runs.  zcol: plane offset of column (i,j) of an M-wide plane;
// zinv: the column at a plane offset.
__host__ __device__ __forceinline__ long zcol(int i, int j, int M) {
  const int h0 = (M + 1 - (j & 1)) >> 1;  // columns with (i+j) even in row j
  return (long)j * M + (((i + j) & 1) ? h0 : 0) + (i >> 1);
}
__host__ __device__ __forceinline__ void zinv(long q, int M, int &i, int &j) {
  j = (int)(q / M);
  const int r = (int)(q - (long)j * M), h0 = (M + 1 - (j & 1)) >> 1;
  i = r < h0 ? 2 * r + (j & 1) : 2 * (r - h0) + ((j + 1) & 1);
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  xcd_block() renumbers them so every XCD works through one
// contiguous run of logical blocks: stencil neighbours in the next and
// previous rows then sit in the same L2 instead of being fetched once per
// XCD.  A bijection of [0, number of blocks), for 1-D and 2-D grids.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned n = gridDim.x * gridDim.y, b = blockIdx.x + gridDim.x * blockIdx.y;
  const unsigned per = n / kXcds, rem = n % kXcds, x = b % kXcds;
  return x * per + (x < rem ? x : rem) + b / kXcds;
}

struct Geo {
  int Mloc, Nloc, Kloc, Mloc1, Nloc1, Kloc1;
  int Ibeg, Iend, Iend1, Jbeg, Jend, Jend1, Kbeg, Kend, Kend1;
  int M, N, K;
  double dx, dy;
  // 0-based linear offsets of the Fortran 1-based index (i,j,k)
  __host__ __device__ long c2(int i, int j) const { return (long)(j - 1) * Mloc + (i - 1); }
  __host__ __device__ long x2(int i, int j) const { return (long)(j - 1) * Mloc1 + (i - 1); }
  __host__ __device__ long y2(int i, int j) const { return (long)(j - 1) * Mloc + (i - 1); }
  __host__ __device__ long c3(int i, int j, int k) const {
    return ((long)(k - 1) * Nloc + (j - 1)) * Mloc + (i - 1);
  }
  __host__ __device__ long x3(int i, int j, int k) const {
    return ((long)(k - 1) * Nloc + (j - 1)) * Mloc1 + (i - 1);
  }
  __host__ __device__ long y3(int i, int j, int k) const {
    return ((long)(k - 1) * Nloc1 + (j - 1)) * Mloc + (i - 1);
  }
  // interior row index of the Poisson unknown (src/pressure.F:506-514)
  __host__ __device__ long row(int i, int j, int k) const {
    return (long)(i - Ibeg) + (long)M * ((j - Jbeg) + (long)N * (k - Kbeg));
  }
  __host__ __device__ long n2() const { return (long)Mloc * Nloc; }
  __host__ __device__ long nx2() const { return (long)Mloc1 * Nloc; }
  __host__ __device__ long ny2() const { return (long)Mloc * Nloc1; }
  __host__ __device__ long n3() const { return n2() * Kloc; }
  __host__ __device__ long nx3() const { return nx2() * Kloc; }
  __host__ __device__ long ny3() const { return ny2() * Kloc; }
  __host__ __device__ long nz3() const { return n2() * Kloc1; }
  __host__ __device__ long neq() const { return (long)M * N * K; }
};

// Thread counts of the x- and y-face flux kernels: the x kernel gives each
// row of faces wavefronts of 64 lanes covering kXFaces faces each (each lane
// walks the sigma levels); the y kernel gives each (i, k) threads walking
// kYFaces faces of increasing j.
constexpr int kXFaces = 62;
constexpr int kYFaces = 16;
__host__ __device__ __forceinline__ long flux_x_threads(const Geo &g) {
  return (long)((g.M + 1 + kXFaces - 1) / kXFaces) * g.N * 64;
}
__host__ __device__ __forceinline__ long flux_y_threads(const Geo &g) {
  return (long)g.M * g.K * ((g.N + 1 + kYFaces - 1) / kYFaces);
}

// inflow and outflow boundaries (BC 3): offset of (j,k) in the (Nloc,Kloc)
// arrays Uin/Vin/Win_X0 (west) and Uin/Vin/Win_Xn, Uin_Xni (east)
__host__ __device__ __forceinline__ long inflow_idx(const Geo &g, int j, int k) {
  return (long)(k - 1) * g.Nloc + (j - 1);
}

// Constants of the west wavemakers and the sponge (src/waves.F:518-903,
// src/nhwave.F:6462-6502), computed once on the host with the host libm:
// the device evaluates only the functions of TIME.
struct Wave {
  int kind;        // NHW_WAVE_*
  int sponge_on;
  double amp, dep;
  double cel, at, xs, c2;                 // solitary: Celerity, Atmp, Xstart, C2
  double segma, wk, shd, drift, cth, sth; // linear/Stokes: Segma, Wave_Number, sinh(k h), Stokes_Drift,
                                          // cos/sin(THETA)
  double eta2, sh4;                       // Stokes: second-order Eta coefficient, sinh(k h)**4
  double cur, tramp;                      // sponge current and ramp
  // cnoidal: modulus, K(m), trough level, wave length, period, the phase
  // offset Xstart*0.5*K and rate 2K, the Landen start 2**(n-1)*a(n) and the
  // recurrence length n (the ratios c(i)/a(i) are the device table wtab)
  double cm, cck, cyt, cwl, cper, cu0, cu1, cysc;
  int cn_n;
  // internal irregular wavemaker: the number of spectrum components (their
  // constants are the device table wtab, five a component, then the row
  // offsets dy/2 + yc(j)), the source box's interior columns [si0, si1] x
  // [sj0, sj1] and its cells over the whole array [bi0, bi1] x [bj0, bj1]
  // (empty: bi0 > bi1), where the k-epsilon state is reset
  int irr_n, irr_nd, si0, si1, sj0, sj1, bi0, bi1, bj0, bj1;  // irr_nd: components of the first frequency
};

// the wavemakers of update_wave_bc's west boundary (BC_X0 = 3 inflow
// values); the internal wavemaker is a mass source instead
__host__ __device__ inline bool west_wavemaker(int kind) {
  return kind >= NHW_WAVE_LEF_SOL && kind <= NHW_WAVE_LEF_CON;
}

struct Par {
  double min_dep, cfl, dt_min, dt_max, froude_cap, visc, fcor, wind_u, wind_v;
  int bc_x0, bc_xn, bc_y0, bc_yn, bc_z0, froude_on, non_hydro, it_order;
  // viscous branch: Schmidt, VISCOUS_NUMBER; with constant viscosities
  // (turb = 1) the current uniform CmuHt, CmuVt (the turbulence seed until
  // eval_turb) are device scalars (Dev::scal[6], [7]), with the k-epsilon
  // model (turb = 10) per-cell device arrays (Dev::CmuVt, CmuHt = CmuVt)
  int viscous, turb, rng_on;  // rng_on: RNG (the k-epsilon constants)
  double schmidt, visc_number;
  // bottom friction of Bc_Z0 = 5: Ibot (1: Cd0; else log law with Zob)
  int ibot;
  double cd0, zob;
  double alpha[3], beta[3];
  Wave wave;
};

// Device pointers of the resident state.  R is the storage type of the 3-D
// velocity/momentum fields, their stage copies, the face fluxes and Omega:
// double in the plain form, float in the mixed form.  Depth, Eta and every
// other 2-D field, P and the Poisson right-hand side stay double.
template <class R>
struct DevT {
  const double *dsig, *sig, *sigc;
  double *Hc, *DelxH, *DelyH, *Eta, *Eta0, *D, *D0, *DelxEta, *DelyEta;
  double *SourceX, *SourceY, *Wsx, *Wsy, *R1;
  int *Mask, *MaskNew, *Mask_Struct, *Mask9;
  double *Hfx, *Hfy, *Hfx0, *Hfy0, *EtaxL, *EtaxR, *EtayL, *EtayR;
  R *U, *V, *W, *DU, *DV, *DW, *DU0, *DV0, *DW0;
  R *Ex, *Fx, *Gx, *Hx, *Ey, *Fy, *Gy, *Hy, *Fz, *Gz, *Hz;
  R *Omega;
  double *P;
  // surface (k = Kend1) and bottom (k = Kbeg) z-face states of U,V,W used by
  // vel_bc; zero outside the interior like construct_3D_z leaves them
  double *ztU, *ztV, *ztW, *zbU, *zbV, *zbW;
  double *scal;   // [0] dt, [1] TIME, [2] dt_old, [3] scratch min, [5] largest |CmuHt| (k-epsilon),
                  // [6] CmuHt, [7] CmuVt (constant viscosities)
  int *iscal;     // [0] RUN_STEP, [1] error flags (bit 0: dt < dt_min)
  double *red;    // reduction partials
  // west inflow values Ein_X0, Din_X0 (Nloc) and Uin/Vin/Win_X0 (Nloc,Kloc),
  // j fastest; wtab: the linear wavemaker's phase lag per row, or the
  // cnoidal wavemaker's Landen ratios c(i)/a(i) (i = 2..n at [i-1]);
  // Sponge (Mloc,Nloc).
  // Null when the configuration has no wavemaker / sponge.
  double *Ein0, *Din0, *Uin0, *Vin0, *Win0, *wtab, *Sponge;
  double *SourceC;  // the internal wavemaker's mass source (Mloc,Nloc); null without one
  // FLUX_LR (FP64 only; null otherwise): the east open boundary's Ein_Xn,
  // Din_Xn (Nloc), Uin_Xn, Vin_Xn, Win_Xn, the outflow velocity Uin_Xni and
  // its previous value Uin_Xni0 (Nloc,Kloc, j fastest), and the previous
  // steps' Eta00 (Mloc,Nloc), U0, V0, W0, U00, V00, W00 (Mloc,Nloc,Kloc) the
  // Orlanski radiation reads
  double *Ein_Xn, *Din_Xn, *Uin_Xn, *Vin_Xn, *Win_Xn, *Uin_Xni, *Uin_Xni0;
  double *Eta00, *U0, *V0, *W0, *U00, *V00, *W00;
  // viscous branch (FP64, single scenario; null otherwise): DelxSc/DelySc
  // (Mloc,Nloc,Kloc) and DelxSl/DelySl (Mloc,Nloc,Kloc1) of the last
  // sigma_transform; Diff: Diffxx, Diffxy, Diffyx, Diffyy, Diffzx, Diffzy
  // back to back, (Mloc,Nloc,Kloc) each
  double *DelxSc, *DelySc, *DelxSl, *DelySl, *Diff;
  // k-epsilon model (null otherwise): the eddy viscosity CmuVt (CmuHt
  // aliases it), Tke, Eps and DTke = D Tke, DEps = D Eps, (Mloc,Nloc,Kloc)
  double *CmuVt, *CmuHt, *Tke, *Eps, *DTke, *DEps;
};
// The context keeps the double-typed view; the 3-D members point at float
// storage when the context is mixed (allocated with sizeof(float) elements).
using Dev = DevT<double>;

template <class R>
inline DevT<R> view(const Dev &d) {
  DevT<R> v;
  v.dsig = d.dsig; v.sig = d.sig; v.sigc = d.sigc;
  v.Hc = d.Hc; v.DelxH = d.DelxH; v.DelyH = d.DelyH; v.Eta = d.Eta; v.Eta0 = d.Eta0; v.D = d.D; v.D0 = d.D0;
  v.DelxEta = d.DelxEta; v.DelyEta = d.DelyEta;
  v.SourceX = d.SourceX; v.SourceY = d.SourceY; v.Wsx = d.Wsx; v.Wsy = d.Wsy; v.R1 = d.R1;
  v.Mask = d.Mask; v.MaskNew = d.MaskNew; v.Mask_Struct = d.Mask_Struct; v.Mask9 = d.Mask9;
  v.Hfx = d.Hfx; v.Hfy = d.Hfy; v.Hfx0 = d.Hfx0; v.Hfy0 = d.Hfy0;
  v.EtaxL = d.EtaxL; v.EtaxR = d.EtaxR; v.EtayL = d.EtayL; v.EtayR = d.EtayR;
#define NHW_R3(f) v.f = reinterpret_cast<R *>(d.f);
  NHW_R3(U) NHW_R3(V) NHW_R3(W) NHW_R3(DU) NHW_R3(DV) NHW_R3(DW) NHW_R3(DU0) NHW_R3(DV0) NHW_R3(DW0)
  NHW_R3(Ex) NHW_R3(Fx) NHW_R3(Gx) NHW_R3(Hx) NHW_R3(Ey) NHW_R3(Fy) NHW_R3(Gy) NHW_R3(Hy)
  NHW_R3(Fz) NHW_R3(Gz) NHW_R3(Hz) NHW_R3(Omega)
#undef NHW_R3
  v.P = d.P;
  v.ztU = d.ztU; v.ztV = d.ztV; v.ztW = d.ztW; v.zbU = d.zbU; v.zbV = d.zbV; v.zbW = d.zbW;
  v.scal = d.scal; v.iscal = d.iscal; v.red = d.red;
  v.Ein0 = d.Ein0; v.Din0 = d.Din0; v.Uin0 = d.Uin0; v.Vin0 = d.Vin0; v.Win0 = d.Win0;
  v.wtab = d.wtab; v.Sponge = d.Sponge; v.SourceC = d.SourceC;
  v.Ein_Xn = d.Ein_Xn; v.Din_Xn = d.Din_Xn; v.Uin_Xn = d.Uin_Xn; v.Vin_Xn = d.Vin_Xn; v.Win_Xn = d.Win_Xn;
  v.Uin_Xni = d.Uin_Xni; v.Uin_Xni0 = d.Uin_Xni0; v.Eta00 = d.Eta00;
  v.U0 = d.U0; v.V0 = d.V0; v.W0 = d.W0; v.U00 = d.U00; v.V00 = d.V00; v.W00 = d.W00;
  v.DelxSc = d.DelxSc; v.DelySc = d.DelySc; v.DelxSl = d.DelxSl; v.DelySl = d.DelySl; v.Diff = d.Diff;
  v.CmuVt = d.CmuVt; v.CmuHt = d.CmuHt; v.Tke = d.Tke; v.Eps = d.Eps; v.DTke = d.DTke; v.DEps = d.DEps;
  return v;
}

// Krylov workspace (interior vectors padded by `pad` zeros on both sides).
struct Krylov {
  long n = 0, pad = 0;
  double *coef = nullptr;   // 15*n, diag-major (Coef(neqns,15), src/pressure.F:536-555)
  double *rhs = nullptr;
  double *x = nullptr, *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr;
  double *s = nullptr, *t = nullptr, *ph = nullptr, *sh = nullptr;
  double *lu_c = nullptr, *lu_d = nullptr;   // line-preconditioner factors
  double *part = nullptr;   // block partial sums (nblk * 4)
  double *sc = nullptr;     // device scalars
  int *flag = nullptr;      // [0] converged/stop, [1] iterations, [2] breakdown
  int nblk = 0;
  // FP32 vectors of the mixed form's inner solve (e: the correction)
  float *fe = nullptr, *fr = nullptr, *frh = nullptr, *fp = nullptr, *fv = nullptr;
  float *fs = nullptr, *ft = nullptr, *fph = nullptr, *fsh = nullptr;
};

// Geometric multigrid hierarchy of the matrix-free Poisson operator.  Level 0
// is the interior grid; each coarser level halves M and N (K is kept).  Column
// fields live in an (M+2) x (N+2) array with a zero halo.
constexpr int kMgKMax = 16;      // largest Kglob the multigrid kernels take
constexpr long kMgCoarseCols = 1024;  // coarsest level: at most this many columns ...
constexpr int kMgCoarseSide = 64;     // ... and at most this many along either axis (a long,
                                      // narrow grid such as a one-row flume keeps coarsening)
constexpr long kMgPairCols = 1L << 18;  // levels up to this many columns restrict on lane pairs
struct MGLevel {
  int M = 0, N = 0, fx = 0, fy = 0;  // fx, fy: coarsening factors to the next level
  double dx = 0.0, dy = 0.0;
  double *ax = nullptr, *bx = nullptr, *ay = nullptr, *by = nullptr, *rd = nullptr;
  double *x = nullptr, *b = nullptr, *r = nullptr;  // x, b unused on level 0; r unused
  void *kc = nullptr;  // device table of per-sigma-level constants (LvlK), in the level's type
};
// a level restricted on lane pairs (k_resid22_pair): 2x2 coarsening, at most
// kMgPairCols columns
inline bool mg_pair_restrict(const MGLevel &lv) {
  return lv.fx == 2 && lv.fy == 2 && (long)lv.M * lv.N <= kMgPairCols;
}
struct MG {
  std::vector<MGLevel> lv;    // FP64 hierarchy (plain form), or FP64 level 0 only (mixed form)
  std::vector<MGLevel> lvf;   // FP32 hierarchy of the mixed form's inner solve
  std::vector<double> dsig_h, sig_h;  // 1-based host copies of dsig, sig
  int coarse_sweeps = 2;
  int last_iters = 4;                 // iteration count of the previous (inner) solve
};

// An asynchronous snapshot in flight (nhw_snapshot_begin/_wait): device
// staging written on the context stream, copied to pinned host staging on a
// stream of its own, unpacked into the caller's arrays at the wait.
struct SnapItem {
  void *host;       // caller's array (FP64 / int32)
  size_t off;       // byte offset in the staging buffers
  size_t dbytes;    // device bytes (FP32 for 3-D fields of the mixed form)
  bool widen;       // FP32 on the device, FP64 on the host
};
struct Snap {
  void *dev = nullptr, *host = nullptr;
  size_t cap = 0;
  hipStream_t stream = nullptr;
  hipEvent_t taken = nullptr, done = nullptr;
  std::vector<SnapItem> items;
  bool busy = false;
};

// Kernels timed by the optional per-launch event instrumentation.
enum KernelId {
  KID_FLUX_X = 0, KID_FLUX_Y, KID_FLUX_Z, KID_EXPLICIT, KID_UVW_A, KID_UVW_B, KID_PROJECTION,
  KID_RK_BLEND, KID_OMEGA_X, KID_OMEGA_Y, KID_OMEGA_COL, KID_DT, KID_MASK, KID_SOURCE,
  KID_ASSEMBLE, KID_SPMV_V, KID_SPMV_T, KID_RESID, KID_PREC_P, KID_PREC_S, KID_LINE_FACTOR,
  KID_KUPDATE, KID_KREDUCE, KID_KMISC, KID_COPY, KID_RHS, KID_MF_V, KID_MF_T, KID_MF_RESID,
  KID_MG_SETUP, KID_MG_SMOOTH, KID_MG_RESID, KID_MG_COARSE, KID_MG_PROLONG, KID_KVEC, KID_WAVE_BC,
  KID_SPONGE, KID_DIFFUSION, KID_VISC_SIGMA, KID_VISC_SOLVE, KID_KEPS, KID_COUNT
};

struct KTimer {
  bool on = false;
  std::vector<std::pair<std::pair<int, double>, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> pool;
  double total_ms[KID_COUNT] = {};
  double total_bytes[KID_COUNT] = {};
  long launches[KID_COUNT] = {};
};

}  // namespace nhw

struct nhw_ctx {
  nhw_config cfg;
  nhw::Geo g;
  nhw::Par p;
  nhw::Dev d;
  nhw::Krylov kr;
  nhw::MG mg;
  hipStream_t stream = nullptr;
  int device = 0;
  bool mixed = false;       // NHW_MIXED: 3-D fields stored in float
  void *pinned = nullptr;   // small pinned host buffer for scalar readback
  nhw_stats stats{};
  bool grid_set = false;
  nhw::KTimer kt;
  nhw::Snap snap;
};

namespace nhw {
// the Poisson vectors of the multigrid solver are in the zebra layout; those
// of the line solver are in natural (i fastest) order
inline int zlayout(const nhw_ctx *c) { return c->cfg.solver != NHW_SOLVER_LINE; }

// RAII bracket of one launch with events on the context stream (no-op when
// timing is off).  `bytes` is the launch's algorithmic HBM traffic; a
// negative value takes the per-kernel model (kernel_bytes).
struct Timed {
  nhw_ctx *c;
  int id;
  double bytes;
  hipEvent_t a = nullptr;
  Timed(nhw_ctx *ctx, int kid, double bytes = -1.0);
  ~Timed();
};
void ktimer_flush(nhw_ctx *c);   // after a stream sync: fold finished pairs in
double kernel_bytes(const nhw_ctx *c, int kid);  // algorithmic bytes per launch
const char *kernel_name(int kid);
}  // namespace nhw

namespace nhw {
// one host <-> device array of the model state (nhw_fields order)
struct Xfer {
  void *host;
  void *dev;
  size_t bytes;     // host bytes (FP64 / int32)
  bool r3 = false;  // 3-D field stored in the context's storage type
};
// nhw_fields holds only pointers, one slot per array, in field_table's order
static_assert(sizeof(nhw_fields) % sizeof(void *) == 0, "nhw_fields holds pointers only");
constexpr size_t kFieldSlots = sizeof(nhw_fields) / sizeof(void *);
std::vector<Xfer> field_table(nhw_ctx *c, const nhw_fields *h);
void field_table_into(nhw_ctx *c, const nhw_fields *h, Xfer *out);  // kFieldSlots entries, no allocation
// NHW_EINVAL when a host array is given for a field the context does not hold
int check_present(const Xfer *t, size_t n, const char *who);
void snap_free(nhw_ctx *c);
void set_error(const std::string &msg);
int hip_check(hipError_t e, const char *what);
// explicit-part launchers (nhw_explicit.hip)
int launch_initial(nhw_ctx *c);
int launch_estimate_dt(nhw_ctx *c);
int launch_update_mask(nhw_ctx *c);
int launch_update_wind(nhw_ctx *c);
int launch_update_vars(nhw_ctx *c);
int launch_source_terms(nhw_ctx *c);
int launch_fluxes(nhw_ctx *c);
int launch_eval_explicit(nhw_ctx *c, int istep);
int launch_get_uvw(nhw_ctx *c);
int launch_projection(nhw_ctx *c);
int launch_rk_blend(nhw_ctx *c, int istep);
int launch_get_omega(nhw_ctx *c);
// boundary forcing (nhw_forcing.hip)
int forcing_setup(nhw_ctx *c);   // constants and device tables; after the grid
void forcing_free(nhw_ctx *c);
int launch_wave_bc(nhw_ctx *c);
int launch_sponge(nhw_ctx *c);
int launch_int_source(nhw_ctx *c);  // the internal wavemaker's SourceC (before source_terms' slopes)
int launch_flux_lr_history(nhw_ctx *c);  // FLUX_LR: Eta00/U00.. <- Eta0/U0.., U0.. <- U.. (update_vars)
int launch_east_ghosts(nhw_ctx *c);      // FLUX_LR: vel_bc's open east ghosts (after k_uvw_a)
int internal_wavemaker_setup(nhw_ctx *c, const double box[4], int nfreq, int ndir, const double *freq,
                             const double *dire, const double *spc, const double *phs);
constexpr int kViscKMax = 32;     // largest Kglob of the viscous branch (LDS of the column solve: 1 KB per level)
int viscous_setup(nhw_ctx *c);    // device arrays of the viscous branch; after the grid
void viscous_free(nhw_ctx *c);
int launch_diffusion(nhw_ctx *c);   // diffusion, at the start of eval_duvw
int launch_visc_solve(nhw_ctx *c);  // sigma_transform + momentum with implicit vertical solves
int eval_turb(nhw_ctx *c);
size_t viscous_state_bytes(const nhw_ctx *c);                        // checkpoint section
int viscous_state(nhw_ctx *c, std::vector<char> &buf, bool save);  // save: device -> buf
size_t flux_lr_state_bytes(const nhw_ctx *c);                      // checkpoint section (0: none)
int flux_lr_state(nhw_ctx *c, std::vector<char> &buf, bool save);  // save: device -> buf
int turb_zero(nhw_ctx *c);        // before initial's vel_bc: CmuHt = CmuVt = 0, as allocated
int turb_seed(nhw_ctx *c);        // after it: the seed viscosities (and Tke, Eps)
int turb_seed_depth(nhw_ctx *c);  // then DTke = D Tke Mask, DEps = D Eps Mask (k-epsilon)
// Poisson (nhw_poisson.hip)
int poisson_alloc(nhw_ctx *c);
void poisson_free(nhw_ctx *c);
int poisson_solve(nhw_ctx *c);
int poisson_system(nhw_ctx *c, double *coef_h, double *rhs_h);  // nhw_generate_coef_rhs
int poisson_apply(nhw_ctx *c, const double *x_h, double *y_h);  // nhw_poisson_apply
// matrix-free operator and multigrid (nhw_mg.hip); V is the vector type
int mg_alloc(nhw_ctx *c, std::vector<MGLevel> &lvs, int esize, int maxlev);
void mg_free(std::vector<MGLevel> &lvs);
int mg_upload_consts(nhw_ctx *c);  // per-level sigma constants, after nhw_set_grid
template <class V> int mg_setup(nhw_ctx *c, std::vector<MGLevel> &lvs);
template <class V> void mg_apply(nhw_ctx *c, const std::vector<MGLevel> &lvs, const V *b, V *x, const int *stop);
template <class V>
void mf_apply(nhw_ctx *c, const std::vector<MGLevel> &lvs, const V *x, V *y, const V *d1, int ndot, double *part,
              const int *stop, int nblk, int kid);
template <class R>
void mf_init(nhw_ctx *c, const std::vector<MGLevel> &lvs, const double *x, const double *b, R *r, R *rh, R *p, R *v,
             R *e, double *part, int nblk);
void mf_true_resid(nhw_ctx *c, const std::vector<MGLevel> &lvs, const double *x, const double *b, double *part,
                   int nblk);
}  // namespace nhw
