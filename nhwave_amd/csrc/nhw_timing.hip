// Per-kernel launch timing with HIP events on the context's stream, and the
// algorithmic byte model of every kernel (what one launch must move through
// HBM at minimum: each persistent array element read or written once per
// pass, FP64 = 8 B, int mask = 4 B; halo re-reads and temporaries excluded).
#include "nhw_impl.h"

namespace nhw {

Timed::Timed(nhw_ctx *ctx, int kid, double nbytes) : c(ctx), id(kid), bytes(nbytes) {
  if (!c->kt.on) return;
  if (c->kt.pool.empty()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    c->kt.pool.push_back(e);
  }
  a = c->kt.pool.back();
  c->kt.pool.pop_back();
  (void)hipEventRecord(a, c->stream);
}

Timed::~Timed() {
  if (!a) return;
  hipEvent_t b;
  if (c->kt.pool.empty()) {
    if (hipEventCreate(&b) != hipSuccess) return;
  } else {
    b = c->kt.pool.back();
    c->kt.pool.pop_back();
  }
  (void)hipEventRecord(b, c->stream);
  if (bytes < 0.0) bytes = kernel_bytes(c, id);
  c->kt.pending.push_back({{id, bytes}, {a, b}});
  if (c->kt.pending.size() > 4096) {
    (void)hipStreamSynchronize(c->stream);
    ktimer_flush(c);
  }
}

void ktimer_flush(nhw_ctx *c) {
  for (auto &p : c->kt.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
      c->kt.total_ms[p.first.first] += ms;
      c->kt.total_bytes[p.first.first] += p.first.second;
      c->kt.launches[p.first.first] += 1;
    }
    c->kt.pool.push_back(p.second.first);
    c->kt.pool.push_back(p.second.second);
  }
  c->kt.pending.clear();
}

const char *kernel_name(int kid) {
  static const char *names[KID_COUNT] = {
      "k_flux_x", "k_flux_y", "k_flux_z", "k_eval_explicit", "k_uvw_a", "k_uvw_b", "k_projection",
      "k_rk_blend", "k_omega_x", "k_omega_y", "k_omega_col", "k_dt_partial", "k_mask", "k_source_terms",
      "k_assemble", "k_spmv_v", "k_spmv_t", "k_resid", "k_prec_p", "k_prec_s", "k_line_factor",
      "k_update", "k_reduce_final", "k_krylov_misc", "k_copy", "k_rhs", "k_mf_apply_v", "k_mf_apply_t",
      "k_mf_resid0", "k_mg_setup", "k_zebra", "k_resid_restrict", "k_coarse", "k_prolong_add", "k_krylov_vec",
      "k_wave_bc", "k_sponge", "k_diffusion", "k_visc_sigma", "k_visc_solve", "k_kepsilon"};
  return (kid >= 0 && kid < KID_COUNT) ? names[kid] : "?";
}

double kernel_bytes(const nhw_ctx *c, int kid) {
  const Geo &g = c->g;
  const double n = (double)g.M * g.N * g.K;          // interior cells
  const double n2 = (double)g.M * g.N;                // interior columns
  const double nfx = (double)(g.M + 1) * g.N * g.K;   // interior x faces
  const double nfy = (double)g.M * (g.N + 1) * g.K;
  const double nfz = (double)g.M * g.N * (g.K + 1);
  const double w = 8.0, mi = 4.0;
  const double w3 = c->mixed ? 4.0 : 8.0;  // 3-D fields in the storage type
  switch (kid) {
    case KID_FLUX_X:  // U,V,W,DU,DV,DW + Eta,D,Hc,Mask + Hfx,Hfx0 -> Ex,Fx,Gx,Hx
      return w3 * (6 * n + 4 * nfx) + w * (3 * n2 + 2 * (g.M + 1.0) * g.N) + mi * n2;
    case KID_FLUX_Y:
      return w3 * (6 * n + 4 * nfy) + w * (3 * n2 + 2 * g.M * (g.N + 1.0)) + mi * n2;
    case KID_FLUX_Z:  // U,V,W + Omega -> Fz,Gz,Hz
      return w3 * (3 * n + 4 * nfz);
    case KID_EXPLICIT:  // Ex..Hy, Fz..Hz, DU,DV,DW (r/w); D,D0,Hc,SourceX/Y,Mask; D,Eta,R1 written
      return w3 * (4 * nfx + 4 * nfy + 3 * nfz + 6 * n) + w * 9 * n2 + mi * n2;
    case KID_UVW_A:  // DU,DV,DW -> U,V,W; face states and KBC 2-D
      return w3 * 6 * n + w * 12 * n2;
    case KID_UVW_B:  // 2-D slopes of the face states + 4 ghost layers of 6 fields
      return w * (8 * n2 + 12 * n2) + w3 * 24 * n2 + mi * n2;
    case KID_PROJECTION:  // P, DU,DV,DW (r/w); D, DelxH/Eta, DelyH/Eta
      return w * n + w3 * 6 * n + w * 5 * n2 + mi * n2;
    case KID_RK_BLEND:  // DU0..DW0, DU..DW -> DU..DW
      return w3 * 9 * n + w * n2 + mi * n2;
    case KID_OMEGA_X:  // U, DU + Eta, D, Hfx -> Ex, EtaxL/R
      return w3 * (2 * n + nfx) + w * 5 * n2 + mi * n2;
    case KID_OMEGA_Y:
      return w3 * (2 * n + nfy) + w * 5 * n2 + mi * n2;
    case KID_OMEGA_COL:  // Ex, Ey, R1 -> Omega
      return w3 * (nfx + nfy + nfz) + w * n2 + mi * n2;
    case KID_DT:  // U,V,W + D
      return w3 * 3 * n + w * n2;
    case KID_MASK:
      return w * 4 * n2 + mi * 3 * n2;
    case KID_SOURCE:
      return w * 10 * n2 + mi * n2;
    case KID_ASSEMBLE:  // U,V,W + 2-D geometry -> 15 diagonals + rhs
      return w * (3 * n + 16 * n) + w * 5 * n2;
    case KID_SPMV_V:  // 15 diagonals + x + rh (dot) -> v
    case KID_SPMV_T:  // 15 diagonals + x + s (dot) -> t
    case KID_RESID:   // 15 diagonals + x + b -> r (+rh)
      return w * 18 * n;
    case KID_PREC_P:  // r, p, v, c13, lu_c, lu_d -> p, ph
      return w * 8 * n;
    case KID_PREC_S:  // r, v, c13, lu_c, lu_d -> s, sh
      return w * 7 * n;
    case KID_LINE_FACTOR:  // c1, c6, c13 -> lu_c, lu_d
      return w * 5 * n;
    case KID_KUPDATE:  // x, ph, sh, s, t, rh -> x, r
      return w * 8 * n;
    case KID_RHS:  // U,V,W + 2-D geometry -> rhs
      return w3 * 3 * n + w * n + w * 5 * n2;
    case KID_MF_V:  // ph, rh -> v; 5 column fields
    case KID_MF_T:  // sh, s -> t
      return w * 3 * n + w * 5 * n2;
    case KID_MF_RESID:  // x, b -> r, rh, p, v
      return w * 6 * n + w * 5 * n2;
    case KID_MG_SETUP:  // D, DelxH/Eta, DelyH/Eta -> 5 column fields (level 0)
      return w * 10 * n2;
    case KID_WAVE_BC:  // Hfx at the west face -> Ein, Din, Uin/Vin/Win per row
      return w * g.Nloc * (3.0 + 3.0 * g.Kloc);
    case KID_SPONGE: {  // every cell: Mask, Sponge, Hc, Eta (r/w), D; U, V (r/w), DU, DV
      const double all2 = (double)g.Mloc * g.Nloc;
      return w * 5 * all2 + mi * all2 + w3 * 6 * all2 * g.Kloc;
    }
    case KID_DIFFUSION:  // DU, DV, DW, DelxSc, DelySc, DelxSl, DelySl, Mask, D -> 6 stress terms
      return w * 13 * n + w * n2 + mi * n2;
    case KID_VISC_SIGMA:  // D, DelxH, DelyH, DelxEta, DelyEta -> DelxSc, DelySc, DelxSl, DelySl
      return w * 5 * (double)g.Mloc * g.Nloc + w * 4 * n;
    case KID_VISC_SOLVE:  // 11 face fluxes, DU, DV, DW (r/w), 6 stress terms, U, V at two levels; 2-D fields
      return w * (11 + 6 + 6) * n + w * 12 * n2 + mi * n2;
    default:
      return 0.0;
  }
}

__global__ __launch_bounds__(256) void k_copy(const double *__restrict__ a, double *__restrict__ b, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

}  // namespace nhw

// Reads and writes `nbytes` once per launch with a plain grid-stride copy:
// the calibration kernel for HBM byte counters (FETCH_SIZE/WRITE_SIZE).
extern "C" int nhw_bench_copy(int64_t nbytes, int32_t reps, double *avg_ms) {
  using namespace nhw;
  const long n = (long)(nbytes / 8);
  double *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  if (n <= 0 || reps <= 0) {
    set_error("nhw_bench_copy: need nbytes >= 8 and reps >= 1");
    return NHW_EINVAL;
  }
  if (hip_check(hipMalloc(&a, n * 8), "bench_copy") || hip_check(hipMalloc(&b, n * 8), "bench_copy")) {
    if (a) (void)hipFree(a);
    return NHW_ENOMEM;
  }
  (void)hipMemset(a, 0, n * 8);
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
  (void)hipEventRecord(e1, 0);
  int rc = hip_check(hipEventSynchronize(e1), "bench_copy");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (avg_ms) *avg_ms = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc ? NHW_EHIP : 0;
}
