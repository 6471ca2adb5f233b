// Device side of the BiCGStab iteration shared by the single-scenario and
// batched Poisson solves: scalar slots, deterministic reductions and the
// per-iteration scalar logic (what the single-thread "fin" steps decide).
#pragma once
#include "nhw_impl.h"

namespace nhw {

constexpr int kBS = 256;  // threads per block of the Krylov kernels
constexpr double kInnerTol = 1e-4;  // relative residual reduction of one FP32 inner solve
constexpr int kMaxRefine = 20;      // refinement rounds of the mixed form

enum { S_RHO = 0, S_ALPHA, S_OMEGA, S_BETA, S_BB, S_RR, S_RHV, S_TS, S_TT, S_TRUE, S_REF, S_ITOL, S_N };

// scalars and flags of one solve.  flag[0]: 0 running, 1 converged,
// 2 breakdown, 3 ITMAX reached, 4 zero right-hand side; flag[1]: iterations;
// flag[2]: breakdown code.  `inner` marks the mixed form's FP32 inner solve,
// which stops on |r| <= itol*|r0| instead of |r| <= tol*|b|.
struct KScal {
  double *sc;
  int *flag;
  int itmax;
  double tol2, itol2;
  int inner;
};

// deterministic block reduction of up to 3 values into part[3*blockIdx+m]
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double *part) {
  __shared__ double sm[NV][kBS];
#pragma unroll
  for (int m = 0; m < NV; ++m) sm[m][threadIdx.x] = v[m];
  __syncthreads();
  for (int s = kBS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int m = 0; m < NV; ++m) sm[m][threadIdx.x] += sm[m][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int m = 0; m < NV; ++m) part[3 * blockIdx.x + m] = sm[m][0];
}

// sum of the partials m of blocks b0 .. b1-1 (fixed order: deterministic)
__device__ __forceinline__ double final_sum(const double *part, int m, long b0, long b1) {
  __shared__ double sm[kBS];
  double a = 0.0;
  for (long b = b0 + threadIdx.x; b < b1; b += kBS) a += part[3 * b + m];
  sm[threadIdx.x] = a;
  __syncthreads();
  for (int s = kBS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  double out = sm[0];
  __syncthreads();
  return out;
}

// start of a solve: rr = |r0|^2, bb = |b|^2
__device__ __forceinline__ void fin0_logic(const KScal &k, double rr, double bb) {
  k.sc[S_RR] = rr; k.sc[S_BB] = bb; k.sc[S_RHO] = rr; k.sc[S_TRUE] = rr;
  k.sc[S_REF] = k.inner ? rr : bb;
  // an inner solve reduces its residual by kInnerTol, or only as far as the
  // outer test needs (|r| to TOL|b|, with a factor 2 margin), whichever is
  // less work
  k.sc[S_ITOL] = fmax(k.itol2, 0.25 * k.tol2 * bb / rr);
  k.sc[S_ALPHA] = 1.0; k.sc[S_OMEGA] = 1.0; k.sc[S_BETA] = 0.0;
  k.flag[1] = 0; k.flag[2] = 0;
  if (bb == 0.0) k.flag[0] = 4;                       // zero right-hand side: x = 0
  else if (!(rr == rr)) { k.flag[0] = 2; k.flag[2] = 1; }
  else if (rr <= k.tol2 * bb) k.flag[0] = 1;          // already converged
  else k.flag[0] = 0;
}

// alpha = rho / (rh, v)
__device__ __forceinline__ void fin1_logic(const KScal &k, double rhv) {
  k.sc[S_RHV] = rhv;
  if (rhv == 0.0 || !(rhv == rhv)) { k.flag[0] = 2; k.flag[2] = 2; return; }
  k.sc[S_ALPHA] = k.sc[S_RHO] / rhv;
}

// omega = (t, s) / (t, t)
__device__ __forceinline__ void fin2_logic(const KScal &k, double ts, double tt) {
  k.sc[S_TS] = ts; k.sc[S_TT] = tt;
  // tt == 0 means s == 0: the half step already solves the system
  k.sc[S_OMEGA] = tt == 0.0 ? 0.0 : ts / tt;
  if (!(tt == tt)) { k.flag[0] = 2; k.flag[2] = 3; }
}

// convergence test and beta after the update; rhr = (rh, r), rr = |r|^2
__device__ __forceinline__ void fin3_logic(const KScal &k, double rhr, double rr) {
  const int it = k.flag[1] + 1;
  k.flag[1] = it;
  k.sc[S_RR] = rr;
  const double omega = k.sc[S_OMEGA];
  if (!(rr == rr)) { k.flag[0] = 2; k.flag[2] = 4; return; }
  if (rr <= (k.inner ? k.sc[S_ITOL] : k.tol2) * k.sc[S_REF]) { k.flag[0] = 1; return; }
  if (it >= k.itmax) { k.flag[0] = 3; return; }
  if (rhr == 0.0 || omega == 0.0) { k.flag[0] = 2; k.flag[2] = 5; return; }
  k.sc[S_BETA] = (rhr / k.sc[S_RHO]) * (k.sc[S_ALPHA] / omega);
  k.sc[S_RHO] = rhr;
}

}  // namespace nhw
