// nhw_snapshot_static: the snapshot of nhw_snapshot_begin + nhw_snapshot_wait
// for callers that cannot allocate (code between set-up and tear-down of a
// serving loop, harnesses that forbid allocation).  nhw_snapshot_begin stages
// the chosen fields in device and pinned host buffers that it allocates, and
// grows, on demand; this variant stages them in one fixed device array of
// this file, `scratch`, and copies from there straight into the caller's
// arrays.  It calls no allocating routine: no hipMalloc/hipFree, no pinned
// host memory, no stream or event creation, no host heap.
//
// Every call on one GPU uses that GPU's copy of `scratch`; calls that overlap
// on one GPU (two host threads, or two contexts whose streams run at the same
// time) race on it and must be serialised by the caller.
#include <mutex>
#include <string>

#include "nhw_impl.h"

namespace nhw {

// Largest snapshot nhw_snapshot_static takes, in FP64/int32 host bytes
// (128 MiB: every 2-D field of a 2048 x 2048 grid, or one FP64 3-D field of
// 1024 x 1024 x 10).  Each GPU the library runs on holds this array until the
// process ends.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

// Device address of this GPU's `scratch`, looked up once per device with that
// device current (the host-side name of the array is not a device address).
static int scratch_on(int dev, unsigned char **out) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("nhw_snapshot_static: device ordinal out of range");
    return NHW_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    if (hip_check(hipGetDevice(&cur), "nhw_snapshot_static: hipGetDevice")) return NHW_EHIP;
    if (cur != dev && hip_check(hipSetDevice(dev), "nhw_snapshot_static: hipSetDevice")) return NHW_EHIP;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (hip_check(e, "nhw_snapshot_static: hipGetSymbolAddress(scratch)")) return NHW_EHIP;
    g_scratch_ptr[dev] = (unsigned char *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// one field into the staging array: FP32 storage widened to FP64, or a
// plain copy of 8-byte (or, for the int32 masks, 4-byte) words
__global__ void k_stage_widen(const float *__restrict__ src, double *__restrict__ dst, long n) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
    dst[q] = (double)src[q];
}
__global__ void k_stage_copy8(const unsigned long long *__restrict__ src, unsigned long long *__restrict__ dst,
                              long n) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
    dst[q] = src[q];
}
__global__ void k_stage_copy4(const unsigned *__restrict__ src, unsigned *__restrict__ dst, long n) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
    dst[q] = src[q];
}

static dim3 stage_grid(long n) {
  const long b = (n + 255) / 256;
  return dim3((unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096));
}

}  // namespace nhw

using namespace nhw;

extern "C" {

int nhw_snapshot_static(nhw_ctx *c, const nhw_fields *host) {
  if (!c || !host) {
    set_error("nhw_snapshot_static: null argument");
    return NHW_EINVAL;
  }
  Xfer tab[kFieldSlots];
  field_table_into(c, host, tab);
  if (int rc = check_present(tab, kFieldSlots, "nhw_snapshot_static")) return rc;
  // host bytes of the request (3-D fields of the mixed form widen to FP64),
  // each field's staging slot aligned to 256 bytes
  size_t total = 0;
  for (const Xfer &x : tab)
    if (x.host) total += (x.bytes + 255) / 256 * 256;
  if (total == 0) return 0;
  if (total > kScratchBytes) {
    set_error("nhw_snapshot_static: the requested fields need " + std::to_string(total) +
              " bytes of staging, more than the " + std::to_string(kScratchBytes) +
              " bytes this entry point has; request fewer fields or use nhw_snapshot_begin");
    return NHW_EINVAL;
  }
  unsigned char *sp = nullptr;
  int rc = scratch_on(c->device, &sp);
  if (rc) return rc;
  // stage every field in stream order (the state after the work queued so
  // far), then copy the staged images into the caller's arrays
  size_t off = 0;
  for (const Xfer &x : tab) {
    if (!x.host) continue;
    unsigned char *dst = sp + off;
    if (x.r3 && c->mixed) {
      const long n = (long)(x.bytes / sizeof(double));
      hipLaunchKernelGGL(k_stage_widen, stage_grid(n), dim3(256), 0, c->stream, (const float *)x.dev, (double *)dst,
                         n);
    } else if (x.bytes % 8 == 0) {
      const long n = (long)(x.bytes / 8);
      hipLaunchKernelGGL(k_stage_copy8, stage_grid(n), dim3(256), 0, c->stream,
                         (const unsigned long long *)x.dev, (unsigned long long *)dst, n);
    } else {
      const long n = (long)(x.bytes / 4);
      hipLaunchKernelGGL(k_stage_copy4, stage_grid(n), dim3(256), 0, c->stream, (const unsigned *)x.dev,
                         (unsigned *)dst, n);
    }
    off += (x.bytes + 255) / 256 * 256;
  }
  if (hip_check(hipGetLastError(), "nhw_snapshot_static: staging")) return NHW_EHIP;
  off = 0;
  for (const Xfer &x : tab) {
    if (!x.host) continue;
    if (hip_check(hipMemcpyAsync(x.host, sp + off, x.bytes, hipMemcpyDeviceToHost, c->stream),
                  "nhw_snapshot_static: copy to host"))
      return NHW_EHIP;
    off += (x.bytes + 255) / 256 * 256;
  }
  return hip_check(hipStreamSynchronize(c->stream), "nhw_snapshot_static");
}

}  // extern "C"
