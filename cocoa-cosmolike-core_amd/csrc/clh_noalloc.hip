// Allocation-free real-space 3x2pt data vector: clh_compute_3x2pt_real with
// every device buffer of the draw taken from a fixed array in device memory.
//
// `scratch` is a file-scope __device__ array of CLH_SCRATCH_BYTES.  Each GPU
// the library runs on holds its own copy until the process ends.  Host code
// never passes the array's name to a kernel or a copy (on the host it is a
// stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, taken once per device with that device current, and
// only that pointer is used.  Every call on one GPU shares that GPU's array
// and finds what the last call left there: the forward writes every word it
// later reads, so nothing has to start at zero.  Calls that overlap on one
// GPU (two host threads, or two contexts on different streams) race on the
// array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/cosmolike_hip.h"
#include "clh_internal.hpp"

#define CLH_SCRATCH_BYTES ((size_t)16 << 20)  // 16 MiB per GPU
#define CLH_SCRATCH_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[CLH_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[CLH_SCRATCH_MAX_DEVICES] = {nullptr};

// device address of `scratch` on `dev` (looked up on the first call there)
static hipError_t scratch_address(int dev, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

extern "C" int clh_compute_3x2pt_real_noalloc(clh_ctx* ctx, const int ord[3], double* dv, int n) {
  if (!ctx) return CLH_ERR_ARG;
  const int dev = clh_context_device(ctx);
  if (dev < 0 || dev >= CLH_SCRATCH_MAX_DEVICES) {
    clh_record_error(ctx, CLH_ERR_ARG, "allocation-free compute: device index out of range");
    return CLH_ERR_ARG;
  }
  void* p = nullptr;
  const hipError_t e = scratch_address(dev, &p);
  if (e != hipSuccess) {
    clh_record_error(ctx, CLH_ERR_HIP, hipGetErrorString(e));
    return CLH_ERR_HIP;
  }
  return clh_compute_3x2pt_real_scratch(ctx, ord, dv, n, p, CLH_SCRATCH_BYTES);
}
