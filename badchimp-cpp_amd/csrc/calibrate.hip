// Plain device copy used to calibrate bandwidth figures and HBM counters:
// the same rocprof counters read on the step kernels are compared with what
// they read for a copy of known size.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/badchimp_amd.h"
#include "errors.hpp"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// one 16-byte element per thread, streaming store: the fastest plain copy
// among the variants tried on MI355X (grid-stride loops were 25 % slower)
__global__ __launch_bounds__(256) void k_copy(const v4f* __restrict__ in, v4f* __restrict__ out,
                                              int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(in[i], &out[i]);
}

}  // namespace

extern "C" int lb_bench_copy(int device, int64_t bytes, int32_t iters, double* gbs, double* ms_per_iter) {
  using lbamd::fail_global;
  if (bytes < 16 || iters < 1) return fail_global(LB_ERR_ARG, "lb_bench_copy: bad argument");
  if (hipSetDevice(device) != hipSuccess) return fail_global(LB_ERR_HIP, "hipSetDevice failed");
  const int64_t n = bytes / 16;
  v4f *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess) return fail_global(LB_ERR_NOMEM, "hipMalloc failed");
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    hipFree(a);
    return fail_global(LB_ERR_NOMEM, "hipMalloc failed");
  }
  hipMemset(a, 0, n * 16);
  hipStream_t s;
  hipStreamCreate(&s);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const int64_t blocks = (n + 255) / 256;
  k_copy<<<dim3(static_cast<unsigned>(blocks)), 256, 0, s>>>(a, b, n);
  hipEventRecord(e0, s);
  for (int32_t k = 0; k < iters; ++k) k_copy<<<dim3(static_cast<unsigned>(blocks)), 256, 0, s>>>(a, b, n);
  hipEventRecord(e1, s);
  hipError_t err = hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamDestroy(s);
  hipFree(a);
  hipFree(b);
  if (err != hipSuccess) return fail_global(LB_ERR_HIP, "copy kernel failed: %s", hipGetErrorString(err));
  const double per = ms / iters;
  if (ms_per_iter) *ms_per_iter = per;
  if (gbs) *gbs = 2.0 * n * 16 / (per * 1e-3) / 1e9;
  return LB_OK;
}
