// Plain device copy used to calibrate bandwidth figures and HBM counters:
// the same rocprof counters read on the step kernels are compared with what
// they read for a copy of known size.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/badchimp_amd.h"
#include "context.hpp"
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

namespace {

// Identity streaming of Q-direction AoSoA populations (the collide-stream's
// data movement with neither gathers nor collision): one node per thread,
// Q float loads and Q streaming float stores at pidx(q, i).  Its rate is the
// ceiling of the population layout itself, below a plain copy.
template <int Q>
__global__ __launch_bounds__(256) void k_stream_aosoa(const float* __restrict__ in, float* __restrict__ out,
                                                      int32_t n) {
  const int32_t i = static_cast<int32_t>(blockIdx.x) * 256 + static_cast<int32_t>(threadIdx.x);
  if (i >= n) return;
  float g[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) g[q] = in[lbamd::pidx<Q>(q, static_cast<uint32_t>(i))];
  float* __restrict__ o = out + lbamd::pidx<Q>(0, static_cast<uint32_t>(i));
#pragma unroll
  for (int q = 0; q < Q; ++q) __builtin_nontemporal_store(g[q], o + (q << lbamd::kPLog));
}

}  // namespace

extern "C" int lb_bench_stream(int device, int32_t lattice, int64_t n_nodes, int32_t iters, double* gbs,
                               double* ms_per_iter) {
  using lbamd::fail_global;
  if (n_nodes < 256 || iters < 1 || (lattice != LB_D2Q9 && lattice != LB_D3Q19) || n_nodes > (1 << 26))
    return fail_global(LB_ERR_ARG, "lb_bench_stream: bad argument");
  if (hipSetDevice(device) != hipSuccess) return fail_global(LB_ERR_HIP, "hipSetDevice failed");
  const int Q = lattice == LB_D2Q9 ? 9 : 19;
  const int64_t NP = (n_nodes + 255) / 256 * 256;
  const size_t bytes = static_cast<size_t>(Q) * NP * sizeof(float);
  float *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return fail_global(LB_ERR_NOMEM, "hipMalloc failed");
  if (hipMalloc(&b, bytes) != hipSuccess) {
    hipFree(a);
    return fail_global(LB_ERR_NOMEM, "hipMalloc failed");
  }
  hipMemset(a, 0, bytes);
  hipMemset(b, 0, bytes);
  hipStream_t s;
  hipStreamCreate(&s);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const dim3 grid(static_cast<unsigned>(NP / 256));
  const int32_t n = static_cast<int32_t>(NP);
  auto launch = [&](const float* x, float* y) {
    if (Q == 9) k_stream_aosoa<9><<<grid, 256, 0, s>>>(x, y, n);
    else k_stream_aosoa<19><<<grid, 256, 0, s>>>(x, y, n);
  };
  launch(a, b);
  hipEventRecord(e0, s);
  for (int32_t k = 0; k < iters; ++k) launch(k & 1 ? b : a, k & 1 ? a : b);
  hipEventRecord(e1, s);
  hipError_t err = hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamDestroy(s);
  hipFree(a);
  hipFree(b);
  if (err != hipSuccess) return fail_global(LB_ERR_HIP, "stream kernel failed: %s", hipGetErrorString(err));
  const double per = ms / iters;
  if (ms_per_iter) *ms_per_iter = per;
  if (gbs) *gbs = 2.0 * static_cast<double>(bytes) / (per * 1e-3) / 1e9;
  return LB_OK;
}
