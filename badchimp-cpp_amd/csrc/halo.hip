// Halo exchange between subdomains: the populations a rank sends to a
// neighbouring rank and the ones it receives from it, the device side of
// MonLatMpi::communicateLbField (LBmonlatmpi.h:236-300).
//
// The reference pushes: after propagation, the values a rank wrote into its
// ghost nodes (another rank's nodes) travel to their owner, which stores them
// into its own nodes.  Here the step pulls, so the same values move the other
// way round in memory: the sender reads its own post-collision value of
// direction q at a boundary node (element pidx(q, i), which the reference
// propagates into the ghost neighbour(q, n)), and the receiver writes it into
// the slot of its ghost node, from which the next gather of the bulk node
// neighbour(q, ghost) reads it.  Element lists come from the host (one
// element index per value).
//
// The values travel as doubles holding what the context stores: f in FP64,
// f - w_q in FP32S and FP16S (the two-phase forms: t and d), so a pack and an
// unpack between two contexts of one precision move the stored bits exactly.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "../../include/badchimp_amd.h"
#include "context.hpp"
#include "kernels_f32.hpp"

namespace lbamd {
namespace {

__device__ __forceinline__ double to_double(double v) { return v; }
__device__ __forceinline__ double to_double(float v) { return static_cast<double>(v); }
__device__ __forceinline__ double to_double(uint16_t v) {
  return static_cast<double>(__half2float(__ushort_as_half(v)) * (1.0f / kHalfScale));
}
__device__ __forceinline__ void from_double(double d, double& v) { v = d; }
__device__ __forceinline__ void from_double(double d, float& v) { v = static_cast<float>(d); }
__device__ __forceinline__ void from_double(double d, uint16_t& v) {
  v = __half_as_ushort(__float2half_rn(static_cast<float>(d) * kHalfScale));
}

// out[k n + j] = field k of element elem[j] (fields of an element adjacent:
// nf_total e + k)
template <typename S>
__global__ __launch_bounds__(kBlock) void k_halo_pack(const S* __restrict__ f, int nf_total, int nf,
                                                      const uint32_t* __restrict__ elem, int64_t n,
                                                      double* __restrict__ out) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (j >= n) return;
  const int64_t e = static_cast<int64_t>(elem[j]) * nf_total;
  for (int k = 0; k < nf; ++k) out[k * n + j] = to_double(f[e + k]);
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_halo_unpack(S* __restrict__ f, int nf_total, int nf,
                                                        const uint32_t* __restrict__ elem, int64_t n,
                                                        const double* __restrict__ in) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (j >= n) return;
  const int64_t e = static_cast<int64_t>(elem[j]) * nf_total;
  for (int k = 0; k < nf; ++k) from_double(in[k * n + j], f[e + k]);
}

template <class F>
hipError_t by_precision(int precision, F&& fn) {
  switch (precision) {
    case LB_FP64: fn(double{}); break;
    case LB_FP32S: fn(float{}); break;
    case LB_FP16S: fn(uint16_t{}); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

dim3 grid_of(int64_t n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

}  // namespace

hipError_t halo_pack(int precision, int nf_total, int nf, const void* f, const uint32_t* elem, int64_t n,
                     double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return by_precision(precision, [&](auto t) {
    using S = decltype(t);
    k_halo_pack<S><<<grid_of(n), dim3(kBlock), 0, s>>>(static_cast<const S*>(f), nf_total, nf, elem, n, out);
  });
}

hipError_t halo_unpack(int precision, int nf_total, int nf, void* f, const uint32_t* elem, int64_t n,
                       const double* in, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return by_precision(precision, [&](auto t) {
    using S = decltype(t);
    k_halo_unpack<S><<<grid_of(n), dim3(kBlock), 0, s>>>(static_cast<S*>(f), nf_total, nf, elem, n, in);
  });
}

}  // namespace lbamd
