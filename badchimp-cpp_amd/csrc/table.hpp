// Viscosity-table lookup shared by the FP64 and reduced-precision kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace lbamd {

// std::upper_bound over the inner rows sr[1 .. n-2] of a table with its
// constant extension rows 0 and n-1 (libstdc++'s loop), returned as the
// index i with sr[i] <= v < sr[i+1] (LBrheology.h:459-460).
template <typename T>
__device__ __forceinline__ int table_index(const T* __restrict__ sr, int n, T v) {
  int first = 1, len = n - 2;
  while (len > 0) {
    const int half = len >> 1, mid = first + half;
    if (v < sr[mid]) {
      len = half;
    } else {
      first = mid + 1;
      len = len - half - 1;
    }
  }
  return first - 1;
}

}  // namespace lbamd
