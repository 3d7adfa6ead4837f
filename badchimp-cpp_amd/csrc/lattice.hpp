// Lattice tables and per-node moment helpers for D2Q9 and D3Q19.
//
// Direction order, weights, reverse map and the B coefficients follow the
// reference's generated lattice headers (src/lbsolver/LBd2q9.h:14-40,
// src/lbsolver/LBd3q19.h:14-40): the rest direction is stored last, and the
// constants are formed by the same double expressions (e.g. w1 = 2.0/36.0,
// w1c2Inv = w1*c2Inv, c2 = 1.0/c2Inv) so that they carry the same bits.
//
// The helpers reproduce the reference's summation order term by term, which
// is what makes FP64 results bit-identical with -ffp-contract=off:
//   qsum      left fold over q                  (LBd3q19.h:173-180)
//   qsumc     signed sum over the directions with c_qd != 0, ascending q,
//             first term positive               (LBd3q19.h:182-190, LBd2q9.h:170-177)
//   cdot      +-v_a +- v_b in dimension order   (LBd3q19.h:129-153, LBd2q9.h:128-142)
//   grad      w1c2Inv*(axis terms) + w2c2Inv*(diagonal terms), ascending q
//                                               (LBd3q19.h:155-163, LBd2q9.h:144-151)
//   dot       a0*b0 + a1*b1 (+ a2*b2)           (LBd3q19.h:111-115)
// The loops below are fully unrolled at compile time; every operation is a
// separate IEEE add/sub/mul, in the order given.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define LB_HD __host__ __device__ __forceinline__
#else
#define LB_HD inline
#endif

namespace lbamd {

struct D2Q9 {
  static constexpr int nD = 2;
  static constexpr int nQ = 9;
  static constexpr int nDirPairs = 4;
  static constexpr double c2Inv = 3.0;
  static constexpr double c4Inv = 9.0;
  static constexpr double c2 = 1.0 / c2Inv;
  static constexpr double c4Inv0_5 = 0.5 * c4Inv;
  static constexpr double w0 = 16.0 / 36.0;
  static constexpr double w1 = 4.0 / 36.0;
  static constexpr double w2 = 1.0 / 36.0;
  static constexpr double w1c2Inv = w1 * c2Inv;
  static constexpr double w2c2Inv = w2 * c2Inv;
  LB_HD static constexpr int c(int q, int d) {
    constexpr int t[18] = {1, 0, 1, 1, 0, 1, -1, 1, -1, 0, -1, -1, 0, -1, 1, -1, 0, 0};
    return t[2 * q + d];
  }
  LB_HD static constexpr double w(int q) {
    return q == 8 ? w0 : ((q & 1) ? w2 : w1);
  }
  LB_HD static constexpr int rev(int q) { return q == 8 ? 8 : (q + 4) % 8; }
  LB_HD static constexpr double B(int q) {
    return q == 8 ? -16.0 / 108.0 : ((q & 1) ? 5.0 / 108.0 : 8.0 / 108.0);
  }
  LB_HD static constexpr double cNorm(int q) {
    return q == 8 ? 0.0 : ((q & 1) ? 1.4142135623730950488 : 1.0);
  }
};

struct D3Q19 {
  static constexpr int nD = 3;
  static constexpr int nQ = 19;
  static constexpr int nDirPairs = 9;
  static constexpr double c2Inv = 3.0;
  static constexpr double c4Inv = 9.0;
  static constexpr double c2 = 1.0 / c2Inv;
  static constexpr double c4Inv0_5 = 0.5 * c4Inv;
  static constexpr double w0 = 12.0 / 36.0;
  static constexpr double w1 = 2.0 / 36.0;
  static constexpr double w2 = 1.0 / 36.0;
  static constexpr double w1c2Inv = w1 * c2Inv;
  static constexpr double w2c2Inv = w2 * c2Inv;
  LB_HD static constexpr int c(int q, int d) {
    constexpr int t[57] = {1, 0, 0,  0, 1, 0,  0, 0, 1,  1, 1, 0,  1, -1, 0,
                           1, 0, 1,  1, 0, -1, 0, 1, 1,  0, 1, -1, -1, 0, 0,
                           0, -1, 0, 0, 0, -1, -1, -1, 0, -1, 1, 0,  -1, 0, -1,
                           -1, 0, 1, 0, -1, -1, 0, -1, 1, 0, 0, 0};
    return t[3 * q + d];
  }
  LB_HD static constexpr bool axis(int q) { return (q % 9) < 3; }
  LB_HD static constexpr double w(int q) {
    return q == 18 ? w0 : (axis(q) ? w1 : w2);
  }
  LB_HD static constexpr int rev(int q) { return q == 18 ? 18 : (q + 9) % 18; }
  LB_HD static constexpr double B(int q) {
    return q == 18 ? -12.0 / 54.0 : (axis(q) ? 1.0 / 54.0 : 2.0 / 54.0);
  }
  LB_HD static constexpr double cNorm(int q) {
    return q == 18 ? 0.0 : (axis(q) ? 1.0 : 1.4142135623730950488);
  }
};

// ---------------------------------------------------------------- helpers
template <int N, typename F>
LB_HD constexpr void unroll(F&& f) {
  if constexpr (N > 0) {
    unroll<N - 1>(f);
    f(N - 1);
  }
}

// Sum of f over all q, left fold starting from 0.0.
template <class L, typename T>
LB_HD T qsum(const T* f) {
  T r = T(0);
#pragma unroll
  for (int q = 0; q < L::nQ; ++q) r += f[q];
  return r;
}

// Signed sum sum_q c_qd f_q in the reference's term order.
template <class L, typename T>
LB_HD T qsumc(const T* f, int d) {
  T r = T(0);
  bool first = true;
#pragma unroll
  for (int q = 0; q < L::nQ; ++q) {
    const int cq = L::c(q, d);
    if (cq == 0) continue;
    if (first) {
      r = cq > 0 ? f[q] : -f[q];
      first = false;
    } else {
      r = cq > 0 ? r + f[q] : r - f[q];
    }
  }
  return r;
}

// c_q . v in the reference's term order; 0 for the rest direction.
template <class L, typename T>
LB_HD T cdot(int q, const T* v) {
  T r = T(0);
  bool first = true;
#pragma unroll
  for (int d = 0; d < L::nD; ++d) {
    const int cq = L::c(q, d);
    if (cq == 0) continue;
    if (first) {
      r = cq > 0 ? v[d] : -v[d];
      first = false;
    } else {
      r = cq > 0 ? r + v[d] : r - v[d];
    }
  }
  return r;
}

template <class L, typename T>
LB_HD T dot(const T* a, const T* b) {
  T r = a[0] * b[0];
#pragma unroll
  for (int d = 1; d < L::nD; ++d) r = r + a[d] * b[d];
  return r;
}

// Lattice gradient of a scalar sampled at the Q neighbours, component d.
template <class L, typename T>
LB_HD T grad(const T* s, int d) {
  T ax = T(0), dg = T(0);
  bool fa = true, fd = true;
#pragma unroll
  for (int q = 0; q < L::nQ - 1; ++q) {
    const int cq = L::c(q, d);
    if (cq == 0) continue;
    int n2 = 0;
#pragma unroll
    for (int e = 0; e < L::nD; ++e) n2 += L::c(q, e) * L::c(q, e);
    T& acc = (n2 == 1) ? ax : dg;
    bool& first = (n2 == 1) ? fa : fd;
    if (first) {
      acc = cq > 0 ? s[q] : -s[q];
      first = false;
    } else {
      acc = cq > 0 ? acc + s[q] : acc - s[q];
    }
  }
  return T(L::w1c2Inv) * ax + T(L::w2c2Inv) * dg;
}

}  // namespace lbamd
