// Reduced-precision BGK + Guo collide-stream on shifted populations.
//
// With g_q = f_q - w_q (reference density 1) the moments are
//   rho = 1 + sum_q g_q,   j = sum_q g_q c_q   (sum_q w_q c_q = 0)
// and the BGK relaxation of the reference (LBcollision.h:41-44) becomes
//   g_q' = g_q - (g_q - geq_q)/tau + dF_q,
//   geq_q = f_eq,q - w_q = w_q (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2)),
// with the Guo term dF_q of LBcollision.h:206-211 unchanged.  Subtracting
// w_q before rounding keeps the float (or half) mantissa for the deviation
// from rest, which is what the flow carries.  Streaming and bounce-back use
// the same address layouts as the FP64 path.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels_f32.hpp"
#include "lattice.hpp"
#include "table.hpp"

namespace lbamd {
namespace {

template <typename S>
struct Store;
template <>
struct Store<float> {
  static __device__ __forceinline__ float get(float v) { return v; }
  static __device__ __forceinline__ float put(float v) { return v; }
  static __device__ __forceinline__ float put_sr(float v, uint32_t) { return v; }
};
// two halves in one register (the operand type of the half conversions)
typedef __fp16 HalfPair __attribute__((ext_vector_type(2)));

// half values travel as their 16-bit patterns
template <>
struct Store<uint16_t> {
  static __device__ __forceinline__ float get(uint16_t v) {
    return __half2float(__ushort_as_half(v)) * (1.0f / kHalfScale);
  }
  static __device__ __forceinline__ uint16_t put(float v) {
    return __half_as_ushort(__float2half_rn(v * kHalfScale));
  }
  // Stochastic rounding: the half below |v| (towards zero) or the next one
  // away from zero, the latter with probability equal to the distance from
  // the lower one in units of their gap, drawn from the uniform word r.
  // Round to nearest freezes a relaxation once its increments drop below
  // half a unit of the last place, which stops a steady-state solve early
  // with a biased result; stochastic rounding applies such increments on
  // average.  gfx950 converts with stochastic rounding in one instruction
  // (v_cvt_sr_f16_f32).  Its behaviour on this hardware: for every
  // normal-range half the result is one of those two neighbours and the mean
  // over all 13-bit random patterns equals the float exactly; zero, inf,
  // NaN and overflow convert as with round to nearest.
  static __device__ __forceinline__ uint16_t put_sr(float v, uint32_t r) {
    HalfPair o = __builtin_bit_cast(HalfPair, r);  // the high half is not used
    o = __builtin_amdgcn_cvt_sr_f16_f32(o, v * kHalfScale, r, false);
    return static_cast<uint16_t>(__builtin_bit_cast(uint32_t, o));
  }
};

// The two stored values (t, d) of one two-phase element, adjacent in memory
// (field k of element e at 2 e + k): one 8-byte (float) or 4-byte (half)
// access moves both.
template <typename S>
struct PairOf;
template <>
struct PairOf<float> {
  typedef float T __attribute__((ext_vector_type(2)));
};
template <>
struct PairOf<uint16_t> {
  typedef uint16_t T __attribute__((ext_vector_type(2)));
};
template <typename S>
using Pair = typename PairOf<S>::T;

// Random words of the stochastic rounding: two hashes of the node (its index
// in the solve) and the solve's own step count, then one word per stored
// value k: a rotation of the first hash, xor the second, plus a per-k
// constant.  Each word is uniform; the rotations differ between the values
// of one node.  The same solve therefore rounds alike alone, in a batch or
// in chunks.
struct SrSeed {
  uint32_t a, b;
};
__device__ __forceinline__ SrSeed sr_node(int32_t i, uint32_t step) {
  uint32_t x = (static_cast<uint32_t>(i) * 0x9E3779B9u) ^ (step * 0x85EBCA6Bu);
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  uint32_t y = (x ^ 0x5bd1e995u) * 0x2C1B3C6Du;
  y ^= y >> 15;
  return SrSeed{x, y};
}
// k is a constant once the loops over directions are unrolled
__device__ __forceinline__ uint32_t sr_bits(SrSeed h, uint32_t k) {
  const uint32_t s = (11u * k + (k >> 5) * 5u) & 31u;  // distinct for the values of one node
  return (__builtin_amdgcn_alignbit(h.a, h.a, s) ^ h.b) + k * 0x9E3779B9u;
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) r += red[w];
  }
  __syncthreads();
  return r;
}

// block_sum for each 256-thread slice of a workgroup of NS slices: the same
// shuffles and the same order over the slice's four waves, so a slice sums
// exactly as a 256-thread workgroup does.  Result valid in the slice's
// first thread.
template <int NS>
__device__ __forceinline__ double slice_sum(double v) {
  __shared__ double red[NS * (kBlock / 64)];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = 0.0;
  if ((threadIdx.x % kBlock) == 0) {
    const int w0 = (threadIdx.x / kBlock) * (kBlock / 64);
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) r += red[w0 + w];
  }
  __syncthreads();
  return r;
}

template <class L, typename S, bool FIRST, class A>
__device__ __forceinline__ void load_g(const S* __restrict__ gin, const A& ad, int32_t i, int64_t NP,
                                       float (&g)[L::nQ]) {
  constexpr bool O = A::kOff32;
  const uint32_t np = static_cast<uint32_t>(NP);
  if (FIRST) {
#pragma unroll
    for (int q = 0; q < L::nQ - 1; ++q) g[q] = Store<S>::get(ld_elem<O>(gin, pidx<L::nQ>(q, i)));
  } else {
    uint32_t src[L::nQ - 1];
    ad.template sources<L>(i, src);
    if constexpr (sizeof(S) == 2) {
      // 2-byte gathers as loads of the aligned 4-byte word that holds the
      // value (the word stays inside the buffer, whose size is a multiple of
      // 4 bytes), then the half it sits in: the same bits, 2 % faster on the
      // 256^3 pack (0.154 against 0.157 ms per FP16S step)
      const uint32_t* w = reinterpret_cast<const uint32_t*>(gin);
#pragma unroll
      for (int q = 0; q < L::nQ - 1; ++q) {
        const uint32_t word = ld_elem<O>(w, src[q] >> 1);
        g[q] = Store<S>::get(static_cast<S>(word >> ((src[q] & 1u) << 4)));
      }
    } else {
#pragma unroll
      for (int q = 0; q < L::nQ - 1; ++q) g[q] = Store<S>::get(ld_elem<O>(gin, src[q]));
    }
  }
  g[L::nQ - 1] = Store<S>::get(ld_elem<O>(gin, pidx<L::nQ>(L::nQ - 1, i)));
}

// Viscosity of a node from its strain rate (GeneralizedNewtonian,
// LBrheology.h:436-462, source 0) on shifted populations:
// f - feq = g - w (a0 + 4.5 rho cu^2 + 3 rho cu).
template <class L>
__device__ __forceinline__ float rheology_visc_lp(const float (&g)[L::nQ], float a0, float r45, float r3,
                                                  const float* u, const float (&F)[3], const float* __restrict__ sr,
                                                  const float* __restrict__ nu, int n) {
  constexpr int Q = L::nQ, D = L::nD;
  float E[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) E[i][j] = 0.0f;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const float cu = cdot<L>(q, u);
    const float neq = g[q] - float(L::w(q)) * (a0 + r45 * cu * cu + r3 * cu);
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const int ci = L::c(q, i);
      E[i][i] += neq * (float(ci * ci) - float(L::c2));
#pragma unroll
      for (int j = i + 1; j < D; ++j) {
        const int cc = ci * L::c(q, j);
        if (cc > 0) E[i][j] += neq;
        else if (cc < 0) E[i][j] -= neq;
      }
    }
  }
  float sq = 0.0f;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    E[i][i] += u[i] * F[i];
    sq += E[i][i] * E[i][i];
#pragma unroll
    for (int j = i + 1; j < D; ++j) {
      E[i][j] += 0.5f * (u[i] * F[j] + u[j] * F[i]);
      sq += 2.0f * E[i][j] * E[i][j];
    }
  }
  const int k = table_index(sr, n, sq);
  return (nu[k + 1] - nu[k]) * (sq - sr[k]) / (sr[k + 1] - sr[k]) + nu[k];
}

// One node of the BGK + Guo step on shifted populations; returns j_x (FP64)
// for the permeability sum when REDUCE.  MODE kTrt: the two-relaxation-time
// operator (LBcollision.h:50-75, 214-233) with 1/tauAnti = tauA_inv and
// tauAnti_factor = tauA_factor; tau_inv, tau_factor are then the symmetric
// ones.  MODE kRheology: tau_inv, tau_factor from the node's viscosity
// (table sr/nu of n rows), stored beside rho (rho_o[NP + i]) when STORE.
template <class L, typename S, class A, bool FIRST, bool STORE, bool REDUCE, int MODE = kBgk>
__device__ __forceinline__ double bgk_node_lp(const S* __restrict__ gin, S* __restrict__ gout, const A& ad,
                                              int32_t i, int64_t NP, float tau_inv, float tau_factor,
                                              const float (&F)[3], double* __restrict__ rho_o,
                                              double* __restrict__ vel_o, float tauA_inv = 0.0f,
                                              float tauA_factor = 0.0f, const float* __restrict__ sr = nullptr,
                                              const float* __restrict__ nu = nullptr, int tab_n = 0,
                                              uint32_t step = 0) {
  constexpr int Q = L::nQ, D = L::nD;
  const SrSeed hn = sr_node(i, step);  // unused (removed) for float storage
  float g[Q];
  load_g<L, S, FIRST>(gin, ad, i, NP, g);
  float drho = 0.0f;
#pragma unroll
  for (int q = 0; q < Q; ++q) drho += g[q];
  const float rho = 1.0f + drho;
  const float rinv = 1.0f / rho;
  float u[D], jd[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    jd[d] = qsumc<L>(g, d) + 0.5f * F[d];
    u[d] = jd[d] * rinv;
  }
  if (STORE) {
    rho_o[i] = static_cast<double>(rho);
#pragma unroll
    for (int d = 0; d < D; ++d) vel_o[d * NP + i] = static_cast<double>(u[d]);
  }
  // Directions q and rev q share |c_q . u| and |c_q . F|: split geq and the
  // Guo term into the parts even and odd in c_q and form both at once.
  //   geq  = w (a0 + 4.5 rho cu^2 +- 3 rho cu),  a0 = drho - 1.5 rho u^2
  //   dF   = w tf (9 cF cu - 3 uF +- 3 cF)
  //   g'   = (1 - 1/tau) g + geq/tau + dF
  const float a0 = drho - 1.5f * rho * dot<L>(u, u);
  const float m3uF = -3.0f * dot<L>(u, F);
  const float r45 = 4.5f * rho, r3 = 3.0f * rho;
  if constexpr (MODE == kRheology) {
    const float visc = rheology_visc_lp<L>(g, a0, r45, r3, u, F, sr, nu, tab_n);
    const float tau = visc * 3.0f / rho + 0.5f;
    tau_inv = 1.0f / tau;
    tau_factor = 1.0f - 0.5f / tau;
    if (STORE) rho_o[NP + i] = static_cast<double>(visc);
  }
  const float omt = 1.0f - tau_inv;
  // one 64-bit address per node; the direction offsets fold into the stores
  S* __restrict__ out = elem_ptr<A::kOff32>(gout, pidx<Q>(0, i));
  if constexpr (MODE == kTrt) {
    // Each pair relaxes its even part (gs = (g_q + g_r)/2 = f_sym - w, towards
    // w(a0 + 4.5 rho cu^2)) with 1/tauSym and its odd part (ga = (g_q - g_r)/2,
    // towards 3 w rho cu) with 1/tauAnti; the Guo term splits the same way.
    const float oma = 1.0f - tauA_inv;
#pragma unroll
    for (int q = 0; q < Q - 1; ++q) {
      const int rq = L::rev(q);
      if (rq < q) continue;
      const float w = float(L::w(q));
      const float cu = cdot<L>(q, u);
      const float cF = cdot<L>(q, F);
      const float gs = 0.5f * (g[q] + g[rq]);
      const float ga = 0.5f * (g[q] - g[rq]);
      const float sym = omt * gs + (w * tau_inv) * (a0 + r45 * cu * cu) + (w * tau_factor) * (9.0f * cF * cu + m3uF);
      const float anti = oma * ga + (w * tauA_inv) * (r3 * cu) + (w * tauA_factor) * (3.0f * cF);
      __builtin_nontemporal_store(Store<S>::put_sr(sym + anti, sr_bits(hn, q)), out + (q << kPLog));
      __builtin_nontemporal_store(Store<S>::put_sr(sym - anti, sr_bits(hn, rq)), out + (rq << kPLog));
    }
    {
      constexpr int q = Q - 1;  // rest: even only
      const float w = float(L::w(q));
      const float gq = omt * g[q] + (w * tau_inv) * a0 + (w * tau_factor) * m3uF;
      __builtin_nontemporal_store(Store<S>::put_sr(gq, sr_bits(hn, q)), out + (q << kPLog));
    }
    return REDUCE ? static_cast<double>(jd[0]) : 0.0;
  }
#pragma unroll
  for (int q = 0; q < Q - 1; ++q) {
    const int rq = L::rev(q);
    if (rq < q) continue;  // each pair once
    const float w = float(L::w(q));
    const float wt = w * tau_inv, wf = w * tau_factor;
    const float cu = cdot<L>(q, u);
    const float cF = cdot<L>(q, F);
    const float ev = a0 + r45 * cu * cu;
    const float od = r3 * cu;
    const float fev = 9.0f * cF * cu + m3uF;
    const float fod = 3.0f * cF;
    const float gq = omt * g[q] + wt * (ev + od) + wf * (fev + fod);
    const float gr = omt * g[rq] + wt * (ev - od) + wf * (fev - fod);
    __builtin_nontemporal_store(Store<S>::put_sr(gq, sr_bits(hn, q)), out + (q << kPLog));
    __builtin_nontemporal_store(Store<S>::put_sr(gr, sr_bits(hn, rq)), out + (rq << kPLog));
  }
  {
    constexpr int q = Q - 1;  // rest: c = 0
    const float w = float(L::w(q));
    const float gq = omt * g[q] + (w * tau_inv) * a0 + (w * tau_factor) * m3uF;
    __builtin_nontemporal_store(Store<S>::put_sr(gq, sr_bits(hn, q)), out + (q << kPLog));
  }
  return REDUCE ? static_cast<double>(jd[0]) : 0.0;
}

template <class L, typename S, class A, bool FIRST, bool STORE, bool REDUCE, int MODE>
__global__ __launch_bounds__(kBlock) void k_bgk_lp(const S* __restrict__ gin, S* __restrict__ gout, A ad,
                                                   int32_t n, int64_t NP, float tau_inv, float tau_factor,
                                                   float F0, float F1, float F2,
                                                   double* __restrict__ rho_o,
                                                   double* __restrict__ vel_o,
                                                   double* __restrict__ partial, float tauA_inv,
                                                   float tauA_factor, const float* __restrict__ sr,
                                                   const float* __restrict__ nu, int tab_n, uint32_t step,
                                                   int32_t accum) {
  const int32_t b = node_block(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t i = b * kBlock + threadIdx.x;
  // the block's running sum, read before the node work (latency overlapped)
  const double prev = REDUCE && accum && threadIdx.x == 0 ? partial[b] : 0.0;
  double jx = 0.0;
  if (i < n) {
    const float F[3] = {F0, F1, F2};
    jx = bgk_node_lp<L, S, A, FIRST, STORE, REDUCE, MODE>(gin, gout, ad, i, NP, tau_inv, tau_factor, F, rho_o,
                                                          vel_o, tauA_inv, tauA_factor, sr, nu, tab_n, step);
  }
  if (REDUCE) {
    const double s = block_sum(jx);
    if (threadIdx.x == 0) partial[b] = accum ? prev + s : s;
  }
}

// Batched form (see k_batch_bgk in the FP64 kernels): the parameters are
// rounded to float exactly as the single-solve launcher rounds them.
template <class L, typename S, bool REDUCE, int NS>
__global__ __launch_bounds__(kBlock * NS) void k_batch_bgk_lp(const BatchSolve* __restrict__ solves,
                                                              const uint32_t* __restrict__ items, int32_t n_items,
                                                              int32_t per_wg, int32_t parity,
                                                              const uint32_t* __restrict__ step0, int32_t koff,
                                                              int32_t accum) {
  const uint32_t gstep = *step0 + static_cast<uint32_t>(koff);  // global step of this launch
  const int32_t g = node_block(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t sub = static_cast<int32_t>(threadIdx.x) / kBlock, tid = static_cast<int32_t>(threadIdx.x) % kBlock;
  const int32_t beg = g * per_wg;
  const int32_t end = beg + per_wg < n_items ? beg + per_wg : n_items;
  for (int32_t base = beg; base < end; base += NS) {
    const int32_t it = base + sub;
    const uint32_t item = items[it < end ? it : base];
    const BatchSolve& s = solves[item >> kItemShift];
    // a slice past the run, or a solve stopped at an earlier check of this
    // launch sequence, computes nothing (but joins the slice sums)
    const bool act = it < end && s.live;
    const int32_t lb = static_cast<int32_t>(item & kItemMask);
    const int p = (parity ^ s.parity) & 1;
    const RowWin ad{s.coord, s.win, s.NP, s.PY, s.W};
    const int32_t i = lb * kBlock + tid;
    // the running sum of this block over the interval, read before the node
    // work so that its latency overlaps the gathers
    const double prev = REDUCE && accum && tid == 0 && act ? s.partial[lb] : 0.0;
    double jx = 0.0;
    if (act && i < s.n) {
      const float F[3] = {static_cast<float>(s.a.F[0]), static_cast<float>(s.a.F[1]), static_cast<float>(s.a.F[2])};
      jx = bgk_node_lp<L, S, RowWin, false, false, REDUCE>(
          static_cast<const S*>(s.f[p]), static_cast<S*>(s.f[p ^ 1]), ad, i, s.NP, static_cast<float>(s.a.tau_inv),
          static_cast<float>(s.a.tau_factor), F, nullptr, nullptr, 0.0f, 0.0f, nullptr, nullptr, 0,
          gstep - s.step_base);
    }
    if (REDUCE) {
      const double t = slice_sum<NS>(jx);
      if (act && tid == 0) s.partial[lb] = accum ? prev + t : t;
    }
  }
}

// Conversion from and to FP64 populations.  One field: g = f - w.  Two fields
// (the two-phase populations f0, f1) are kept as the shifted total and the
// colour difference,
//   t = f0 + f1 - w,   d = f0 - f1,
// so that the flow, which lives in the total, keeps the shifted precision and
// the colour keeps the full relative precision of d.
// Rest state of a two-phase solve placed where its first pull reads it:
// f_k = w rho_k of node i goes to the source element of each direction of
// node i (the pull is a permutation of the elements), stored as
// (f0 + f1 - w, f0 - f1) with k_store's arithmetic.  r0: phase-0 density per
// compact node (phase 1: 1 - r0).
template <class L, typename S>
__global__ __launch_bounds__(kBlock) void k_prestream(RowWin ad, const double* __restrict__ r0, int32_t n, int64_t NP,
                                                      S* __restrict__ g) {
  constexpr int Q = L::nQ;
  const int32_t i = static_cast<int32_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t src[Q - 1];
  ad.template sources<L>(i, src);
  const double a = r0[i], b = 1.0 - a;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int64_t o = q < Q - 1 ? static_cast<int64_t>(src[q]) : static_cast<int64_t>(pidx<Q>(Q - 1, i));
    const double f0 = L::w(q) * a, f1 = L::w(q) * b;
    reinterpret_cast<Pair<S>*>(g)[o] =
        Pair<S>{Store<S>::put(static_cast<float>((f0 + f1) - L::w(q))), Store<S>::put(static_cast<float>(f0 - f1))};
  }
}

template <class L, typename S>
__global__ __launch_bounds__(kBlock) void k_store(const double* __restrict__ f, S* __restrict__ g,
                                                  int32_t n, int64_t NP, int nf) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int q = 0; q < L::nQ; ++q) {
    const int64_t o = pidx<L::nQ>(q, i);
    if (nf == 2) {
      const double f0 = f[2 * o], f1 = f[2 * o + 1];
      reinterpret_cast<Pair<S>*>(g)[o] =
          Pair<S>{Store<S>::put(static_cast<float>((f0 + f1) - L::w(q))), Store<S>::put(static_cast<float>(f0 - f1))};
    } else {
      g[o] = Store<S>::put(static_cast<float>(f[o] - L::w(q)));
    }
  }
}

// rank == null: f in the same AoSoA layout as g (two fields: element e at
// 2 e, 2 e + 1).  Otherwise node i goes to
// f[(rank[i] nf + k) Q + q], the LbField layout over the bulk nodes in
// ascending node order (rank[i]: the position of node i among them).
template <class L, typename S>
__global__ __launch_bounds__(kBlock) void k_load(const S* __restrict__ g, double* __restrict__ f,
                                                 int32_t n, int64_t NP, int nf, const int32_t* __restrict__ rank) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int64_t r = rank ? static_cast<int64_t>(rank[i]) * nf * L::nQ : 0;
#pragma unroll
  for (int q = 0; q < L::nQ; ++q) {
    const int64_t o = pidx<L::nQ>(q, i);
    if (nf == 2) {
      const Pair<S> v = reinterpret_cast<const Pair<S>*>(g)[o];
      const double t = static_cast<double>(Store<S>::get(v.x)) + L::w(q);
      const double d = static_cast<double>(Store<S>::get(v.y));
      f[rank ? r + q : 2 * o] = 0.5 * (t + d);
      f[rank ? r + L::nQ + q : 2 * o + 1] = 0.5 * (t - d);
    } else {
      f[rank ? r + q : o] = static_cast<double>(Store<S>::get(g[o])) + L::w(q);
    }
  }
}

// Both values (t, d) of each direction of a two-phase node: one pair load
// per direction.
template <class L, typename S, bool FIRST, class A>
__device__ __forceinline__ void load_td(const S* __restrict__ gin, const A& ad, int32_t i, double (&t)[L::nQ],
                                        double (&d)[L::nQ]) {
  constexpr bool O = A::kOff32;
  const Pair<S>* __restrict__ g2 = reinterpret_cast<const Pair<S>*>(gin);
  uint32_t src[L::nQ - 1];
  if (FIRST) {
#pragma unroll
    for (int q = 0; q < L::nQ - 1; ++q) src[q] = pidx<L::nQ>(q, i);
  } else {
    ad.template sources<L>(i, src);
  }
#pragma unroll
  for (int q = 0; q < L::nQ - 1; ++q) {
    const Pair<S> v = ld_elem<O>(g2, src[q]);
    t[q] = Store<S>::get(v.x);
    d[q] = Store<S>::get(v.y);
  }
  const Pair<S> v = ld_elem<O>(g2, pidx<L::nQ>(L::nQ - 1, i));
  t[L::nQ - 1] = Store<S>::get(v.x);
  d[L::nQ - 1] = Store<S>::get(v.y);
}

// (t, d) of a node from row-window links, and the links themselves for the
// neighbour phi lookup: bit q of fluid set when the source of direction q is
// the fluid node j[q], otherwise it is a solid (bounce-back).
template <class L, typename S>
__device__ __forceinline__ void load_td_links(const S* __restrict__ gin, const RowWin& ad, int32_t i,
                                              double (&t)[L::nQ], double (&d)[L::nQ], uint32_t (&j)[L::nQ - 1],
                                              uint32_t& fluid) {
  const Pair<S>* __restrict__ g2 = reinterpret_cast<const Pair<S>*>(gin);
  rowwin_links<L>(ad, i, j, fluid);
#pragma unroll
  for (int q = 0; q < L::nQ - 1; ++q) {
    const uint32_t m = 0u - ((fluid >> q) & 1u);
    const uint32_t e = (pidx<L::nQ>(q, j[q]) & m) | (pidx<L::nQ>(L::rev(q), i) & ~m);
    const Pair<S> v = ld_elem<true>(g2, e);
    t[q] = Store<S>::get(v.x);
    d[q] = Store<S>::get(v.y);
  }
  const Pair<S> v = ld_elem<true>(g2, pidx<L::nQ>(L::nQ - 1, i));
  t[L::nQ - 1] = Store<S>::get(v.x);
  d[L::nQ - 1] = Store<S>::get(v.y);
}

// phi = (r0 - r1) / (r0 + r1) = sum d / (1 + sum t)  (main_TWOPHASE.cpp:246-262)
template <class L>
__device__ __forceinline__ double phi_td(const double (&t)[L::nQ], const double (&d)[L::nQ]) {
  double st = 0.0, sd = 0.0;
#pragma unroll
  for (int q = 0; q < L::nQ; ++q) {
    st += t[q];
    sd += d[q];
  }
  return sd / (1.0 + st);
}

// Two-phase flux sums of one node on (t, d): sum_q (f0 + f1) c = sum_q t c,
// and the phase shares r_k / rho of it (LBglobalforcing.h:61-78).
template <class L>
__device__ __forceinline__ void flux_td(const double (&t)[L::nQ], const double (&d)[L::nQ], int dir, double& s0,
                                        double& s1, double& s2) {
  double st = 0.0, sd = 0.0;
#pragma unroll
  for (int q = 0; q < L::nQ; ++q) {
    st += t[q];
    sd += d[q];
  }
  const double rho = 1.0 + st;
  const double fc = qsumc<L>(t, dir);
  s0 = fc;
  s1 = (0.5 * (rho + sd) / rho) * fc;
  s2 = (0.5 * (rho - sd) / rho) * fc;
}

// Two-phase pass 3 on (t, d) (main_TWOPHASE.cpp:318-392): BGK with the mixed
// relaxation time and the Guo force on the total, the surface-tension
// perturbation (LBcollision2phase.h:7-20) and the recolouring
// (LBcollision2phase.h:70-86).  With base_q = f_tot,q + Omega_q + dOmegaF_q +
// dOmegaST_q and c_k = rho_k / rho,
//   f0' = c0 base + RC,  f1' = c1 base - RC
// give t' = base - w and d' = phi (base) + 2 RC with phi = c0 - c1.
// pn[qq]: phi of neighbor(qq, n) (LButilities.h:12-22).
// The two-phase arithmetic runs in FP64 on the stored values (float or half
// converted exactly), and only the stores round.  With float arithmetic the
// interfaces drift from the FP64 run: after 30 000 steps on three packs
// (64^3-128^3) the phase permeabilities differed from FP64 by 4e-4-1.2e-3 of
// k0 + k1, against 2e-6-3.5e-5 with FP64 arithmetic on the same FP32
// storage; the step time is the same (the passes are memory-bound).
template <class L, typename S, bool O, bool STORE>
__device__ __forceinline__ void cg_collide(const double (&t)[L::nQ], const double (&dd)[L::nQ],
                                           const float (&pn)[L::nQ], const double (&F)[3], double nu0Inv,
                                           double nu1Inv, double sigma, double beta, S* __restrict__ gout,
                                           int32_t i, int64_t NP, double* __restrict__ rho_o,
                                           double* __restrict__ vel_o, uint32_t step) {
  constexpr int Q = L::nQ, D = L::nD;
  const SrSeed hn = sr_node(i, step);  // unused (removed) for float storage
  double drho = 0.0, sd = 0.0;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    drho += t[q];
    sd += dd[q];
  }
  const double rho = 1.0 + drho;
  const double rinv = 1.0 / rho;
  const double r0 = 0.5 * (rho + sd), r1 = 0.5 * (rho - sd);
  const double phi_n = sd * rinv;
  double u[D];
#pragma unroll
  for (int d = 0; d < D; ++d) u[d] = (qsumc<L>(t, d) + 0.5 * F[d]) * rinv;
  if (STORE) {
    rho_o[i] = r0;
    rho_o[NP + i] = r1;
#pragma unroll
    for (int d = 0; d < D; ++d) vel_o[d * NP + i] = u[d];
  }
  const double tau = (3.0 * rho) / (r0 * nu0Inv + r1 * nu1Inv) + 0.5;
  const double ti = 1.0 / tau;
  const double tf = 1.0 - 0.5 * ti;
  double pr[Q];
#pragma unroll
  for (int q = 0; q < Q - 1; ++q) pr[q] = pn[q];
  pr[Q - 1] = 0.0;
  double g[D];
#pragma unroll
  for (int d = 0; d < D; ++d) g[d] = grad<L>(pr, d);
  const double cgn = sqrt(dot<L>(g, g));
  const double ginv = 1.0 / (cgn + (cgn < 2.220446049250313080847e-16 ? 1.0 : 0.0));
#pragma unroll
  for (int d = 0; d < D; ++d) g[d] = g[d] * ginv;
  const double af = 1.125 * cgn * sigma * ti;
  const double rfb = beta * r0 * r1 * rinv;
  const double a0 = drho - 1.5 * rho * dot<L>(u, u);
  const double m3uF = -3.0 * dot<L>(u, F);
  const double omt = 1.0 - ti;
  const double r45 = 4.5 * rho, r3 = 3.0 * rho;
  Pair<S>* __restrict__ out = elem_ptr<O>(reinterpret_cast<Pair<S>*>(gout), pidx<Q>(0, i));
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const double w = L::w(q);
    const double cu = cdot<L>(q, u);
    const double cF = cdot<L>(q, F);
    const double cg = cdot<L>(q, g);
    const double geq = w * (a0 + r45 * cu * cu + r3 * cu);
    const double dF = (w * tf) * (9.0 * cF * cu + m3uF + 3.0 * cF);
    double st, rc;
    if (q < Q - 1) {
      st = af * (w * cg * cg - L::B(q));
      rc = rfb * w * cg * (1.0 / L::cNorm(q));
    } else {
      st = -af * L::B(q);
      rc = 0.0;
    }
    const double bs = omt * t[q] + ti * geq + dF + st;
    const float bt = static_cast<float>(bs);
    const float bd = static_cast<float>(phi_n * (w + bs) + 2.0 * rc);
    const Pair<S> v{Store<S>::put_sr(bt, sr_bits(hn, q)), Store<S>::put_sr(bd, sr_bits(hn, q + 32))};
    __builtin_nontemporal_store(v, out + (q << kPLog));
  }
}

// Two-phase pass 1 (+ the momentum sums of the flux controller, in FP64:
// sum_q (f0 + f1) c_x = sum_q t c_x, main_TWOPHASE.cpp:292-308).
template <class L, typename S, bool FIRST, bool MOMX, class A>
__global__ __launch_bounds__(kBlock) void k_cg_phi_lp(const S* __restrict__ gin, A ad, int32_t n,
                                                      int64_t NP, float* __restrict__ phi,
                                                      double* __restrict__ partial) {
  constexpr int Q = L::nQ;
  const int32_t b = node_block(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t i = b * kBlock + threadIdx.x;
  double mx = 0.0;
  if (i < n) {
    double t[Q], d[Q];
    load_td<L, S, FIRST>(gin, ad, i, t, d);
    phi[i] = static_cast<float>(phi_td<L>(t, d));
    if (MOMX) mx = qsumc<L>(t, 0);
  }
  if (MOMX) {
    const double s = block_sum(mx);
    if (threadIdx.x == 0) partial[b] = s;
  }
}

// Two-phase pass 3; solid neighbours take their wall colour (neighbour_phi).
template <class L, typename S, bool FIRST, bool STORE, class A>
__global__ __launch_bounds__(kBlock) void k_cg_lp(const S* __restrict__ gin, S* __restrict__ gout, A ad,
                                                  int32_t n, int64_t NP, double nu0Inv, double nu1Inv,
                                                  double sigma, double beta, double F0, double F1, double F2,
                                                  int32_t momx_mode, const float* __restrict__ phi, WallPhi wp,
                                                  const double* __restrict__ dscalar,
                                                  double* __restrict__ rho_o, double* __restrict__ vel_o,
                                                  uint32_t step) {
  constexpr int Q = L::nQ;
  const int32_t b = node_block_rev(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t i = b * kBlock + threadIdx.x;
  if (i >= n) return;
  double t[Q], dd[Q];
  load_td<L, S, FIRST>(gin, ad, i, t, dd);
  double F[3] = {F0, F1, F2};
  if (momx_mode) F[0] = dscalar[0];
  float pn[Q];
  neighbour_phi<L>(ad, i, phi, wp, pn);
  cg_collide<L, S, A::kOff32, STORE>(t, dd, pn, F, nu0Inv, nu1Inv, sigma, beta, gout, i, NP, rho_o, vel_o,
                                     step);
}

// Batched two-phase passes (items as k_batch_bgk_lp; fixed force; every solid
// neighbour carries the solve's wall colour phi_w).  Pass 1 also writes the
// three flux partials per block on the interval's last step.
template <class L, typename S, bool REDUCE, int NS>
__global__ __launch_bounds__(kBlock * NS) void k_batch_cg_phi_lp(const BatchSolve* __restrict__ solves,
                                                                 const uint32_t* __restrict__ items,
                                                                 int32_t n_items, int32_t per_wg, int32_t parity) {
  constexpr int Q = L::nQ;
  const int32_t g = node_block(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t sub = static_cast<int32_t>(threadIdx.x) / kBlock, tid = static_cast<int32_t>(threadIdx.x) % kBlock;
  const int32_t beg = g * per_wg;
  const int32_t end = beg + per_wg < n_items ? beg + per_wg : n_items;
  for (int32_t base = beg; base < end; base += NS) {
    const int32_t it = base + sub;
    const uint32_t item = items[it < end ? it : base];
    const BatchSolve& s = solves[item >> kItemShift];
    const bool act = it < end && s.live;
    const int32_t lb = static_cast<int32_t>(item & kItemMask);
    const int p = (parity ^ s.parity) & 1;
    const RowWin ad{s.coord, s.win, s.NP, s.PY, s.W};
    const int32_t i = lb * kBlock + tid;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (act && i < s.n) {
      double t[Q], d[Q];
      load_td<L, S, false>(static_cast<const S*>(s.f[p]), ad, i, t, d);
      static_cast<float*>(s.phi)[i] = static_cast<float>(phi_td<L>(t, d));
      if (REDUCE) flux_td<L>(t, d, 0, s0, s1, s2);
    }
    if (REDUCE) {
      const int64_t nb = (s.n + kBlock - 1) / kBlock;
      s0 = slice_sum<NS>(s0);
      s1 = slice_sum<NS>(s1);
      s2 = slice_sum<NS>(s2);
      if (act && tid == 0) {
        s.partial[lb] = s0;
        s.partial[nb + lb] = s1;
        s.partial[2 * nb + lb] = s2;
      }
    }
  }
}

// At least 4 waves per SIMD (at most 128 VGPRs, no spills): left to
// itself the compiler takes 132 VGPRs (3 waves) for the FP64 two-phase
// arithmetic, and the pass ran 6 % slower (192^3 pack, FP32S).
template <class L, typename S, int NS>
__global__ __launch_bounds__(kBlock * NS, 4) void k_batch_cg_lp(const BatchSolve* __restrict__ solves,
                                                             const uint32_t* __restrict__ items, int32_t n_items,
                                                             int32_t per_wg, int32_t parity,
                                                             const uint32_t* __restrict__ step0, int32_t koff) {
  constexpr int Q = L::nQ;
  const uint32_t gstep = *step0 + static_cast<uint32_t>(koff);  // global step of this launch
  const int32_t g = node_block_rev(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t sub = static_cast<int32_t>(threadIdx.x) / kBlock, tid = static_cast<int32_t>(threadIdx.x) % kBlock;
  const int32_t beg = g * per_wg;
  const int32_t end = beg + per_wg < n_items ? beg + per_wg : n_items;
  for (int32_t base = beg; base < end; base += NS) {
    const int32_t it = base + sub;
    if (it >= end) continue;
    const uint32_t item = items[it];
    const BatchSolve& s = solves[item >> kItemShift];
    if (!s.live) continue;  // stopped at an earlier check of this launch sequence
    const int32_t lb = static_cast<int32_t>(item & kItemMask);
    const int p = (parity ^ s.parity) & 1;
    const RowWin ad{s.coord, s.win, s.NP, s.PY, s.W};
    const int32_t i = lb * kBlock + tid;
    if (i >= s.n) continue;
    double t[Q], dd[Q];
    uint32_t j[Q - 1], fluid;
    load_td_links<L, S>(static_cast<const S*>(s.f[p]), ad, i, t, dd, j, fluid);
    const float pw = static_cast<float>(s.phi_w);
    float pn[Q];
#pragma unroll
    for (int qq = 0; qq < Q - 1; ++qq) {
      const int q = L::rev(qq);
      const bool f = (fluid >> q) & 1u;
      const float v = static_cast<const float*>(s.phi)[f ? j[q] : static_cast<uint32_t>(i)];
      pn[qq] = f ? v : pw;
    }
    const double F[3] = {s.a.F[0], s.a.F[1], s.a.F[2]};
    cg_collide<L, S, true, false>(t, dd, pn, F, s.nu0Inv, s.nu1Inv, s.sigma, s.beta, static_cast<S*>(s.f[p ^ 1]),
                                  i, s.NP, nullptr, nullptr, gstep - s.step_base);
  }
}

template <class L, typename S, class A>
__global__ __launch_bounds__(kBlock) void k_gather_lp(const S* __restrict__ gin, S* __restrict__ gout,
                                                      A ad, int32_t n, int64_t NP, int nf) {
  constexpr int Q = L::nQ;
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t src[Q - 1];
  ad.template sources<L>(i, src);
  const int64_t m = nf;  // fields per element
  for (int k = 0; k < nf; ++k) {
#pragma unroll
    for (int q = 0; q < Q - 1; ++q) gout[m * pidx<Q>(q, i) + k] = gin[m * src[q] + k];
    gout[m * pidx<Q>(Q - 1, i) + k] = gin[m * pidx<Q>(Q - 1, i) + k];
  }
}

template <class L, typename S, class A, bool FIRST>
__global__ __launch_bounds__(kBlock) void k_flux_lp(const S* __restrict__ gin, A ad, int32_t n,
                                                    int64_t NP, int dir, double* __restrict__ partial,
                                                    int64_t nblocks) {
  constexpr int Q = L::nQ;
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  double s0 = 0.0;
  if (i < n) {
    float g[Q];
    load_g<L, S, FIRST>(gin, ad, i, NP, g);
    s0 = static_cast<double>(qsumc<L>(g, dir));
  }
  s0 = block_sum(s0);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s0;
    partial[nblocks + blockIdx.x] = 0.0;
    partial[2 * nblocks + blockIdx.x] = 0.0;
  }
}

template <class L, typename S, bool FIRST>
__global__ __launch_bounds__(kBlock) void k_flux_td(const S* __restrict__ gin, TableAddr ad, int32_t n,
                                                    int64_t NP, int dir, double* __restrict__ partial,
                                                    int64_t nblocks) {
  constexpr int Q = L::nQ;
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (i < n) {
    double t[Q], d[Q];
    load_td<L, S, FIRST>(gin, ad, i, t, d);
    flux_td<L>(t, d, dir, s0, s1, s2);
  }
  s0 = block_sum(s0);
  s1 = block_sum(s1);
  s2 = block_sum(s2);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s0;
    partial[nblocks + blockIdx.x] = s1;
    partial[2 * nblocks + blockIdx.x] = s2;
  }
}

inline dim3 grid_for(int32_t n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

// f(lattice tag, storage tag, address layout) for the runtime choices
template <typename F>
hipError_t dispatch(int precision, int lattice, const Addr& ad, F&& f) {
  auto by_addr = [&](auto l, auto st) {
    if (ad.layout == 2) f(l, st, ad.win);
    else if (ad.layout == 1) f(l, st, ad.rows);
    else f(l, st, ad.table);
  };
  if (lattice == 0) {
    if (precision == 1) by_addr(D2Q9{}, float{});
    else by_addr(D2Q9{}, uint16_t{});
  } else {
    if (precision == 1) by_addr(D3Q19{}, float{});
    else by_addr(D3Q19{}, uint16_t{});
  }
  return hipGetLastError();
}

}  // namespace

hipError_t prestream_lowp(int precision, int lattice, const RowWin& ad, const double* r0, int32_t n, int64_t NP,
                          void* g, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return dispatch(precision, lattice, Addr{}, [&](auto l, auto st, auto) {
    using L = decltype(l);
    using S = decltype(st);
    k_prestream<L, S><<<dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(
        ad, r0, n, NP, static_cast<S*>(g));
  });
}

hipError_t store_lowp(int precision, int lattice, int nf, const double* f, void* g, int32_t n,
                      int64_t NP, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return dispatch(precision, lattice, Addr{}, [&](auto l, auto st, auto) {
    using L = decltype(l);
    using S = decltype(st);
    k_store<L, S><<<grid_for(n), dim3(kBlock), 0, s>>>(f, static_cast<S*>(g), n, NP, nf);
  });
}

hipError_t load_lowp(int precision, int lattice, int nf, const void* g, double* f, int32_t n,
                     int64_t NP, hipStream_t s, const int32_t* rank) {
  if (n <= 0) return hipSuccess;
  return dispatch(precision, lattice, Addr{}, [&](auto l, auto st, auto) {
    using L = decltype(l);
    using S = decltype(st);
    k_load<L, S><<<grid_for(n), dim3(kBlock), 0, s>>>(static_cast<const S*>(g), f, n, NP, nf, rank);
  });
}

hipError_t gather_lowp(int precision, int lattice, int nf, const void* gin, void* gout,
                       const Addr& ad, int32_t n, int64_t NP, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return dispatch(precision, lattice, ad, [&](auto l, auto st, auto a) {
    using L = decltype(l);
    using S = decltype(st);
    using A = decltype(a);
    k_gather_lp<L, S, A><<<grid_for(n), dim3(kBlock), 0, s>>>(static_cast<const S*>(gin), static_cast<S*>(gout),
                                                             a, n, NP, nf);
  });
}

hipError_t bgk_lowp(int precision, int lattice, bool first, bool store, bool reduce,
                    const void* gin, void* gout, const Addr& ad, int32_t n, int64_t NP,
                    const BgkArgs& a, double* rho, double* vel, double* partial, hipStream_t s, bool accum) {
  if (n <= 0) return hipSuccess;
  const int32_t ac = accum ? 1 : 0;
  const float ti = static_cast<float>(a.tau_inv), tf = static_cast<float>(a.tau_factor);
  const float F0 = static_cast<float>(a.F[0]), F1 = static_cast<float>(a.F[1]), F2 = static_cast<float>(a.F[2]);
  const float tai = static_cast<float>(a.tauA_inv), taf = static_cast<float>(a.tauA_factor);  // TRT only
  // (first, store, reduce)
  const int key = (first ? 4 : 0) | (store ? 2 : 0) | (reduce ? 1 : 0);
  return dispatch(precision, lattice, ad, [&](auto l, auto st, auto addr) {
    using L = decltype(l);
    using S = decltype(st);
    using A = decltype(addr);
    const S* gi = static_cast<const S*>(gin);
    S* go = static_cast<S*>(gout);
    const dim3 g = grid_for(n), b(kBlock);
#define LB_K(F_, S_, R_, T_)                                                                             \
  k_bgk_lp<L, S, A, F_, S_, R_, T_><<<g, b, 0, s>>>(gi, go, addr, n, NP, ti, tf, F0, F1, F2, rho, vel, partial, \
                                                    tai, taf, a.tab_srf, a.tab_nuf, a.tab_n, a.step, ac)
#define LB_KEYS(T_)                                 \
  switch (key) {                                    \
    case 0: LB_K(false, false, false, T_); break;   \
    case 1: LB_K(false, false, true, T_); break;    \
    case 2: LB_K(false, true, false, T_); break;    \
    case 3: LB_K(false, true, true, T_); break;     \
    case 4: LB_K(true, false, false, T_); break;    \
    case 5: LB_K(true, false, true, T_); break;     \
    case 6: LB_K(true, true, false, T_); break;     \
    default: LB_K(true, true, true, T_); break;     \
  }
    if (a.mode == kTrt) {
      LB_KEYS(kTrt)
    } else if (a.mode == kRheology) {
      LB_KEYS(kRheology)
    } else {
      LB_KEYS(kBgk)
    }
#undef LB_KEYS
#undef LB_K
  });
}

hipError_t flux_lowp(int precision, int lattice, int nf, bool first, const void* gin,
                     const Addr& ad, int32_t n, int64_t NP, int dir, double* partial,
                     hipStream_t s) {
  if (nf != 1 && nf != 2) return hipErrorInvalidValue;
  const dim3 g = grid_for(n > 0 ? n : 1), b(kBlock);
  const int64_t nb = g.x;
  if (nf == 2) {
    return dispatch(precision, lattice, Addr{}, [&](auto l, auto st, auto) {
      using L = decltype(l);
      using S = decltype(st);
      const S* gi = static_cast<const S*>(gin);
      if (first) k_flux_td<L, S, true><<<g, b, 0, s>>>(gi, ad.table, n, NP, dir, partial, nb);
      else k_flux_td<L, S, false><<<g, b, 0, s>>>(gi, ad.table, n, NP, dir, partial, nb);
    });
  }
  return dispatch(precision, lattice, ad, [&](auto l, auto st, auto addr) {
    using L = decltype(l);
    using S = decltype(st);
    using A = decltype(addr);
    if (first) k_flux_lp<L, S, A, true><<<g, b, 0, s>>>(static_cast<const S*>(gin), addr, n, NP, dir, partial, nb);
    else k_flux_lp<L, S, A, false><<<g, b, 0, s>>>(static_cast<const S*>(gin), addr, n, NP, dir, partial, nb);
  });
}

hipError_t batch_bgk_lowp(int precision, int lattice, bool reduce, const BatchSolve* solves, const uint32_t* items,
                          int32_t n_items, int32_t grid, int32_t per_wg, int32_t parity, const uint32_t* step0,
                          int32_t koff, int ns, bool accum, hipStream_t s) {
  return dispatch(precision, lattice, Addr{}, [&](auto l, auto st, auto) {
    using L = decltype(l);
    using S = decltype(st);
    const dim3 g(grid), b(kBlock);
    if (ns == 2) {
      const dim3 b2(2 * kBlock);
      if (reduce) k_batch_bgk_lp<L, S, true, 2><<<g, b2, 0, s>>>(solves, items, n_items, per_wg, parity, step0, koff, accum ? 1 : 0);
      else k_batch_bgk_lp<L, S, false, 2><<<g, b2, 0, s>>>(solves, items, n_items, per_wg, parity, step0, koff, accum ? 1 : 0);
    } else {
      if (reduce) k_batch_bgk_lp<L, S, true, 1><<<g, b, 0, s>>>(solves, items, n_items, per_wg, parity, step0, koff, accum ? 1 : 0);
      else k_batch_bgk_lp<L, S, false, 1><<<g, b, 0, s>>>(solves, items, n_items, per_wg, parity, step0, koff, accum ? 1 : 0);
    }
  });
}

hipError_t batch_cg_lowp(int precision, int lattice, bool reduce, const BatchSolve* solves, const uint32_t* items,
                         int32_t n_items, int32_t grid, int32_t per_wg, int32_t parity, const uint32_t* step0,
                         int32_t koff, int ns, hipStream_t s) {
  return dispatch(precision, lattice, Addr{}, [&](auto l, auto st, auto) {
    using L = decltype(l);
    using S = decltype(st);
    const dim3 g(grid), b(kBlock);
    if (ns == 2) {
      const dim3 b2(2 * kBlock);
      if (reduce) k_batch_cg_phi_lp<L, S, true, 2><<<g, b2, 0, s>>>(solves, items, n_items, per_wg, parity);
      else k_batch_cg_phi_lp<L, S, false, 2><<<g, b2, 0, s>>>(solves, items, n_items, per_wg, parity);
      k_batch_cg_lp<L, S, 2><<<g, b2, 0, s>>>(solves, items, n_items, per_wg, parity, step0, koff);
    } else {
      if (reduce) k_batch_cg_phi_lp<L, S, true, 1><<<g, b, 0, s>>>(solves, items, n_items, per_wg, parity);
      else k_batch_cg_phi_lp<L, S, false, 1><<<g, b, 0, s>>>(solves, items, n_items, per_wg, parity);
      k_batch_cg_lp<L, S, 1><<<g, b, 0, s>>>(solves, items, n_items, per_wg, parity, step0, koff);
    }
  });
}

hipError_t cg_phi_lowp(int precision, int lattice, bool first, bool momx, const void* gin, const Addr& ad,
                       bool rowwin, int32_t n, int64_t NP, float* phi, double* partial, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  Addr sel = ad;
  sel.layout = rowwin ? 2 : 0;
  return dispatch(precision, lattice, sel, [&](auto l, auto st, auto addr) {
    using L = decltype(l);
    using S = decltype(st);
    using A = decltype(addr);
    if constexpr (std::is_same<A, RowAddr>::value) {
      return;  // not used for two-phase
    } else {
      const S* gi = static_cast<const S*>(gin);
      const dim3 g = grid_for(n), b(kBlock);
      if (first) {
        if (momx) k_cg_phi_lp<L, S, true, true, A><<<g, b, 0, s>>>(gi, addr, n, NP, phi, partial);
        else k_cg_phi_lp<L, S, true, false, A><<<g, b, 0, s>>>(gi, addr, n, NP, phi, partial);
      } else {
        if (momx) k_cg_phi_lp<L, S, false, true, A><<<g, b, 0, s>>>(gi, addr, n, NP, phi, partial);
        else k_cg_phi_lp<L, S, false, false, A><<<g, b, 0, s>>>(gi, addr, n, NP, phi, partial);
      }
    }
  });
}

hipError_t cg_lowp(int precision, int lattice, bool first, bool store, const void* gin, void* gout, const Addr& ad,
                   bool rowwin, int32_t n, int64_t NP, const CgArgs& a, const float* phi, const WallPhi& wp,
                   const double* dscalar, double* rho, double* vel, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (a.q0 != 0.0 || a.q1 != 0.0) return hipErrorInvalidValue;  // no mass sources in this form
  const double n0 = a.nu0Inv, n1 = a.nu1Inv, sg = a.sigma, be = a.beta;
  const double F0 = a.F[0], F1 = a.F[1], F2 = a.F[2];
  Addr sel = ad;
  sel.layout = rowwin ? 2 : 0;
  return dispatch(precision, lattice, sel, [&](auto l, auto st, auto addr) {
    using L = decltype(l);
    using S = decltype(st);
    using A = decltype(addr);
    if constexpr (std::is_same<A, RowAddr>::value) {
      return;  // not used for two-phase
    } else {
      const S* gi = static_cast<const S*>(gin);
      S* go = static_cast<S*>(gout);
      const dim3 g = grid_for(n), b(kBlock);
#define LB_K(F_, S_)                                                                                        \
  k_cg_lp<L, S, F_, S_, A><<<g, b, 0, s>>>(gi, go, addr, n, NP, n0, n1, sg, be, F0, F1, F2, a.momx_mode, phi, \
                                           wp, dscalar, rho, vel, a.step)
      if (first) {
        if (store) LB_K(true, true);
        else LB_K(true, false);
      } else {
        if (store) LB_K(false, true);
        else LB_K(false, false);
      }
#undef LB_K
    }
  });
}

}  // namespace lbamd
