// FP64 kernels: the reference's arithmetic, one IEEE operation at a time.
//
// This file is compiled with -ffp-contract=off: no multiply-add fusion, so
// every product and sum rounds exactly where the reference's x86-64 build
// (SSE2, no FMA) rounds it, and the results are bit-identical.
//
// Streaming is in pull form.  The reference pushes the post-collision values
// to neighbor(q,n) (LBfield.h:350-357) and then copies back the values that a
// node pushed into a solid neighbour (LBhalfwaybb.h:37-63).  Reading instead
// f_post(q, neighbor(rev q, n)) for a fluid source, and f_post(rev q, n) for a
// solid one, yields the same f_{t+1}(q,n) for every bulk node; the buffer keeps
// post-collision values and each step starts with the gather.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "lattice.hpp"
#include "layout_select.hpp"
#include "table.hpp"

namespace lbamd {
namespace {

// Both fields of one element (two-phase populations, adjacent in memory).
typedef double Double2 __attribute__((ext_vector_type(2)));

// Load the pre-collision populations of node i for NF fields (two fields:
// one 16-byte load per direction).
template <class L, int NF, bool FIRST, class A>
__device__ __forceinline__ void load_node(const double* __restrict__ fin, const A& ad, int32_t i,
                                          int64_t /*NP*/, double (&f)[NF][L::nQ]) {
  static_assert(NF == 1 || NF == 2, "one or two fields");
  constexpr bool O = A::kOff32;
  uint32_t src[L::nQ - 1];
  if (FIRST) {
#pragma unroll
    for (int q = 0; q < L::nQ - 1; ++q) src[q] = pidx<L::nQ>(q, i);
  } else {
    ad.template sources<L>(i, src);
  }
  auto load = [&](int q, uint32_t e) {
    if constexpr (NF == 1) {
      f[0][q] = ld_elem<O>(fin, e);
    } else {
      const Double2 v = ld_elem<O>(reinterpret_cast<const Double2*>(fin), e);
      f[0][q] = v.x;
      f[NF - 1][q] = v.y;
    }
  };
#pragma unroll
  for (int q = 0; q < L::nQ - 1; ++q) load(q, src[q]);
  load(L::nQ - 1, pidx<L::nQ>(L::nQ - 1, i));
}

// Deterministic block sum of 256 values; result valid in thread 0.
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
// shuffles and the same order over the slice's four waves.
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

// ------------------------------------------------------------------ BGK
// GeneralizedNewtonian::omegaBGK's viscosity (LBrheology.h:436-462, source
// 0): strain-rate tensor from the non-equilibrium moments and the Guo
// terms, its square sum, table lookup and linear interpolation.  Terms
// that are +-0 (c_i c_j = 0 off the diagonal, the source terms) are left
// out: a sum that starts at +0 never holds -0, so adding +-0 never changes it.
template <class L>
__device__ __forceinline__ double rheology_visc(const double (&f)[L::nQ], double rho, const double* u,
                                                double c2u2, const BgkArgs& a) {
  constexpr int Q = L::nQ, D = L::nD;
  double S[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) S[i][j] = 0.0;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const double cu = cdot<L>(q, u);
    const double feq = (L::w(q) * rho) * ((1 + L::c2Inv * cu) + L::c4Inv0_5 * (cu * cu - c2u2));
    const double neq = f[q] - feq;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const int ci = L::c(q, i);
      S[i][i] += neq * (static_cast<double>(ci * ci) - L::c2);
#pragma unroll
      for (int j = i + 1; j < D; ++j) {
        const int cc = ci * L::c(q, j);
        if (cc > 0) S[i][j] += neq;
        else if (cc < 0) S[i][j] += -neq;
      }
    }
  }
  double sq = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    S[i][i] += (u[i] * a.F[i]);
    sq += S[i][i] * S[i][i];
#pragma unroll
    for (int j = i + 1; j < D; ++j) {
      S[i][j] += 0.5 * (u[i] * a.F[j] + u[j] * a.F[i]);
      sq += (2 * S[i][j]) * S[i][j];
    }
  }
  const double* __restrict__ sr = a.tab_sr;
  const double* __restrict__ nu = a.tab_nu;
  const int k = table_index(sr, a.tab_n, sq);
  return ((nu[k + 1] - nu[k]) * (sq - sr[k])) / (sr[k + 1] - sr[k]) + nu[k];
}

// One node of the std_case loop body (std_case/main.cpp:110-145); returns the
// node's j_x = sum_q f c_x + F_x/2 for the permeability sum when REDUCE.
// MODE kTrt: the loop body with the two-relaxation-time operator in place of
// calcOmegaBGK / calcDeltaOmegaF (LBcollision.h:50-75, 214-233).  MODE
// kRheology: polymer_viscosity_tabulation/main.cpp:105-155, the relaxation
// time from the node's strain rate and the viscosity table; the viscosity
// is stored beside rho (rho_o[NP + i]) when STORE.
template <class L, class A, bool FIRST, bool STORE, bool REDUCE, int MODE = kBgk>
__device__ __forceinline__ double bgk_node(const double* __restrict__ fin, double* __restrict__ fout, const A& ad,
                                           int32_t i, int64_t NP, const BgkArgs& a, double* __restrict__ rho_o,
                                           double* __restrict__ vel_o) {
  constexpr int Q = L::nQ, D = L::nD;
  double jx = 0.0;
  double f[1][Q];
  load_node<L, 1, FIRST>(fin, ad, i, NP, f);
  // calcRho, calcVel (LBmacroscopic.h:11-45)
  const double rho = qsum<L>(f[0]);
  double u[D];
#pragma unroll
  for (int d = 0; d < D; ++d) u[d] = (qsumc<L>(f[0], d) + 0.5 * a.F[d]) / rho;
  if (STORE) {
    rho_o[i] = rho;
#pragma unroll
    for (int d = 0; d < D; ++d) vel_o[d * NP + i] = u[d];
  }
  if (REDUCE) jx = qsumc<L>(f[0], 0) + 0.5 * a.F[0];
  const double u2 = dot<L>(u, u);
  const double uF = dot<L>(u, a.F);
  const double c2u2 = L::c2 * u2;
  const double c2uF = L::c2 * uF;
  double* __restrict__ out = elem_ptr<A::kOff32>(fout, pidx<Q>(0, i));
  if constexpr (MODE == kRheology) {
    const double visc = rheology_visc<L>(f[0], rho, u, c2u2, a);
    // tau_ = visc_ * c2Inv / rho + 0.5 (LBrheology.h:463); calcDeltaOmegaF's
    // tau_factor at that tau (LBcollision.h:206)
    const double tau = (visc * L::c2Inv) / rho + 0.5;
    const double tau_inv = 1.0 / tau;
    const double tau_factor = 1 - 0.5 / tau;
    if (STORE) rho_o[NP + i] = visc;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double cu = cdot<L>(q, u);
      const double cF = cdot<L>(q, a.F);
      const double feq = (L::w(q) * rho) * ((1 + L::c2Inv * cu) + L::c4Inv0_5 * (cu * cu - c2u2));
      const double om = (-tau_inv) * (f[0][q] - feq);
      const double dF = (L::w(q) * tau_factor) * (L::c2Inv * cF + L::c4Inv * (cF * cu - c2uF));
      __builtin_nontemporal_store((f[0][q] + om) + dF, out + (q << kPLog));
    }
    return jx;
  }
  if constexpr (MODE == kTrt) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int r = L::rev(q);
      const double cu = cdot<L>(q, u);
      const double cF = cdot<L>(q, a.F);
      // calcOmegaBGKTRT (LBcollision.h:66-70)
      const double fqSym = 0.5 * (f[0][q] + f[0][r]);
      const double fqAnti = 0.5 * (f[0][q] - f[0][r]);
      const double rw = rho * L::w(q);
      const double om = (-a.tau_inv) * (fqSym - rw * (1.0 + L::c4Inv0_5 * (cu * cu - c2u2))) -
                        a.tauA_inv * (fqAnti - (rw * L::c2Inv) * cu);
      // calcDeltaOmegaFTRT, phi = 1 (LBcollision.h:229): w * 1.0 == w
      const double dF = L::w(q) * ((a.tauA_factor * L::c2Inv) * cF + (a.tau_factor * L::c4Inv) * (cF * cu - c2uF));
      __builtin_nontemporal_store((f[0][q] + om) + dF, out + (q << kPLog));
    }
    return jx;
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const double cu = cdot<L>(q, u);
    const double cF = cdot<L>(q, a.F);
    // calcOmegaBGK (LBcollision.h:41-44)
    const double feq = (rho * L::w(q)) * ((1.0 + L::c2Inv * cu) + L::c4Inv0_5 * (cu * cu - c2u2));
    const double om = (-a.tau_inv) * (f[0][q] - feq);
    // calcDeltaOmegaF (LBcollision.h:206-211)
    const double dF = (L::w(q) * a.tau_factor) * (L::c2Inv * cF + L::c4Inv * (cF * cu - c2uF));
    __builtin_nontemporal_store((f[0][q] + om) + dF, out + (q << kPLog));
  }
  return jx;
}

// REDUCE: the block's sum of j_x goes to partial[b], or is added to it
// (accum), so that a check interval's partials hold the sum over its steps.
template <class L, class A, bool FIRST, bool STORE, bool REDUCE, int MODE>
__global__ __launch_bounds__(kBlock) void k_bgk(const double* __restrict__ fin,
                                                double* __restrict__ fout, A ad, int32_t n,
                                                int64_t NP, BgkArgs a, double* __restrict__ rho_o,
                                                double* __restrict__ vel_o,
                                                double* __restrict__ partial, int32_t accum) {
  const int32_t b = node_block(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t i = b * kBlock + threadIdx.x;
  // the block's running sum, read before the node work (latency overlapped)
  const double prev = REDUCE && accum && threadIdx.x == 0 ? partial[b] : 0.0;
  double jx = 0.0;
  if (i < n) jx = bgk_node<L, A, FIRST, STORE, REDUCE, MODE>(fin, fout, ad, i, NP, a, rho_o, vel_o);
  if (REDUCE) {
    const double s = block_sum(jx);
    if (threadIdx.x == 0) partial[b] = accum ? prev + s : s;
  }
}

// Batched form: one launch advances every resident solve one step.  Work
// items are 256-node blocks of the solves (item = slot << kItemShift | block);
// each workgroup takes a contiguous run of per_wg items, so the resident
// workgroups sweep the items as one compact front.  A solve's buffers, address table and parameters come from its
// descriptor (uniform per item), and the per-node arithmetic is the single
// solve's (bgk_node), so a solve gets the same bits alone or in any mix.
// NS: items a workgroup works on at a time, one per 256-thread slice; every
// slice reduces its item exactly as a 256-thread workgroup would.
template <class L, bool REDUCE, int NS>
__global__ __launch_bounds__(kBlock * NS) void k_batch_bgk(const BatchSolve* __restrict__ solves,
                                                           const uint32_t* __restrict__ items, int32_t n_items,
                                                           int32_t per_wg, int32_t parity, int32_t accum) {
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
    if (act && i < s.n)
      jx = bgk_node<L, RowWin, false, false, REDUCE>(static_cast<const double*>(s.f[p]),
                                                     static_cast<double*>(s.f[p ^ 1]), ad, i, s.NP, s.a,
                                                     nullptr, nullptr);
    if (REDUCE) {
      const double t = slice_sum<NS>(jx);
      if (act && tid == 0) s.partial[lb] = accum ? prev + t : t;
    }
  }
}

// Per-solve sums of the block partials (1 component single-phase, 3
// two-phase), in the order k_sum_partials uses: out[3 slot + c].
__global__ __launch_bounds__(kBlock) void k_batch_sum(const BatchSolve* __restrict__ solves,
                                                      const int32_t* __restrict__ slots,
                                                      double* __restrict__ out) {
  const int32_t slot = slots[blockIdx.x];
  const BatchSolve& s = solves[slot];
  if (!s.live) return;
  const int64_t nb = s.n > 0 ? (s.n + kBlock - 1) / kBlock : 1;
  const int ncomp = s.kind == 1 ? 3 : 1;
  for (int c = 0; c < ncomp; ++c) {
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += kBlock) v += s.partial[c * nb + b];
    v = block_sum(v);
    if (threadIdx.x == 0) out[3 * slot + c] = v;
  }
}

// The stop test of every listed solve after a check interval of M steps
// (one thread per solve): the arithmetic of lb_solve_permeability and
// lb_solve_relperm, so a solve stops at the same step in a batch and alone.
// A decided solve is marked not live; the step kernels then skip it.
__global__ __launch_bounds__(kBlock) void k_batch_check(BatchSolve* __restrict__ solves,
                                                        SolveCheck* __restrict__ chk,
                                                        const int32_t* __restrict__ slots, int32_t n_slots,
                                                        const double* __restrict__ out, int32_t M) {
  const int32_t t = static_cast<int32_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= n_slots) return;
  const int32_t slot = slots[t];
  SolveCheck& c = chk[slot];
  if (c.status >= 0) return;
  c.steps += M;
  const double* o = out + 3 * slot;
  bool conv;
  if (c.kind == 0) {
    // k = nu <j_x> / F_x with j_x averaged over the states the interval's
    // steps started from
    c.k[0] = c.nu[0] * ((o[0] / static_cast<double>(M)) / c.n_box) / c.F0;
    conv = c.have_prev && fabs(c.k[0] - c.kp[0]) <= c.tol * fabs(c.k[0]);
  } else {
    // phase permeabilities k_i = nu_i <phi_i sum_q f c_x> / F_x (LBglobalforcing.h:61-78)
    c.k[1] = c.nu[1] * (o[1] / c.n_box) / c.F0;
    c.k[2] = c.nu[2] * (o[2] / c.n_box) / c.F0;
    c.k[0] = c.k[1] + c.k[2];
    const double scale = fabs(c.k[1]) + fabs(c.k[2]);
    conv = c.have_prev && fabs(c.k[1] - c.kp[1]) <= c.tol * scale && fabs(c.k[2] - c.kp[2]) <= c.tol * scale;
  }
  if (!isfinite(c.k[0])) {
    c.status = 2;
  } else {
    c.held = conv ? c.held + 1 : 0;
    if (c.held >= kStopHold) {
      c.status = 0;
    } else {
      for (int j = 0; j < 3; ++j) c.kp[j] = c.k[j];
      c.have_prev = 1;
      if (c.steps >= c.max_steps) c.status = 1;
    }
  }
  if (c.status >= 0) solves[slot].live = 0;
}

// The global step at the start of the next interval (read by the FP16
// stochastic rounding): one thread, inside the captured interval.
__global__ void k_batch_next_step(uint32_t* __restrict__ step, int32_t M) { *step += static_cast<uint32_t>(M); }

// Rest state f_k = w rho_k of a two-phase solve (rho_0 = r0 per compact
// node, rho_1 = 1 - r0) written where the first pull through ad reads it: the
// value of direction q of node i goes to the element its gather of q reads
// (the pull is a permutation of the elements).
template <class L>
__global__ __launch_bounds__(kBlock) void k_prestream(RowWin ad, const double* __restrict__ r0, int32_t n,
                                                      int64_t NP, double* __restrict__ f) {
  constexpr int Q = L::nQ;
  const int32_t i = static_cast<int32_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t src[Q - 1];
  ad.template sources<L>(i, src);
  const double a = r0[i], b = 1.0 - a;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int64_t o = q < Q - 1 ? static_cast<int64_t>(src[q]) : static_cast<int64_t>(pidx<Q>(Q - 1, i));
    reinterpret_cast<Double2*>(f)[o] = Double2{L::w(q) * a, L::w(q) * b};
  }
}

// Rest state f = w of a solve's pre-collision buffer (FP64 form).
template <class L>
__global__ __launch_bounds__(kBlock) void k_batch_rest(double* __restrict__ f, int64_t NP) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= static_cast<int64_t>(L::nQ) * NP) return;
  const int q = static_cast<int>((e >> kPLog) % L::nQ);
  f[e] = L::w(q);
}

// ------------------------------------------------------------------ tagged boundaries
// std_one_phase/main.cpp:513-575 with its boundary lists (main.cpp:27-204):
// the pull table t.link carries the links the tags make bounce-back (own
// reversed slot) beside the streaming ones; abb[i] marks the outgoing
// directions whose post-collision value becomes the anti-bounce-back value
// -f + 2 w rho_w (1 + 0.5 (9 cu^2 - 3 u^2)) (main.cpp:170-173), which the
// node then reads back through its bounce-back link.  info[i]: interior
// domain label (bits 0-30) and the mass-source flag (bit 31); fon[i]: the
// force attribute.

// Per-domain sums of 1 - rho over the pre-collision state, one block partial
// per domain label 1..n_label (main.cpp:527-535).
template <class L, bool FIRST>
__global__ __launch_bounds__(kBlock) void k_tag_mass(const double* __restrict__ fin, TableAddr ad, int32_t n,
                                                     int64_t NP, const uint32_t* __restrict__ info, int n_label,
                                                     double* __restrict__ partial) {
  const int32_t b = node_block(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t i = b * kBlock + threadIdx.x;
  double v = 0.0;
  int lab = 0;
  if (i < n) {
    double f[1][L::nQ];
    load_node<L, 1, FIRST>(fin, ad, i, NP, f);
    v = 1.0 - qsum<L>(f[0]);
    lab = static_cast<int>(info[i] & 0x7fffffffu);
  }
  for (int l = 1; l <= n_label; ++l) {
    const double s = block_sum(lab == l ? v : 0.0);
    if (threadIdx.x == 0) partial[static_cast<int64_t>(l - 1) * gridDim.x + b] = s;
  }
}

template <class L, bool FIRST, bool STORE>
__global__ __launch_bounds__(kBlock) void k_tagged(const double* __restrict__ fin, double* __restrict__ fout,
                                                   TableAddr ad, int32_t n, int64_t NP, BgkArgs a, double rho_w,
                                                   const uint32_t* __restrict__ info,
                                                   const double* __restrict__ fon,
                                                   const uint32_t* __restrict__ abb,
                                                   const double* __restrict__ scale,
                                                   const double* __restrict__ mass, double* __restrict__ rho_o,
                                                   double* __restrict__ vel_o) {
  constexpr int Q = L::nQ, D = L::nD;
  const int32_t i = static_cast<int32_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  double f[1][Q];
  load_node<L, 1, FIRST>(fin, ad, i, NP, f);
  // calcRho plus half the mass source (main.cpp:546-552)
  const uint32_t inf = info[i];
  const int lab = static_cast<int>(inf & 0x7fffffffu);
  const double add = (inf >> 31) ? 1.0 : 0.0;
  const double qm = (((0.9 * 2) * scale[lab]) * mass[lab]) * add;
  const double rho = qsum<L>(f[0]) + 0.5 * qm;
  // force = bodyforce * force attribute (main.cpp:554)
  const double fo = fon[i];
  double F[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) F[d] = a.F[d] * fo;
  double u[D];
#pragma unroll
  for (int d = 0; d < D; ++d) u[d] = (qsumc<L>(f[0], d) + 0.5 * F[d]) / rho;
  if (STORE) {
    rho_o[i] = rho;
#pragma unroll
    for (int d = 0; d < D; ++d) vel_o[d * NP + i] = u[d];
  }
  const double u2 = dot<L>(u, u);
  const double uF = dot<L>(u, F);
  const double c2u2 = L::c2 * u2;
  const double c2uF = L::c2 * uF;
  const uint32_t anti = abb[i];
  double* __restrict__ out = elem_ptr<false>(fout, pidx<Q>(0, i));
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const double cu = cdot<L>(q, u);
    const double cF = cdot<L>(q, F);
    const double feq = (rho * L::w(q)) * ((1.0 + L::c2Inv * cu) + L::c4Inv0_5 * (cu * cu - c2u2));
    const double om = (-a.tau_inv) * (f[0][q] - feq);
    const double dF = (L::w(q) * a.tau_factor) * (L::c2Inv * cF + L::c4Inv * (cF * cu - c2uF));
    // calcDeltaOmegaQ (LBcollision.h:91-95)
    const double dQ = ((a.tau_factor * qm) * L::w(q)) * ((1.0 + L::c2Inv * cu) + L::c4Inv0_5 * (cu * cu - c2u2));
    double v = ((f[0][q] + om) + dF) + dQ;
    if ((anti >> q) & 1u) {
      // -f(q, wall) + 2*w*rho_w*(1 + 0.5*(c4Inv*cu*cu - c2Inv*u2)) (main.cpp:173)
      const double bnd = ((2 * L::w(q)) * rho_w) * (1 + 0.5 * ((L::c4Inv * cu) * cu - L::c2Inv * u2));
      v = (-v) + bnd;
    }
    __builtin_nontemporal_store(v, out + (q << kPLog));
  }
}

// ------------------------------------------------------------------ two-phase pass 1
template <class L, bool FIRST, bool MOMX, class A>
__global__ __launch_bounds__(kBlock) void k_cg_phi(const double* __restrict__ fin, A ad,
                                                   int32_t n, int64_t NP, double* __restrict__ phi,
                                                   double* __restrict__ partial) {
  constexpr int Q = L::nQ;
  const int32_t b = node_block(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t i = b * kBlock + threadIdx.x;
  double mx = 0.0;
  if (i < n) {
    double f[2][Q];
    load_node<L, 2, FIRST>(fin, ad, i, NP, f);
    const double r0 = qsum<L>(f[0]);
    const double r1 = qsum<L>(f[1]);
    phi[i] = (r0 - r1) / (r0 + r1);
    if (MOMX) {
      double ft[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) ft[q] = f[0][q] + f[1][q];
      mx = qsumc<L>(ft, 0);
    }
  }
  if (MOMX) {
    const double s = block_sum(mx);
    if (threadIdx.x == 0) partial[b] = s;
  }
}

// ------------------------------------------------------------------ two-phase pass 3
// One node of the colour-gradient collision (main_TWOPHASE.cpp:318-392) with
// the body force F; phi: the colour field of this step (pass 1).
template <class L, bool FIRST, bool STORE, class A>
__device__ __forceinline__ void cg_node(const double* __restrict__ fin, double* __restrict__ fout, const A& ad,
                                        int32_t i, int64_t NP, const CgArgs& a, const double* __restrict__ phi,
                                        const WallPhi& wp, const double (&F)[L::nD], double* __restrict__ rho_o,
                                        double* __restrict__ vel_o) {
  constexpr int Q = L::nQ, D = L::nD;
  double f[2][Q];
  load_node<L, 2, FIRST>(fin, ad, i, NP, f);
  double r0 = qsum<L>(f[0]);
  double r1 = qsum<L>(f[1]);
  double ft[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) ft[q] = f[0][q] + f[1][q];
  const double rho = r0 + r1;
  double u[D];
#pragma unroll
  for (int d = 0; d < D; ++d) u[d] = (qsumc<L>(ft, d) + 0.5 * F[d]) / rho;
  // mass-source correction of the densities (main_TWOPHASE.cpp:332-335);
  // the stored rho fields carry the corrected values
  r0 = r0 + 0.5 * a.q0;
  r1 = r1 + 0.5 * a.q1;
  if (STORE) {
    rho_o[i] = r0;
    rho_o[NP + i] = r1;
#pragma unroll
    for (int d = 0; d < D; ++d) vel_o[d * NP + i] = u[d];
  }
  // viscosity mix (main_TWOPHASE.cpp:340)
  const double tau = (L::c2Inv * rho) / (r0 * a.nu0Inv + r1 * a.nu1Inv) + 0.5;
  const double tau_inv = 1.0 / tau;
  const double tau_factor = 1.0 - 0.5 / tau;
  const double uu = dot<L>(u, u);
  const double uF = dot<L>(u, F);
  const double c2uu = L::c2 * uu;
  const double c2uF = L::c2 * uF;
  // colour gradient from the neighbours' phi (LButilities.h:12-22, :62-66)
  double pn[Q];
  neighbour_phi<L>(ad, i, phi, wp, pn);
  double g[D];
#pragma unroll
  for (int d = 0; d < D; ++d) g[d] = grad<L>(pn, d);
  const double cgn = sqrt(dot<L>(g, g));
  const double ginv = 1.0 / (cgn + (cgn < 2.220446049250313080847e-16 ? 1.0 : 0.0));
#pragma unroll
  for (int d = 0; d < D; ++d) g[d] = g[d] * ginv;
  // calcDeltaOmegaST / RC prefactors (LBcollision2phase.h:12,75)
  const double af = ((1.125 * cgn) * a.sigma) / tau;
  const double rfb = ((a.beta * r0) * r1) / rho;
  const double c0 = r0 / rho;
  const double c1 = r1 / rho;
  Double2* __restrict__ out = elem_ptr<A::kOff32>(reinterpret_cast<Double2*>(fout), pidx<Q>(0, i));
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const double cu = cdot<L>(q, u);
    const double cF = cdot<L>(q, F);
    const double poly = (1.0 + L::c2Inv * cu) + L::c4Inv0_5 * (cu * cu - c2uu);
    const double om = (-tau_inv) * (ft[q] - (rho * L::w(q)) * poly);
    const double dF = (L::w(q) * tau_factor) * (L::c2Inv * cF + L::c4Inv * (cF * cu - c2uF));
    const double dQ0 = ((tau_factor * a.q0) * L::w(q)) * poly;
    const double dQ1 = ((tau_factor * a.q1) * L::w(q)) * poly;
    const double cg = cdot<L>(q, g);
    double st, rc;
    if (q < Q - 1) {
      st = af * ((L::w(q) * cg) * cg - L::B(q));
      rc = ((rfb * L::w(q)) * cg) / L::cNorm(q);
    } else {
      st = (-af) * L::B(q);
      rc = 0.0;
    }
    const double base = ((ft[q] + om) + dF) + st;
    const Double2 v{((c0 * base) + rc) + dQ0, ((c1 * base) - rc) + dQ1};
    __builtin_nontemporal_store(v, out + (q << kPLog));
  }
}

template <class L, bool FIRST, bool STORE, class A>
__global__ __launch_bounds__(kBlock) void k_cg(const double* __restrict__ fin,
                                               double* __restrict__ fout, A ad, int32_t n,
                                               int64_t NP, CgArgs a,
                                               const double* __restrict__ phi, WallPhi wp,
                                               const double* __restrict__ dscalar,
                                               double* __restrict__ rho_o,
                                               double* __restrict__ vel_o) {
  constexpr int D = L::nD;
  const int32_t b = node_block_rev(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t i = b * kBlock + threadIdx.x;
  if (i >= n) return;
  double F[D];
#pragma unroll
  for (int d = 0; d < D; ++d) F[d] = a.F[d];
  if (a.momx_mode) F[0] = dscalar[0];
  cg_node<L, FIRST, STORE>(fin, fout, ad, i, NP, a, phi, wp, F, rho_o, vel_o);
}

// Batched colour-gradient step in FP64 (items as k_batch_bgk; fixed force;
// every solid neighbour carries the solve's wall colour phi_w).  Pass 1
// writes the colour field and, on the interval's last step, the three flux
// partials per block with k_flux's arithmetic; pass 2 is cg_node.
template <class L, bool REDUCE, int NS>
__global__ __launch_bounds__(kBlock * NS) void k_batch_cg_phi(const BatchSolve* __restrict__ solves,
                                                              const uint32_t* __restrict__ items, int32_t n_items,
                                                              int32_t per_wg, int32_t parity) {
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
      double f[2][Q];
      load_node<L, 2, false>(static_cast<const double*>(s.f[p]), ad, i, s.NP, f);
      const double r0 = qsum<L>(f[0]);
      const double r1 = qsum<L>(f[1]);
      static_cast<double*>(s.phi)[i] = (r0 - r1) / (r0 + r1);
      if (REDUCE) {
        double ft[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) ft[q] = f[0][q] + f[1][q];
        const double fc = qsumc<L>(ft, 0);
        const double rt = r0 + r1;
        s0 = fc;
        s1 = (r0 / rt) * fc;
        s2 = (r1 / rt) * fc;
      }
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

template <class L, int NS>
__global__ __launch_bounds__(kBlock * NS) void k_batch_cg(const BatchSolve* __restrict__ solves,
                                                          const uint32_t* __restrict__ items, int32_t n_items,
                                                          int32_t per_wg, int32_t parity) {
  constexpr int D = L::nD;
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
    CgArgs a{};
    a.nu0Inv = s.nu0Inv;
    a.nu1Inv = s.nu1Inv;
    a.sigma = s.sigma;
    a.beta = s.beta;
    double F[D];
#pragma unroll
    for (int d = 0; d < D; ++d) F[d] = s.a.F[d];
    const WallPhi wp{nullptr, s.phi_w};
    cg_node<L, false, false>(static_cast<const double*>(s.f[p]), static_cast<double*>(s.f[p ^ 1]), ad, i, s.NP, a,
                             static_cast<const double*>(s.phi), wp, F, nullptr, nullptr);
  }
}

// ------------------------------------------------------------------ utilities
template <class L, class A>
__global__ __launch_bounds__(kBlock) void k_gather(const double* __restrict__ fin,
                                                   double* __restrict__ fout, A ad, int32_t n,
                                                   int64_t NP, int nf) {
  constexpr int Q = L::nQ;
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t src[Q - 1];
  ad.template sources<L>(i, src);
  const int64_t m = nf;  // fields per element
  for (int k = 0; k < nf; ++k) {
#pragma unroll
    for (int q = 0; q < Q - 1; ++q) fout[m * pidx<Q>(q, i) + k] = fin[m * src[q] + k];
    fout[m * pidx<Q>(Q - 1, i) + k] = fin[m * pidx<Q>(Q - 1, i) + k];
  }
}

template <class L, class A, int NF, bool FIRST>
__global__ __launch_bounds__(kBlock) void k_flux(const double* __restrict__ fin, A ad, int32_t n,
                                                 int64_t NP, int dir, double* __restrict__ partial,
                                                 int64_t nblocks) {
  constexpr int Q = L::nQ;
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (i < n) {
    double f[NF][Q];
    load_node<L, NF, FIRST>(fin, ad, i, NP, f);
    double ft[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) ft[q] = NF == 2 ? f[0][q] + f[NF - 1][q] : f[0][q];
    const double fc = qsumc<L>(ft, dir);
    s0 = fc;
    if (NF == 2) {
      const double r0 = qsum<L>(f[0]);
      const double r1 = qsum<L>(f[NF - 1]);
      const double rt = r0 + r1;
      s1 = (r0 / rt) * fc;
      s2 = (r1 / rt) * fc;
    }
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

template <class L>
__global__ __launch_bounds__(kBlock) void k_init_eq(double* __restrict__ f, int nf, int field,
                                                    int32_t n, int64_t NP,
                                                    const double* __restrict__ rho,
                                                    const double* __restrict__ vel) {
  constexpr int Q = L::nQ, D = L::nD;
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double u[D];
#pragma unroll
  for (int d = 0; d < D; ++d) u[d] = vel[d * NP + i];
  const double uu = dot<L>(u, u);
  const double r = rho[i];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const double cu = cdot<L>(q, u);
    f[static_cast<int64_t>(nf) * pidx<Q>(q, i) + field] = (L::w(q) * r) * ((1.0 + L::c2Inv * cu) + L::c4Inv0_5 * (cu * cu - L::c2 * uu));
  }
}

// Populations of nf fields in the LbField layout over the bulk nodes in
// ascending node order: node i goes to f[(rank[i] nf + k) Q + q] (from
// g[nf pidx(q, i) + k]).
template <class L>
__global__ __launch_bounds__(kBlock) void k_unstage(const double* __restrict__ g, double* __restrict__ f,
                                                    int32_t n, int64_t NP, int nf,
                                                    const int32_t* __restrict__ rank) {
  constexpr int Q = L::nQ;
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int64_t r = static_cast<int64_t>(rank[i]) * nf * Q;
  for (int k = 0; k < nf; ++k)
#pragma unroll
    for (int q = 0; q < Q; ++q) f[r + k * Q + q] = g[static_cast<int64_t>(nf) * pidx<Q>(q, i) + k];
}

__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* __restrict__ partial,
                                                         int64_t nblocks, int ncomp,
                                                         double* __restrict__ out) {
  for (int c = 0; c < ncomp; ++c) {
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += kBlock) s += partial[c * nblocks + b];
    s = block_sum(s);
    if (threadIdx.x == 0) out[c] = s;
  }
}

__global__ __launch_bounds__(kBlock) void k_finish_momx(const double* __restrict__ partial,
                                                        int64_t nblocks, double momx,
                                                        int64_t n_global,
                                                        double* __restrict__ dscalar) {
  double s = 0.0;
  for (int64_t b = threadIdx.x; b < nblocks; b += kBlock) s += partial[b];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    const double mean = s / static_cast<double>(n_global);
    dscalar[0] = 2.0 * (momx - mean);
    dscalar[1] = s;
  }
}

inline dim3 grid_for(int32_t n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

}  // namespace

// ------------------------------------------------------------------ launchers
#define LB_DISPATCH_LATTICE(lat, L, ...) \
  do {                                   \
    if ((lat) == 0) {                    \
      using L = D2Q9;                    \
      __VA_ARGS__;                       \
    } else {                             \
      using L = D3Q19;                   \
      __VA_ARGS__;                       \
    }                                    \
  } while (0)

template <class L, class A>
static void launch_bgk(bool first, bool store, bool reduce, dim3 g, hipStream_t s, const double* fin,
                       double* fout, const A& ad, int32_t n, int64_t NP, const BgkArgs& a, double* rho,
                       double* vel, double* partial, bool accum) {
  const dim3 b(kBlock);
  // (first, store, reduce)
  const int key = (first ? 4 : 0) | (store ? 2 : 0) | (reduce ? 1 : 0);
  const int32_t ac = accum ? 1 : 0;
#define LB_K(F_, S_, R_, T_) \
  k_bgk<L, A, F_, S_, R_, T_><<<g, b, 0, s>>>(fin, fout, ad, n, NP, a, rho, vel, partial, ac)
#define LB_KEYS(T_)                            \
  switch (key) {                               \
    case 0: LB_K(false, false, false, T_); break; \
    case 1: LB_K(false, false, true, T_); break;  \
    case 2: LB_K(false, true, false, T_); break;  \
    case 3: LB_K(false, true, true, T_); break;   \
    case 4: LB_K(true, false, false, T_); break;  \
    case 5: LB_K(true, false, true, T_); break;   \
    case 6: LB_K(true, true, false, T_); break;   \
    default: LB_K(true, true, true, T_); break;   \
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
}

hipError_t bgk_f64(int lattice, bool first, bool store, bool reduce, const double* fin,
                   double* fout, const Addr& ad, int32_t n, int64_t NP, const BgkArgs& a,
                   double* rho, double* vel, double* partial, hipStream_t s, bool accum) {
  if (n <= 0) return hipSuccess;
  const dim3 g = grid_for(n);
  LB_DISPATCH_LATTICE(lattice, L, {
    if (ad.layout == 2)
      launch_bgk<L>(first, store, reduce, g, s, fin, fout, ad.win, n, NP, a, rho, vel, partial, accum);
    else if (ad.layout == 1)
      launch_bgk<L>(first, store, reduce, g, s, fin, fout, ad.rows, n, NP, a, rho, vel, partial, accum);
    else launch_bgk<L>(first, store, reduce, g, s, fin, fout, ad.table, n, NP, a, rho, vel, partial, accum);
  });
  return hipGetLastError();
}

hipError_t cg_phi_f64(int lattice, bool first, bool momx, const double* fin, const Addr& ad, bool rowwin,
                      int32_t n, int64_t NP, double* phi, double* partial, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g = grid_for(n), b(kBlock);
  auto go = [&](auto addr) {
    using A = decltype(addr);
    LB_DISPATCH_LATTICE(lattice, L, {
      if (first && momx) k_cg_phi<L, true, true, A><<<g, b, 0, s>>>(fin, addr, n, NP, phi, partial);
      else if (first) k_cg_phi<L, true, false, A><<<g, b, 0, s>>>(fin, addr, n, NP, phi, partial);
      else if (momx) k_cg_phi<L, false, true, A><<<g, b, 0, s>>>(fin, addr, n, NP, phi, partial);
      else k_cg_phi<L, false, false, A><<<g, b, 0, s>>>(fin, addr, n, NP, phi, partial);
    });
  };
  if (rowwin) go(ad.win);
  else go(ad.table);
  return hipGetLastError();
}

hipError_t cg_f64(int lattice, bool first, bool store, const double* fin, double* fout, const Addr& ad,
                  bool rowwin, int32_t n, int64_t NP, const CgArgs& a, const double* phi, const WallPhi& wp,
                  const double* dscalar, double* rho, double* vel, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g = grid_for(n), b(kBlock);
  auto go = [&](auto addr) {
    using A = decltype(addr);
    LB_DISPATCH_LATTICE(lattice, L, {
      if (first && store) k_cg<L, true, true, A><<<g, b, 0, s>>>(fin, fout, addr, n, NP, a, phi, wp, dscalar, rho, vel);
      else if (first) k_cg<L, true, false, A><<<g, b, 0, s>>>(fin, fout, addr, n, NP, a, phi, wp, dscalar, rho, vel);
      else if (store) k_cg<L, false, true, A><<<g, b, 0, s>>>(fin, fout, addr, n, NP, a, phi, wp, dscalar, rho, vel);
      else k_cg<L, false, false, A><<<g, b, 0, s>>>(fin, fout, addr, n, NP, a, phi, wp, dscalar, rho, vel);
    });
  };
  if (rowwin) go(ad.win);
  else go(ad.table);
  return hipGetLastError();
}

hipError_t gather_f64(int lattice, int nfields, const double* fin, double* fout,
                      const Addr& ad, int32_t n, int64_t NP, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  LB_DISPATCH_LATTICE(lattice, L, {
    if (ad.layout == 2) k_gather<L, RowWin><<<grid_for(n), dim3(kBlock), 0, s>>>(fin, fout, ad.win, n, NP, nfields);
    else if (ad.layout == 1) k_gather<L, RowAddr><<<grid_for(n), dim3(kBlock), 0, s>>>(fin, fout, ad.rows, n, NP, nfields);
    else k_gather<L, TableAddr><<<grid_for(n), dim3(kBlock), 0, s>>>(fin, fout, ad.table, n, NP, nfields);
  });
  return hipGetLastError();
}

template <class L, class A>
static void launch_flux(int nfields, bool first, dim3 g, hipStream_t s, const double* fin, const A& ad,
                        int32_t n, int64_t NP, int dir, double* partial, int64_t nb) {
  const dim3 b(kBlock);
  if (nfields == 2) {
    if (first) k_flux<L, A, 2, true><<<g, b, 0, s>>>(fin, ad, n, NP, dir, partial, nb);
    else k_flux<L, A, 2, false><<<g, b, 0, s>>>(fin, ad, n, NP, dir, partial, nb);
  } else {
    if (first) k_flux<L, A, 1, true><<<g, b, 0, s>>>(fin, ad, n, NP, dir, partial, nb);
    else k_flux<L, A, 1, false><<<g, b, 0, s>>>(fin, ad, n, NP, dir, partial, nb);
  }
}

hipError_t flux_f64(int lattice, int nfields, bool first, const double* fin, const Addr& ad,
                    int32_t n, int64_t NP, int dir, double* partial, hipStream_t s) {
  const dim3 g = grid_for(n > 0 ? n : 1);
  const int64_t nb = g.x;
  LB_DISPATCH_LATTICE(lattice, L, {
    // two fields may not fit the row windows' 32-bit offsets
    const int lay = launch_layout(ad.layout, L::nQ, NP, nfields, sizeof(double));
    if (lay == 2) launch_flux<L>(nfields, first, g, s, fin, ad.win, n, NP, dir, partial, nb);
    else if (lay == 1) launch_flux<L>(nfields, first, g, s, fin, ad.rows, n, NP, dir, partial, nb);
    else launch_flux<L>(nfields, first, g, s, fin, ad.table, n, NP, dir, partial, nb);
  });
  return hipGetLastError();
}

hipError_t init_eq_f64(int lattice, double* f, int nf, int field, int32_t n, int64_t NP,
                       const double* rho, const double* vel, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  LB_DISPATCH_LATTICE(lattice, L, {
    k_init_eq<L><<<grid_for(n), dim3(kBlock), 0, s>>>(f, nf, field, n, NP, rho, vel);
  });
  return hipGetLastError();
}

hipError_t unstage_f64(int lattice, int nf, const double* g, double* f, int32_t n, int64_t NP,
                       const int32_t* rank, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  LB_DISPATCH_LATTICE(lattice, L, {
    k_unstage<L><<<grid_for(n), dim3(kBlock), 0, s>>>(g, f, n, NP, nf, rank);
  });
  return hipGetLastError();
}

hipError_t batch_check(BatchSolve* solves, SolveCheck* chk, const int32_t* slots, int32_t n_slots, const double* out,
                       int32_t M, uint32_t* step, hipStream_t s) {
  if (n_slots > 0)
    k_batch_check<<<dim3(static_cast<unsigned>((n_slots + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(
        solves, chk, slots, n_slots, out, M);
  k_batch_next_step<<<dim3(1), dim3(1), 0, s>>>(step, M);
  return hipGetLastError();
}

hipError_t prestream_f64(int lattice, const RowWin& ad, const double* r0, int32_t n, int64_t NP, double* f,
                         hipStream_t s) {
  if (n <= 0) return hipSuccess;
  LB_DISPATCH_LATTICE(lattice, L, {
    k_prestream<L><<<grid_for(n), dim3(kBlock), 0, s>>>(ad, r0, n, NP, f);
  });
  return hipGetLastError();
}

hipError_t batch_cg_f64(int lattice, bool reduce, const BatchSolve* solves, const uint32_t* items,
                        int32_t n_items, int32_t grid, int32_t per_wg, int32_t parity, int ns, hipStream_t s) {
  LB_DISPATCH_LATTICE(lattice, L, {
    const dim3 g(grid), b(kBlock), b2(2 * kBlock);
    if (ns == 2) {
      if (reduce) k_batch_cg_phi<L, true, 2><<<g, b2, 0, s>>>(solves, items, n_items, per_wg, parity);
      else k_batch_cg_phi<L, false, 2><<<g, b2, 0, s>>>(solves, items, n_items, per_wg, parity);
      k_batch_cg<L, 2><<<g, b2, 0, s>>>(solves, items, n_items, per_wg, parity);
    } else {
      if (reduce) k_batch_cg_phi<L, true, 1><<<g, b, 0, s>>>(solves, items, n_items, per_wg, parity);
      else k_batch_cg_phi<L, false, 1><<<g, b, 0, s>>>(solves, items, n_items, per_wg, parity);
      k_batch_cg<L, 1><<<g, b, 0, s>>>(solves, items, n_items, per_wg, parity);
    }
  });
  return hipGetLastError();
}

hipError_t batch_bgk_f64(int lattice, bool reduce, const BatchSolve* solves, const uint32_t* items,
                         int32_t n_items, int32_t grid, int32_t per_wg, int32_t parity, int ns, bool accum, hipStream_t s) {
  LB_DISPATCH_LATTICE(lattice, L, {
    const dim3 b(kBlock * (ns == 2 ? 2 : 1));
    if (ns == 2) {
      if (reduce) k_batch_bgk<L, true, 2><<<dim3(grid), b, 0, s>>>(solves, items, n_items, per_wg, parity, accum ? 1 : 0);
      else k_batch_bgk<L, false, 2><<<dim3(grid), b, 0, s>>>(solves, items, n_items, per_wg, parity, accum ? 1 : 0);
    } else {
      if (reduce) k_batch_bgk<L, true, 1><<<dim3(grid), b, 0, s>>>(solves, items, n_items, per_wg, parity, accum ? 1 : 0);
      else k_batch_bgk<L, false, 1><<<dim3(grid), b, 0, s>>>(solves, items, n_items, per_wg, parity, accum ? 1 : 0);
    }
  });
  return hipGetLastError();
}

hipError_t batch_sum(const BatchSolve* solves, const int32_t* slots, int32_t n_slots, double* out,
                     hipStream_t s) {
  if (n_slots <= 0) return hipSuccess;
  k_batch_sum<<<dim3(n_slots), dim3(kBlock), 0, s>>>(solves, slots, out);
  return hipGetLastError();
}

hipError_t batch_rest_f64(int lattice, double* f, int64_t NP, hipStream_t s) {
  LB_DISPATCH_LATTICE(lattice, L, {
    const int64_t n = static_cast<int64_t>(L::nQ) * NP;
    k_batch_rest<L><<<dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(f, NP);
  });
  return hipGetLastError();
}

hipError_t tagged_f64(int lattice, bool first, bool store, const double* fin, double* fout, const TableAddr& ad,
                      int32_t n, int64_t NP, const BgkArgs& a, double rho_w, const uint32_t* info,
                      const double* fon, const uint32_t* abb, const double* scale, const double* mass, double* rho,
                      double* vel, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g = grid_for(n), b(kBlock);
  LB_DISPATCH_LATTICE(lattice, L, {
    if (first && store) k_tagged<L, true, true><<<g, b, 0, s>>>(fin, fout, ad, n, NP, a, rho_w, info, fon, abb, scale, mass, rho, vel);
    else if (first) k_tagged<L, true, false><<<g, b, 0, s>>>(fin, fout, ad, n, NP, a, rho_w, info, fon, abb, scale, mass, rho, vel);
    else if (store) k_tagged<L, false, true><<<g, b, 0, s>>>(fin, fout, ad, n, NP, a, rho_w, info, fon, abb, scale, mass, rho, vel);
    else k_tagged<L, false, false><<<g, b, 0, s>>>(fin, fout, ad, n, NP, a, rho_w, info, fon, abb, scale, mass, rho, vel);
  });
  return hipGetLastError();
}

hipError_t tag_mass_f64(int lattice, bool first, const double* fin, const TableAddr& ad, int32_t n, int64_t NP,
                        const uint32_t* info, int n_label, double* partial, hipStream_t s) {
  if (n <= 0 || n_label <= 0) return hipSuccess;
  const dim3 g = grid_for(n), b(kBlock);
  LB_DISPATCH_LATTICE(lattice, L, {
    if (first) k_tag_mass<L, true><<<g, b, 0, s>>>(fin, ad, n, NP, info, n_label, partial);
    else k_tag_mass<L, false><<<g, b, 0, s>>>(fin, ad, n, NP, info, n_label, partial);
  });
  return hipGetLastError();
}

hipError_t sum_partials(const double* partial, int64_t nblocks, int ncomp, double* out,
                        hipStream_t s) {
  k_sum_partials<<<dim3(1), dim3(kBlock), 0, s>>>(partial, nblocks, ncomp, out);
  return hipGetLastError();
}

hipError_t finish_momx(const double* partial, int64_t nblocks, double momx,
                       int64_t n_global, double* dscalar, hipStream_t s) {
  k_finish_momx<<<dim3(1), dim3(kBlock), 0, s>>>(partial, nblocks, momx, n_global, dscalar);
  return hipGetLastError();
}

}  // namespace lbamd
