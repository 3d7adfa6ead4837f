// Device context of one rank's subdomain and the shared device-side layout.
//
// Layout in HBM (over the compact bulk order i = 0..n_bulk-1: the
// reference's findBulkNodes order, or the tiled row order of the row
// layouts; padded to NP = multiple of 256):
//   f[k](q,i)     populations of field k at pidx(q, i) (blocks of 64 nodes,
//                 directions one after another), storage type per precision,
//                 two buffers (ping-pong)
//   link[q][i]    q = 0..Q-2 (the rest direction Q-1 always reads itself):
//                 bits 30-31 kind, bits 0-29 index
//                   kind 0  stream     read f[k][q][j]       j = compact(neighbor(rev q, n))
//                   kind 1  bounce     read f[k][rev q][i]   index = wall slot of the solid
//                   kind 2  ghost 0    read f[k][rev q][j]   j = last bulk node pushing rev q
//                                      into node 0 (the reference's shared slot)
//   rho[k][i], vel[d][i], phi[i]       macroscopic fields of the last step
//   phi_wall[s]                        colour field of the solid neighbours
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace lbamd {

constexpr uint32_t kLinkStream = 0u;
constexpr uint32_t kLinkBounce = 1u;
constexpr uint32_t kLinkGhost0 = 2u;
constexpr uint32_t kLinkIndexMask = (1u << 30) - 1u;
constexpr int kBlock = 256;
// Steady-state stop of the permeability solves: the relative change of k
// between two checks is at most the tolerance at this many consecutive
// checks.  One check alone can pass by chance while k still drifts (the
// stochastically rounded FP16 storage adds noise of about the tolerance).
constexpr int kStopHold = 2;

// Population element (direction q, node i) of one field: blocks of kPB
// consecutive nodes keep their Q directions one after another (AoSoA), so a
// wavefront's loads and stores of one node block fall into one DRAM region
// instead of Q separate planes (NP is a multiple of kPB).  With nf fields
// (two for the two-phase populations) the fields of one element are
// adjacent: field k of element e = pidx(q, i) is at nf * e + k, so one
// vector access moves both fields of a direction.
constexpr int kPLog = 6;
constexpr uint32_t kPB = 1u << kPLog;
// Written as i + 64 (Q - 1) (i >> 6) + 64 q (the same value), so that the
// per-node part is one 24-bit multiply-add (i >> 6 < 2^24 for every element
// index below 2^32, which a context guarantees) and the direction part is a
// constant that folds into the memory instruction's offset.
template <int Q>
__device__ __forceinline__ uint32_t pidx(uint32_t q, uint32_t i) {
  return __umul24(i >> kPLog, static_cast<uint32_t>((Q - 1) << kPLog)) + i + (q << kPLog);
}
inline uint32_t pidx_host(int Q, uint32_t q, uint32_t i) {
  return (((i >> kPLog) * static_cast<uint32_t>(Q) + q) << kPLog) | (i & (kPB - 1u));
}

struct BgkArgs {
  double tau_inv;     // 1.0 / tau                (LBcollision.h:40); 1 / tauSym for TRT (:62)
  double tau_factor;  // 1 - 0.5 / tau            (LBcollision.h:206); tauSym_factor for TRT (:223)
  double F[3];        // body force
  // collision operator: kBgk; kTrt, the two-relaxation-time operator
  // (calcOmegaBGKTRT / calcDeltaOmegaFTRT with phi = 1, LBcollision.h:50-75,
  // 214-233); kRheology, GeneralizedNewtonian::omegaBGK with the node's
  // relaxation time from a viscosity table (LBrheology.h:321-476)
  int32_t mode;
  double tauA_inv;     // 1.0 / tauAnti                  (LBcollision.h:63)
  double tauA_factor;  // 1 - 0.5 / tauAnti              (LBcollision.h:224)
  // viscosity table with its constant extension, tab_n = rows + 2
  // (LBrheology.h:370-384); FP64 for the FP64 kernels, float otherwise
  const double* tab_sr;
  const double* tab_nu;
  const float* tab_srf;
  const float* tab_nuf;
  int32_t tab_n;
  uint32_t step;       // steps since the state was set: seeds the FP16 stochastic rounding
};
constexpr int kBgk = 0, kTrt = 1, kRheology = 2;

// One solve of the batched form: its two population buffers (each Q * NP
// elements, AoSoA as above), its row-window address table, its block
// partials and its parameters.  parity: which buffer holds the current state
// at even global steps.
struct BatchSolve {
  const uint32_t* coord;
  const uint4* win;
  void* f[2];
  double* partial;  // [ncomp][blocks]: 1 component single-phase, 3 two-phase
  void* phi;        // two-phase colour field (NP; float, double in FP64), null for single-phase
  uint32_t NP;
  int32_t PY, W, n;
  int32_t parity;
  int32_t kind;     // 0 single-phase BGK, 1 colour-gradient two-phase
  BgkArgs a;        // tau_inv, tau_factor (single-phase) and the body force
  double nu0Inv, nu1Inv, sigma, beta, phi_w;  // two-phase parameters, wall colour
  uint32_t step_base;  // global batch step at admission: the solve's own step count is global - step_base
  int32_t live;        // 1 while the solve advances; 0 once its stop test decided (set on the device)
};

// Stop test of one batched solve, kept and evaluated on the device after each
// check interval: the relative change of k (single-phase, k[0]) or of both
// phase permeabilities k[1], k[2] (two-phase, against |k[1]| + |k[2]|) is at
// most tol at kStopHold consecutive checks (status 0), or the step limit is
// reached (status 1), or k is not finite (status 2); -1 while running.
struct SolveCheck {
  double k[3];   // at the last check: k (k0 + k1 for two-phase), k0, k1
  double kp[3];  // at the check before
  double nu[3];  // nu, nu0, nu1
  double n_box, F0, tol;
  int64_t steps, max_steps;
  int32_t held, have_prev, status, kind;
};
// Work item of a batched launch: slot << kItemShift | 256-node block of the solve.
constexpr int kItemShift = 18;
constexpr uint32_t kItemMask = (1u << kItemShift) - 1u;

struct CgArgs {
  double nu0Inv, nu1Inv;  // 1.0 / (c2 * (tau_k - 0.5))  (main_TWOPHASE.cpp:128-129)
  double sigma, beta;
  double q0, q1;          // mass sources (0 in the reference case)
  double F[3];            // fixed force; F[0] replaced by the device value in MOMX mode
  double momx;
  int32_t momx_mode;
  uint32_t step;          // steps since the state was set (FP16 stochastic rounding)
};

}  // namespace lbamd
