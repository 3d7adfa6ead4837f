// Device context of one rank's subdomain and the shared device-side layout.
//
// Layout in HBM (SoA over the compact bulk order i = 0..n_bulk-1, which is
// the reference's findBulkNodes order, padded to NP = multiple of 256):
//   f[k][q][i]    populations of field k, storage type per precision,
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

struct BgkArgs {
  double tau_inv;     // 1.0 / tau                (LBcollision.h:40)
  double tau_factor;  // 1 - 0.5 / tau            (LBcollision.h:206)
  double F[3];        // body force
};

struct CgArgs {
  double nu0Inv, nu1Inv;  // 1.0 / (c2 * (tau_k - 0.5))  (main_TWOPHASE.cpp:128-129)
  double sigma, beta;
  double q0, q1;          // mass sources (0 in the reference case)
  double F[3];            // fixed force; F[0] replaced by the device value in MOMX mode
  double momx;
  int32_t momx_mode;
};

}  // namespace lbamd

struct lb_ctx {
  int lattice = 0, precision = 0, device = 0;
  int Q = 0, nD = 0;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double kernel_ms = 0.0;
  int64_t launches = 0;

  bool has_topo = false;
  int32_t n_nodes = 0, n_bulk = 0;
  int64_t NP = 0;
  int64_t n_global = 0;
  std::vector<int32_t> bulk;        // compact -> node number
  std::vector<int32_t> wall_nodes;  // wall slot -> node number
  std::vector<int16_t> node_type;
  uint32_t* d_link = nullptr;

  int n_fields = 0;  // populations allocated for this many fields
  void* d_f[2] = {nullptr, nullptr};
  int cur = 0;
  bool post = false;  // buffer cur holds post-collision values (needs a gather)
  double* d_rho = nullptr;   // [2][NP]
  double* d_vel = nullptr;   // [nD][NP]
  double* d_phi = nullptr;   // [NP]
  double* d_phi_wall = nullptr;  // [n_wall + 1]
  double* d_partial = nullptr;   // [3][n_blocks]
  double* d_scalar = nullptr;    // [8]: 0 force x, 1..3 flux sums
  int64_t n_blocks = 0;
  bool have_macro = false;
  double last_force[3] = {0, 0, 0};
  int64_t bytes = 0;
};
