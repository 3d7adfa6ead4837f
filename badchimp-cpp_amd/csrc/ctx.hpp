// The context object behind the C ABI handle: one rank's subdomain on one GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "kernels.hpp"

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
  std::vector<int32_t> node_of;     // compact index -> reference node number
  // bulk nodes in ascending node order: rank[i] is the position of compact
  // node i among them; runs of consecutive node numbers {first node, first
  // position, length} (lb_download_f_noalloc copies each run in one piece)
  int32_t* d_rank = nullptr;
  std::vector<std::array<int32_t, 3>> bulk_runs;
  std::vector<int32_t> wall_nodes;  // wall slot -> node number
  // other-rank fluid nodes (type 4, the halo): population slots after the
  // bulk nodes in compact order; slot_of[node] = compact index (bulk, then
  // halo) or -1
  int32_t n_ghost = 0;
  std::vector<int32_t> slot_of;
  // exchanges with neighbouring ranks (lb_halo_add): element index per value
  // sent / received, compact index per node sent / received (the colour),
  // and a device staging buffer of both fields
  struct Halo {
    int32_t rank = -1;
    int64_t n_send = 0, n_recv = 0;
    int64_t n_send_nodes = 0, n_recv_nodes = 0;
    uint32_t* d_send = nullptr;
    uint32_t* d_recv = nullptr;
    uint32_t* d_send_nodes = nullptr;
    uint32_t* d_recv_nodes = nullptr;
    double* d_buf = nullptr;
  };
  std::vector<Halo> halos;
  std::vector<int16_t> node_type;
  uint32_t* d_link = nullptr;       // pull table in compact order (both layouts)
  uint32_t* d_coord = nullptr;      // row layout: packed cell coordinates
  uint2* d_rows = nullptr;          // row layout: {base, mask} per (row, word)
  uint4* d_win = nullptr;           // row windows: {base, mask lo, mask hi, 0} per (row, word)
  lbamd::Addr addr;                 // layout used by the single-phase kernels
  std::string layout_note;          // "rows", or why the table layout is used

  int n_fields = 0;  // populations allocated for this many fields
  void* d_f[2] = {nullptr, nullptr};
  int cur = 0;
  bool post = false;  // buffer cur holds post-collision values (needs a gather)
  double* d_rho = nullptr;       // [2][NP]
  double* d_vel = nullptr;       // [nD][NP]
  double* d_phi = nullptr;       // [NP]
  double* d_phi_wall = nullptr;  // [n_wall + 1]
  bool wall_uniform = true;      // every wall slot has the colour wall_phi
  double wall_phi = 0.0;
  double* d_partial = nullptr;   // [3][n_blocks]
  double* d_scalar = nullptr;    // [8]: 0 force x, 1 flux sum, 4..6 reductions
  int64_t n_blocks = 0;
  bool have_macro = false;
  uint32_t step_count = 0;       // steps since the populations were set (seeds FP16 stochastic rounding)
  bool cg_colour_done = false;   // lb_step_cg_colour ran for the iteration at step_count
  bool have_visc = false;        // the last stored step was a rheology step
  // tagged boundaries (std_one_phase): pull table, anti-bounce-back masks,
  // domain labels with mass-source flags, force attribute, per-domain
  // scale factors and mass changes, partials of the mass pass
  bool tag_ready = false;
  int tag_labels = 0;            // highest interior-domain label
  uint32_t* d_tlink = nullptr;   // [Q-1][NP]
  uint32_t* d_tabb = nullptr;    // [NP]
  uint32_t* d_tinfo = nullptr;   // [NP]
  double* d_tfon = nullptr;      // [NP]
  double* d_tscale = nullptr;    // [labels + 1], then mass changes [labels + 1]
  double* d_tpartial = nullptr;  // [labels][n_blocks]
  std::vector<std::pair<int32_t, int>> tag_pressure;  // pressure-boundary nodes (ascending) and phase - 1
  double* d_table = nullptr;     // viscosity table: sr[tab_n], nu[tab_n] (FP64), then the same in float
  int32_t tab_n = 0;             // rows + 2 (constant extension at both ends)
  double last_force[3] = {0, 0, 0};
  int64_t bytes = 0;
};
