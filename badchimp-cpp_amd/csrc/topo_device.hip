// Node list and neighbour index built on the GPU from a voxel geometry, with
// the numbering of the reference's geometry writer (PythonScripts/vtklb.py:
// 6-44, 70-124, 182-218) -- the same topology lb_topo_from_geometry builds on
// the host, for a single-rank geometry (0 solid, 1 fluid).
//
// The writer pads the geometry by one cell per side, copies the opposite
// layer into the rim of each periodic axis (vtklb.py:97-114), numbers the
// interior fluid cells in C order, then the solid cells that a fluid cell
// reaches through one lattice link, again in C order (vtklb.py:184-197).  A
// neighbour that is a periodic copy of a fluid cell resolves to that cell's
// node; one outside the padded box to node 0 (vtklb.py:23-39).
//
// Steps (one thread per padded cell unless noted):
//   k_pad       rim copies: wrap the coordinates of periodic axes, -1 elsewhere
//   k_flags     fluid / wall flag per cell, per-tile counts
//   k_scan      exclusive scan of the tile counts (one workgroup)
//   k_number    node number per cell (tile offset + in-tile scan), cell of node
//   k_nodes     positions, neighbour index, nodetype attribute (thread per node)
//   k_types     Nodes::setupNodeType (LBnodes.h:156-209) (thread per node)
// Integer work only; every step is a deterministic function of the input.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <new>
#include <vector>

#include "../../include/badchimp_amd.h"
#include "errors.hpp"
#include "lattice.hpp"
#include "topology.hpp"

using lbamd::fail_global;

namespace {

constexpr int kTB = 256;            // threads per workgroup
constexpr int kPer = 16;            // cells per thread in the numbering passes
constexpr int kTile = kTB * kPer;   // cells per tile
constexpr int kScanTB = 1024;       // threads of the tile-offset scan

struct Pad {
  int nd, Q, periodic;
  int32_t dims[3];     // geometry dimensions
  int32_t pd[3];       // padded
  int64_t stride[3];   // padded strides, C order
  int64_t M;           // padded cells
  int64_t shift[19];   // flat offset of c_q
  int32_t c[19][3];
};

__device__ __forceinline__ void coords(const Pad& P, int64_t i, int32_t (&k)[3]) {
  for (int d = P.nd - 1; d >= 0; --d) {
    k[d] = static_cast<int32_t>(i % P.pd[d]);
    i /= P.pd[d];
  }
}

__device__ __forceinline__ bool interior(const Pad& P, const int32_t (&k)[3]) {
  bool in = true;
  for (int d = 0; d < P.nd; ++d) in = in && k[d] >= 1 && k[d] <= P.pd[d] - 2;
  return in;
}

// Rim copies: for each axis a rim coordinate of a periodic axis takes the
// opposite interior layer, any other rim coordinate makes the cell -1.
__global__ __launch_bounds__(kTB) void k_pad(const int32_t* __restrict__ geo, int8_t* __restrict__ g, Pad P,
                                             int* __restrict__ bad) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x; i < P.M;
       i += static_cast<int64_t>(gridDim.x) * kTB) {
    int32_t k[3] = {0, 0, 0};
    coords(P, i, k);
    bool rim = false;
    int64_t src = 0;
    for (int d = 0; d < P.nd; ++d) {
      int32_t u = k[d] - 1;  // unpadded coordinate
      if (k[d] == 0 || k[d] == P.pd[d] - 1) {
        if (P.periodic & (1 << d)) u = k[d] == 0 ? P.dims[d] - 1 : 0;
        else rim = true;
      }
      src = src * P.dims[d] + u;
    }
    int8_t v = -1;
    if (!rim) {
      const int32_t x = geo[src];
      if (x < 0) atomicOr(bad, 1);
      else if (x > 1) atomicOr(bad, 2);
      v = static_cast<int8_t>(x > 1 ? 1 : (x < 0 ? 0 : x));
    }
    g[i] = v;
  }
}

// bit 0: own fluid (interior, value 1); bit 1: wall (value 0, one link from
// an own fluid cell).
__device__ __forceinline__ uint8_t cell_flags(const int8_t* __restrict__ g, const Pad& P, int64_t i) {
  int32_t k[3] = {0, 0, 0};
  coords(P, i, k);
  const int8_t v = g[i];
  if (v == 1) return interior(P, k) ? 1 : 0;
  if (v != 0) return 0;
  for (int q = 0; q < P.Q; ++q) {
    int32_t kk[3] = {k[0] - P.c[q][0], k[1] - P.c[q][1], k[2] - P.c[q][2]};
    if (interior(P, kk) && g[i - P.shift[q]] == 1) return 2;
  }
  return 0;
}

// Inclusive scan of two counters over the workgroup (NT threads, waves of
// 64); returns the exclusive prefix and sets the workgroup totals.
template <int NT>
__device__ __forceinline__ void block_scan2(uint32_t a, uint32_t b, uint32_t& ea, uint32_t& eb, uint32_t& ta,
                                            uint32_t& tb) {
  __shared__ uint32_t wa[NT / 64], wb[NT / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t ia = a, ib = b;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t xa = __shfl_up(ia, off, 64), xb = __shfl_up(ib, off, 64);
    if (lane >= off) {
      ia += xa;
      ib += xb;
    }
  }
  if (lane == 63) {
    wa[wave] = ia;
    wb[wave] = ib;
  }
  __syncthreads();
  uint32_t pa = 0, pb = 0, sa = 0, sb = 0;
  for (int w = 0; w < NT / 64; ++w) {
    if (w < wave) {
      pa += wa[w];
      pb += wb[w];
    }
    sa += wa[w];
    sb += wb[w];
  }
  ea = pa + ia - a;
  eb = pb + ib - b;
  ta = sa;
  tb = sb;
  __syncthreads();
}

__global__ __launch_bounds__(kTB) void k_flags(const int8_t* __restrict__ g, uint8_t* __restrict__ flags,
                                               uint32_t* __restrict__ tile_cnt, Pad P) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + static_cast<int64_t>(threadIdx.x) * kPer;
  uint32_t nf = 0, nw = 0;
  for (int e = 0; e < kPer; ++e) {
    const int64_t i = base + e;
    if (i >= P.M) break;
    const uint8_t f = cell_flags(g, P, i);
    flags[i] = f;
    nf += f & 1;
    nw += f >> 1;
  }
  uint32_t ea, eb, ta, tb;
  block_scan2<kTB>(nf, nw, ea, eb, ta, tb);
  if (threadIdx.x == 0) {
    tile_cnt[2 * blockIdx.x] = ta;
    tile_cnt[2 * blockIdx.x + 1] = tb;
  }
}

// Exclusive offsets of the tiles; totals[0..1] = fluid and wall cells.
__global__ __launch_bounds__(kScanTB) void k_scan(const uint32_t* __restrict__ tile_cnt,
                                                  uint32_t* __restrict__ tile_off, int64_t ntiles,
                                                  uint32_t* __restrict__ totals) {
  uint32_t ca = 0, cb = 0;
  for (int64_t s = 0; s < ntiles; s += kScanTB) {
    const int64_t t = s + threadIdx.x;
    const uint32_t a = t < ntiles ? tile_cnt[2 * t] : 0u, b = t < ntiles ? tile_cnt[2 * t + 1] : 0u;
    uint32_t ea, eb, ta, tb;
    block_scan2<kScanTB>(a, b, ea, eb, ta, tb);
    if (t < ntiles) {
      tile_off[2 * t] = ca + ea;
      tile_off[2 * t + 1] = cb + eb;
    }
    ca += ta;
    cb += tb;
  }
  if (threadIdx.x == 0) {
    totals[0] = ca;
    totals[1] = cb;
  }
}

__global__ __launch_bounds__(kTB) void k_number(const uint8_t* __restrict__ flags,
                                                const uint32_t* __restrict__ tile_off, uint32_t n_fluid,
                                                int32_t* __restrict__ local, int64_t* __restrict__ cellof, Pad P) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + static_cast<int64_t>(threadIdx.x) * kPer;
  uint32_t nf = 0, nw = 0;
  for (int e = 0; e < kPer; ++e) {
    const int64_t i = base + e;
    if (i >= P.M) break;
    const uint8_t f = flags[i];
    nf += f & 1;
    nw += f >> 1;
  }
  uint32_t ea, eb, ta, tb;
  block_scan2<kTB>(nf, nw, ea, eb, ta, tb);
  uint32_t fnext = tile_off[2 * blockIdx.x] + ea;
  uint32_t wnext = n_fluid + tile_off[2 * blockIdx.x + 1] + eb;
  for (int e = 0; e < kPer; ++e) {
    const int64_t i = base + e;
    if (i >= P.M) break;
    const uint8_t f = flags[i];
    int32_t node = 0;
    if (f & 1) node = static_cast<int32_t>(++fnext);
    else if (f & 2) node = static_cast<int32_t>(++wnext);
    local[i] = node;
    if (node) cellof[node - 1] = i;
  }
}

__global__ __launch_bounds__(kTB) void k_nodes(const int8_t* __restrict__ g, const int32_t* __restrict__ local,
                                               const int64_t* __restrict__ cellof, int32_t npts, Pad P,
                                               int32_t* __restrict__ nbr, int32_t* __restrict__ pos,
                                               int32_t* __restrict__ attr) {
  const int32_t n = static_cast<int32_t>(blockIdx.x) * kTB + static_cast<int32_t>(threadIdx.x) + 1;
  if (n > npts) return;
  const int64_t i = cellof[n - 1];
  int32_t k[3] = {0, 0, 0};
  coords(P, i, k);
  for (int d = 0; d < P.nd; ++d) pos[static_cast<int64_t>(n) * P.nd + d] = k[d] - 1;
  attr[n] = g[i] > 0 ? 1 : 0;
  for (int q = 0; q < P.Q; ++q) {
    int32_t kk[3] = {k[0] + P.c[q][0], k[1] + P.c[q][1], k[2] + P.c[q][2]};
    bool inside = true;
    for (int d = 0; d < P.nd; ++d) inside = inside && kk[d] >= 0 && kk[d] < P.pd[d];
    int32_t v = 0;
    if (inside) {
      const int64_t j = i + P.shift[q];
      v = local[j];
      if (v == 0 && g[j] == 1) {
        // a periodic copy of a fluid cell: the node of the cell it copies
        int64_t jj = 0;
        for (int d = 0; d < P.nd; ++d) {
          int32_t u = kk[d];
          if (u == 0) u = P.pd[d] - 2;
          else if (u == P.pd[d] - 1) u = 1;
          jj = jj * P.pd[d] + u;
        }
        v = local[jj];
      }
    }
    nbr[static_cast<int64_t>(n) * P.Q + q] = v;
  }
}

// Nodes::setupNodeType for one rank: solid 1 next to fluid else 0; fluid 2
// next to a solid node (node 0 counts as solid) else 3.
__global__ __launch_bounds__(kTB) void k_types(const int32_t* __restrict__ nbr, const int32_t* __restrict__ attr,
                                               int32_t n_nodes, int Q, int16_t* __restrict__ type) {
  const int32_t n = static_cast<int32_t>(blockIdx.x) * kTB + static_cast<int32_t>(threadIdx.x);
  if (n >= n_nodes) return;
  if (n == 0) {
    type[0] = -1;
    return;
  }
  const bool fluid = attr[n] != 0;
  bool other = false;
  for (int q = 0; q < Q; ++q) {
    const int32_t m = nbr[static_cast<int64_t>(n) * Q + q];
    const bool mf = m != 0 && attr[m] != 0;
    other = other || (fluid ? !mf : mf);
  }
  type[n] = fluid ? (other ? 2 : 3) : (other ? 1 : 0);
}

struct DevBuf {
  std::vector<void*> p;
  ~DevBuf() {
    for (void* x : p) hipFree(x);
  }
  template <class T>
  hipError_t alloc(T** out, size_t n) {
    void* x = nullptr;
    hipError_t e = hipMalloc(&x, n ? n * sizeof(T) : 1);
    if (e == hipSuccess) p.push_back(x);
    *out = static_cast<T*>(x);
    return e;
  }
};

#define LB_DEV(expr)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ == hipErrorOutOfMemory) return fail_global(LB_ERR_NOMEM, "device builder: out of device memory"); \
    if (e_ != hipSuccess) return fail_global(LB_ERR_HIP, "device builder: %s", hipGetErrorString(e_)); \
  } while (0)

template <class L>
void fill_lattice(Pad& P) {
  P.Q = L::nQ;
  for (int q = 0; q < L::nQ; ++q) {
    P.shift[q] = 0;
    for (int d = 0; d < 3; ++d) P.c[q][d] = d < L::nD ? L::c(q, d) : 0;
    for (int d = 0; d < L::nD; ++d) P.shift[q] += P.c[q][d] * P.stride[d];
  }
}

int build_device(int device, int lattice, int nd, const int32_t* dims, const int32_t* geo, int periodic,
                 lb_topo& t) {
  Pad P{};
  P.nd = nd;
  P.periodic = periodic;
  int64_t n_in = 1;
  for (int d = 0; d < nd; ++d) {
    P.dims[d] = dims[d];
    P.pd[d] = dims[d] + 2;
    n_in *= dims[d];
  }
  P.M = 1;
  for (int d = nd - 1; d >= 0; --d) {
    P.stride[d] = P.M;
    P.M *= P.pd[d];
  }
  if (lattice == LB_D2Q9) fill_lattice<lbamd::D2Q9>(P);
  else fill_lattice<lbamd::D3Q19>(P);
  const int Q = P.Q;

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail_global(LB_ERR_HIP, "device builder: no HIP device");
  if (device < 0 || device >= ndev) return fail_global(LB_ERR_ARG, "device builder: no device %d", device);
  LB_DEV(hipSetDevice(device));
  DevBuf B;
  int32_t* d_geo;
  int8_t* d_g;
  uint8_t* d_flags;
  uint32_t *d_cnt, *d_off, *d_tot;
  int* d_bad;
  const int64_t ntiles = (P.M + kTile - 1) / kTile;
  LB_DEV(B.alloc(&d_geo, n_in));
  LB_DEV(B.alloc(&d_g, P.M));
  LB_DEV(B.alloc(&d_flags, P.M));
  LB_DEV(B.alloc(&d_cnt, 2 * ntiles));
  LB_DEV(B.alloc(&d_off, 2 * ntiles));
  LB_DEV(B.alloc(&d_tot, 2));
  LB_DEV(B.alloc(&d_bad, 1));
  LB_DEV(hipMemcpy(d_geo, geo, n_in * sizeof(int32_t), hipMemcpyHostToDevice));
  LB_DEV(hipMemset(d_bad, 0, sizeof(int)));
  const int64_t gpad = std::min<int64_t>((P.M + kTB - 1) / kTB, 1 << 16);
  k_pad<<<dim3(static_cast<uint32_t>(gpad)), dim3(kTB)>>>(d_geo, d_g, P, d_bad);
  k_flags<<<dim3(static_cast<uint32_t>(ntiles)), dim3(kTB)>>>(d_g, d_flags, d_cnt, P);
  k_scan<<<dim3(1), dim3(kScanTB)>>>(d_cnt, d_off, ntiles, d_tot);
  LB_DEV(hipGetLastError());
  uint32_t tot[2] = {0, 0};
  int bad = 0;
  LB_DEV(hipMemcpy(tot, d_tot, sizeof(tot), hipMemcpyDeviceToHost));
  LB_DEV(hipMemcpy(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost));
  if (bad & 1) return fail_global(LB_ERR_ARG, "geometry values must be >= 0");
  if (bad & 2)
    return fail_global(LB_ERR_UNSUPPORTED,
                       "device builder: single-rank geometries only (values 0 and 1); use lb_topo_from_geometry");
  if (tot[0] == 0) return fail_global(LB_ERR_ARG, "rank 1 not present in the geometry");
  const int64_t npts64 = static_cast<int64_t>(tot[0]) + tot[1];
  if (npts64 >= (int64_t(1) << 31) - 1) return fail_global(LB_ERR_ARG, "device builder: more than 2^31 nodes");
  const int32_t npts = static_cast<int32_t>(npts64), N = npts + 1;

  int32_t *d_local, *d_nbr, *d_pos, *d_attr;
  int64_t* d_cellof;
  int16_t* d_type;
  LB_DEV(B.alloc(&d_local, P.M));
  LB_DEV(B.alloc(&d_cellof, npts));
  LB_DEV(B.alloc(&d_nbr, static_cast<size_t>(N) * Q));
  LB_DEV(B.alloc(&d_pos, static_cast<size_t>(N) * nd));
  LB_DEV(B.alloc(&d_attr, N));
  LB_DEV(B.alloc(&d_type, N));
  LB_DEV(hipMemset(d_nbr, 0, sizeof(int32_t) * Q));
  LB_DEV(hipMemset(d_pos, 0xff, sizeof(int32_t) * nd));
  LB_DEV(hipMemset(d_attr, 0, sizeof(int32_t)));
  k_number<<<dim3(static_cast<uint32_t>(ntiles)), dim3(kTB)>>>(d_flags, d_off, tot[0], d_local, d_cellof, P);
  k_nodes<<<dim3(static_cast<uint32_t>((npts + kTB - 1) / kTB)), dim3(kTB)>>>(d_g, d_local, d_cellof, npts, P,
                                                                              d_nbr, d_pos, d_attr);
  k_types<<<dim3(static_cast<uint32_t>((N + kTB - 1) / kTB)), dim3(kTB)>>>(d_nbr, d_attr, N, Q, d_type);
  LB_DEV(hipGetLastError());

  t.lattice = lattice;
  t.nD = nd;
  t.Q = Q;
  t.rank = 0;
  t.zero_ghost = true;
  t.global_dims.assign(P.pd, P.pd + nd);
  t.n_nodes = N;
  t.pos.resize(static_cast<size_t>(N) * nd);
  t.neighbor.resize(static_cast<size_t>(N) * Q);
  t.nodetype_attr.resize(N);
  t.node_type.resize(N);
  LB_DEV(hipMemcpy(t.neighbor.data(), d_nbr, t.neighbor.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  LB_DEV(hipMemcpy(t.pos.data(), d_pos, t.pos.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  LB_DEV(hipMemcpy(t.nodetype_attr.data(), d_attr, N * sizeof(int32_t), hipMemcpyDeviceToHost));
  LB_DEV(hipMemcpy(t.node_type.data(), d_type, N * sizeof(int16_t), hipMemcpyDeviceToHost));
  t.node_rank.assign(N, 0);
  t.node_rank[0] = -1;
  // findBulkNodes: the own fluid nodes, numbered first
  t.bulk.resize(tot[0]);
  for (uint32_t k = 0; k < tot[0]; ++k) t.bulk[k] = static_cast<int32_t>(k + 1);
  t.processors.clear();
  t.attributes.clear();
  return LB_OK;
}

}  // namespace

extern "C" int lb_topo_from_geometry_device(int device, int lattice, int32_t nd, const int32_t* dims,
                                            const int32_t* geo, int32_t periodic, lb_topo** out) {
  if (!dims || !geo || !out || (lattice != LB_D2Q9 && lattice != LB_D3Q19))
    return fail_global(LB_ERR_ARG, "lb_topo_from_geometry_device: bad argument");
  if (nd != (lattice == LB_D2Q9 ? 2 : 3)) return fail_global(LB_ERR_ARG, "nd does not match lattice");
  for (int d = 0; d < nd; ++d)
    if (dims[d] < 1) return fail_global(LB_ERR_ARG, "dims must be positive");
  try {
    auto t = std::make_unique<lb_topo>();
    const int rc = build_device(device, lattice, nd, dims, geo, periodic, *t);
    if (rc != LB_OK) return rc;
    *out = t.release();
  } catch (const std::bad_alloc&) {
    return fail_global(LB_ERR_NOMEM, "lb_topo_from_geometry_device: out of host memory");
  }
  return LB_OK;
}
