// Where a bulk node pulls each population from: three address layouts.
//
// Every layout provides sources<L>(i, src): the population element index
// (pidx, context.hpp) of the post-collision value that becomes f_q of node i,
// for q = 0..Q-2 (the rest direction always reads itself).  A stream link
// gives (q, j) with j the compact index of the upwind neighbour; a link from
// a solid gives (rev q, i), the half-way bounce-back.  Each layout first
// issues all its addressing loads and then does the arithmetic, so that the
// loads are in flight together.
//
// TableAddr  one 32-bit link per (direction, node): kind + index (see
//            context.hpp).  Works for any node list, including fluid links to
//            the ghost node 0.  4 B per link per step.
//
// RowAddr    the bulk nodes are cells of a Cartesian box, cut into rows along
//            the last axis (z in 3-D, y in 2-D).  Per row and 32-cell word it
//            keeps {base, mask}: the fluid bit mask and the compact index of
//            the word's first fluid cell.  A node stores only its packed cell
//            coordinate (4 B); the source cell r - c_q (periodic axes wrap) is
//            fluid if its bit is set, and then its compact index is base +
//            popcount(mask below the bit).  The table is a few MB and stays in
//            L2 / MALL.  One lookup per direction.
//
// RowWin     as RowAddr for boxes that are not periodic along the row axis:
//            per row and word w it keeps a 64-bit window of the row's fluid
//            bits over cells [32w-16, 32w+48) and the compact index of the
//            window's first fluid cell.  The node's own word then covers the
//            cells z-1, z, z+1 of every neighbouring row, so the 9 (D3Q19) or
//            3 (D2Q9) neighbouring rows are one 16-byte load each, and the
//            three sources in a row follow from one popcount: with b0 the
//            index of the first fluid cell at or after z, the source at z is
//            b0, at z-1 it is b0-1, at z+1 it is b0 + bit(z).  The window
//            table has a halo of rows, so no wrap logic runs per node.
//
// Rows are ordered in tiles of 2^lx x 2^ly rows, which keeps the neighbour
// rows of a tile close in the compact order.  The row layouts are validated on
// the host against the neighbour index and switched on only when every link
// they compute equals the table's.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "context.hpp"

namespace lbamd {

// Block order: the hardware hands consecutive workgroups to the eight XCDs
// round-robin.  node_block() gives each XCD runs of kXcdChunk consecutive
// 256-node blocks, so neighbouring rows are mostly read through one L2, while
// all CUs still sweep one compact front whose halo stays in the shared MALL.
// On the 256^3 pack runs of 4 blocks are 2.4 % faster than runs of 1 (the
// plain order) in FP32S and 2 % in FP64; runs of 8-32 give about the same,
// one contiguous range per XCD (eight fronts) is 3 % slower.  The batched
// kernels map their workgroups the same way (+0.7 % on the 512-solve
// ensemble).  The mapping only changes which workgroup computes a block:
// per-block partial sums keep their block index, so results do not change.
constexpr int32_t kXcdChunk = 4;
__device__ __forceinline__ int32_t node_block(int32_t b, int32_t nb) {
  const int32_t full = (nb / (8 * kXcdChunk)) * (8 * kXcdChunk);
  if (b >= full) return b;  // the tail keeps the plain order
  const int32_t xcd = b & 7, k = b >> 3;
  return ((k / kXcdChunk) * 8 + xcd) * kXcdChunk + (k % kXcdChunk);
}

// The same mapping with the dispatch order reversed: a pass that follows
// another one over the same nodes starts where the first one ended, so its
// first reads find the lines that the first pass fetched last still in the
// MALL.  The two-phase collision pass runs this way after the colour pass.
__device__ __forceinline__ int32_t node_block_rev(int32_t b, int32_t nb) { return node_block(nb - 1 - b, nb); }

// Element indices are 32-bit: a context keeps Q * NP * fields below 2^32.
// kOff32: the layout is only used for population buffers below 4 GiB, so the
// kernels may address them with 32-bit byte offsets from the buffer base.
struct TableAddr {
  static constexpr bool kOff32 = false;
  const uint32_t* __restrict__ link;
  uint32_t NP;
  template <class L>
  __device__ __forceinline__ void sources(int32_t i, uint32_t (&src)[L::nQ - 1]) const {
    uint32_t l[L::nQ - 1];
#pragma unroll
    for (int q = 0; q < L::nQ - 1; ++q) l[q] = link[static_cast<uint64_t>(q) * NP + i];
#pragma unroll
    for (int q = 0; q < L::nQ - 1; ++q) {
      const uint32_t kind = l[q] >> 30;
      const uint32_t j = l[q] & kLinkIndexMask;
      // stream: (q, j); bounce: (rev q, i); ghost 0: (rev q, j) -- one index
      // computation on selected operands, no branch
      const uint32_t jj = kind == kLinkBounce ? static_cast<uint32_t>(i) : j;
      const uint32_t qq = kind == kLinkStream ? static_cast<uint32_t>(q) : static_cast<uint32_t>(L::rev(q));
      src[q] = pidx<L::nQ>(qq, jj);
    }
  }
};

struct RowGeom {
  int32_t E[3];      // box extent along X, Y, Z (2-D: X = x, Y = 1, Z = y)
  int32_t periodic;  // bit 0 X, bit 1 Y, bit 2 Z
  int32_t lx, ly;    // tile of 2^lx x 2^ly rows
  int32_t nTy, W;    // tiles along Y, 32-bit words per row
  // row index of cell column (X, Y)
  __host__ __device__ __forceinline__ int32_t row(int32_t X, int32_t Y) const {
    return (((X >> lx) * nTy + (Y >> ly)) << (lx + ly)) | ((X & ((1 << lx) - 1)) << ly) |
           (Y & ((1 << ly) - 1));
  }
  // wrap one coordinate into [0, E); false when it leaves a non-periodic box
  __host__ __device__ static __forceinline__ bool wrap(int32_t& v, int32_t E, bool periodic) {
    if (v >= 0 && v < E) return true;
    v += v < 0 ? E : -E;
    return periodic;
  }
};

// lattice direction components mapped to (X, Y, Z)
template <class L>
__host__ __device__ __forceinline__ constexpr int cx(int q) { return L::c(q, 0); }
template <class L>
__host__ __device__ __forceinline__ constexpr int cy(int q) { return L::nD == 3 ? L::c(q, 1) : 0; }
template <class L>
__host__ __device__ __forceinline__ constexpr int cz(int q) { return L::nD == 3 ? L::c(q, 2) : L::c(q, 1); }

struct RowAddr {
  static constexpr bool kOff32 = false;
  const uint32_t* __restrict__ coord;  // X << 22 | Y << 12 | Z
  const uint2* __restrict__ rows;      // {base, mask} per (row, word)
  uint32_t NP;
  RowGeom g;
  template <class L>
  __device__ __forceinline__ void sources(int32_t i, uint32_t (&src)[L::nQ - 1]) const {
    const uint32_t c = coord[i];
    const int32_t X0 = static_cast<int32_t>(c >> 22), Y0 = static_cast<int32_t>((c >> 12) & 1023u),
                  Z0 = static_cast<int32_t>(c & 4095u);
    uint2 e[L::nQ - 1];
    uint32_t bit[L::nQ - 1];
#pragma unroll
    for (int q = 0; q < L::nQ - 1; ++q) {
      int32_t X = X0 - cx<L>(q), Y = Y0 - cy<L>(q), Z = Z0 - cz<L>(q);
      bool inside = true;
      if (cx<L>(q) != 0) inside = RowGeom::wrap(X, g.E[0], g.periodic & 1) && inside;
      if (cy<L>(q) != 0) inside = RowGeom::wrap(Y, g.E[1], g.periodic & 2) && inside;
      if (cz<L>(q) != 0) inside = RowGeom::wrap(Z, g.E[2], g.periodic & 4) && inside;
      const int64_t slot = inside ? static_cast<int64_t>(g.row(X, Y)) * g.W + (Z >> 5) : 0;
      e[q] = rows[slot];
      bit[q] = inside ? (1u << (Z & 31)) : 0u;
    }
#pragma unroll
    for (int q = 0; q < L::nQ - 1; ++q) {
      const uint32_t m = 0u - static_cast<uint32_t>((e[q].y & bit[q]) != 0u);
      const uint32_t j = e[q].x + __popc(e[q].y & (bit[q] - 1u));
      src[q] = (pidx<L::nQ>(q, j) & m) | (pidx<L::nQ>(L::rev(q), i) & ~m);
    }
  }
};

struct RowWin {
  static constexpr bool kOff32 = true;
  const uint32_t* __restrict__ coord;  // X << 22 | Y << 12 | Z
  const uint4* __restrict__ win;       // {base, mask lo, mask hi, 0} per (padded row, word)
  uint32_t NP;
  int32_t PY;  // padded rows along Y: E_Y + 2
  int32_t W;   // words per row

  // The window table is indexed by the halo-padded row (X+1, Y+1): halo rows
  // of a periodic axis repeat the rows they wrap to, those of a closed axis
  // are empty (all solid).  The 9 (D3Q19) or 3 (D2Q9) neighbouring rows of a
  // node are then one vector offset plus a per-direction uniform base.
  template <class L>
  __device__ __forceinline__ void sources(int32_t i, uint32_t (&src)[L::nQ - 1]) const {
    const uint32_t c = coord[i];
    const uint32_t X = c >> 22, Y = (c >> 12) & 1023u, Z = c & 4095u;
    const uint32_t w = Z >> 5;
    const uint32_t t = Z - 32u * w + 16u;  // bit of z in the window, 16..47
    const uint32_t base = ((X + 1u) * static_cast<uint32_t>(PY) + (Y + 1u)) * static_cast<uint32_t>(W) + w;
    constexpr int NY = L::nD == 3 ? 3 : 1;
    uint4 e[3 * NY];
#pragma unroll
    for (int k = 0; k < 3 * NY; ++k) {
      const int dx = k / NY - 1, dy = NY > 1 ? k % NY - 1 : 0;
      const uint4* __restrict__ wk = win + (dx * PY + dy) * W;
      e[k] = *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(wk) + base * 16u);
    }
    // per row: index of the first fluid cell at or after z, fluid bits of z-1, z, z+1
    uint32_t b0[3 * NY];
    uint32_t bm[3 * NY];
    const uint64_t below = (1ull << t) - 1ull;
#pragma unroll
    for (int k = 0; k < 3 * NY; ++k) {
      const uint64_t m = (static_cast<uint64_t>(e[k].z) << 32) | e[k].y;
      b0[k] = e[k].x + static_cast<uint32_t>(__popcll(m & below));
      bm[k] = static_cast<uint32_t>(m >> (t - 1u)) & 7u;  // bit0: z-1, bit1: z, bit2: z+1
    }
#pragma unroll
    for (int q = 0; q < L::nQ - 1; ++q) {
      // source cell r - c_q
      const int k = (1 - cx<L>(q)) * NY + (NY > 1 ? 1 - cy<L>(q) : 0);
      const int dz = -cz<L>(q);
      const uint32_t fluid = (bm[k] >> (dz + 1)) & 1u;
      const uint32_t j = dz == 0 ? b0[k] : (dz < 0 ? b0[k] - 1u : b0[k] + ((bm[k] >> 1) & 1u));
      // both candidates, combined by a mask: a per-lane select, not a branch
      const uint32_t m = 0u - fluid;
      src[q] = (pidx<L::nQ>(q, j) & m) | (pidx<L::nQ>(L::rev(q), i) & ~m);
    }
  }
};

// RowWin::links: the source node of every direction: bit q of fluid is set
// when the source cell of direction q is the fluid node j[q] (else a solid,
// and j[q] is undefined).
template <class L>
__device__ __forceinline__ void rowwin_links(const RowWin& ad, int32_t i, uint32_t (&j)[L::nQ - 1], uint32_t& fluid) {
  const uint32_t c = ad.coord[i];
  const uint32_t X = c >> 22, Y = (c >> 12) & 1023u, Z = c & 4095u;
  const uint32_t w = Z >> 5;
  const uint32_t t = Z - 32u * w + 16u;
  const uint32_t base = ((X + 1u) * static_cast<uint32_t>(ad.PY) + (Y + 1u)) * static_cast<uint32_t>(ad.W) + w;
  constexpr int NY = L::nD == 3 ? 3 : 1;
  uint4 e[3 * NY];
#pragma unroll
  for (int k = 0; k < 3 * NY; ++k) {
    const int dx = k / NY - 1, dy = NY > 1 ? k % NY - 1 : 0;
    const uint4* __restrict__ wk = ad.win + (dx * ad.PY + dy) * ad.W;
    e[k] = *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(wk) + base * 16u);
  }
  const uint64_t below = (1ull << t) - 1ull;
  fluid = 0u;
#pragma unroll
  for (int q = 0; q < L::nQ - 1; ++q) {
    const int k = (1 - cx<L>(q)) * NY + (NY > 1 ? 1 - cy<L>(q) : 0);
    const int dz = -cz<L>(q);
    const uint64_t m = (static_cast<uint64_t>(e[k].z) << 32) | e[k].y;
    const uint32_t b0 = e[k].x + static_cast<uint32_t>(__popcll(m & below));
    const uint32_t bm = static_cast<uint32_t>(m >> (t - 1u)) & 7u;
    fluid |= ((bm >> (dz + 1)) & 1u) << q;
    j[q] = dz == 0 ? b0 : (dz < 0 ? b0 - 1u : b0 + ((bm >> 1) & 1u));
  }
}

// Colour of the neighbours neighbor(qq, n) for the colour gradient
// (LButilities.h:12-22): a fluid neighbour's phi; a solid neighbour's wall
// colour, per wall slot with the link table or one value for the whole node
// list with the row windows; 0 for the ghost node 0.  pn[Q-1] = 0.
struct WallPhi {
  const double* slot;  // per wall slot (link table)
  double uniform;      // every solid neighbour (row windows)
};

template <class L, typename T>
__device__ __forceinline__ void neighbour_phi(const TableAddr& ad, int32_t i, const T* __restrict__ phi,
                                              const WallPhi& wp, T (&pn)[L::nQ]) {
#pragma unroll
  for (int qq = 0; qq < L::nQ - 1; ++qq) {
    const int q = L::rev(qq);  // neighbor(qq, n) is the stream source of rev(qq)
    const uint32_t l = ad.link[static_cast<uint32_t>(q) * ad.NP + i];
    const uint32_t kind = l >> 30;
    const uint32_t j = l & kLinkIndexMask;
    pn[qq] = kind == kLinkStream ? phi[j] : (kind == kLinkBounce ? static_cast<T>(wp.slot[j]) : static_cast<T>(0));
  }
  pn[L::nQ - 1] = static_cast<T>(0);
}

template <class L, typename T>
__device__ __forceinline__ void neighbour_phi(const RowWin& ad, int32_t i, const T* __restrict__ phi,
                                              const WallPhi& wp, T (&pn)[L::nQ]) {
  uint32_t j[L::nQ - 1], fluid;
  rowwin_links<L>(ad, i, j, fluid);
  const T pw = static_cast<T>(wp.uniform);
#pragma unroll
  for (int qq = 0; qq < L::nQ - 1; ++qq) {
    const int q = L::rev(qq);
    // load from the node itself where the neighbour is solid (always in
    // bounds), then select: no branch
    const bool f = (fluid >> q) & 1u;
    const T v = phi[f ? j[q] : static_cast<uint32_t>(i)];
    pn[qq] = f ? v : pw;
  }
  pn[L::nQ - 1] = static_cast<T>(0);
}

// Element access of a population buffer: 32-bit byte offsets from the base
// when the layout guarantees the buffer is below 4 GiB (global loads with a
// scalar base and a 32-bit vector offset), 64-bit addresses otherwise.
template <bool OFF32, typename T>
__device__ __forceinline__ T ld_elem(const T* __restrict__ base, uint32_t e) {
  if (OFF32) return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + e * static_cast<uint32_t>(sizeof(T)));
  return base[e];
}
template <bool OFF32, typename T>
__device__ __forceinline__ T* elem_ptr(T* __restrict__ base, uint32_t e) {
  if (OFF32) return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + e * static_cast<uint32_t>(sizeof(T)));
  return base + e;
}

}  // namespace lbamd
