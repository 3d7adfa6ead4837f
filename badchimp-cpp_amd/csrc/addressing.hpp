// Where a bulk node pulls each population from: two address layouts.
//
// TableAddr  one 32-bit link per (direction, node): kind + index (see
//            context.hpp).  Works for any node list, including fluid links to
//            the ghost node 0.  Costs 4 B per link per step.
//
// RowAddr    the Cartesian box of the bulk nodes is cut into rows along the
//            last axis (z in 3-D, y in 2-D).  Per row and 32-cell word it
//            keeps {base, mask}: the fluid bit mask and the compact index of
//            the word's first fluid cell.  A node stores only its packed
//            cell coordinate (4 B); the source of direction q is the cell
//            r - c_q (periodic axes wrap): if its bit is set the source is
//            compact index base + popcount(mask below the bit) (stream),
//            otherwise the node's own reverse population (bounce-back).  The
//            row table is a few MB and stays in L2 / MALL, so the index
//            traffic per node and step drops from 72 B (D3Q19) to 4 B.
//            Rows are ordered in tiles of TX x TY rows, which keeps the 18
//            neighbour rows of a tile close in the compact order.
//
// Both are validated on the host against the neighbour index: RowAddr is
// switched on only when every link it computes equals the table's.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "context.hpp"

namespace lbamd {

// Logical block of a launch, so that each XCD (blocks are dispatched
// round-robin over the 8 XCDs) works on one contiguous range of nodes and its
// L2 sees the neighbours' rows.
__device__ __forceinline__ int32_t xcd_block(int32_t b, int32_t nblocks) {
  const int32_t q = nblocks >> 3, r = nblocks & 7;
  const int32_t xcd = b & 7, slot = b >> 3;
  return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// Each layout splits the address computation in two phases, so that a
// kernel first issues all Q-1 loads of addressing data (pre) and only then
// the Q-1 population gathers (src): both batches are in flight together
// instead of one dependent pair at a time.
struct TableAddr {
  const uint32_t* __restrict__ link;
  int64_t NP;
  struct Node {};
  using Pre = uint32_t;
  __device__ __forceinline__ Node node(int32_t) const { return {}; }
  template <class L>
  __device__ __forceinline__ Pre pre(int q, int32_t i, Node) const {
    return link[q * NP + i];
  }
  template <class L>
  __device__ __forceinline__ int64_t src(int q, int32_t i, Node, Pre l) const {
    const uint32_t kind = l >> 30;
    const int64_t j = static_cast<int64_t>(l & kLinkIndexMask);
    const int rq = L::rev(q);
    return kind == kLinkStream ? q * NP + j : (kind == kLinkBounce ? rq * NP + i : rq * NP + j);
  }
  template <class L>
  __device__ __forceinline__ int64_t src(int q, int32_t i, Node n) const {
    return src<L>(q, i, n, pre<L>(q, i, n));
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
};

struct RowAddr {
  const uint32_t* __restrict__ coord;  // X << 22 | Y << 12 | Z
  const uint2* __restrict__ rows;      // {base, mask} per (row, word)
  int64_t NP;
  RowGeom g;
  struct Node {
    int32_t X, Y, Z;
  };
  __device__ __forceinline__ Node node(int32_t i) const {
    const uint32_t c = coord[i];
    return {static_cast<int32_t>(c >> 22), static_cast<int32_t>((c >> 12) & 1023u),
            static_cast<int32_t>(c & 4095u)};
  }
  // lattice direction components mapped to (X, Y, Z)
  template <class L>
  static __device__ __forceinline__ int cx(int q) { return L::c(q, 0); }
  template <class L>
  static __device__ __forceinline__ int cy(int q) { return L::nD == 3 ? L::c(q, 1) : 0; }
  template <class L>
  static __device__ __forceinline__ int cz(int q) { return L::nD == 3 ? L::c(q, 2) : L::c(q, 1); }

  // wrap one coordinate; false when it leaves a non-periodic box
  static __device__ __forceinline__ bool wrap(int32_t& v, int32_t E, bool periodic) {
    if (v >= 0 && v < E) return true;
    v += v < 0 ? E : -E;
    return periodic;
  }
  // pre: {base, mask restricted to the source bit} of the source cell's word
  // (mask 0 when the cell is solid or outside the box)
  using Pre = uint2;
  template <class L>
  __device__ __forceinline__ Pre pre(int q, int32_t, Node n) const {
    int32_t X = n.X - cx<L>(q), Y = n.Y - cy<L>(q), Z = n.Z - cz<L>(q);
    bool inside = true;
    if (cx<L>(q) != 0) inside = wrap(X, g.E[0], g.periodic & 1) && inside;
    if (cy<L>(q) != 0) inside = wrap(Y, g.E[1], g.periodic & 2) && inside;
    if (cz<L>(q) != 0) inside = wrap(Z, g.E[2], g.periodic & 4) && inside;
    const int64_t slot = inside ? static_cast<int64_t>(g.row(X, Y)) * g.W + (Z >> 5) : 0;
    const uint2 e = rows[slot];
    const uint32_t bit = inside ? (1u << (Z & 31)) : 0u;
    // {compact index if the cell is fluid, fluid flag}
    return uint2{e.x + __popc(e.y & (bit - 1u)), (e.y & bit) ? 1u : 0u};
  }
  template <class L>
  __device__ __forceinline__ int64_t src(int q, int32_t i, Node, Pre p) const {
    return p.y ? q * NP + static_cast<int64_t>(p.x) : L::rev(q) * NP + i;
  }
  template <class L>
  __device__ __forceinline__ int64_t src(int q, int32_t i, Node n) const {
    return src<L>(q, i, n, pre<L>(q, i, n));
  }
};

}  // namespace lbamd
