// Host-side construction of the row layouts (addressing.hpp: RowAddr, RowWin).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/badchimp_amd.h"
#include "addressing.hpp"
#include "lattice.hpp"

namespace lbamd {

// Rows are ordered in tiles of 2^kTileLogX x 2^kTileLogY (X, Y) columns.
// Runs of 32 rows along Y measured best on the 256^3 pack (33.0 GLUPS FP32S,
// against 31.6 for 8 x 8 tiles with blocks remapped per XCD and 32.9 with no
// tiles); the differences between tile shapes are small.
constexpr int kTileLogX = 0;
constexpr int kTileLogY = 5;

// Cartesian row layout of the bulk nodes (see addressing.hpp): the box, the
// compact order (tiled row order), the packed cell of each compact node and
// {base, mask} per (row, 32-cell word).
struct RowBuild {
  RowGeom g{};
  int32_t lo[3] = {0, 0, 0};    // rows_from_geometry: geometry cell of box coordinate 0
  std::vector<int32_t> order;   // compact index -> bulk index
  std::vector<uint32_t> coord;  // compact index -> packed cell
  std::vector<uint2> rows;
};

// The row layout straight from a geometry (C order over x, y[, z]; cells > 0
// are fluid), without a neighbour index: the fluid cells in tiled row order,
// which is the order build_rows gives the node list that the vtklb writer
// makes of the same geometry.  Periodic axes span the whole box; closed axes
// span the fluid cells' extent.
inline bool rows_from_geometry(int lattice, const int32_t* dims, const int32_t* geo, int32_t periodic_mask,
                               RowBuild& rb, std::string& why) {
  const int nD = lattice == LB_D2Q9 ? 2 : 3;
  // geometry axis of row axis a (X, Y, Z), -1 for the missing Y of 2-D
  const int ax[3] = {0, nD == 3 ? 1 : -1, nD == 3 ? 2 : 1};
  int64_t ncell = 1;
  for (int d = 0; d < nD; ++d) {
    if (dims[d] < 1) { why = "empty geometry"; return false; }
    ncell *= dims[d];
  }
  int32_t D[3], lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    D[a] = ax[a] < 0 ? 1 : dims[ax[a]];
    lo[a] = D[a];
    hi[a] = -1;
  }
  const int64_t sy = nD == 3 ? dims[2] : 1, sx = nD == 3 ? static_cast<int64_t>(dims[1]) * dims[2] : dims[1];
  auto at = [&](int32_t X, int32_t Y, int32_t Z) -> int64_t {  // geometry index of row-axis cell
    return nD == 3 ? X * sx + Y * sy + Z : X * sx + Z;
  };
  int64_t nfluid = 0;
  for (int32_t x = 0; x < D[0]; ++x)
    for (int32_t y = 0; y < D[1]; ++y)
      for (int32_t z = 0; z < D[2]; ++z)
        if (geo[at(x, y, z)] > 0) {
          ++nfluid;
          const int32_t p[3] = {x, y, z};
          for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
          }
        }
  if (nfluid == 0) { why = "no fluid cells"; return false; }
  for (int a = 0; a < 3; ++a) rb.lo[a] = lo[a];
  RowGeom& g = rb.g;
  g.periodic = 0;
  for (int a = 0; a < 3; ++a) {
    const bool per = ax[a] >= 0 && ((periodic_mask >> ax[a]) & 1);
    g.E[a] = per ? D[a] : hi[a] - lo[a] + 1;
    if (per) g.periodic |= 1 << a;
  }
  if (g.E[0] > 1024 || g.E[1] > 1024 || g.E[2] > 4096) { why = "box too large for packed coordinates"; return false; }
  g.lx = kTileLogX;
  g.ly = g.E[1] >= (1 << kTileLogY) ? kTileLogY : 0;
  const int32_t nTx = (g.E[0] + (1 << g.lx) - 1) >> g.lx;
  g.nTy = (g.E[1] + (1 << g.ly) - 1) >> g.ly;
  g.W = (g.E[2] + 31) / 32;
  const int64_t R = static_cast<int64_t>(nTx) * g.nTy << (g.lx + g.ly);
  rb.order.clear();
  rb.coord.clear();
  rb.coord.reserve(static_cast<size_t>(nfluid));
  rb.rows.assign(static_cast<size_t>(R) * g.W, uint2{0u, 0u});
  // cell of the box coordinate v along row axis a
  auto cellc = [&](int a, int32_t v) { return (g.periodic >> a) & 1 ? (v + lo[a]) % D[a] : v + lo[a]; };
  for (int32_t tx = 0; tx < nTx; ++tx)
    for (int32_t ty = 0; ty < g.nTy; ++ty)
      for (int32_t ix = 0; ix < (1 << g.lx); ++ix)
        for (int32_t iy = 0; iy < (1 << g.ly); ++iy) {
          const int32_t X = (tx << g.lx) | ix, Y = (ty << g.ly) | iy;
          if (X >= g.E[0] || Y >= g.E[1]) continue;
          const size_t r = static_cast<size_t>(g.row(X, Y));
          const int32_t cx0 = cellc(0, X), cy0 = cellc(1, Y);
          for (int32_t Z = 0; Z < g.E[2]; ++Z) {
            if (geo[at(cx0, cy0, cellc(2, Z))] <= 0) continue;
            rb.coord.push_back((static_cast<uint32_t>(X) << 22) | (static_cast<uint32_t>(Y) << 12) |
                               static_cast<uint32_t>(Z));
            rb.rows[r * g.W + (Z >> 5)].y |= 1u << (Z & 31);
          }
        }
  uint32_t run = 0;
  for (auto& e : rb.rows) {
    e.x = run;
    run += static_cast<uint32_t>(__builtin_popcount(e.y));
  }
  return true;
}

// Row windows (addressing.hpp, RowWin): for the halo-padded row (X+1, Y+1)
// and word w, the fluid bits of cells [32w-16, 32w+48) and the compact index
// of the first fluid cell at or after 32w-16.  Halo rows of a periodic axis
// repeat the rows they wrap to; those of a closed axis stay empty.  Needs a
// box that is not periodic along the row axis.
inline bool build_windows(const RowBuild& rb, std::vector<uint4>& win, int32_t& PY, std::string& why) {
  const RowGeom& g = rb.g;
  if (g.periodic & 4) {
    why = "periodic along the row axis";
    return false;
  }
  const int32_t PX = g.E[0] + 2;
  PY = g.E[1] + 2;
  win.assign(static_cast<size_t>(PX) * PY * g.W, uint4{0u, 0u, 0u, 0u});
  auto first_at_or_after = [&](size_t r, int32_t z) -> uint32_t {
    // compact index of the first fluid cell of row r at or after z (or the
    // index one past the row's last fluid cell)
    const int32_t zc = std::max(z, 0);
    if (zc >= 32 * g.W) {
      const uint2 e = rb.rows[r * g.W + g.W - 1];
      return e.x + static_cast<uint32_t>(__builtin_popcount(e.y));
    }
    const uint2 e = rb.rows[r * g.W + (zc >> 5)];
    return e.x + static_cast<uint32_t>(__builtin_popcount(e.y & ((1u << (zc & 31)) - 1u)));
  };
  for (int32_t px = 0; px < PX; ++px)
    for (int32_t py = 0; py < PY; ++py) {
      int32_t X = px - 1, Y = py - 1;
      bool inside = RowGeom::wrap(X, g.E[0], g.periodic & 1);
      inside = RowGeom::wrap(Y, g.E[1], g.periodic & 2) && inside;
      if (!inside) continue;
      const size_t r = static_cast<size_t>(g.row(X, Y));
      const uint2* words = &rb.rows[r * g.W];
      for (int32_t w = 0; w < g.W; ++w) {
        // cells 32w - 16 .. 32w + 47 of the row: the upper half of word
        // w - 1, word w, the lower half of word w + 1 (no cells before z = 0
        // or after the row's last word; a word holds no bits beyond E_Z)
        const uint64_t prev = w > 0 ? words[w - 1].y : 0u;
        const uint64_t cur = words[w].y;
        const uint64_t next = w + 1 < g.W ? words[w + 1].y : 0u;
        const uint64_t m = (prev >> 16) | (cur << 16) | (next << 48);
        win[(static_cast<size_t>(px) * PY + py) * g.W + w] =
            uint4{first_at_or_after(r, 32 * w - 16), static_cast<uint32_t>(m), static_cast<uint32_t>(m >> 32), 0u};
      }
    }
  return true;
}

// Source of direction q for compact node i under the row windows, as RowWin
// computes it on the device.
inline std::pair<bool, int64_t> win_source(int lattice, const RowBuild& rb, const std::vector<uint4>& win, int32_t PY,
                                    int q, int32_t i) {
  const int nD = lattice == LB_D2Q9 ? 2 : 3;
  auto C = [&](int a) -> int {
    const int d = nD == 3 ? a : (a == 0 ? 0 : (a == 2 ? 1 : -1));
    if (d < 0) return 0;
    return lattice == LB_D2Q9 ? D2Q9::c(q, d) : D3Q19::c(q, d);
  };
  const RowGeom& g = rb.g;
  const uint32_t cc = rb.coord[i];
  const int32_t X = static_cast<int32_t>(cc >> 22), Y = static_cast<int32_t>((cc >> 12) & 1023u),
                Z = static_cast<int32_t>(cc & 4095u);
  const int32_t w = Z >> 5;
  const uint32_t t = static_cast<uint32_t>(Z - 32 * w + 16);
  const size_t slot = (static_cast<size_t>(X + 1 - C(0)) * PY + (Y + 1 - C(1))) * g.W + w;
  const uint4 e = win[slot];
  const uint64_t m = (static_cast<uint64_t>(e.z) << 32) | e.y;
  const int64_t b0 = static_cast<int64_t>(e.x) + __builtin_popcountll(m & ((1ull << t) - 1ull));
  const uint32_t bm = static_cast<uint32_t>(m >> (t - 1)) & 7u;
  const int dz = -C(2);
  if (!((bm >> (dz + 1)) & 1u)) return {false, i};
  return {true, dz == 0 ? b0 : (dz < 0 ? b0 - 1 : b0 + ((bm >> 1) & 1u))};
}


}  // namespace lbamd
