// C ABI: context, topology upload (pull table), field transfer, steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/badchimp_amd.h"
#include "ctx.hpp"
#include "errors.hpp"
#include "host_threads.hpp"
#include "kernels.hpp"
#include "kernels_f32.hpp"
#include "lattice.hpp"
#include "layout_select.hpp"
#include "rowbuild.hpp"

using namespace lbamd;

namespace {

// Staging memory of lb_download_f_noalloc: a fixed array in each device's
// memory, so that the entry point allocates nothing.  Every device the
// library runs on holds its copy until the process ends.  Host code never
// uses the name `scratch` as an address (it stands for the host-side shadow
// of the symbol): the device address comes from hipGetSymbolAddress, taken
// once per device with that device current, at the first context creation
// on the device.
constexpr size_t kScratchBytes = size_t(128) << 20;
__device__ __align__(256) static double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
struct ScratchSlot {
  bool tried = false;
  double* ptr = nullptr;
  std::string why;
};
std::mutex g_scratch_mu;
ScratchSlot g_scratch[kMaxDevices];

// Resolve the device address of `scratch` on `device` (current).  A failure
// is remembered and reported by the entry point; contexts work without it.
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  ScratchSlot& sl = g_scratch[device];
  if (sl.tried) return;
  sl.tried = true;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    sl.why = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
    return;
  }
  sl.ptr = static_cast<double*>(p);
}

double* scratch_on(int device, std::string& why) {
  if (device < 0 || device >= kMaxDevices) {
    why = "device number beyond the scratch table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  const ScratchSlot& sl = g_scratch[device];
  if (!sl.ptr) why = sl.tried ? sl.why : "scratch not set up on this device";
  return sl.ptr;
}

int fail(lb_ctx* c, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  c->err = vformat(fmt, ap);
  va_end(ap);
  return code;
}

#define LB_HIP(c, expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((c), LB_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

int rev_of(int lattice, int q) { return lattice == LB_D2Q9 ? D2Q9::rev(q) : D3Q19::rev(q); }
double w_of(int lattice, int q) { return lattice == LB_D2Q9 ? D2Q9::w(q) : D3Q19::w(q); }

size_t elem_size(int precision) {
  return precision == LB_FP64 ? 8 : (precision == LB_FP32S ? 4 : 2);
}

template <typename T>
int dev_alloc(lb_ctx* c, T** p, size_t count) {
  if (*p) {
    hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
  if (e != hipSuccess) return fail(c, LB_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
  c->bytes += static_cast<int64_t>(count * sizeof(T));
  return LB_OK;
}

void free_tags(lb_ctx* c) {
  for (void* p : {static_cast<void*>(c->d_tlink), static_cast<void*>(c->d_tabb), static_cast<void*>(c->d_tinfo),
                  static_cast<void*>(c->d_tfon), static_cast<void*>(c->d_tscale), static_cast<void*>(c->d_tpartial)})
    if (p) hipFree(p);
  c->d_tlink = c->d_tabb = c->d_tinfo = nullptr;
  c->d_tfon = c->d_tscale = c->d_tpartial = nullptr;
  c->tag_ready = false;
  c->tag_labels = 0;
  c->tag_pressure.clear();
}

void free_halos(lb_ctx* c) {
  for (auto& h : c->halos)
    for (void* p : {static_cast<void*>(h.d_send), static_cast<void*>(h.d_recv), static_cast<void*>(h.d_send_nodes),
                    static_cast<void*>(h.d_recv_nodes), static_cast<void*>(h.d_buf)})
      if (p) hipFree(p);
  c->halos.clear();
}

void free_fields(lb_ctx* c) {
  for (auto& p : c->d_f) {
    if (p) hipFree(p);
    p = nullptr;
  }
  c->n_fields = 0;
}

int alloc_fields(lb_ctx* c, int nf) {
  if (c->n_fields == nf) return LB_OK;
  free_fields(c);
  const size_t bytes = static_cast<size_t>(nf) * c->Q * c->NP * elem_size(c->precision);
  for (auto& p : c->d_f) {
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(c, LB_ERR_NOMEM, "hipMalloc(%zu bytes) for populations: %s", bytes, hipGetErrorString(e));
    c->bytes += static_cast<int64_t>(bytes);
  }
  c->n_fields = nf;
  c->cur = 0;
  c->post = false;
  c->step_count = 0;
  c->cg_colour_done = false;
  return LB_OK;
}

int check_ready(lb_ctx* c, int nf) {
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology: call lb_set_topology first");
  if (c->n_fields != nf)
    return fail(c, LB_ERR_STATE, "populations hold %d field(s), %d needed: upload or initialise first", c->n_fields, nf);
  return LB_OK;
}

// Time a sequence of launches on the context stream.
struct Timer {
  lb_ctx* c;
  explicit Timer(lb_ctx* ctx) : c(ctx) { hipEventRecord(c->ev0, c->stream); }
  int stop(int64_t n_launch) {
    hipEventRecord(c->ev1, c->stream);
    hipError_t e = hipEventSynchronize(c->ev1);
    if (e != hipSuccess) return fail(c, LB_ERR_HIP, "kernel execution failed: %s", hipGetErrorString(e));
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->kernel_ms += ms;
    c->launches += n_launch;
    return LB_OK;
  }
};

// Ensure d_f[cur] holds post-collision values or pre-collision ones; for
// transfers we need the pre-collision state in some buffer: returns it.
int pre_state(lb_ctx* c, const void** out) {
  if (!c->post) {
    *out = c->d_f[c->cur];
    return LB_OK;
  }
  void* dst = c->d_f[1 - c->cur];
  hipError_t e;
  if (c->precision == LB_FP64)
    e = gather_f64(c->lattice, c->n_fields, static_cast<const double*>(c->d_f[c->cur]),
                   static_cast<double*>(dst), c->addr, c->n_bulk, c->NP, c->stream);
  else
    e = gather_lowp(c->precision, c->lattice, c->n_fields, c->d_f[c->cur], dst, c->addr,
                    c->n_bulk, c->NP, c->stream);
  if (e != hipSuccess) return fail(c, LB_ERR_HIP, "gather: %s", hipGetErrorString(e));
  *out = dst;
  return LB_OK;
}

// Row layout of a node list (positions and neighbour index).  Returns false
// with a reason when the node list is not a set of cells of one box that the
// layout can address.
bool build_rows(int lattice, int32_t n_nodes, const int32_t* nb, const int32_t* bulk, int32_t n_bulk,
                const std::vector<int32_t>& bidx, const int32_t* pos, RowBuild& rb, std::string& why) {
  const int nD = lattice == LB_D2Q9 ? 2 : 3, Q = lattice == LB_D2Q9 ? 9 : 19;
  (void)n_nodes;
  auto P = [&](int32_t n, int a) -> int32_t {
    if (nD == 3) return pos[static_cast<size_t>(n) * 3 + a];
    return a == 0 ? pos[static_cast<size_t>(n) * 2] : (a == 1 ? 0 : pos[static_cast<size_t>(n) * 2 + 1]);
  };
  auto C = [&](int q, int a) -> int {
    const int d = nD == 3 ? a : (a == 0 ? 0 : (a == 2 ? 1 : -1));
    if (d < 0) return 0;
    return lattice == LB_D2Q9 ? D2Q9::c(q, d) : D3Q19::c(q, d);
  };
  if (n_bulk == 0) { why = "no bulk nodes"; return false; }
  // extent and periods over chunks of the bulk list (one slot per thread):
  // the periods come from the fluid links that wrap
  int32_t tlo[kMaxHostThreads][3], thi[kMaxHostThreads][3], tper[kMaxHostThreads][3];
  bool tbad[kMaxHostThreads];
  for (int t = 0; t < kMaxHostThreads; ++t) {
    tbad[t] = false;
    for (int a = 0; a < 3; ++a) { tlo[t][a] = P(bulk[0], a); thi[t][a] = tlo[t][a]; tper[t][a] = 0; }
  }
  parallel_ranges(n_bulk, [&](int64_t b, int64_t e, int t) {
    for (int64_t k = b; k < e; ++k) {
      const int32_t n = bulk[k];
      for (int a = 0; a < 3; ++a) {
        tlo[t][a] = std::min(tlo[t][a], P(n, a));
        thi[t][a] = std::max(thi[t][a], P(n, a));
      }
      for (int q = 0; q < Q - 1; ++q) {
        const int32_t m = nb[static_cast<size_t>(n) * Q + q];
        if (m <= 0 || bidx[m] < 0) continue;
        for (int a = 0; a < 3; ++a) {
          const int32_t d = std::abs(P(m, a) - P(n, a) - C(q, a));
          if (d == 0) continue;
          if (tper[t][a] == 0) tper[t][a] = d;
          else if (tper[t][a] != d) tbad[t] = true;
        }
      }
    }
  });
  int32_t lo[3], hi[3], period[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a) { lo[a] = tlo[0][a]; hi[a] = thi[0][a]; }
  for (int t = 0; t < kMaxHostThreads; ++t) {
    if (tbad[t]) { why = "inconsistent periodic wrap"; return false; }
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], tlo[t][a]);
      hi[a] = std::max(hi[a], thi[t][a]);
      if (tper[t][a] == 0) continue;
      if (period[a] == 0) period[a] = tper[t][a];
      else if (period[a] != tper[t][a]) { why = "inconsistent periodic wrap"; return false; }
    }
  }
  int32_t E[3];
  for (int a = 0; a < 3; ++a) {
    E[a] = period[a] ? period[a] : hi[a] - lo[a] + 1;
    if (hi[a] - lo[a] + 1 > E[a]) { why = "bulk nodes wider than the period"; return false; }
  }
  if (E[0] > 1024 || E[1] > 1024 || E[2] > 4096) { why = "box too large for packed coordinates"; return false; }
  RowGeom& g = rb.g;
  for (int a = 0; a < 3; ++a) g.E[a] = E[a];
  g.periodic = (period[0] ? 1 : 0) | (period[1] ? 2 : 0) | (period[2] ? 4 : 0);
  g.lx = kTileLogX;
  g.ly = E[1] >= (1 << kTileLogY) ? kTileLogY : 0;
  const int32_t nTx = (E[0] + (1 << g.lx) - 1) >> g.lx;
  g.nTy = (E[1] + (1 << g.ly) - 1) >> g.ly;
  g.W = (E[2] + 31) / 32;
  const int64_t R = static_cast<int64_t>(nTx) * g.nTy << (g.lx + g.ly);
  auto cell = [&](int32_t n, int32_t* X) {
    for (int a = 0; a < 3; ++a) {
      int32_t v = P(n, a) - lo[a];
      v %= E[a];
      if (v < 0) v += E[a];
      X[a] = v;
    }
  };
  // compact order: by (row, cell along the row), i.e. bucketed by row, each
  // row's nodes then ordered by their cell
  std::vector<int32_t> row_of(n_bulk), cz(n_bulk);
  parallel_ranges(n_bulk, [&](int64_t b, int64_t e, int) {
    for (int64_t k = b; k < e; ++k) {
      int32_t X[3];
      cell(bulk[k], X);
      row_of[k] = g.row(X[0], X[1]);
      cz[k] = X[2];
    }
  });
  std::vector<int64_t> start(static_cast<size_t>(R) + 1, 0);
  for (int32_t k = 0; k < n_bulk; ++k) ++start[static_cast<size_t>(row_of[k]) + 1];
  for (int64_t r = 0; r < R; ++r) start[r + 1] += start[r];
  rb.order.resize(n_bulk);
  {
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int32_t k = 0; k < n_bulk; ++k) rb.order[fill[row_of[k]]++] = k;
  }
  bool dup[kMaxHostThreads] = {};
  parallel_ranges(R, [&](int64_t b, int64_t e, int t) {
    for (int64_t r = b; r < e; ++r) {
      int32_t* o0 = rb.order.data() + start[r];
      int32_t* o1 = rb.order.data() + start[r + 1];
      std::sort(o0, o1, [&](int32_t x, int32_t y) { return cz[x] < cz[y]; });
      for (int32_t* o = o0 + 1; o < o1; ++o)
        if (cz[*o] == cz[*(o - 1)]) dup[t] = true;
    }
  });
  for (int t = 0; t < kMaxHostThreads; ++t)
    if (dup[t]) { why = "two bulk nodes share a cell"; return false; }
  rb.coord.resize(n_bulk);
  rb.rows.assign(static_cast<size_t>(R) * g.W, uint2{0u, 0u});
  // packed cells per compact node; the row bit masks row by row (a row's
  // nodes are consecutive in the compact order)
  parallel_ranges(R, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r)
      for (int64_t i = start[r]; i < start[r + 1]; ++i) {
        int32_t X[3];
        cell(bulk[rb.order[i]], X);
        rb.coord[i] = (static_cast<uint32_t>(X[0]) << 22) | (static_cast<uint32_t>(X[1]) << 12) |
                      static_cast<uint32_t>(X[2]);
        rb.rows[static_cast<size_t>(r) * g.W + (X[2] >> 5)].y |= 1u << (X[2] & 31);
      }
  });
  uint32_t run = 0;
  for (auto& e : rb.rows) {
    e.x = run;
    run += static_cast<uint32_t>(__builtin_popcount(e.y));
  }
  return true;
}

// Source of direction q for compact node i under the row layout, as the
// kernels compute it: {stream, j} or {bounce, i}.
std::pair<bool, int64_t> row_source(int lattice, const RowBuild& rb, int q, int32_t i) {
  const int nD = lattice == LB_D2Q9 ? 2 : 3;
  auto C = [&](int a) -> int {
    const int d = nD == 3 ? a : (a == 0 ? 0 : (a == 2 ? 1 : -1));
    if (d < 0) return 0;
    return lattice == LB_D2Q9 ? D2Q9::c(q, d) : D3Q19::c(q, d);
  };
  const uint32_t cc = rb.coord[i];
  int32_t X[3] = {static_cast<int32_t>(cc >> 22), static_cast<int32_t>((cc >> 12) & 1023u), static_cast<int32_t>(cc & 4095u)};
  for (int a = 0; a < 3; ++a) {
    if (C(a) == 0) continue;
    X[a] -= C(a);
    if (X[a] < 0 || X[a] >= rb.g.E[a]) {
      if (rb.g.periodic & (1 << a)) X[a] += X[a] < 0 ? rb.g.E[a] : -rb.g.E[a];
      else return {false, i};
    }
  }
  const uint2 e = rb.rows[static_cast<size_t>(rb.g.row(X[0], X[1])) * rb.g.W + (X[2] >> 5)];
  const uint32_t bit = 1u << (X[2] & 31);
  if (e.y & bit) return {true, static_cast<int64_t>(e.x + __builtin_popcount(e.y & (bit - 1u)))};
  return {false, i};
}

int setup_topology(lb_ctx* c, int32_t n_nodes, const int32_t* neighbor, const int16_t* node_type,
                   const int32_t* bulk, int32_t n_bulk, const int32_t* pos, int32_t flags) {
  if (n_nodes < 1 || n_bulk < 0 || !neighbor || !node_type || (n_bulk > 0 && !bulk))
    return fail(c, LB_ERR_ARG, "lb_set_topology: bad argument");
  if (n_bulk >= static_cast<int32_t>(kLinkIndexMask)) return fail(c, LB_ERR_ARG, "too many bulk nodes");
  hipSetDevice(c->device);
  const int Q = c->Q;
  std::vector<int32_t> bidx(n_nodes, -1);  // node -> bulk index
  for (int32_t k = 0; k < n_bulk; ++k) {
    const int32_t n = bulk[k];
    if (n <= 0 || n >= n_nodes) return fail(c, LB_ERR_ARG, "bulk node %d out of range", n);
    if (k > 0 && bulk[k - 1] >= n) return fail(c, LB_ERR_ARG, "bulk list must be strictly ascending");
    if (node_type[n] < 2) return fail(c, LB_ERR_ARG, "bulk node %d is not fluid (type %d)", n, node_type[n]);
    bidx[n] = k;
  }
  for (size_t k = 0; k < static_cast<size_t>(n_nodes) * Q; ++k)
    if (neighbor[k] < 0 || neighbor[k] >= n_nodes) return fail(c, LB_ERR_ARG, "neighbour index out of range");
  // other-rank fluid nodes (the halo): a population slot each, after the
  // bulk nodes, in ascending node order; the bulk nodes gather from them
  // through the link table (lb_halo_unpack fills them between steps)
  std::vector<int32_t> ghosts;
  for (int32_t n = 1; n < n_nodes; ++n)
    if (node_type[n] == 4) ghosts.push_back(n);
  const int32_t n_ghost = static_cast<int32_t>(ghosts.size());
  if (static_cast<int64_t>(n_bulk) + n_ghost >= static_cast<int64_t>(kLinkIndexMask))
    return fail(c, LB_ERR_ARG, "too many bulk and halo nodes");
  // compact order: the bulk order, or the tiled row order of the row layout
  RowBuild rb;
  std::string why = "no positions given";
  bool rows_ok = false;
  if (n_ghost > 0)
    why = "other-rank neighbours (the halo slots are addressed through the link table)";
  else
    rows_ok = pos && !(flags & LB_TOPO_TABLE_ONLY) &&
              build_rows(c->lattice, n_nodes, neighbor, bulk, n_bulk, bidx, pos, rb, why);
  std::vector<int32_t> order(n_bulk);
  for (int32_t k = 0; k < n_bulk; ++k) order[k] = rows_ok ? rb.order[k] : k;
  std::vector<int32_t> compact(n_nodes, -1);
  for (int32_t i = 0; i < n_bulk; ++i) compact[bulk[order[i]]] = i;
  for (int32_t k = 0; k < n_ghost; ++k) compact[ghosts[k]] = n_bulk + k;
  // The push of the reference writes f_post(p, m) into slot (p, neighbor(p,m)).
  // For node 0 the slot is shared: the last bulk node (ascending) wins.
  std::vector<int32_t> winner(Q, -1);
  for (int32_t k = 0; k < n_bulk; ++k) {
    const int32_t* nb = neighbor + static_cast<size_t>(bulk[k]) * Q;
    for (int p = 0; p < Q - 1; ++p)
      if (nb[p] == 0) winner[p] = compact[bulk[k]];
  }
  const int64_t n_slots = static_cast<int64_t>(n_bulk) + n_ghost;
  const int64_t NP = ((n_slots + kBlock - 1) / kBlock) * kBlock + (n_slots == 0 ? kBlock : 0);
  // element indices of both fields of a buffer are 32-bit
  if (static_cast<uint64_t>(Q) * static_cast<uint64_t>(NP) * 2u >= (1ull << 32))
    return fail(c, LB_ERR_UNSUPPORTED, "%d bulk nodes exceed one context (limit %llu); split the domain", n_bulk,
                static_cast<unsigned long long>((1ull << 32) / (2u * Q)));
  // the row windows address one population buffer with 32-bit byte offsets
  if (rows_ok && static_cast<uint64_t>(Q) * static_cast<uint64_t>(NP) * elem_size(c->precision) >= (1ull << 32))
    why = "population buffer above 4 GiB";
  const bool win_allowed = static_cast<uint64_t>(Q) * static_cast<uint64_t>(NP) * elem_size(c->precision) < (1ull << 32);
  // The pull table, built over chunks of the compact nodes on host threads.
  // A wall link first holds the wall's node number; the walls then get their
  // slots in the order of first use (compact node, then direction), which
  // one sequential pass over the nodes with wall links assigns.
  if (n_nodes > static_cast<int32_t>(kLinkIndexMask))
    return fail(c, LB_ERR_UNSUPPORTED, "node list of %d nodes: at most %u", n_nodes, kLinkIndexMask);
  std::vector<uint32_t> link(static_cast<size_t>(Q - 1) * NP, 0u);  // padding nodes link nowhere
  std::vector<uint32_t> wall_dirs(n_bulk);  // bit q: direction q links to a wall
  struct LinkError {
    int64_t at = INT64_MAX;  // i (Q - 1) + q of the first refused link of a chunk
    int32_t node = 0, src = 0, q = 0;
  };
  LinkError err[kMaxHostThreads];
  bool ghost0[kMaxHostThreads] = {};
  const bool allow0 = (flags & LB_TOPO_ALLOW_GHOST0) != 0;
  parallel_ranges(n_bulk, [&](int64_t b, int64_t e, int t) {
    for (int64_t i = b; i < e; ++i) {
      const int32_t n = bulk[order[i]];
      const int32_t* nb = neighbor + static_cast<size_t>(n) * Q;
      uint32_t wd = 0u;
      for (int q = 0; q < Q - 1; ++q) {
        const int32_t s = nb[rev_of(c->lattice, q)];
        uint32_t l = 0u;
        if (compact[s] >= 0) {
          l = (kLinkStream << 30) | static_cast<uint32_t>(compact[s]);
        } else if (node_type[s] > 1 || (s == 0 && !allow0)) {
          if (err[t].at == INT64_MAX) err[t] = LinkError{i * (Q - 1) + q, n, s, q};
        } else if (s == 0) {
          l = (kLinkGhost0 << 30) | static_cast<uint32_t>(winner[rev_of(c->lattice, q)]);
          ghost0[t] = true;
        } else {
          l = (kLinkBounce << 30) | static_cast<uint32_t>(s);
          wd |= 1u << q;
        }
        link[static_cast<size_t>(q) * NP + i] = l;
      }
      wall_dirs[i] = wd;
    }
  });
  // the first refused link in (compact node, direction) order, as one pass
  // over the nodes would find it
  const LinkError* first = std::min_element(err, err + kMaxHostThreads,
                                            [](const LinkError& x, const LinkError& y) { return x.at < y.at; });
  if (first->at != INT64_MAX) {
    if (first->src != 0)
      return fail(c, LB_ERR_UNSUPPORTED,
                  "bulk node %d has a fluid neighbour %d on another rank: halo exchange is not available on one GPU",
                  first->node, first->src);
    return fail(c, LB_ERR_GHOST0_LINK,
                "bulk node %d links to the ghost node 0 (direction %d); make the geometry periodic or walled, "
                "or pass LB_TOPO_ALLOW_GHOST0 to reproduce the reference's shared-slot result", first->node, first->q);
  }
  if (std::any_of(ghost0, ghost0 + kMaxHostThreads, [](bool x) { return x; })) {
    rows_ok = false;
    why = "fluid links to the ghost node 0";
  }
  std::vector<int32_t> wall_slot(n_nodes, -1);
  std::vector<int32_t> walls;
  for (int32_t i = 0; i < n_bulk; ++i)
    for (uint32_t wd = wall_dirs[i]; wd != 0u; wd &= wd - 1u) {
      const int q = __builtin_ctz(wd);
      uint32_t& l = link[static_cast<size_t>(q) * NP + i];
      const int32_t s = static_cast<int32_t>(l & kLinkIndexMask);
      if (wall_slot[s] < 0) {
        wall_slot[s] = static_cast<int32_t>(walls.size());
        walls.push_back(s);
      }
      l = (kLinkBounce << 30) | static_cast<uint32_t>(wall_slot[s]);
    }
  // a layout is used only when every link it yields is the table's (checked
  // over chunks of the compact nodes on host threads)
  auto agrees = [&](auto&& source) {
    bool bad[kMaxHostThreads] = {};
    parallel_ranges(n_bulk, [&](int64_t b, int64_t e, int t) {
      for (int64_t i = b; i < e && !bad[t]; ++i)
        for (int q = 0; q < Q - 1; ++q) {
          const uint32_t l = link[static_cast<size_t>(q) * NP + i];
          const auto r = source(q, static_cast<int32_t>(i));
          const bool stream = (l >> 30) == kLinkStream;
          if (r.first != stream || (stream && r.second != static_cast<int64_t>(l & kLinkIndexMask))) {
            bad[t] = true;
            break;
          }
        }
    });
    return std::none_of(bad, bad + kMaxHostThreads, [](bool x) { return x; });
  };
  if (rows_ok && !agrees([&](int q, int32_t i) { return row_source(c->lattice, rb, q, i); })) {
    rows_ok = false;
    why = "row layout disagrees with the neighbour index";
  }
  // row windows when the row axis is not periodic, checked link by link
  std::vector<uint4> win;
  int32_t PY = 0;
  bool win_ok = rows_ok && win_allowed && build_windows(rb, win, PY, why);
  if (win_ok && !agrees([&](int q, int32_t i) { return win_source(c->lattice, rb, win, PY, q, i); })) {
    win_ok = false;
    why = "row windows disagree with the neighbour index";
  }
  // device allocation: the context holds no topology until every buffer
  // below is in place, so a failure leaves no pointer to freed memory
  hipStreamSynchronize(c->stream);
  c->has_topo = false;
  c->have_macro = false;
  c->have_visc = false;
  c->addr = Addr{};
  c->n_bulk = 0;
  c->NP = 0;
  free_fields(c);
  free_tags(c);
  free_halos(c);
  c->n_ghost = 0;
  c->slot_of.clear();
  for (void* p : {static_cast<void*>(c->d_coord), static_cast<void*>(c->d_rows), static_cast<void*>(c->d_win)})
    if (p) hipFree(p);
  c->d_coord = nullptr;
  c->d_rows = nullptr;
  c->d_win = nullptr;
  c->bytes = 0;
  int rc;
  if ((rc = dev_alloc(c, &c->d_link, link.size())) != LB_OK) return rc;
  LB_HIP(c, hipMemcpy(c->d_link, link.data(), link.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  c->addr = Addr{};
  c->addr.layout = 0;
  c->addr.table = TableAddr{c->d_link, static_cast<uint32_t>(NP)};
  if (rows_ok) {
    std::vector<uint32_t> coord(NP, 0u);
    std::copy(rb.coord.begin(), rb.coord.end(), coord.begin());
    if ((rc = dev_alloc(c, &c->d_coord, coord.size())) != LB_OK) return rc;
    if ((rc = dev_alloc(c, &c->d_rows, rb.rows.size())) != LB_OK) return rc;
    LB_HIP(c, hipMemcpy(c->d_coord, coord.data(), coord.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    LB_HIP(c, hipMemcpy(c->d_rows, rb.rows.data(), rb.rows.size() * sizeof(uint2), hipMemcpyHostToDevice));
    c->addr.layout = 1;
    c->addr.rows = RowAddr{c->d_coord, c->d_rows, static_cast<uint32_t>(NP), rb.g};
    if (win_ok) {
      if ((rc = dev_alloc(c, &c->d_win, win.size())) != LB_OK) return rc;
      LB_HIP(c, hipMemcpy(c->d_win, win.data(), win.size() * sizeof(uint4), hipMemcpyHostToDevice));
      c->addr.layout = 2;
      c->addr.win = RowWin{c->d_coord, c->d_win, static_cast<uint32_t>(NP), PY, rb.g.W};
    }
  }
  c->layout_note = win_ok ? "row windows" : (rows_ok ? "rows (" + why + ")" : why);
  c->n_blocks = (n_bulk + kBlock - 1) / kBlock;
  if (c->n_blocks == 0) c->n_blocks = 1;
  if ((rc = dev_alloc(c, &c->d_rho, 2 * NP)) != LB_OK) return rc;
  if ((rc = dev_alloc(c, &c->d_vel, static_cast<size_t>(c->nD) * NP)) != LB_OK) return rc;
  if ((rc = dev_alloc(c, &c->d_phi, NP)) != LB_OK) return rc;
  if ((rc = dev_alloc(c, &c->d_phi_wall, walls.size() + 1)) != LB_OK) return rc;
  if ((rc = dev_alloc(c, &c->d_partial, 3 * c->n_blocks)) != LB_OK) return rc;
  if ((rc = dev_alloc(c, &c->d_scalar, 8)) != LB_OK) return rc;
  LB_HIP(c, hipMemset(c->d_phi_wall, 0, (walls.size() + 1) * sizeof(double)));
  c->wall_uniform = true;
  c->wall_phi = 0.0;
  LB_HIP(c, hipMemset(c->d_rho, 0, 2 * NP * sizeof(double)));
  LB_HIP(c, hipMemset(c->d_vel, 0, c->nD * NP * sizeof(double)));
  LB_HIP(c, hipMemset(c->d_phi, 0, NP * sizeof(double)));
  LB_HIP(c, hipMemset(c->d_scalar, 0, 8 * sizeof(double)));
  // partials stay 0 where no block writes them (an empty bulk list)
  LB_HIP(c, hipMemset(c->d_partial, 0, 3 * c->n_blocks * sizeof(double)));
  c->n_nodes = n_nodes;
  c->n_bulk = n_bulk;
  c->NP = NP;
  c->n_global = n_bulk;
  c->node_of.resize(n_bulk);
  for (int32_t i = 0; i < n_bulk; ++i) c->node_of[i] = bulk[order[i]];
  {
    // positions of the bulk nodes in ascending node order, and their runs
    std::vector<int32_t> by_node(n_bulk), rank(NP, 0);
    for (int32_t i = 0; i < n_bulk; ++i) by_node[i] = i;
    std::sort(by_node.begin(), by_node.end(),
              [&](int32_t a, int32_t b) { return c->node_of[a] < c->node_of[b]; });
    c->bulk_runs.clear();
    for (int32_t r = 0; r < n_bulk; ++r) {
      const int32_t i = by_node[r], n = c->node_of[i];
      rank[i] = r;
      if (!c->bulk_runs.empty() && c->bulk_runs.back()[0] + c->bulk_runs.back()[2] == n)
        ++c->bulk_runs.back()[2];
      else
        c->bulk_runs.push_back({n, r, 1});
    }
    if ((rc = dev_alloc(c, &c->d_rank, rank.size())) != LB_OK) return rc;
    LB_HIP(c, hipMemcpy(c->d_rank, rank.data(), rank.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  c->wall_nodes = walls;
  c->node_type.assign(node_type, node_type + n_nodes);
  c->n_ghost = n_ghost;
  c->slot_of = std::move(compact);
  c->has_topo = true;
  c->have_macro = false;
  return LB_OK;
}

}  // namespace

extern "C" {

const char* lb_version(void) { return "badchimp-amd 0.1 (gfx950)"; }

const char* lb_last_error(const lb_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int lb_ctx_create(int lattice, int precision, int device, lb_ctx** out) {
  if (!out) return fail_global(LB_ERR_ARG, "lb_ctx_create: null output");
  if (lattice != LB_D2Q9 && lattice != LB_D3Q19) return fail_global(LB_ERR_ARG, "unknown lattice %d", lattice);
  if (precision != LB_FP64 && precision != LB_FP32S && precision != LB_FP16S)
    return fail_global(LB_ERR_ARG, "unknown precision %d", precision);
  int n_dev = 0;
  hipError_t e = hipGetDeviceCount(&n_dev);
  if (e != hipSuccess || n_dev == 0) return fail_global(LB_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n_dev) return fail_global(LB_ERR_ARG, "device %d out of range (%d devices)", device, n_dev);
  e = hipSetDevice(device);
  if (e != hipSuccess) return fail_global(LB_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  lb_ctx* c = new (std::nothrow) lb_ctx();
  if (!c) return fail_global(LB_ERR_NOMEM, "out of host memory");
  c->lattice = lattice;
  c->precision = precision;
  c->device = device;
  c->Q = lattice == LB_D2Q9 ? 9 : 19;
  c->nD = lattice == LB_D2Q9 ? 2 : 3;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return fail_global(LB_ERR_HIP, "stream/event creation failed");
  }
  scratch_setup(device);
  *out = c;
  return LB_OK;
}

int lb_ctx_destroy(lb_ctx* c) {
  if (!c) return LB_OK;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  free_fields(c);
  free_tags(c);
  free_halos(c);
  for (void* p : {static_cast<void*>(c->d_link), static_cast<void*>(c->d_rho), static_cast<void*>(c->d_vel),
                  static_cast<void*>(c->d_phi), static_cast<void*>(c->d_phi_wall), static_cast<void*>(c->d_partial),
                  static_cast<void*>(c->d_scalar), static_cast<void*>(c->d_coord),
                  static_cast<void*>(c->d_rows), static_cast<void*>(c->d_win), static_cast<void*>(c->d_table),
                  static_cast<void*>(c->d_rank)})
    if (p) hipFree(p);
  hipEventDestroy(c->ev0);
  hipEventDestroy(c->ev1);
  hipStreamDestroy(c->stream);
  delete c;
  return LB_OK;
}

int lb_set_topology(lb_ctx* c, int32_t n_nodes, const int32_t* neighbor, const int16_t* node_type,
                    const int32_t* bulk, int32_t n_bulk, int32_t flags) {
  if (!c) return LB_ERR_ARG;
  return setup_topology(c, n_nodes, neighbor, node_type, bulk, n_bulk, nullptr, flags);
}

int lb_set_topology_grid(lb_ctx* c, int32_t n_nodes, const int32_t* neighbor, const int16_t* node_type,
                         const int32_t* bulk, int32_t n_bulk, const int32_t* pos, int32_t flags) {
  if (!c) return LB_ERR_ARG;
  if (!pos) return fail(c, LB_ERR_ARG, "lb_set_topology_grid: null positions");
  return setup_topology(c, n_nodes, neighbor, node_type, bulk, n_bulk, pos, flags);
}

int32_t lb_layout(const lb_ctx* c) { return c && c->has_topo ? c->addr.layout : -1; }

const char* lb_layout_note(const lb_ctx* c) { return c ? c->layout_note.c_str() : "null context"; }

int lb_set_global_bulk_count(lb_ctx* c, int64_t n) {
  if (!c) return LB_ERR_ARG;
  if (n <= 0) return fail(c, LB_ERR_ARG, "global bulk count must be positive");
  c->n_global = n;
  return LB_OK;
}

// ------------------------------------------------------------ fields
int lb_upload_f(lb_ctx* c, int32_t nf, const double* f) {
  if (!c) return LB_ERR_ARG;
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology");
  if ((nf != 1 && nf != 2) || !f) return fail(c, LB_ERR_ARG, "lb_upload_f: bad argument");
  hipSetDevice(c->device);
  int rc = alloc_fields(c, nf);
  if (rc != LB_OK) return rc;
  const int Q = c->Q;
  const int64_t NP = c->NP;
  const size_t count = static_cast<size_t>(nf) * Q * NP;
  // the device layout of the populations staged on the host: every slot of
  // the NP padded nodes written once, on host threads (padding nodes zero)
  std::unique_ptr<double[]> h(new (std::nothrow) double[count]);
  if (!h) return fail(c, LB_ERR_NOMEM, "lb_upload_f: no host staging memory");
  double* hp = h.get();
  parallel_ranges(NP, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i) {
      if (i < c->n_bulk) {
        const double* src = f + static_cast<size_t>(c->node_of[i]) * nf * Q;
        for (int k = 0; k < nf; ++k)
          for (int q = 0; q < Q; ++q) hp[static_cast<size_t>(nf) * pidx_host(Q, q, i) + k] = src[k * Q + q];
      } else {
        for (int k = 0; k < nf; ++k)
          for (int q = 0; q < Q; ++q) hp[static_cast<size_t>(nf) * pidx_host(Q, q, i) + k] = 0.0;
      }
    }
  });
  if (c->precision == LB_FP64) {
    LB_HIP(c, hipMemcpy(c->d_f[0], hp, count * sizeof(double), hipMemcpyHostToDevice));
  } else {
    double* tmp = nullptr;
    LB_HIP(c, hipMalloc(&tmp, count * sizeof(double)));
    hipError_t e = hipMemcpy(tmp, hp, count * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = store_lowp(c->precision, c->lattice, nf, tmp, c->d_f[0], c->n_bulk, NP, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(tmp);
    if (e != hipSuccess) return fail(c, LB_ERR_HIP, "upload conversion: %s", hipGetErrorString(e));
  }
  c->cur = 0;
  c->post = false;
  c->step_count = 0;
  c->cg_colour_done = false;
  c->have_macro = false;
  c->have_visc = false;  // the viscosity slot holds no rheology result now
  return LB_OK;
}

int lb_download_f(lb_ctx* c, int32_t nf, double* f) {
  if (!c) return LB_ERR_ARG;
  if (!f) return fail(c, LB_ERR_ARG, "null output");
  int rc = check_ready(c, nf);
  if (rc != LB_OK) return rc;
  hipSetDevice(c->device);
  const void* src = nullptr;
  if ((rc = pre_state(c, &src)) != LB_OK) return rc;
  const int Q = c->Q;
  const int64_t NP = c->NP;
  const size_t count = static_cast<size_t>(nf) * Q * NP;
  std::unique_ptr<double[]> h(new (std::nothrow) double[count]);  // every element copied in below
  if (!h) return fail(c, LB_ERR_NOMEM, "lb_download_f: no host staging memory");
  double* hp = h.get();
  {
    // fault the staging pages in on host threads, not inside the copy
    const size_t bytes = count * sizeof(double);
    char* base = reinterpret_cast<char*>(hp);
    parallel_ranges(static_cast<int64_t>((bytes + 4095) / 4096), [&](int64_t b, int64_t e, int) {
      for (int64_t p = b; p < e; ++p) base[std::min(static_cast<size_t>(p) * 4096, bytes - 1)] = 0;
    });
  }
  if (c->precision == LB_FP64) {
    LB_HIP(c, hipMemcpyAsync(hp, src, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LB_HIP(c, hipStreamSynchronize(c->stream));
  } else {
    double* tmp = nullptr;
    LB_HIP(c, hipMalloc(&tmp, count * sizeof(double)));
    hipError_t e = load_lowp(c->precision, c->lattice, nf, src, tmp, c->n_bulk, NP, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hp, tmp, count * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(tmp);
    if (e != hipSuccess) return fail(c, LB_ERR_HIP, "download conversion: %s", hipGetErrorString(e));
  }
  parallel_ranges(c->n_bulk, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i) {
      double* dst = f + static_cast<size_t>(c->node_of[i]) * nf * Q;
      for (int k = 0; k < nf; ++k)
        for (int q = 0; q < Q; ++q) dst[k * Q + q] = hp[static_cast<size_t>(nf) * pidx_host(Q, q, i) + k];
    }
  });
  return LB_OK;
}

int lb_download_f_noalloc(lb_ctx* c, int32_t nf, double* f) {
  if (!c) return LB_ERR_ARG;
  if (!f) return fail(c, LB_ERR_ARG, "null output");
  int rc = check_ready(c, nf);
  if (rc != LB_OK) return rc;
  const int Q = c->Q;
  const size_t bytes = static_cast<size_t>(c->n_bulk) * nf * Q * sizeof(double);
  if (bytes > kScratchBytes)
    return fail(c, LB_ERR_UNSUPPORTED,
                "lb_download_f_noalloc: %d bulk nodes x %d field(s) x %d directions need %zu bytes of staging, "
                "above its fixed %zu MiB; use lb_download_f",
                c->n_bulk, nf, Q, bytes, kScratchBytes >> 20);
  std::string why;
  double* stage = scratch_on(c->device, why);
  if (!stage) return fail(c, LB_ERR_HIP, "lb_download_f_noalloc: no staging memory: %s", why.c_str());
  hipSetDevice(c->device);
  const void* src = nullptr;
  if ((rc = pre_state(c, &src)) != LB_OK) return rc;
  hipError_t e;
  if (c->precision == LB_FP64)
    e = unstage_f64(c->lattice, nf, static_cast<const double*>(src), stage, c->n_bulk, c->NP, c->d_rank,
                    c->stream);
  else
    e = load_lowp(c->precision, c->lattice, nf, src, stage, c->n_bulk, c->NP, c->stream, c->d_rank);
  const size_t rec = static_cast<size_t>(nf) * Q;
  for (size_t k = 0; k < c->bulk_runs.size() && e == hipSuccess; ++k) {
    const auto& r = c->bulk_runs[k];
    e = hipMemcpyAsync(f + static_cast<size_t>(r[0]) * rec, stage + static_cast<size_t>(r[1]) * rec,
                       static_cast<size_t>(r[2]) * rec * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, LB_ERR_HIP, "lb_download_f_noalloc: %s", hipGetErrorString(e));
  return LB_OK;
}

int lb_init_equilibrium(lb_ctx* c, int32_t n_fields, int32_t field, int32_t rho_nfields,
                        int32_t rho_field, const double* rho, const double* vel) {
  if (!c) return LB_ERR_ARG;
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology");
  if ((n_fields != 1 && n_fields != 2) || field < 0 || field >= n_fields || rho_nfields < 1 ||
      rho_field < 0 || rho_field >= rho_nfields || !rho || !vel)
    return fail(c, LB_ERR_ARG, "lb_init_equilibrium: bad argument");
  hipSetDevice(c->device);
  c->cg_colour_done = false;
  if (c->n_fields != n_fields || c->post) {
    // (re)start from an unset state: all fields are to be initialised; the
    // fresh buffer starts at zero, so that a field not yet initialised holds
    // finite values (the reduced-precision forms store two fields jointly)
    free_fields(c);
    int rc = alloc_fields(c, n_fields);
    if (rc != LB_OK) return rc;
    LB_HIP(c, hipMemset(c->d_f[c->cur], 0,
                        static_cast<size_t>(n_fields) * c->Q * c->NP * elem_size(c->precision)));
  }
  const int64_t NP = c->NP;
  std::vector<double> hr(NP, 0.0), hv(static_cast<size_t>(c->nD) * NP, 0.0);
  for (int32_t i = 0; i < c->n_bulk; ++i) {
    const int32_t n = c->node_of[i];
    hr[i] = rho[static_cast<size_t>(n) * rho_nfields + rho_field];
    for (int d = 0; d < c->nD; ++d) hv[d * NP + i] = vel[static_cast<size_t>(n) * c->nD + d];
  }
  double *dr = nullptr, *dv = nullptr;
  hipError_t e = hipMalloc(&dr, NP * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dv, c->nD * NP * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(dr, hr.data(), NP * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dv, hv.data(), c->nD * NP * sizeof(double), hipMemcpyHostToDevice);
  if (c->precision == LB_FP64) {
    if (e == hipSuccess) e = init_eq_f64(c->lattice, static_cast<double*>(c->d_f[c->cur]), c->n_fields, field, c->n_bulk, NP, dr, dv,
                                         c->stream);
  } else {
    double* tmp = nullptr;
    const size_t cnt = static_cast<size_t>(c->n_fields) * c->Q * NP;
    if (e == hipSuccess) e = hipMalloc(&tmp, cnt * sizeof(double));
    if (e == hipSuccess) e = load_lowp(c->precision, c->lattice, c->n_fields, c->d_f[c->cur], tmp, c->n_bulk, NP, c->stream);
    if (e == hipSuccess) e = init_eq_f64(c->lattice, tmp, c->n_fields, field, c->n_bulk, NP, dr, dv, c->stream);
    if (e == hipSuccess) e = store_lowp(c->precision, c->lattice, c->n_fields, tmp, c->d_f[c->cur], c->n_bulk, NP, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (tmp) hipFree(tmp);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (dr) hipFree(dr);
  if (dv) hipFree(dv);
  if (e != hipSuccess) return fail(c, LB_ERR_HIP, "init_equilibrium: %s", hipGetErrorString(e));
  c->have_macro = false;
  c->have_visc = false;  // the viscosity slot holds no rheology result now
  return LB_OK;
}

int lb_download_scalar(lb_ctx* c, int32_t which, int32_t nf, double* s) {
  if (!c) return LB_ERR_ARG;
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology");
  if (!s) return fail(c, LB_ERR_ARG, "null output");
  if (!c->have_macro) return fail(c, LB_ERR_STATE, "no step has stored macroscopic fields yet");
  hipSetDevice(c->device);
  const int64_t NP = c->NP;
  if (which == LB_RHO) {
    if (nf != c->n_fields) return fail(c, LB_ERR_ARG, "rho has %d field(s)", c->n_fields);
    std::vector<double> h(2 * NP);
    LB_HIP(c, hipMemcpy(h.data(), c->d_rho, 2 * NP * sizeof(double), hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < c->n_bulk; ++i)
      for (int k = 0; k < nf; ++k) s[static_cast<size_t>(c->node_of[i]) * nf + k] = h[k * NP + i];
  } else if (which == LB_PHI) {
    if (nf != 1) return fail(c, LB_ERR_ARG, "phi has one field");
    if (c->n_fields != 2) return fail(c, LB_ERR_STATE, "phi exists only for two-phase fields");
    std::vector<double> h(NP);
    LB_HIP(c, hipMemcpy(h.data(), c->d_phi, NP * sizeof(double), hipMemcpyDeviceToHost));
    if (c->precision == LB_FP64) {
      for (int32_t i = 0; i < c->n_bulk; ++i) s[c->node_of[i]] = h[i];
    } else {
      // float phi values in the reduced-precision forms
      const float* hf = reinterpret_cast<const float*>(h.data());
      for (int32_t i = 0; i < c->n_bulk; ++i) s[c->node_of[i]] = static_cast<double>(hf[i]);
    }
  } else if (which == LB_VISC) {
    if (nf != 1) return fail(c, LB_ERR_ARG, "the viscosity has one field");
    if (!c->have_visc) return fail(c, LB_ERR_STATE, "the last stored step was not a rheology step");
    std::vector<double> h(NP);
    LB_HIP(c, hipMemcpy(h.data(), c->d_rho + NP, NP * sizeof(double), hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < c->n_bulk; ++i) s[c->node_of[i]] = h[i];
  } else {
    return fail(c, LB_ERR_ARG, "unknown scalar %d", which);
  }
  return LB_OK;
}

int lb_download_vector(lb_ctx* c, int32_t which, int32_t nf, double* v) {
  if (!c) return LB_ERR_ARG;
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology");
  if (!v || which != LB_VEL || nf != 1) return fail(c, LB_ERR_ARG, "lb_download_vector: bad argument");
  if (!c->have_macro) return fail(c, LB_ERR_STATE, "no step has stored macroscopic fields yet");
  hipSetDevice(c->device);
  const int64_t NP = c->NP;
  std::vector<double> h(static_cast<size_t>(c->nD) * NP);
  LB_HIP(c, hipMemcpy(h.data(), c->d_vel, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int32_t i = 0; i < c->n_bulk; ++i)
    for (int d = 0; d < c->nD; ++d) v[static_cast<size_t>(c->node_of[i]) * c->nD + d] = h[d * NP + i];
  return LB_OK;
}

int lb_set_wall_densities(lb_ctx* c, const double* rho2) {
  if (!c) return LB_ERR_ARG;
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology");
  if (!rho2) return fail(c, LB_ERR_ARG, "null argument");
  hipSetDevice(c->device);
  std::vector<double> h(c->wall_nodes.size() + 1, 0.0);
  for (size_t k = 0; k < c->wall_nodes.size(); ++k) {
    const int32_t n = c->wall_nodes[k];
    // cgField is written only for solid-boundary nodes (main_TWOPHASE.cpp:280-284)
    if (c->node_type[n] == 1) {
      const double r0 = rho2[2 * static_cast<size_t>(n)];
      const double r1 = rho2[2 * static_cast<size_t>(n) + 1];
      h[k] = (r0 - r1) / (r0 + r1);
    }
  }
  LB_HIP(c, hipMemcpy(c->d_phi_wall, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  // one colour on every wall slot lets the two-phase steps use the row windows
  c->wall_uniform = true;
  c->wall_phi = c->wall_nodes.empty() ? 0.0 : h[0];
  for (size_t k = 0; k < c->wall_nodes.size(); ++k)
    if (!(h[k] == c->wall_phi)) c->wall_uniform = false;
  return LB_OK;
}

// ------------------------------------------------------------ steps
}  // extern "C"

namespace {
BgkArgs with_step(BgkArgs a, uint32_t step) {
  a.step = step;
  return a;
}

// Step parameters of the single-phase loop body: BGK (tau_anti <= 0) or the
// two-relaxation-time operator, with the reference's constants
// (LBcollision.h:40, 62-63, 206, 223-224).
BgkArgs single_phase_args(const lb_ctx* c, double tau, double tau_anti, const double* force) {
  BgkArgs a{};
  a.tau_inv = 1.0 / tau;
  a.tau_factor = 1 - 0.5 / tau;
  for (int d = 0; d < 3; ++d) a.F[d] = d < c->nD ? force[d] : 0.0;
  if (tau_anti > 0.0) {
    a.mode = kTrt;
    a.tauA_inv = 1.0 / tau_anti;
    a.tauA_factor = 1 - 0.5 / tau_anti;
  }
  return a;
}

// One single-phase step from buffer `from` into the other one.
hipError_t launch_single(lb_ctx* c, const BgkArgs& a, int from, bool first, bool last) {
  if (c->precision == LB_FP64)
    return bgk_f64(c->lattice, first, last, false, static_cast<const double*>(c->d_f[from]),
                   static_cast<double*>(c->d_f[1 - from]), c->addr, c->n_bulk, c->NP, a, c->d_rho, c->d_vel,
                   c->d_partial, c->stream);
  return bgk_lowp(c->precision, c->lattice, first, last, false, c->d_f[from], c->d_f[1 - from], c->addr,
                  c->n_bulk, c->NP, with_step(a, c->step_count), c->d_rho, c->d_vel, c->d_partial, c->stream);
}

// Long runs replay a captured graph of kGraphSteps steps (an even count, so
// a replay returns to the buffer it started from): on small node lists the
// launches, not the kernels, set the step time.  On the 256 x 128 D2Q9
// channel (32 k nodes) a plain launch costs 3.8 us per step, a step inside a
// graph 3.1 us, and each graph launch about 12 us more: graphs of 2 steps
// ran at 8.1 us per step, of 8 at 3.9, of 32 at 3.4, of 128 at 3.2.  The
// first step after an upload and the last one (it stores rho and u) are
// launched on their own, and so are the steps left over.  FP16S steps pass
// the step count as their rounding seed and are always launched one by one.
// The kernels and their arguments are those of the plain launches, so the
// results are the same bits.
constexpr int32_t kGraphSteps = 128;
constexpr int32_t kGraphMinSteps = 2 * kGraphSteps;

int step_single(lb_ctx* c, int32_t nsteps, const BgkArgs& a) {
  hipSetDevice(c->device);
  Timer t(c);
  auto advance = [c](int32_t n) {
    if (n & 1) c->cur = 1 - c->cur;
    c->post = true;
    c->step_count += static_cast<uint32_t>(n);
    c->cg_colour_done = false;
  };
  int32_t s = 0;
  hipGraphExec_t run = nullptr;
  // (no bulk node: the launches enqueue nothing, so there is nothing to capture)
  if (nsteps >= kGraphMinSteps && c->precision != LB_FP16S && c->n_bulk > 0) {
    if (!c->post) {
      hipError_t e = launch_single(c, a, c->cur, true, false);
      if (e != hipSuccess) return fail(c, LB_ERR_HIP, "bgk launch: %s", hipGetErrorString(e));
      advance(1);
      ++s;
    }
    const int32_t replays = (nsteps - s - 1) / kGraphSteps;
    if (replays > 0) {
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return fail(c, LB_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
      for (int32_t k = 0; k < kGraphSteps && e == hipSuccess; ++k)
        e = launch_single(c, a, (k & 1) ? 1 - c->cur : c->cur, false, false);
      const hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
      if (e == hipSuccess) e = e2;
      if (e == hipSuccess) e = hipGraphInstantiate(&run, graph, nullptr, nullptr, 0);
      if (graph) hipGraphDestroy(graph);
      for (int32_t k = 0; e == hipSuccess && k < replays; ++k) e = hipGraphLaunch(run, c->stream);
      if (e != hipSuccess) {
        if (run) {
          hipStreamSynchronize(c->stream);
          hipGraphExecDestroy(run);
        }
        return fail(c, LB_ERR_HIP, "step graph: %s", hipGetErrorString(e));
      }
      advance(kGraphSteps * replays);
      s += kGraphSteps * replays;
    }
  }
  for (; s < nsteps; ++s) {
    hipError_t e = launch_single(c, a, c->cur, !c->post, s == nsteps - 1);
    if (e != hipSuccess) {
      if (run) {
        hipStreamSynchronize(c->stream);
        hipGraphExecDestroy(run);
      }
      return fail(c, LB_ERR_HIP, "bgk launch: %s", hipGetErrorString(e));
    }
    advance(1);
  }
  const int rc = t.stop(nsteps);
  if (run) hipGraphExecDestroy(run);  // the stream has drained (t.stop waits for it)
  if (rc != LB_OK) return rc;
  c->have_macro = true;
  c->have_visc = a.mode == kRheology;
  return LB_OK;
}
}  // namespace

extern "C" {

int lb_step_bgk(lb_ctx* c, int32_t nsteps, double tau, const double* force) {
  if (!c) return LB_ERR_ARG;
  int rc = check_ready(c, 1);
  if (rc != LB_OK) return rc;
  if (nsteps < 0 || !force || !(tau > 0.0)) return fail(c, LB_ERR_ARG, "lb_step_bgk: bad argument");
  if (nsteps == 0) return LB_OK;
  return step_single(c, nsteps, single_phase_args(c, tau, 0.0, force));
}

int lb_set_viscosity_table(lb_ctx* c, int32_t n, const double* strain_rate_sq, const double* viscosity) {
  if (!c) return LB_ERR_ARG;
  if (n < 1 || !strain_rate_sq || !viscosity) return fail(c, LB_ERR_ARG, "lb_set_viscosity_table: bad argument");
  for (int32_t k = 0; k < n; ++k) {
    if (!std::isfinite(strain_rate_sq[k]) || !std::isfinite(viscosity[k]))
      return fail(c, LB_ERR_ARG, "viscosity table row %d is not finite", k);
    if (k > 0 && !(strain_rate_sq[k] > strain_rate_sq[k - 1]))
      return fail(c, LB_ERR_ARG, "viscosity table: strain-rate squares must increase (row %d)", k);
  }
  // GeneralizedNewtonian's table with its constant extension (LBrheology.h:370-384)
  const int32_t m = n + 2;
  std::vector<double> h(2 * static_cast<size_t>(m));
  double* sr = h.data();
  double* nu = h.data() + m;
  for (int32_t k = 0; k < n; ++k) {
    sr[k + 1] = strain_rate_sq[k];
    nu[k + 1] = viscosity[k];
  }
  sr[0] = sr[1] - 1;
  nu[0] = nu[1];
  sr[n + 1] = sr[n] + 1;
  nu[n + 1] = nu[n];
  std::vector<float> hf(h.begin(), h.end());
  hipSetDevice(c->device);
  if (c->d_table) hipFree(c->d_table);
  c->d_table = nullptr;
  c->tab_n = 0;
  const size_t bytes = h.size() * sizeof(double) + hf.size() * sizeof(float);
  LB_HIP(c, hipMalloc(reinterpret_cast<void**>(&c->d_table), bytes));
  LB_HIP(c, hipMemcpy(c->d_table, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  LB_HIP(c, hipMemcpy(c->d_table + h.size(), hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice));
  c->tab_n = m;
  return LB_OK;
}

int lb_step_rheology(lb_ctx* c, int32_t nsteps, const double* force) {
  if (!c) return LB_ERR_ARG;
  int rc = check_ready(c, 1);
  if (rc != LB_OK) return rc;
  if (nsteps < 0 || !force) return fail(c, LB_ERR_ARG, "lb_step_rheology: bad argument");
  if (!c->d_table) return fail(c, LB_ERR_STATE, "lb_step_rheology: no viscosity table (lb_set_viscosity_table)");
  if (nsteps == 0) return LB_OK;
  BgkArgs a = single_phase_args(c, 1.0, 0.0, force);
  a.mode = kRheology;
  a.tab_n = c->tab_n;
  a.tab_sr = c->d_table;
  a.tab_nu = c->d_table + c->tab_n;
  a.tab_srf = reinterpret_cast<const float*>(c->d_table + 2 * c->tab_n);
  a.tab_nuf = a.tab_srf + c->tab_n;
  return step_single(c, nsteps, a);
}

// ------------------------------------------------------------ tagged boundaries
int lb_set_node_tags(lb_ctx* c, int32_t n_nodes, const int32_t* neighbor, const int32_t* tags,
                     const int32_t* interior_domains, const double* force_on) {
  if (!c) return LB_ERR_ARG;
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology: call lb_set_topology first");
  if (c->precision != LB_FP64)
    return fail(c, LB_ERR_UNSUPPORTED, "tagged boundaries are implemented for LB_FP64 contexts");
  if (c->n_ghost > 0)
    return fail(c, LB_ERR_UNSUPPORTED, "tagged boundaries are implemented for one-rank node lists (no halo)");
  if (!neighbor || !tags || !interior_domains || !force_on || n_nodes != c->n_nodes)
    return fail(c, LB_ERR_ARG, "lb_set_node_tags: bad argument (arrays over the %d nodes of the topology)",
                c->n_nodes);
  const int Q = c->Q;
  const int64_t NP = c->NP;
  std::vector<int32_t> compact(n_nodes, -1);
  for (int32_t i = 0; i < c->n_bulk; ++i) compact[c->node_of[i]] = i;
  auto tag = [&](int32_t n) { return static_cast<int>(static_cast<int16_t>(tags[n])); };  // Nodes keeps a short
  std::vector<uint32_t> link(static_cast<size_t>(Q - 1) * NP, 0u), abb(NP, 0u), info(NP, 0u);
  std::vector<double> fon(NP, 0.0);
  for (int32_t i = 0; i < c->n_bulk; ++i) {
    const int32_t n = c->node_of[i];
    const int32_t* nb = neighbor + static_cast<size_t>(n) * Q;
    const int t = tag(n);
    for (int q = 0; q < Q - 1; ++q) {
      const int qo = rev_of(c->lattice, q);  // outgoing direction whose neighbour sends into q
      const int32_t s = nb[qo];
      if (s < 0 || s >= n_nodes) return fail(c, LB_ERR_ARG, "neighbour index out of range");
      const int ts = tag(s);
      // the lists of std_one_phase/main.cpp:27-126, seen from the node that receives
      const bool solid_fluid = ((t >> 3) & 1) && (ts & 3) == 0;
      const bool pressure = ((t >> 4) & 1) && (ts & 3) == 3;
      const bool fluid_fluid = (((t >> 2) & 1) && (t & 3) == 1 && (ts & 3) == 2) ||
                               ((t & 3) == 2 && ((ts >> 2) & 1) && (ts & 3) == 1);
      uint32_t l;
      if (s == 0 && (solid_fluid || pressure || fluid_fluid || compact[s] < 0))
        return fail(c, LB_ERR_UNSUPPORTED, "node %d links to the ghost node 0 (direction %d)", n, qo);
      if (pressure) {
        if (compact[s] >= 0)
          return fail(c, LB_ERR_UNSUPPORTED, "pressure link of node %d to node %d, which is a bulk node", n, s);
        abb[i] |= 1u << qo;
        l = kLinkBounce << 30;
      } else if (solid_fluid || fluid_fluid) {
        l = kLinkBounce << 30;
      } else if (compact[s] >= 0) {
        l = (kLinkStream << 30) | static_cast<uint32_t>(compact[s]);
      } else {
        return fail(c, LB_ERR_UNSUPPORTED,
                    "node %d (tag %d) receives direction %d from non-bulk node %d (tag %d) with no boundary tag: "
                    "the reference leaves that population stale",
                    n, t, q, s, ts);
      }
      link[static_cast<size_t>(q) * NP + i] = l;
    }
  }
  // interior domains and their mass-source scale factors (main.cpp:322-354)
  int lmax = 0;
  for (int32_t n = 1; n < n_nodes; ++n) {
    if (interior_domains[n] < 0) return fail(c, LB_ERR_ARG, "negative interior domain label at node %d", n);
    lmax = std::max(lmax, interior_domains[n]);
  }
  std::vector<double> scale(2 * static_cast<size_t>(lmax + 1), 0.0);
  for (int32_t n = 1; n < n_nodes; ++n)
    if (interior_domains[n] > 0 && tag(n) < 3) scale[interior_domains[n]] += 1;
  for (int l = 1; l <= lmax; ++l) {
    if (scale[l] == 0.0) return fail(c, LB_ERR_ARG, "Interior domain %d has no interior nodes", l);
    scale[l] = 1.0 / scale[l];
  }
  for (int32_t i = 0; i < c->n_bulk; ++i) {
    const int32_t n = c->node_of[i];
    const bool add = interior_domains[n] > 0 && tag(n) < 3;
    info[i] = static_cast<uint32_t>(interior_domains[n]) | (add ? 0x80000000u : 0u);
    fon[i] = force_on[n];
  }
  std::vector<std::pair<int32_t, int>> pres;
  for (int32_t n = 1; n < n_nodes; ++n)
    if ((tag(n) >> 4) & 1) pres.emplace_back(n, (tag(n) & 3) - 1);
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  free_tags(c);
  int rc;
  if ((rc = dev_alloc(c, &c->d_tlink, link.size())) != LB_OK) return rc;
  if ((rc = dev_alloc(c, &c->d_tabb, abb.size())) != LB_OK) return rc;
  if ((rc = dev_alloc(c, &c->d_tinfo, info.size())) != LB_OK) return rc;
  if ((rc = dev_alloc(c, &c->d_tfon, fon.size())) != LB_OK) return rc;
  if ((rc = dev_alloc(c, &c->d_tscale, scale.size())) != LB_OK) return rc;
  if ((rc = dev_alloc(c, &c->d_tpartial, static_cast<size_t>(std::max(lmax, 1)) * c->n_blocks)) != LB_OK) return rc;
  LB_HIP(c, hipMemcpy(c->d_tlink, link.data(), link.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  LB_HIP(c, hipMemcpy(c->d_tabb, abb.data(), abb.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  LB_HIP(c, hipMemcpy(c->d_tinfo, info.data(), info.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  LB_HIP(c, hipMemcpy(c->d_tfon, fon.data(), fon.size() * sizeof(double), hipMemcpyHostToDevice));
  LB_HIP(c, hipMemcpy(c->d_tscale, scale.data(), scale.size() * sizeof(double), hipMemcpyHostToDevice));
  c->tag_labels = lmax;
  c->tag_pressure = std::move(pres);
  c->tag_ready = true;
  return LB_OK;
}

int lb_step_tagged(lb_ctx* c, int32_t nsteps, double tau, const double* force, double rho_w) {
  if (!c) return LB_ERR_ARG;
  int rc = check_ready(c, 1);
  if (rc != LB_OK) return rc;
  if (nsteps < 0 || !force || !(tau > 0.0)) return fail(c, LB_ERR_ARG, "lb_step_tagged: bad argument");
  if (!c->tag_ready) return fail(c, LB_ERR_STATE, "lb_step_tagged: no node tags (lb_set_node_tags)");
  if (nsteps == 0) return LB_OK;
  hipSetDevice(c->device);
  const BgkArgs a = single_phase_args(c, tau, 0.0, force);
  const TableAddr ad{c->d_tlink, static_cast<uint32_t>(c->NP)};
  const int L = c->tag_labels;
  double* mass = c->d_tscale + (L + 1);  // mass[0] stays 0 (its scale factor is 0)
  Timer t(c);
  int64_t launches = 0;
  for (int32_t s = 0; s < nsteps; ++s) {
    const bool last = s == nsteps - 1;
    const double* fin = static_cast<const double*>(c->d_f[c->cur]);
    hipError_t e = hipSuccess;
    if (L > 0) {
      // per-domain mass changes of the state this step starts from (main.cpp:527-535)
      e = tag_mass_f64(c->lattice, !c->post, fin, ad, c->n_bulk, c->NP, c->d_tinfo, L, c->d_tpartial, c->stream);
      if (e == hipSuccess) e = sum_partials(c->d_tpartial, c->n_blocks, L, mass + 1, c->stream);
      launches += 2;
    }
    if (e == hipSuccess)
      e = tagged_f64(c->lattice, !c->post, last, fin, static_cast<double*>(c->d_f[1 - c->cur]), ad, c->n_bulk,
                     c->NP, a, rho_w, c->d_tinfo, c->d_tfon, c->d_tabb, c->d_tscale, mass, c->d_rho, c->d_vel,
                     c->stream);
    if (e != hipSuccess) return fail(c, LB_ERR_HIP, "tagged step launch: %s", hipGetErrorString(e));
    c->cur = 1 - c->cur;
    c->post = true;
    ++c->step_count;
    c->cg_colour_done = false;
    ++launches;
  }
  rc = t.stop(launches);
  if (rc != LB_OK) return rc;
  c->have_macro = true;
  c->have_visc = false;
  return LB_OK;
}

int lb_pressure_flux(lb_ctx* c, double* q) {
  if (!c) return LB_ERR_ARG;
  if (!q) return fail(c, LB_ERR_ARG, "null output");
  if (!c->tag_ready) return fail(c, LB_ERR_STATE, "no node tags (lb_set_node_tags)");
  if (!c->have_macro) return fail(c, LB_ERR_STATE, "no step has stored macroscopic fields yet");
  if (c->nD != 3) return fail(c, LB_ERR_UNSUPPORTED, "the mass flux is the z component (D3Q19)");
  hipSetDevice(c->device);
  const int64_t NP = c->NP;
  std::vector<double> rho(NP), vz(NP);
  LB_HIP(c, hipMemcpy(rho.data(), c->d_rho, NP * sizeof(double), hipMemcpyDeviceToHost));
  LB_HIP(c, hipMemcpy(vz.data(), c->d_vel + 2 * NP, NP * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<int32_t> compact(c->n_nodes, -1);
  for (int32_t i = 0; i < c->n_bulk; ++i) compact[c->node_of[i]] = i;
  // main.cpp:600-612: ascending node order, per phase, halved
  double m[2] = {0.0, 0.0};
  for (const auto& pr : c->tag_pressure) {
    if (pr.second != 0 && pr.second != 1)
      return fail(c, LB_ERR_STATE, "Fluid phase = %d in write mass flux (node %d)", pr.second, pr.first);
    const int32_t i = compact[pr.first];
    if (i < 0) return fail(c, LB_ERR_STATE, "pressure-boundary node %d is not a bulk node", pr.first);
    m[pr.second] += vz[i] * rho[i];
  }
  q[0] = 0.5 * m[0];
  q[1] = 0.5 * m[1];
  return LB_OK;
}

int lb_step_trt(lb_ctx* c, int32_t nsteps, double tau_sym, double tau_anti, const double* force) {
  if (!c) return LB_ERR_ARG;
  int rc = check_ready(c, 1);
  if (rc != LB_OK) return rc;
  if (nsteps < 0 || !force || !(tau_sym > 0.5) || !(tau_anti > 0.5))
    return fail(c, LB_ERR_ARG, "lb_step_trt: bad argument (tau_sym and tau_anti must exceed 1/2)");
  if (nsteps == 0) return LB_OK;
  return step_single(c, nsteps, single_phase_args(c, tau_sym, tau_anti, force));
}

}  // extern "C"

namespace {

bool cg_args_ok(double tau0, double tau1, int32_t force_mode, const double* force) {
  return force && tau0 > 0.5 && tau1 > 0.5 && (force_mode == LB_FORCE_FIXED || force_mode == LB_FORCE_MOMX);
}

CgArgs cg_args(const lb_ctx* c, double tau0, double tau1, double sigma, double beta, int32_t force_mode,
               const double* force) {
  CgArgs a{};
  const double c2 = c->lattice == LB_D2Q9 ? D2Q9::c2 : D3Q19::c2;
  a.nu0Inv = 1.0 / (c2 * (tau0 - 0.5));
  a.nu1Inv = 1.0 / (c2 * (tau1 - 0.5));
  a.sigma = sigma;
  a.beta = beta;
  a.q0 = 0.0;
  a.q1 = 0.0;
  for (int d = 0; d < 3; ++d) a.F[d] = d < c->nD ? force[d] : 0.0;
  a.momx_mode = force_mode == LB_FORCE_MOMX;
  a.momx = force[0];
  return a;
}

// Row windows when every solid neighbour has one colour and both fields fit
// the 32-bit byte offsets of that layout; else the link table.
bool cg_rowwin(const lb_ctx* c) {
  return c->addr.layout == 2 && c->wall_uniform &&
         rowwin_fits(c->Q, c->NP, 2, static_cast<int64_t>(elem_size(c->precision)));
}

// The colour pass of the iteration at step_count: phi of every bulk node,
// and with momx the per-block partial sums of the x-momentum.  The phi
// buffer holds float values in the reduced-precision forms.
hipError_t colour_pass(lb_ctx* c, bool rowwin, bool momx) {
  if (c->precision == LB_FP64)
    return cg_phi_f64(c->lattice, !c->post, momx, static_cast<const double*>(c->d_f[c->cur]), c->addr, rowwin,
                      c->n_bulk, c->NP, c->d_phi, c->d_partial, c->stream);
  return cg_phi_lowp(c->precision, c->lattice, !c->post, momx, c->d_f[c->cur], c->addr, rowwin, c->n_bulk, c->NP,
                     reinterpret_cast<float*>(c->d_phi), c->d_partial, c->stream);
}

// The collision pass of the same iteration; macroscopic fields stored when
// `store`.  The caller then advances the state (advance_cg).
hipError_t collision_pass(lb_ctx* c, bool rowwin, bool store, CgArgs a) {
  const WallPhi wp{c->d_phi_wall, c->wall_phi};
  if (c->precision == LB_FP64)
    return cg_f64(c->lattice, !c->post, store, static_cast<const double*>(c->d_f[c->cur]),
                  static_cast<double*>(c->d_f[1 - c->cur]), c->addr, rowwin, c->n_bulk, c->NP, a, c->d_phi, wp,
                  c->d_scalar, c->d_rho, c->d_vel, c->stream);
  a.step = c->step_count;
  return cg_lowp(c->precision, c->lattice, !c->post, store, c->d_f[c->cur], c->d_f[1 - c->cur], c->addr, rowwin,
                 c->n_bulk, c->NP, a, reinterpret_cast<const float*>(c->d_phi), wp, c->d_scalar, c->d_rho, c->d_vel,
                 c->stream);
}

void advance_cg(lb_ctx* c) {
  c->cur = 1 - c->cur;
  c->post = true;
  ++c->step_count;
  c->cg_colour_done = false;
}

void finish_cg(lb_ctx* c, double force_x, const CgArgs& a) {
  c->last_force[0] = force_x;
  c->last_force[1] = a.F[1];
  c->last_force[2] = a.F[2];
  c->have_macro = true;
  c->have_visc = false;  // the viscosity slot holds no rheology result now
}

}  // namespace

extern "C" {

int lb_step_cg(lb_ctx* c, int32_t nsteps, double tau0, double tau1, double sigma, double beta,
               int32_t force_mode, const double* force) {
  if (!c) return LB_ERR_ARG;
  int rc = check_ready(c, 2);
  if (rc != LB_OK) return rc;
  if (c->n_ghost > 0)
    return fail(c, LB_ERR_UNSUPPORTED,
                "colour-gradient steps with other-rank neighbours exchange the colour of the halo nodes between "
                "the two passes: use lb_step_cg_colour, lb_halo_pack_colour / lb_halo_unpack_colour and "
                "lb_step_cg_collide");
  if (nsteps < 0 || !cg_args_ok(tau0, tau1, force_mode, force)) return fail(c, LB_ERR_ARG, "lb_step_cg: bad argument");
  if (nsteps == 0) return LB_OK;
  hipSetDevice(c->device);
  const CgArgs a = cg_args(c, tau0, tau1, sigma, beta, force_mode, force);
  const bool momx = a.momx_mode != 0;
  const bool rowwin = cg_rowwin(c);
  Timer t(c);
  int64_t launches = 0;
  for (int32_t s = 0; s < nsteps; ++s) {
    hipError_t e = colour_pass(c, rowwin, momx);
    if (e == hipSuccess && momx)
      e = finish_momx(c->d_partial, c->n_blocks, a.momx, c->n_global, c->d_scalar, c->stream);
    if (e == hipSuccess) e = collision_pass(c, rowwin, s == nsteps - 1, a);
    if (e != hipSuccess) return fail(c, LB_ERR_HIP, "two-phase launch: %s", hipGetErrorString(e));
    launches += momx ? 3 : 2;
    advance_cg(c);
  }
  rc = t.stop(launches);
  if (rc != LB_OK) return rc;
  double fx = a.F[0];
  if (momx) {
    double h[2];
    LB_HIP(c, hipMemcpy(h, c->d_scalar, 2 * sizeof(double), hipMemcpyDeviceToHost));
    fx = h[0];
  }
  finish_cg(c, fx, a);
  return LB_OK;
}

int lb_step_cg_colour(lb_ctx* c, double* flux_x) {
  if (!c) return LB_ERR_ARG;
  int rc = check_ready(c, 2);
  if (rc != LB_OK) return rc;
  hipSetDevice(c->device);
  const bool momx = flux_x != nullptr;
  Timer t(c);
  hipError_t e = colour_pass(c, cg_rowwin(c), momx);
  // the block partials summed as the momentum-controlled step sums them
  // (slot 1 of the scalar buffer; slot 0 receives -2 sum, unused)
  if (e == hipSuccess && momx) e = finish_momx(c->d_partial, c->n_blocks, 0.0, 1, c->d_scalar, c->stream);
  if (e != hipSuccess) return fail(c, LB_ERR_HIP, "colour pass: %s", hipGetErrorString(e));
  rc = t.stop(momx ? 2 : 1);
  if (rc != LB_OK) return rc;
  if (momx) LB_HIP(c, hipMemcpy(flux_x, c->d_scalar + 1, sizeof(double), hipMemcpyDeviceToHost));
  c->cg_colour_done = true;
  return LB_OK;
}

int lb_step_cg_collide(lb_ctx* c, double tau0, double tau1, double sigma, double beta, int32_t force_mode,
                       const double* force, double flux_x_total) {
  if (!c) return LB_ERR_ARG;
  int rc = check_ready(c, 2);
  if (rc != LB_OK) return rc;
  if (!cg_args_ok(tau0, tau1, force_mode, force)) return fail(c, LB_ERR_ARG, "lb_step_cg_collide: bad argument");
  if (!c->cg_colour_done)
    return fail(c, LB_ERR_STATE, "no colour pass for this iteration: call lb_step_cg_colour first");
  hipSetDevice(c->device);
  CgArgs a = cg_args(c, tau0, tau1, sigma, beta, force_mode, force);
  if (a.momx_mode) {
    // main_TWOPHASE.cpp:300-307, the arithmetic of k_finish_momx
    const double mean = flux_x_total / static_cast<double>(c->n_global);
    a.F[0] = 2.0 * (a.momx - mean);
    a.momx_mode = 0;
  }
  Timer t(c);
  hipError_t e = collision_pass(c, cg_rowwin(c), true, a);
  if (e != hipSuccess) return fail(c, LB_ERR_HIP, "two-phase launch: %s", hipGetErrorString(e));
  advance_cg(c);
  rc = t.stop(1);
  if (rc != LB_OK) return rc;
  finish_cg(c, a.F[0], a);
  return LB_OK;
}

int lb_solve_relperm(lb_ctx* c, double tau0, double tau1, double sigma, double beta, const double* force,
                     int32_t check_every, double tol, int64_t max_steps, double n_box, double* k_phase,
                     int64_t* steps_out) {
  if (!c) return LB_ERR_ARG;
  int rc = check_ready(c, 2);
  if (rc != LB_OK) return rc;
  if (c->n_ghost > 0)
    return fail(c, LB_ERR_UNSUPPORTED,
                "lb_solve_relperm: a rank of a split domain exchanges colour and populations between the passes "
                "and sums the flux over the ranks; step with lb_step_cg_colour / lb_step_cg_collide and the "
                "lb_halo_* calls, and reduce with lb_reduce_flux");
  if (!force || !(tau0 > 0.5) || !(tau1 > 0.5) || check_every < 1 || !(tol > 0.0) || max_steps < 1 ||
      !(n_box > 0.0) || force[0] == 0.0)
    return fail(c, LB_ERR_ARG, "lb_solve_relperm: bad argument");
  const double nu0 = (tau0 - 0.5) / 3.0, nu1 = (tau1 - 0.5) / 3.0;
  double k[2] = {0.0, 0.0}, kp[2] = {0.0, 0.0};
  bool have_prev = false;
  int held = 0;
  int64_t steps = 0;
  while (steps < max_steps) {
    const int64_t chunk = std::min<int64_t>(check_every, max_steps - steps);
    // the phase fluxes of the state the chunk's last step starts from
    if ((rc = lb_step_cg(c, static_cast<int32_t>(chunk - 1), tau0, tau1, sigma, beta, LB_FORCE_FIXED, force)) !=
        LB_OK)
      return rc;
    double s1 = 0.0, s2 = 0.0;
    if ((rc = lb_reduce_flux(c, 0, nullptr, &s1, &s2)) != LB_OK) return rc;
    if ((rc = lb_step_cg(c, 1, tau0, tau1, sigma, beta, LB_FORCE_FIXED, force)) != LB_OK) return rc;
    steps += chunk;
    // phase permeabilities k_i = nu_i <phi_i sum_q f c_x> / F_x (LBglobalforcing.h:61-78)
    k[0] = nu0 * (s1 / n_box) / force[0];
    k[1] = nu1 * (s2 / n_box) / force[0];
    if (!std::isfinite(k[0]) || !std::isfinite(k[1]))
      return fail(c, LB_ERR_STATE, "phase permeability is not finite after %lld steps", (long long)steps);
    const double scale = std::fabs(k[0]) + std::fabs(k[1]);
    const bool ok = have_prev && std::fabs(k[0] - kp[0]) <= tol * scale && std::fabs(k[1] - kp[1]) <= tol * scale;
    held = ok ? held + 1 : 0;
    if (held >= kStopHold) break;
    kp[0] = k[0];
    kp[1] = k[1];
    have_prev = true;
  }
  if (k_phase) {
    k_phase[0] = k[0];
    k_phase[1] = k[1];
  }
  if (steps_out) *steps_out = steps;
  return LB_OK;
}

int lb_last_force(lb_ctx* c, double* force) {
  if (!c) return LB_ERR_ARG;
  if (!force) return fail(c, LB_ERR_ARG, "null output");
  for (int d = 0; d < c->nD; ++d) force[d] = c->last_force[d];
  return LB_OK;
}

int lb_reduce_flux(lb_ctx* c, int32_t dir, double* sum_fc, double* sum_phi0_fc, double* sum_phi1_fc) {
  if (!c) return LB_ERR_ARG;
  if (!c->has_topo || c->n_fields == 0) return fail(c, LB_ERR_STATE, "no populations");
  if (dir < 0 || dir >= c->nD) return fail(c, LB_ERR_ARG, "dir out of range");
  hipSetDevice(c->device);
  hipError_t e;
  if (c->precision == LB_FP64)
    e = flux_f64(c->lattice, c->n_fields, !c->post, static_cast<const double*>(c->d_f[c->cur]), c->addr,
                 c->n_bulk, c->NP, dir, c->d_partial, c->stream);
  else
    e = flux_lowp(c->precision, c->lattice, c->n_fields, !c->post, c->d_f[c->cur], c->addr, c->n_bulk,
                  c->NP, dir, c->d_partial, c->stream);
  if (e == hipSuccess) e = sum_partials(c->d_partial, c->n_blocks, 3, c->d_scalar + 4, c->stream);
  double h[3] = {0, 0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(h, c->d_scalar + 4, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, LB_ERR_HIP, "flux reduction: %s", hipGetErrorString(e));
  if (sum_fc) *sum_fc = h[0];
  if (sum_phi0_fc) *sum_phi0_fc = h[1];
  if (sum_phi1_fc) *sum_phi1_fc = h[2];
  return LB_OK;
}

}  // extern "C"

namespace {
int solve_permeability(lb_ctx* c, const char* name, double tau, double tau_anti, const double* force,
                       int32_t check_every, double tol, int64_t max_steps, double n_box, double* k_out,
                       int64_t* steps_out) {
  int rc = check_ready(c, 1);
  if (rc != LB_OK) return rc;
  if (c->n_ghost > 0)
    return fail(c, LB_ERR_UNSUPPORTED,
                "%s: a rank of a split domain exchanges its halo between steps and sums the flux over the ranks; "
                "step with lb_step_bgk / lb_step_trt and the lb_halo_* calls, and reduce with lb_reduce_flux",
                name);
  if (!force || !(tau > 0.5) || check_every < 1 || !(tol > 0.0) || max_steps < 1 || !(n_box > 0.0) ||
      force[0] == 0.0)
    return fail(c, LB_ERR_ARG, "%s: bad argument", name);
  hipSetDevice(c->device);
  const BgkArgs a = single_phase_args(c, tau, tau_anti, force);
  const double nu = (tau - 0.5) / 3.0;
  double k_prev = 0.0, k = 0.0;
  int64_t steps = 0;
  bool have_prev = false;
  int held = 0;
  Timer t(c);
  int64_t launches = 0;
  while (steps < max_steps) {
    const int64_t chunk = std::min<int64_t>(check_every, max_steps - steps);
    for (int64_t s = 0; s < chunk; ++s) {
      // every step adds its block sums of j_x to the partials (the first
      // step of the chunk overwrites them); the last one stores rho and u
      const bool last = s == chunk - 1;
      hipError_t e;
      if (c->precision == LB_FP64)
        e = bgk_f64(c->lattice, !c->post, last, true, static_cast<const double*>(c->d_f[c->cur]),
                    static_cast<double*>(c->d_f[1 - c->cur]), c->addr, c->n_bulk, c->NP, a, c->d_rho,
                    c->d_vel, c->d_partial, c->stream, s > 0);
      else
        e = bgk_lowp(c->precision, c->lattice, !c->post, last, true, c->d_f[c->cur], c->d_f[1 - c->cur],
                     c->addr, c->n_bulk, c->NP, with_step(a, c->step_count), c->d_rho, c->d_vel, c->d_partial,
                     c->stream, s > 0);
      if (e != hipSuccess) return fail(c, LB_ERR_HIP, "bgk launch: %s", hipGetErrorString(e));
      c->cur = 1 - c->cur;
      c->post = true;
      ++c->step_count;
      c->cg_colour_done = false;
      ++launches;
    }
    steps += chunk;
    hipError_t e = sum_partials(c->d_partial, c->n_blocks, 1, c->d_scalar + 4, c->stream);
    double jx = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(&jx, c->d_scalar + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, LB_ERR_HIP, "permeability reduction: %s", hipGetErrorString(e));
    // j_x summed over the states the chunk's steps started from;
    // k = nu <j_x> / F_x with <j_x> the chunk's average per node of the box
    k = nu * ((jx / static_cast<double>(chunk)) / n_box) / a.F[0];
    if (!std::isfinite(k)) return fail(c, LB_ERR_STATE, "permeability is not finite after %lld steps", (long long)steps);
    held = have_prev && std::fabs(k - k_prev) <= tol * std::fabs(k) ? held + 1 : 0;
    if (held >= kStopHold) break;
    k_prev = k;
    have_prev = true;
  }
  rc = t.stop(launches);
  if (rc != LB_OK) return rc;
  c->have_macro = true;
  c->have_visc = false;  // the viscosity slot holds no rheology result now
  if (k_out) *k_out = k;
  if (steps_out) *steps_out = steps;
  return LB_OK;
}
}  // namespace

extern "C" {

int lb_solve_permeability(lb_ctx* c, double tau, const double* force, int32_t check_every, double tol,
                          int64_t max_steps, double n_box, double* k_out, int64_t* steps_out) {
  if (!c) return LB_ERR_ARG;
  return solve_permeability(c, "lb_solve_permeability", tau, 0.0, force, check_every, tol, max_steps, n_box,
                            k_out, steps_out);
}

int lb_solve_permeability_trt(lb_ctx* c, double tau_sym, double tau_anti, const double* force,
                              int32_t check_every, double tol, int64_t max_steps, double n_box, double* k_out,
                              int64_t* steps_out) {
  if (!c) return LB_ERR_ARG;
  if (!(tau_anti > 0.5)) return fail(c, LB_ERR_ARG, "lb_solve_permeability_trt: tau_anti must exceed 1/2");
  return solve_permeability(c, "lb_solve_permeability_trt", tau_sym, tau_anti, force, check_every, tol,
                            max_steps, n_box, k_out, steps_out);
}

// ------------------------------------------------------------ timing
int lb_sync(lb_ctx* c) {
  if (!c) return LB_ERR_ARG;
  hipSetDevice(c->device);
  LB_HIP(c, hipStreamSynchronize(c->stream));
  return LB_OK;
}

int lb_kernel_time(lb_ctx* c, double* ms, int64_t* launches) {
  if (!c) return LB_ERR_ARG;
  if (ms) *ms = c->kernel_ms;
  if (launches) *launches = c->launches;
  return LB_OK;
}

int lb_kernel_time_reset(lb_ctx* c) {
  if (!c) return LB_ERR_ARG;
  c->kernel_ms = 0.0;
  c->launches = 0;
  return LB_OK;
}

int64_t lb_device_bytes(const lb_ctx* c) { return c ? c->bytes : -1; }

// ------------------------------------------------------------ field files
// The reference's binary field files (LBfield.h:102-138, 233-276, 378-421):
// int32 header, then the doubles in the field's host layout over all nodes
// of the node list.  Non-bulk nodes are written as 0 (their values are not
// part of the state between steps).
namespace {

int write_file(lb_ctx* c, const std::string& path, const std::vector<int32_t>& header, const std::vector<double>& v) {
  FILE* fh = std::fopen(path.c_str(), "wb");
  if (!fh) return fail(c, LB_ERR_IO, "cannot open %s for writing", path.c_str());
  bool ok = std::fwrite(header.data(), sizeof(int32_t), header.size(), fh) == header.size() &&
            std::fwrite(v.data(), sizeof(double), v.size(), fh) == v.size();
  ok = (std::fclose(fh) == 0) && ok;
  if (!ok) return fail(c, LB_ERR_IO, "short write to %s", path.c_str());
  return LB_OK;
}

}  // namespace

int lb_write_lbfield(lb_ctx* c, const char* prefix) {
  if (!c) return LB_ERR_ARG;
  if (!prefix) return fail(c, LB_ERR_ARG, "null path");
  if (!c->has_topo || c->n_fields == 0) return fail(c, LB_ERR_STATE, "no populations");
  const int nf = c->n_fields;
  std::vector<double> f(static_cast<size_t>(c->n_nodes) * nf * c->Q, 0.0);
  int rc = lb_download_f(c, nf, f.data());
  if (rc != LB_OK) return rc;
  return write_file(c, std::string(prefix) + ".lblbf", {nf, c->Q, c->n_nodes}, f);
}

int lb_read_lbfield(lb_ctx* c, const char* prefix) {
  if (!c) return LB_ERR_ARG;
  if (!prefix) return fail(c, LB_ERR_ARG, "null path");
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology");
  const std::string path = std::string(prefix) + ".lblbf";
  FILE* fh = std::fopen(path.c_str(), "rb");
  if (!fh) return fail(c, LB_ERR_IO, "cannot open %s", path.c_str());
  int32_t hdr[3] = {0, 0, 0};
  if (std::fread(hdr, sizeof(int32_t), 3, fh) != 3) {
    std::fclose(fh);
    return fail(c, LB_ERR_IO, "%s: short header", path.c_str());
  }
  if ((hdr[0] != 1 && hdr[0] != 2) || hdr[1] != c->Q || hdr[2] != c->n_nodes) {
    std::fclose(fh);
    return fail(c, LB_ERR_ARG, "%s holds %d field(s), %d directions, %d nodes; the context has %d directions, %d nodes",
                path.c_str(), hdr[0], hdr[1], hdr[2], c->Q, c->n_nodes);
  }
  std::vector<double> f(static_cast<size_t>(hdr[0]) * hdr[1] * hdr[2]);
  const size_t got = std::fread(f.data(), sizeof(double), f.size(), fh);
  std::fclose(fh);
  if (got != f.size()) return fail(c, LB_ERR_IO, "%s: %zu of %zu values", path.c_str(), got, f.size());
  return lb_upload_f(c, hdr[0], f.data());
}

int lb_write_scalar(lb_ctx* c, int32_t which, int32_t n_fields, const char* prefix) {
  if (!c) return LB_ERR_ARG;
  if (!prefix) return fail(c, LB_ERR_ARG, "null path");
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology");
  if (n_fields < 1 || n_fields > 2) return fail(c, LB_ERR_ARG, "bad field count");
  std::vector<double> s(static_cast<size_t>(c->n_nodes) * n_fields, 0.0);
  int rc = lb_download_scalar(c, which, n_fields, s.data());
  if (rc != LB_OK) return rc;
  return write_file(c, std::string(prefix) + ".lbsca", {n_fields, c->n_nodes}, s);
}

int lb_write_vector(lb_ctx* c, int32_t which, const char* prefix) {
  if (!c) return LB_ERR_ARG;
  if (!prefix) return fail(c, LB_ERR_ARG, "null path");
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology");
  std::vector<double> v(static_cast<size_t>(c->n_nodes) * c->nD, 0.0);
  int rc = lb_download_vector(c, which, 1, v.data());
  if (rc != LB_OK) return rc;
  return write_file(c, std::string(prefix) + ".lbvec", {1, c->nD, c->n_nodes}, v);
}

// ------------------------------------------------------------ halo exchange
int lb_halo_add(lb_ctx* c, int32_t neighbor_rank, int32_t n_send, const int32_t* nodes_to_send,
                const int32_t* n_dir_send, const int32_t* dir_send, int32_t n_recv, const int32_t* nodes_received,
                const int32_t* n_dir_recv, const int32_t* dir_recv, int32_t* index) {
  if (!c) return LB_ERR_ARG;
  if (!c->has_topo) return fail(c, LB_ERR_STATE, "no topology: call lb_set_topology first");
  if (n_send < 0 || n_recv < 0 || (n_send > 0 && (!nodes_to_send || !n_dir_send)) ||
      (n_recv > 0 && (!nodes_received || !n_dir_recv)) || !index)
    return fail(c, LB_ERR_ARG, "lb_halo_add: bad argument");
  const int Q = c->Q;
  // element index of each value: the node's own slot of the direction; and
  // the compact index of each node (its colour)
  auto build = [&](int32_t n, const int32_t* nodes, const int32_t* ndir, const int32_t* dirs, bool own,
                   std::vector<uint32_t>& out, std::vector<uint32_t>& slots) -> int {
    int64_t total = 0;
    for (int32_t k = 0; k < n; ++k) {
      if (ndir[k] < 0 || ndir[k] > Q - 1) return fail(c, LB_ERR_ARG, "lb_halo_add: %d directions for one node", ndir[k]);
      total += ndir[k];
    }
    if (total > 0 && !dirs) return fail(c, LB_ERR_ARG, "lb_halo_add: null direction list");
    out.clear();
    out.reserve(static_cast<size_t>(total));
    slots.clear();
    slots.reserve(static_cast<size_t>(n));
    int64_t at = 0;
    for (int32_t k = 0; k < n; ++k) {
      const int32_t node = nodes[k];
      if (node <= 0 || node >= c->n_nodes)
        return fail(c, LB_ERR_ARG, "lb_halo_add: node %d out of range", node);
      const int32_t slot = c->slot_of[node];
      if (own ? !(slot >= 0 && slot < c->n_bulk) : !(slot >= c->n_bulk))
        return fail(c, LB_ERR_ARG, own ? "lb_halo_add: node %d to send is not a bulk node of this rank"
                                       : "lb_halo_add: node %d to receive is not an other-rank (halo) node",
                    node);
      slots.push_back(static_cast<uint32_t>(slot));
      for (int32_t m = 0; m < ndir[k]; ++m, ++at) {
        const int32_t q = dirs[at];
        if (q < 0 || q >= Q - 1) return fail(c, LB_ERR_ARG, "lb_halo_add: direction %d", q);
        out.push_back(pidx_host(Q, static_cast<uint32_t>(q), static_cast<uint32_t>(slot)));
      }
    }
    return LB_OK;
  };
  std::vector<uint32_t> es, er, ns, nr;
  int rc = build(n_send, nodes_to_send, n_dir_send, dir_send, true, es, ns);
  if (rc != LB_OK) return rc;
  rc = build(n_recv, nodes_received, n_dir_recv, dir_recv, false, er, nr);
  if (rc != LB_OK) return rc;
  hipSetDevice(c->device);
  lb_ctx::Halo h;
  h.rank = neighbor_rank;
  h.n_send = static_cast<int64_t>(es.size());
  h.n_recv = static_cast<int64_t>(er.size());
  h.n_send_nodes = n_send;
  h.n_recv_nodes = n_recv;
  // both fields of the populations, or one colour per node
  const size_t nbuf = static_cast<size_t>(std::max<int64_t>(
      {2 * h.n_send, 2 * h.n_recv, h.n_send_nodes, h.n_recv_nodes, int64_t{1}}));
  auto upload = [](uint32_t** d, const std::vector<uint32_t>& v) {
    hipError_t e = hipMalloc(d, std::max<size_t>(v.size(), 1) * sizeof(uint32_t));
    if (e == hipSuccess && !v.empty()) e = hipMemcpy(*d, v.data(), v.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    return e;
  };
  hipError_t e = upload(&h.d_send, es);
  if (e == hipSuccess) e = upload(&h.d_recv, er);
  if (e == hipSuccess) e = upload(&h.d_send_nodes, ns);
  if (e == hipSuccess) e = upload(&h.d_recv_nodes, nr);
  if (e == hipSuccess) e = hipMalloc(&h.d_buf, nbuf * sizeof(double));
  if (e != hipSuccess) {
    for (void* p : {static_cast<void*>(h.d_send), static_cast<void*>(h.d_recv), static_cast<void*>(h.d_send_nodes),
                    static_cast<void*>(h.d_recv_nodes), static_cast<void*>(h.d_buf)})
      if (p) hipFree(p);
    return fail(c, LB_ERR_NOMEM, "lb_halo_add: %s", hipGetErrorString(e));
  }
  c->halos.push_back(h);
  *index = static_cast<int32_t>(c->halos.size()) - 1;
  return LB_OK;
}

int lb_halo_sizes(const lb_ctx* c, int32_t index, int32_t* neighbor_rank, int64_t* n_send, int64_t* n_recv) {
  if (!c || index < 0 || index >= static_cast<int32_t>(c->halos.size())) return LB_ERR_ARG;
  const auto& h = c->halos[index];
  if (neighbor_rank) *neighbor_rank = h.rank;
  if (n_send) *n_send = h.n_send;
  if (n_recv) *n_recv = h.n_recv;
  return LB_OK;
}

namespace {
int halo_check(lb_ctx* c, int32_t index, int32_t nf, const void* buf) {
  if (index < 0 || index >= static_cast<int32_t>(c->halos.size()))
    return fail(c, LB_ERR_ARG, "no halo exchange %d", index);
  if (!buf) return fail(c, LB_ERR_ARG, "null buffer");
  if (c->n_fields == 0) return fail(c, LB_ERR_STATE, "no populations");
  if (nf < 1 || nf > c->n_fields) return fail(c, LB_ERR_ARG, "%d fields of %d", nf, c->n_fields);
  if (!c->post)
    return fail(c, LB_ERR_STATE,
                "the halo carries post-collision values: exchange after a step (the first step after an upload "
                "reads no halo)");
  return LB_OK;
}
}  // namespace

int lb_halo_pack(lb_ctx* c, int32_t index, int32_t nf, double* send) {
  if (!c) return LB_ERR_ARG;
  int rc = halo_check(c, index, nf, send);
  if (rc != LB_OK) return rc;
  hipSetDevice(c->device);
  const auto& h = c->halos[index];
  hipError_t e = halo_pack(c->precision, c->n_fields, nf, c->d_f[c->cur], h.d_send, h.n_send, h.d_buf, c->stream);
  if (e == hipSuccess && h.n_send > 0)
    e = hipMemcpyAsync(send, h.d_buf, static_cast<size_t>(nf) * h.n_send * sizeof(double), hipMemcpyDefault,
                       c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, LB_ERR_HIP, "lb_halo_pack: %s", hipGetErrorString(e));
  return LB_OK;
}

int lb_halo_unpack(lb_ctx* c, int32_t index, int32_t nf, const double* recv) {
  if (!c) return LB_ERR_ARG;
  int rc = halo_check(c, index, nf, recv);
  if (rc != LB_OK) return rc;
  hipSetDevice(c->device);
  const auto& h = c->halos[index];
  hipError_t e = hipSuccess;
  if (h.n_recv > 0)
    e = hipMemcpyAsync(h.d_buf, recv, static_cast<size_t>(nf) * h.n_recv * sizeof(double), hipMemcpyDefault,
                       c->stream);
  if (e == hipSuccess)
    e = halo_unpack(c->precision, c->n_fields, nf, c->d_f[c->cur], h.d_recv, h.n_recv, h.d_buf, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, LB_ERR_HIP, "lb_halo_unpack: %s", hipGetErrorString(e));
  return LB_OK;
}

}  // extern "C"

namespace {
int colour_check(lb_ctx* c, int32_t index, const void* buf) {
  if (index < 0 || index >= static_cast<int32_t>(c->halos.size()))
    return fail(c, LB_ERR_ARG, "no halo exchange %d", index);
  if (!buf) return fail(c, LB_ERR_ARG, "null buffer");
  if (!c->cg_colour_done)
    return fail(c, LB_ERR_STATE, "the colour is exchanged between lb_step_cg_colour and lb_step_cg_collide");
  return LB_OK;
}
// the colour buffer holds double in LB_FP64 and float in the other forms
int colour_precision(const lb_ctx* c) { return c->precision == LB_FP64 ? LB_FP64 : LB_FP32S; }
}  // namespace

extern "C" {

int lb_halo_pack_colour(lb_ctx* c, int32_t index, double* send) {
  if (!c) return LB_ERR_ARG;
  int rc = colour_check(c, index, send);
  if (rc != LB_OK) return rc;
  hipSetDevice(c->device);
  const auto& h = c->halos[index];
  hipError_t e = halo_pack(colour_precision(c), 1, 1, c->d_phi, h.d_send_nodes, h.n_send_nodes, h.d_buf, c->stream);
  if (e == hipSuccess && h.n_send_nodes > 0)
    e = hipMemcpyAsync(send, h.d_buf, h.n_send_nodes * sizeof(double), hipMemcpyDefault, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, LB_ERR_HIP, "lb_halo_pack_colour: %s", hipGetErrorString(e));
  return LB_OK;
}

int lb_halo_unpack_colour(lb_ctx* c, int32_t index, const double* recv) {
  if (!c) return LB_ERR_ARG;
  int rc = colour_check(c, index, recv);
  if (rc != LB_OK) return rc;
  hipSetDevice(c->device);
  const auto& h = c->halos[index];
  hipError_t e = hipSuccess;
  if (h.n_recv_nodes > 0)
    e = hipMemcpyAsync(h.d_buf, recv, h.n_recv_nodes * sizeof(double), hipMemcpyDefault, c->stream);
  if (e == hipSuccess)
    e = halo_unpack(colour_precision(c), 1, 1, c->d_phi, h.d_recv_nodes, h.n_recv_nodes, h.d_buf, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, LB_ERR_HIP, "lb_halo_unpack_colour: %s", hipGetErrorString(e));
  return LB_OK;
}

int32_t lb_halo_count(const lb_ctx* c) { return c ? static_cast<int32_t>(c->halos.size()) : -1; }
int32_t lb_halo_num_nodes(const lb_ctx* c) { return c && c->has_topo ? c->n_ghost : -1; }

}  // extern "C"
