// Batched and serving form: many independent permeability solves resident in
// one device pool, advanced together.
//
// A solve is a geometry (cells > 0 fluid), a box with periodic axes, and the
// std_case parameters (tau, body force).  It is admitted when its footprint
// fits into the pool: two population buffers, coordinates, the row-window
// table and the block partials, carved from one pre-allocated device region by
// a first-fit allocator.  Every step is one launch over all resident solves
// (k_batch_bgk / k_batch_bgk_lp); every check_every steps the last launch also
// writes block partials of j_x, one small launch sums them per solve, the host
// reads one double per solve, evaluates k = nu <j_x> / F_x and retires the
// solves whose |dk| <= tol |k| (the stop of lb_solve_permeability).  Retired
// solves free their memory; waiting requests can then be admitted, at a check
// boundary, so every solve sees whole check intervals.
//
// The step launches of one check interval are captured in a HIP graph and
// replayed while the resident set is unchanged.  The launch shape (workgroups
// per CU of the persistent launch, or one workgroup per 256-node block) is
// chosen by timing the candidates on live steps after each change of the
// resident set; the shape does not change any result, so tuning costs nothing
// but time.
//
// Each solve's per-node arithmetic, block partition and reduction order are
// those of a single context with row windows (lb_solve_permeability), so a
// solve gives the same k and step count alone, in any mix, and through the
// single-solve path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/badchimp_amd.h"
#include "errors.hpp"
#include "kernels.hpp"
#include "kernels_f32.hpp"
#include "lattice.hpp"
#include "rowbuild.hpp"

using namespace lbamd;

namespace {

constexpr int kMaxSlots = 1 << (32 - kItemShift);  // 16384
constexpr int64_t kAlign = 256;

int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

size_t elem_bytes(int precision) { return precision == LB_FP64 ? 8 : (precision == LB_FP32S ? 4 : 2); }

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, size) in kAlign units, coalescing on free.
struct RangeAlloc {
  std::map<int64_t, int64_t> free_;  // offset -> size
  int64_t total = 0, used = 0;
  void init(int64_t size) {
    free_.clear();
    total = size;
    used = 0;
    if (size > 0) free_[0] = size;
  }
  int64_t alloc(int64_t size) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < size) continue;
      const int64_t off = it->first, rest = it->second - size;
      free_.erase(it);
      if (rest > 0) free_[off + size] = rest;
      used += size;
      return off;
    }
    return -1;
  }
  void release(int64_t off, int64_t size) {
    used -= size;
    auto it = free_.emplace(off, size).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t largest() const {
    int64_t m = 0;
    for (auto& kv : free_) m = std::max(m, kv.second);
    return m;
  }
};

// Sizes of one solve's device region (nf = 1 single-phase, 2 two-phase).
struct Footprint {
  int64_t n = 0, NP = 0, nblk = 0, nwin = 0;
  int nf = 1;
  int64_t off_f0 = 0, off_f1 = 0, off_coord = 0, off_win = 0, off_part = 0, off_phi = 0, bytes = 0;
};

Footprint footprint_of(int Q, size_t esz, int64_t n, int64_t nwin, int nf) {
  Footprint fp;
  fp.n = n;
  fp.nf = nf;
  fp.NP = std::max<int64_t>((n + kBlock - 1) / kBlock, 1) * kBlock;
  fp.nblk = std::max<int64_t>((n + kBlock - 1) / kBlock, 1);
  fp.nwin = nwin;
  const int64_t fb = align_up(static_cast<int64_t>(nf) * Q * fp.NP * static_cast<int64_t>(esz));
  fp.off_f0 = 0;
  fp.off_f1 = fb;
  fp.off_coord = 2 * fb;
  fp.off_win = fp.off_coord + align_up(fp.NP * 4);
  fp.off_part = fp.off_win + align_up(nwin * 16);
  fp.off_phi = fp.off_part + align_up((nf == 2 ? 3 : 1) * fp.nblk * 8);
  // colour field: double in FP64, float otherwise
  fp.bytes = fp.off_phi + (nf == 2 ? align_up(fp.NP * (esz == 8 ? 8 : 4)) : 0);
  return fp;
}

struct Slot {
  bool used = false;
  int kind = 0;  // 0 single-phase, 1 two-phase
  int64_t tag = 0;
  int64_t off = 0;
  Footprint fp;
  int64_t steps_seen = 0;  // steps of the solve at the last read-back of its stop state
  double t_admit = 0.0;
};

struct Result {
  int64_t tag, steps;
  double k, k0, k1, seconds;
  int32_t status;  // 0 converged, 1 max_steps reached
};

}  // namespace

struct lb_batch {
  int lattice = 0, precision = 0, device = 0, Q = 0, nD = 0;
  // precision of the solves of each kind (0 single-phase, 1 two-phase)
  int prec[2] = {0, 0};
  int32_t check_every = 100;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int n_cu = 256;
  // pool
  char* pool = nullptr;
  RangeAlloc alloc;
  // solves
  std::vector<Slot> slots;
  BatchSolve* d_solves = nullptr;
  std::vector<BatchSolve> h_solves;  // host copies, the source of the descriptor uploads
  uint32_t* d_items = nullptr;
  int64_t items_cap = 0;
  int32_t* d_slots = nullptr;
  double* d_out = nullptr;
  SolveCheck* d_chk = nullptr;  // stop state of each slot, advanced on the device
  SolveCheck* h_chk = nullptr;  // pinned: read back at the end of each launch sequence
  uint32_t* h_step = nullptr;  // pinned: global step at the start of an interval
  uint32_t* d_step = nullptr;
  // admissions: their uploads and rest-state kernels run on a stream of their
  // own, so that they overlap a launch sequence in flight; the host buffers
  // they read stay here until the next wait
  hipStream_t up = nullptr;
  hipEvent_t up_ev = nullptr;
  bool up_pending = false;
  std::vector<std::shared_ptr<void>> staged;
  bool in_flight = false;
  int32_t seq_intervals = 0;
  std::vector<uint32_t> items;  // single-phase items, then two-phase items
  int32_t n_items_bgk = 0, n_items_cg = 0;
  std::vector<int32_t> active;
  bool dirty = true;
  int64_t global_step = 0;
  int64_t resident_nodes = 0;
  // launch shape: wg_per_cu > 0 persistent, 0 one workgroup per item
  int32_t shape_mode = -1;  // -1 autotune, else fixed
  int32_t shape = 4;
  int32_t ns = 1;           // 256-thread slices per workgroup (workgroups of 256 ns threads)
  int64_t tuned_nodes = -1;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  // results and counters
  std::vector<Result> done;
  double kernel_ms = 0.0;
  int64_t launches = 0, lups = 0;
};

namespace {

int bfail(lb_batch* b, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  b->err = vformat(fmt, ap);
  va_end(ap);
  return code;
}

#define LB_BHIP(b, call)                                                         \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) return bfail(b, LB_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

void drop_graphs(lb_batch* b) {
  for (auto& g : b->graph)
    if (g) {
      hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// grid and items per workgroup of a shape (workgroups per CU; 0: one
// workgroup per ns items) with ns items in flight per workgroup
void shape_dims(const lb_batch* b, int32_t shape, int32_t ns, int32_t n_items, int32_t* grid, int32_t* per) {
  if (shape <= 0 || n_items <= b->n_cu * shape * ns) {
    *per = ns;
    *grid = (n_items + ns - 1) / ns;
    return;
  }
  const int32_t g = b->n_cu * shape;
  *per = (n_items + g - 1) / g;
  *grid = (n_items + *per - 1) / *per;
}

hipError_t launch_step(lb_batch* b, bool reduce, int32_t shape, int32_t ns, int32_t koff) {
  const int32_t parity = static_cast<int32_t>(b->global_step & 1);
  int32_t grid, per;
  hipError_t e = hipSuccess;
  if (b->n_items_bgk > 0) {
    // single-phase: j_x summed over every step of the interval (the first
    // step of the interval overwrites the partials)
    const int32_t n = b->n_items_bgk;
    shape_dims(b, shape, ns, n, &grid, &per);
    if (b->prec[0] == LB_FP64)
      e = batch_bgk_f64(b->lattice, true, b->d_solves, b->d_items, n, grid, per, parity, ns, koff > 0, b->stream);
    else
      e = batch_bgk_lowp(b->prec[0], b->lattice, true, b->d_solves, b->d_items, n, grid, per, parity,
                         b->d_step, koff, ns, koff > 0, b->stream);
  }
  if (e == hipSuccess && b->n_items_cg > 0) {
    const int32_t n = b->n_items_cg;
    const uint32_t* it = b->d_items + b->n_items_bgk;
    shape_dims(b, shape, ns, n, &grid, &per);
    if (b->prec[1] == LB_FP64)
      e = batch_cg_f64(b->lattice, reduce, b->d_solves, it, n, grid, per, parity, ns, b->stream);
    else
      e = batch_cg_lowp(b->prec[1], b->lattice, reduce, b->d_solves, it, n, grid, per, parity, b->d_step, koff,
                        ns, b->stream);
  }
  return e;
}

// rebuild the item and slot lists after admissions / retirements (at a
// point where no launch sequence is in flight)
int refresh(lb_batch* b) {
  b->items.clear();
  b->active.clear();
  b->resident_nodes = 0;
  for (int kind = 0; kind < 2; ++kind) {
    for (int32_t s = 0; s < static_cast<int32_t>(b->slots.size()); ++s) {
      const Slot& sl = b->slots[s];
      if (!sl.used || sl.kind != kind) continue;
      b->active.push_back(s);
      b->resident_nodes += sl.fp.n;
      for (int64_t k = 0; k < sl.fp.nblk; ++k)
        b->items.push_back((static_cast<uint32_t>(s) << kItemShift) | static_cast<uint32_t>(k));
    }
    if (kind == 0) b->n_items_bgk = static_cast<int32_t>(b->items.size());
  }
  b->n_items_cg = static_cast<int32_t>(b->items.size()) - b->n_items_bgk;
  if (static_cast<int64_t>(b->items.size()) > b->items_cap) {
    if (b->d_items) hipFree(b->d_items);
    b->d_items = nullptr;
    b->items_cap = std::max<int64_t>(static_cast<int64_t>(b->items.size()) * 2, 4096);
    LB_BHIP(b, hipMalloc(&b->d_items, b->items_cap * sizeof(uint32_t)));
  }
  // the steps read what the admissions uploaded on their own stream
  if (b->up_pending) LB_BHIP(b, hipStreamWaitEvent(b->stream, b->up_ev, 0));
  if (!b->items.empty())
    LB_BHIP(b, hipMemcpyAsync(b->d_items, b->items.data(), b->items.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, b->stream));
  if (!b->active.empty())
    LB_BHIP(b, hipMemcpyAsync(b->d_slots, b->active.data(), b->active.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, b->stream));
  // the global step at the start of the next interval, read by the kernels;
  // the intervals advance it on the device from here
  *b->h_step = static_cast<uint32_t>(b->global_step);
  LB_BHIP(b, hipMemcpyAsync(b->d_step, b->h_step, sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
  LB_BHIP(b, hipStreamSynchronize(b->stream));
  drop_graphs(b);
  b->dirty = false;
  return LB_OK;
}

// The per-solve sums and stop tests that close an interval.
hipError_t close_interval(lb_batch* b) {
  const int32_t n = static_cast<int32_t>(b->active.size());
  hipError_t e = batch_sum(b->d_solves, b->d_slots, n, b->d_out, b->stream);
  if (e == hipSuccess) e = batch_check(b->d_solves, b->d_chk, b->d_slots, n, b->d_out, b->check_every, b->d_step,
                                       b->stream);
  return e;
}

// One check interval as individual launches, timing each candidate launch
// shape on live steps (the shape never changes a result); keeps the fastest.
// A candidate is (workgroups per CU, 0 for one workgroup per ns items; ns,
// the 256-node items a workgroup works on at a time, i.e. workgroups of 256
// or 512 threads).
int tune_interval(lb_batch* b) {
  const int M = b->check_every;
  const int32_t cand[][2] = {{0, 1}, {1, 1}, {2, 1}, {4, 1}, {8, 1}, {0, 2}, {1, 2}, {2, 2}, {4, 2}};
  constexpr int kCand = 9;
  std::vector<hipEvent_t> ev(2 * M);
  for (auto& e : ev) LB_BHIP(b, hipEventCreate(&e));
  const int64_t gs = b->global_step;
  for (int s = 0; s < M; ++s) {
    hipEventRecord(ev[2 * s], b->stream);
    hipError_t e = launch_step(b, s == M - 1, cand[s % kCand][0], cand[s % kCand][1], s);
    hipEventRecord(ev[2 * s + 1], b->stream);
    if (e != hipSuccess) return bfail(b, LB_ERR_HIP, "batch step: %s", hipGetErrorString(e));
    ++b->global_step;
  }
  b->global_step = gs;
  hipError_t e = close_interval(b);
  if (e != hipSuccess) return bfail(b, LB_ERR_HIP, "batch check: %s", hipGetErrorString(e));
  LB_BHIP(b, hipStreamSynchronize(b->stream));
  double t[kCand] = {};
  int cnt[kCand] = {};
  for (int s = 0; s < M; ++s) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]);
    // the first launch of each shape warms up; skip it when there are others
    if (s >= kCand || M <= kCand) {
      t[s % kCand] += ms;
      ++cnt[s % kCand];
    }
    b->kernel_ms += ms;
  }
  for (auto& e2 : ev) hipEventDestroy(e2);
  int best = 3;
  double bt = 1e300;
  for (int c = 0; c < kCand; ++c)
    if (cnt[c] > 0 && t[c] / cnt[c] < bt) {
      bt = t[c] / cnt[c];
      best = c;
    }
  b->shape = cand[best][0];
  b->ns = cand[best][1];
  b->tuned_nodes = b->resident_nodes;
  b->launches += M + 3;
  return LB_OK;
}

// The graph of one interval starting at this parity: M steps, the sums, the
// stop tests and the step counter, so that a sequence of intervals runs
// without the host.
int interval_graph(lb_batch* b, int parity, hipGraphExec_t* out) {
  hipGraphExec_t& g = b->graph[parity];
  if (!g) {
    const int M = b->check_every;
    hipGraph_t graph = nullptr;
    LB_BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    const int64_t gs = b->global_step;
    b->global_step = parity;  // launch_step reads only the parity of the step
    for (int s = 0; s < M && e == hipSuccess; ++s) {
      e = launch_step(b, s == M - 1, b->shape, b->ns, s);
      ++b->global_step;
    }
    b->global_step = gs;
    if (e == hipSuccess) e = close_interval(b);
    hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph) hipGraphDestroy(graph);
      return bfail(b, LB_ERR_HIP, "graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
      g = nullptr;
      return bfail(b, LB_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
  }
  *out = g;
  return LB_OK;
}

// Enqueue `intervals` check intervals of every resident solve and the read-back
// of their stop states; the host does not wait.
int launch_sequence(lb_batch* b, int32_t intervals) {
  if (b->dirty) {
    const int rc = refresh(b);
    if (rc != LB_OK) return rc;
  }
  if (b->active.empty() || intervals <= 0) return LB_OK;
  const int M = b->check_every;
  int32_t k = 0;
  const bool tune = b->shape_mode < 0 &&
                    (b->tuned_nodes < 0 || std::llabs(b->resident_nodes - b->tuned_nodes) * 5 > b->tuned_nodes);
  if (b->shape_mode >= 0) {
    b->shape = b->shape_mode;
    b->ns = 1;
  }
  if (tune) {
    const int rc = tune_interval(b);
    if (rc != LB_OK) return rc;
    b->global_step += M;
    ++k;
  }
  hipEventRecord(b->ev0, b->stream);
  for (; k < intervals; ++k) {
    hipGraphExec_t g = nullptr;
    const int rc = interval_graph(b, static_cast<int>(b->global_step & 1), &g);
    if (rc != LB_OK) return rc;
    LB_BHIP(b, hipGraphLaunch(g, b->stream));
    b->global_step += M;
    b->launches += M + 3;
  }
  hipEventRecord(b->ev1, b->stream);
  LB_BHIP(b, hipMemcpyAsync(b->h_chk, b->d_chk, b->slots.size() * sizeof(SolveCheck), hipMemcpyDeviceToHost,
                            b->stream));
  b->in_flight = true;
  b->seq_intervals = intervals - (tune ? 1 : 0);
  return LB_OK;
}

void retire(lb_batch* b, int32_t s) {
  Slot& sl = b->slots[s];
  const SolveCheck& c = b->h_chk[s];
  b->done.push_back(Result{sl.tag, c.steps, c.k[0], c.k[1], c.k[2], now_s() - sl.t_admit, c.status});
  b->alloc.release(sl.off, sl.fp.bytes);
  sl.used = false;
  b->dirty = true;
}

// Wait for the sequence in flight and for the admissions' uploads, then
// retire the solves whose stop test decided.
int wait_sequence(lb_batch* b, int32_t* n_finished) {
  int32_t fin = 0;
  if (n_finished) *n_finished = 0;
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e == hipSuccess && b->up_pending) e = hipEventSynchronize(b->up_ev);
  if (b->up_pending) {
    b->staged.clear();
    b->up_pending = false;
  }
  const bool was = b->in_flight;
  b->in_flight = false;
  if (e != hipSuccess) return bfail(b, LB_ERR_HIP, "batch step: %s", hipGetErrorString(e));
  if (!was) return LB_OK;
  if (b->seq_intervals > 0) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, b->ev0, b->ev1);
    b->kernel_ms += ms;
  }
  for (int32_t s : b->active) {
    Slot& sl = b->slots[s];
    if (!sl.used) continue;
    const SolveCheck& c = b->h_chk[s];
    b->lups += (c.steps - sl.steps_seen) * sl.fp.n;
    sl.steps_seen = c.steps;
    if (c.status >= 0) {
      retire(b, s);
      ++fin;
    }
  }
  if (n_finished) *n_finished = fin;
  return LB_OK;
}

}  // namespace

extern "C" {

int lb_batch_create(int lattice, int precision, int device, int64_t pool_bytes, int32_t check_every,
                    lb_batch** out) {
  if (!out) return fail_global(LB_ERR_ARG, "lb_batch_create: null output");
  *out = nullptr;
  if ((lattice != LB_D2Q9 && lattice != LB_D3Q19) ||
      (precision != LB_FP64 && precision != LB_FP32S && precision != LB_FP16S) || pool_bytes < kAlign ||
      check_every < 1)
    return fail_global(LB_ERR_ARG, "lb_batch_create: bad argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail_global(LB_ERR_HIP, "lb_batch_create: no HIP device available");
  if (device < 0 || device >= ndev) return fail_global(LB_ERR_ARG, "lb_batch_create: device %d out of range", device);
  lb_batch* b = new (std::nothrow) lb_batch;
  if (!b) return fail_global(LB_ERR_NOMEM, "lb_batch_create: out of host memory");
  b->lattice = lattice;
  b->precision = precision;
  b->prec[0] = b->prec[1] = precision;
  b->device = device;
  b->Q = lattice == LB_D2Q9 ? 9 : 19;
  b->nD = lattice == LB_D2Q9 ? 2 : 3;
  b->check_every = check_every;
  hipSetDevice(device);
  hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, device);
  if (b->n_cu <= 0) b->n_cu = 256;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->up, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->up_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipMalloc(&b->pool, static_cast<size_t>(pool_bytes));
  if (e == hipSuccess) e = hipMalloc(&b->d_solves, kMaxSlots * sizeof(BatchSolve));
  if (e == hipSuccess) e = hipMalloc(&b->d_slots, kMaxSlots * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&b->d_out, 3 * kMaxSlots * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&b->d_chk, kMaxSlots * sizeof(SolveCheck));
  if (e == hipSuccess) e = hipHostMalloc(&b->h_chk, kMaxSlots * sizeof(SolveCheck), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&b->h_step, sizeof(uint32_t), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&b->d_step, sizeof(uint32_t));
  if (e != hipSuccess) {
    const int rc = fail_global(e == hipErrorOutOfMemory ? LB_ERR_NOMEM : LB_ERR_HIP, "lb_batch_create: %s",
                               hipGetErrorString(e));
    lb_batch_destroy(b);
    return rc;
  }
  b->alloc.init(pool_bytes / kAlign * kAlign);
  *out = b;
  return LB_OK;
}

int lb_batch_destroy(lb_batch* b) {
  if (!b) return LB_OK;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->up) hipStreamSynchronize(b->up);
  b->staged.clear();
  drop_graphs(b);
  if (b->pool) hipFree(b->pool);
  if (b->d_solves) hipFree(b->d_solves);
  if (b->d_items) hipFree(b->d_items);
  if (b->d_slots) hipFree(b->d_slots);
  if (b->d_out) hipFree(b->d_out);
  if (b->d_chk) hipFree(b->d_chk);
  if (b->h_chk) hipHostFree(b->h_chk);
  if (b->h_step) hipHostFree(b->h_step);
  if (b->d_step) hipFree(b->d_step);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->up_ev) hipEventDestroy(b->up_ev);
  if (b->up) hipStreamDestroy(b->up);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return LB_OK;
}

const char* lb_batch_last_error(const lb_batch* b) { return b ? b->err.c_str() : lb_global_error(); }

int lb_batch_footprint(int lattice, int precision, int32_t n_fields, int32_t nd, const int32_t* dims,
                       const int32_t* geo, int32_t periodic_mask, int64_t* bytes, int64_t* n_fluid) {
  const int want = lattice == LB_D2Q9 ? 2 : 3;
  if ((lattice != LB_D2Q9 && lattice != LB_D3Q19) || nd != want || !dims || !geo)
    return fail_global(LB_ERR_ARG, "lb_batch_footprint: bad argument");
  if (precision != LB_FP64 && precision != LB_FP32S && precision != LB_FP16S)
    return fail_global(LB_ERR_ARG, "lb_batch_footprint: bad precision");
  if (n_fields != 1 && n_fields != 2) return fail_global(LB_ERR_ARG, "lb_batch_footprint: n_fields must be 1 or 2");
  RowBuild rb;
  std::string why;
  if (!rows_from_geometry(lattice, dims, geo, periodic_mask, rb, why))
    return fail_global(LB_ERR_UNSUPPORTED, "lb_batch_footprint: %s", why.c_str());
  const int64_t nwin = static_cast<int64_t>(rb.g.E[0] + 2) * (rb.g.E[1] + 2) * rb.g.W;
  const Footprint fp = footprint_of(lattice == LB_D2Q9 ? 9 : 19, elem_bytes(precision),
                                    static_cast<int64_t>(rb.coord.size()), nwin, n_fields);
  if (bytes) *bytes = fp.bytes;
  if (n_fluid) *n_fluid = fp.n;
  return LB_OK;
}

}  // extern "C"

namespace {

// Parameters of a solve being admitted.
struct AddArgs {
  int kind = 0;                 // 0 single-phase, 1 two-phase
  double tau = 1.0;             // single-phase
  double tau0 = 1.0, tau1 = 1.0, sigma = 0.0, beta = 0.0, wett = 1.0;  // two-phase
  const double* rho0 = nullptr;  // two-phase: phase-0 density per geometry cell
  const double* force = nullptr;
  double tol = 0.0;
  int64_t max_steps = 0;
  int64_t tag = 0;
};

// The host side of an admission: row layout and row windows of a geometry,
// checked node by node.  Pure host work, so many run in parallel.
struct Prepared {
  RowBuild rb;
  std::vector<uint4> win;
  int32_t PY = 0;
  int code = LB_OK;
  std::string why;
};

void prepare(int lattice, const int32_t* dims, const int32_t* geo, int32_t periodic_mask, Prepared& pr) {
  if (!rows_from_geometry(lattice, dims, geo, periodic_mask, pr.rb, pr.why) ||
      !build_windows(pr.rb, pr.win, pr.PY, pr.why)) {
    pr.code = LB_ERR_UNSUPPORTED;
    return;
  }
  const RowBuild& rb = pr.rb;
  const int64_t n = static_cast<int64_t>(rb.coord.size());
  // every node's own window must name it (a cheap check of the construction)
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t c = rb.coord[i];
    const uint32_t X = c >> 22, Y = (c >> 12) & 1023u, Z = c & 4095u, w = Z >> 5, t = Z - 32 * w + 16;
    const uint4 e = pr.win[(static_cast<size_t>(X + 1) * pr.PY + (Y + 1)) * rb.g.W + w];
    const uint64_t m = (static_cast<uint64_t>(e.z) << 32) | e.y;
    if (!((m >> t) & 1u) || e.x + static_cast<uint32_t>(__builtin_popcountll(m & ((1ull << t) - 1ull))) != i) {
      pr.code = LB_ERR_STATE;
      pr.why = "row windows inconsistent at node " + std::to_string(i);
      return;
    }
  }
}

// Host data that an admission's asynchronous uploads read: kept until the
// next wait.
struct Staged {
  std::shared_ptr<Prepared> pre;
  std::vector<double> r0c;
  SolveCheck chk;
};

// Admit a solve.  pre: its prepared rows (built here when null).  The
// uploads and the rest-state kernel run on the admission stream; with sync
// (and no launch sequence in flight) the call waits for them.
int add_solve(lb_batch* b, int32_t nd, const int32_t* dims, const int32_t* geo, int32_t periodic_mask,
              const AddArgs& p, std::shared_ptr<Prepared> pre = nullptr, bool sync = true) {
  hipSetDevice(b->device);
  if (!pre) {
    pre = std::make_shared<Prepared>();
    prepare(b->lattice, dims, geo, periodic_mask, *pre);
  }
  if (pre->code != LB_OK) return bfail(b, pre->code, "lb_batch_add: %s", pre->why.c_str());
  const RowBuild& rb = pre->rb;
  const std::vector<uint4>& win = pre->win;
  const int32_t PY = pre->PY;
  const int64_t n = static_cast<int64_t>(rb.coord.size());
  const int nf = p.kind == 1 ? 2 : 1;
  const int prec = b->prec[p.kind];
  const size_t esz = elem_bytes(prec);
  const Footprint fp = footprint_of(b->Q, esz, n, static_cast<int64_t>(win.size()), nf);
  if (static_cast<uint64_t>(nf) * b->Q * static_cast<uint64_t>(fp.NP) * esz >= (1ull << 32) ||
      fp.nblk > static_cast<int64_t>(kItemMask) + 1)
    return bfail(b, LB_ERR_UNSUPPORTED, "lb_batch_add: %lld fluid nodes exceed one solve", (long long)n);
  int32_t slot = -1;
  for (int32_t s = 0; s < static_cast<int32_t>(b->slots.size()); ++s)
    if (!b->slots[s].used) {
      slot = s;
      break;
    }
  if (slot < 0 && b->slots.size() >= static_cast<size_t>(kMaxSlots))
    return bfail(b, LB_ERR_NOMEM, "lb_batch_add: all %d slots in use", kMaxSlots);
  const int64_t off = b->alloc.alloc(fp.bytes);
  if (off < 0)
    return bfail(b, LB_ERR_NOMEM, "lb_batch_add: %lld bytes do not fit (largest free range %lld of %lld free)",
                 (long long)fp.bytes, (long long)b->alloc.largest(), (long long)(b->alloc.total - b->alloc.used));
  if (slot < 0) {
    b->slots.emplace_back();
    slot = static_cast<int32_t>(b->slots.size()) - 1;
  }
  char* base = b->pool + off;
  BatchSolve d{};
  d.coord = reinterpret_cast<const uint32_t*>(base + fp.off_coord);
  d.win = reinterpret_cast<const uint4*>(base + fp.off_win);
  d.f[0] = base + fp.off_f0;
  d.f[1] = base + fp.off_f1;
  d.partial = reinterpret_cast<double*>(base + fp.off_part);
  d.phi = nf == 2 ? static_cast<void*>(base + fp.off_phi) : nullptr;
  d.NP = static_cast<uint32_t>(fp.NP);
  d.PY = PY;
  d.W = rb.g.W;
  d.n = static_cast<int32_t>(n);
  // the solve starts at the global step of the next launch sequence (the
  // host counter already covers a sequence in flight)
  d.parity = static_cast<int32_t>(b->global_step & 1);  // f[0] holds the state at the next step
  d.step_base = static_cast<uint32_t>(b->global_step);   // the solve's own steps start here
  d.kind = p.kind;
  d.live = 1;
  const double tau = p.kind == 1 ? p.tau0 : p.tau;
  d.a.tau_inv = 1.0 / tau;
  d.a.tau_factor = 1 - 0.5 / tau;
  for (int k = 0; k < 3; ++k) d.a.F[k] = k < b->nD ? p.force[k] : 0.0;
  const double c2 = b->lattice == LB_D2Q9 ? D2Q9::c2 : D3Q19::c2;
  if (p.kind == 1) {
    // as lb_step_cg and lb_set_wall_densities compute them
    d.nu0Inv = 1.0 / (c2 * (p.tau0 - 0.5));
    d.nu1Inv = 1.0 / (c2 * (p.tau1 - 0.5));
    d.sigma = p.sigma;
    d.beta = p.beta;
    const double r0 = p.wett, r1 = 1.0 - p.wett;
    d.phi_w = (r0 - r1) / (r0 + r1);
  }
  auto st = std::make_shared<Staged>();
  st->pre = pre;
  // the stop state, as lb_solve_permeability / lb_solve_relperm set it up
  SolveCheck& c = st->chk;
  c = SolveCheck{};
  c.nu[0] = (tau - 0.5) / 3.0;
  c.nu[1] = (p.tau0 - 0.5) / 3.0;
  c.nu[2] = (p.tau1 - 0.5) / 3.0;
  int64_t nbox = 1;
  for (int k = 0; k < nd; ++k) nbox *= dims[k];
  c.n_box = static_cast<double>(nbox);
  c.F0 = p.force[0];
  c.tol = p.tol;
  c.max_steps = p.max_steps;
  c.status = -1;
  c.kind = p.kind;
  hipStream_t up = b->up;
  hipError_t e = hipMemcpyAsync(base + fp.off_coord, rb.coord.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice,
                                up);
  if (e == hipSuccess)
    e = hipMemcpyAsync(base + fp.off_win, win.data(), win.size() * sizeof(uint4), hipMemcpyHostToDevice, up);
  if (e == hipSuccess) {
    if (p.kind == 0) {
      // rest state f = w: zero in the shifted forms; a pull from it is the
      // identity, so the first step needs no special read
      if (prec == LB_FP64)
        e = batch_rest_f64(b->lattice, reinterpret_cast<double*>(base + fp.off_f0), fp.NP, up);
      else
        e = hipMemsetAsync(base + fp.off_f0, 0, static_cast<size_t>(b->Q) * fp.NP * esz, up);
    } else {
      // f_k = w rho_k at rest, placed on the device where the first pull
      // reads it (k_prestream): the host only maps the phase-0 density of
      // each geometry cell to its compact node
      std::vector<double>& r0c = st->r0c;
      r0c.resize(static_cast<size_t>(n));
      for (int64_t i = 0; i < n; ++i) {
        const uint32_t cw = rb.coord[i];
        const int32_t X = static_cast<int32_t>(cw >> 22), Y = static_cast<int32_t>((cw >> 12) & 1023u),
                      Z = static_cast<int32_t>(cw & 4095u);
        // geometry cell of the compact node (inverse of rows_from_geometry)
        int32_t P[3] = {X, Y, Z};
        for (int a = 0; a < 3; ++a) {
          const int32_t Dm = a == 1 && b->nD == 2 ? 1 : dims[b->nD == 3 ? a : (a == 0 ? 0 : 1)];
          P[a] = (rb.g.periodic >> a) & 1 ? (P[a] + rb.lo[a]) % Dm : P[a] + rb.lo[a];
        }
        const int64_t cell = b->nD == 3 ? (static_cast<int64_t>(P[0]) * dims[1] + P[1]) * dims[2] + P[2]
                                        : static_cast<int64_t>(P[0]) * dims[1] + P[2];
        r0c[i] = p.rho0[cell];
      }
      // staged in the second population buffer (2 Q NP elements, at least
      // 8 n bytes), which the first step overwrites
      double* r0d = reinterpret_cast<double*>(base + fp.off_f1);
      e = hipMemcpyAsync(r0d, r0c.data(), r0c.size() * sizeof(double), hipMemcpyHostToDevice, up);
      if (e == hipSuccess) {
        const RowWin ad{reinterpret_cast<const uint32_t*>(base + fp.off_coord),
                        reinterpret_cast<const uint4*>(base + fp.off_win), static_cast<uint32_t>(fp.NP), PY, rb.g.W};
        e = prec == LB_FP64 ? prestream_f64(b->lattice, ad, r0d, static_cast<int32_t>(n), fp.NP,
                                            reinterpret_cast<double*>(base + fp.off_f0), up)
                            : prestream_lowp(prec, b->lattice, ad, r0d, static_cast<int32_t>(n), fp.NP,
                                             base + fp.off_f0, up);
      }
    }
  }
  // one fixed array (never reallocated: earlier uploads may still read it)
  if (b->h_solves.empty()) b->h_solves.resize(kMaxSlots);
  b->h_solves[slot] = d;
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_solves + slot, &b->h_solves[slot], sizeof(BatchSolve), hipMemcpyHostToDevice, up);
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_chk + slot, &st->chk, sizeof(SolveCheck), hipMemcpyHostToDevice, up);
  if (e == hipSuccess) e = hipEventRecord(b->up_ev, up);
  b->staged.push_back(st);
  b->up_pending = true;
  if (e == hipSuccess && sync && !b->in_flight) e = hipStreamSynchronize(up);
  if (e != hipSuccess) {
    hipStreamSynchronize(up);
    b->alloc.release(off, fp.bytes);
    return bfail(b, LB_ERR_HIP, "lb_batch_add: %s", hipGetErrorString(e));
  }
  Slot& sl = b->slots[slot];
  sl = Slot{};
  sl.used = true;
  sl.kind = p.kind;
  sl.tag = p.tag;
  sl.off = off;
  sl.fp = fp;
  sl.t_admit = now_s();
  b->dirty = true;
  return LB_OK;
}

}  // namespace

extern "C" {

int lb_batch_add(lb_batch* b, int32_t nd, const int32_t* dims, const int32_t* geo, int32_t periodic_mask,
                 double tau, const double* force, double tol, int64_t max_steps, int64_t tag) {
  if (!b) return LB_ERR_ARG;
  if (nd != b->nD || !dims || !geo || !force || !(tau > 0.5) || !(tol > 0.0) || max_steps < 1 || force[0] == 0.0)
    return bfail(b, LB_ERR_ARG, "lb_batch_add: bad argument");
  AddArgs p;
  p.kind = 0;
  p.tau = tau;
  p.force = force;
  p.tol = tol;
  p.max_steps = max_steps;
  p.tag = tag;
  return add_solve(b, nd, dims, geo, periodic_mask, p);
}

int lb_batch_add_many(lb_batch* b, int32_t count, int32_t nd, const int32_t* dims, const int32_t* const* geos,
                      int32_t periodic_mask, double tau, const double* force, double tol, int64_t max_steps,
                      const int64_t* tags, int32_t* n_added) {
  if (!b) return LB_ERR_ARG;
  if (n_added) *n_added = 0;
  if (count < 0 || nd != b->nD || (count > 0 && (!dims || !geos || !tags)) || !force || !(tau > 0.5) ||
      !(tol > 0.0) || max_steps < 1 || force[0] == 0.0)
    return bfail(b, LB_ERR_ARG, "lb_batch_add_many: bad argument");
  for (int32_t k = 0; k < count; ++k)
    if (!geos[k]) return bfail(b, LB_ERR_ARG, "lb_batch_add_many: geometry %d is null", k);
  // row layouts of all geometries on the host cores, then admission in order
  std::vector<std::shared_ptr<Prepared>> pre(count);
  for (auto& x : pre) x = std::make_shared<Prepared>();
  std::atomic<int32_t> next(0);
  const int nthr = std::max(1, std::min<int>(16, std::min<int>(count, static_cast<int>(std::thread::hardware_concurrency()))));
  auto work = [&]() {
    for (int32_t k; (k = next.fetch_add(1)) < count;)
      prepare(b->lattice, dims + static_cast<size_t>(k) * nd, geos[k], periodic_mask, *pre[k]);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nthr; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  AddArgs p;
  p.kind = 0;
  p.tau = tau;
  p.force = force;
  p.tol = tol;
  p.max_steps = max_steps;
  int rc = LB_OK;
  int32_t added = 0;
  for (int32_t k = 0; k < count; ++k) {
    p.tag = tags[k];
    rc = add_solve(b, nd, dims + static_cast<size_t>(k) * nd, geos[k], periodic_mask, p, pre[k], false);
    if (rc != LB_OK) break;
    ++added;
  }
  // report upload errors here unless a launch sequence is in flight (the
  // staged host buffers live until the next wait either way)
  const hipError_t e = b->in_flight ? hipSuccess : hipStreamSynchronize(b->up);
  if (n_added) *n_added = added;
  if (rc == LB_ERR_NOMEM) rc = LB_OK;  // the rest waits for retirements
  if (rc == LB_OK && e != hipSuccess) return bfail(b, LB_ERR_HIP, "lb_batch_add_many: %s", hipGetErrorString(e));
  return rc;
}

int lb_batch_add_twophase(lb_batch* b, int32_t nd, const int32_t* dims, const int32_t* geo, int32_t periodic_mask,
                          const double* rho0, double wettability, double tau0, double tau1, double sigma,
                          double beta, const double* force, double tol, int64_t max_steps, int64_t tag) {
  if (!b) return LB_ERR_ARG;
  if (nd != b->nD || !dims || !geo || !rho0 || !force || !(tau0 > 0.5) || !(tau1 > 0.5) || !(tol > 0.0) ||
      max_steps < 1 || force[0] == 0.0 || !(wettability >= 0.0 && wettability <= 1.0))
    return bfail(b, LB_ERR_ARG, "lb_batch_add_twophase: bad argument");
  AddArgs p;
  p.kind = 1;
  p.tau0 = tau0;
  p.tau1 = tau1;
  p.sigma = sigma;
  p.beta = beta;
  p.wett = wettability;
  p.rho0 = rho0;
  p.force = force;
  p.tol = tol;
  p.max_steps = max_steps;
  p.tag = tag;
  return add_solve(b, nd, dims, geo, periodic_mask, p);
}

int lb_batch_advance(lb_batch* b, int32_t intervals, int32_t* n_finished) {
  if (!b) return LB_ERR_ARG;
  if (n_finished) *n_finished = 0;
  if (intervals < 0) return bfail(b, LB_ERR_ARG, "lb_batch_advance: bad argument");
  if (b->in_flight) return bfail(b, LB_ERR_STATE, "lb_batch_advance: a launch sequence is in flight (lb_batch_wait)");
  hipSetDevice(b->device);
  int rc = launch_sequence(b, intervals);
  if (rc != LB_OK) return rc;
  return wait_sequence(b, n_finished);
}

int lb_batch_launch(lb_batch* b, int32_t intervals) {
  if (!b) return LB_ERR_ARG;
  if (intervals < 0) return bfail(b, LB_ERR_ARG, "lb_batch_launch: bad argument");
  if (b->in_flight) return bfail(b, LB_ERR_STATE, "lb_batch_launch: a launch sequence is in flight (lb_batch_wait)");
  hipSetDevice(b->device);
  return launch_sequence(b, intervals);
}

int lb_batch_wait(lb_batch* b, int32_t* n_finished) {
  if (!b) return LB_ERR_ARG;
  hipSetDevice(b->device);
  return wait_sequence(b, n_finished);
}

int lb_batch_collect(lb_batch* b, int32_t max_n, int64_t* tags, double* k, double* k_phase, int64_t* steps,
                     double* seconds, int32_t* status, int32_t* n_out) {
  if (!b) return LB_ERR_ARG;
  if (max_n < 0 || !n_out) return bfail(b, LB_ERR_ARG, "lb_batch_collect: bad argument");
  const int32_t n = std::min<int32_t>(max_n, static_cast<int32_t>(b->done.size()));
  for (int32_t i = 0; i < n; ++i) {
    const Result& r = b->done[i];
    if (tags) tags[i] = r.tag;
    if (k) k[i] = r.k;
    if (k_phase) {
      k_phase[2 * i] = r.k0;
      k_phase[2 * i + 1] = r.k1;
    }
    if (steps) steps[i] = r.steps;
    if (seconds) seconds[i] = r.seconds;
    if (status) status[i] = r.status;
  }
  b->done.erase(b->done.begin(), b->done.begin() + n);
  *n_out = n;
  return LB_OK;
}

int lb_batch_status(const lb_batch* b, int32_t* n_resident, int64_t* resident_nodes, int64_t* bytes_used,
                    int64_t* bytes_free, int64_t* largest_free) {
  if (!b) return LB_ERR_ARG;
  int32_t nr = 0;
  int64_t nodes = 0;
  for (const Slot& s : b->slots)
    if (s.used) {
      ++nr;
      nodes += s.fp.n;
    }
  if (n_resident) *n_resident = nr;
  if (resident_nodes) *resident_nodes = nodes;
  if (bytes_used) *bytes_used = b->alloc.used;
  if (bytes_free) *bytes_free = b->alloc.total - b->alloc.used;
  if (largest_free) *largest_free = b->alloc.largest();
  return LB_OK;
}

int lb_batch_set_precision(lb_batch* b, int32_t kind, int32_t precision) {
  if (!b) return LB_ERR_ARG;
  if (b->in_flight) return bfail(b, LB_ERR_STATE, "lb_batch_set_precision: a launch sequence is in flight");
  if ((kind != 0 && kind != 1) || (precision != LB_FP64 && precision != LB_FP32S && precision != LB_FP16S))
    return bfail(b, LB_ERR_ARG, "lb_batch_set_precision: bad argument");
  for (const Slot& sl : b->slots)
    if (sl.used && sl.kind == kind)
      return bfail(b, LB_ERR_STATE, "lb_batch_set_precision: solves of this kind are resident");
  b->prec[kind] = precision;
  drop_graphs(b);
  return LB_OK;
}

int lb_batch_set_shape(lb_batch* b, int32_t wg_per_cu) {
  if (!b) return LB_ERR_ARG;
  if (b->in_flight) return bfail(b, LB_ERR_STATE, "lb_batch_set_shape: a launch sequence is in flight");
  if (wg_per_cu < -1 || wg_per_cu > 64) return bfail(b, LB_ERR_ARG, "lb_batch_set_shape: bad argument");
  b->shape_mode = wg_per_cu;
  if (wg_per_cu >= 0) b->shape = wg_per_cu;
  b->tuned_nodes = -1;
  drop_graphs(b);
  return LB_OK;
}

int lb_batch_shape(const lb_batch* b, int32_t* wg_per_cu) {
  if (!b || !wg_per_cu) return LB_ERR_ARG;
  *wg_per_cu = b->shape;
  return LB_OK;
}

int lb_batch_tuning(const lb_batch* b, int32_t* wg_per_cu, int32_t* wg_threads, int64_t* tuned_nodes) {
  if (!b) return LB_ERR_ARG;
  if (wg_per_cu) *wg_per_cu = b->shape;
  if (wg_threads) *wg_threads = kBlock * b->ns;
  if (tuned_nodes) *tuned_nodes = b->tuned_nodes;
  return LB_OK;
}

int lb_batch_kernel_time(const lb_batch* b, double* ms, int64_t* launches, int64_t* lups) {
  if (!b) return LB_ERR_ARG;
  if (ms) *ms = b->kernel_ms;
  if (launches) *launches = b->launches;
  if (lups) *lups = b->lups;
  return LB_OK;
}

}  // extern "C"
