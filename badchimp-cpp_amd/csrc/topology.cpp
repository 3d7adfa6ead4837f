// Host topology: .vtklb reader and writer, native node-list builder with the
// reference writer's numbering, node types and bulk list, sphere packs.
#include "topology.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <set>
#include <sstream>

#include "../../include/badchimp_amd.h"
#include "errors.hpp"
#include "host_threads.hpp"
#include "lattice.hpp"

using lbamd::fail_global;

namespace {

template <class L>
int c2q(const std::vector<int>& v) {
  for (int q = 0; q < L::nQ; ++q) {
    bool eq = true;
    for (int d = 0; d < L::nD; ++d) eq = eq && (L::c(q, d) == v[d]);
    if (eq) return q;
  }
  return -1;
}

int lattice_c(int lattice, int q, int d) {
  return lattice == LB_D2Q9 ? lbamd::D2Q9::c(q, d) : lbamd::D3Q19::c(q, d);
}

// ------------------------------------------------------------ tokenizer
// Mirrors the reads LBvtk does with `ifs >> x` (whitespace-separated tokens)
// and std::getline (rest of the current line).
struct Cursor {
  const std::string& s;
  size_t p = 0;
  explicit Cursor(const std::string& str) : s(str) {}
  bool token(std::string& out) {
    while (p < s.size() && std::isspace(static_cast<unsigned char>(s[p]))) ++p;
    if (p >= s.size()) return false;
    size_t b = p;
    while (p < s.size() && !std::isspace(static_cast<unsigned char>(s[p]))) ++p;
    out.assign(s, b, p - b);
    return true;
  }
  bool integer(long& v) {
    std::string t;
    if (!token(t)) return false;
    char* end = nullptr;
    errno = 0;
    v = std::strtol(t.c_str(), &end, 10);
    return errno == 0 && end && *end == '\0';
  }
  void rest_of_line() {
    while (p < s.size() && s[p] != '\n') ++p;
    if (p < s.size()) ++p;
  }
  size_t tell() const { return p; }
  void seek(size_t q) { p = q; }
};

}  // namespace

namespace lbamd {

void finish_topology(lb_topo& t) {
  const int32_t N = t.n_nodes;
  const int Q = t.Q;
  t.node_type.assign(N, -1);
  // Nodes ctor: types from the 'nodetype' attribute (LBnodes.h:130-135)
  const int begin = t.zero_ghost ? 1 : 0;
  for (int32_t n = begin; n < N; ++n) t.node_type[n] = static_cast<int16_t>(t.nodetype_attr[n]);
  // setupNodeType (LBnodes.h:156-209)
  t.node_type[0] = -1;
  for (int32_t n = 1; n < N; ++n) t.node_type[n] = (t.node_type[n] == 0) ? 0 : 3;
  auto is_solid = [&](int32_t n) { return t.node_type[n] < 2; };
  auto is_fluid = [&](int32_t n) { return t.node_type[n] > 1; };
  for (int32_t n = 1; n < N; ++n) {
    const int32_t* nb = &t.neighbor[static_cast<size_t>(n) * Q];
    if (is_solid(n)) {
      bool has_fluid = false;
      for (int q = 0; q < Q; ++q) has_fluid = has_fluid || is_fluid(nb[q]);
      t.node_type[n] = has_fluid ? 1 : 0;
    } else {
      if (t.rank != t.node_rank[n]) {
        t.node_type[n] = 4;
      } else {
        bool has_solid = false;
        for (int q = 0; q < Q; ++q) has_solid = has_solid || is_solid(nb[q]);
        t.node_type[n] = has_solid ? 2 : 3;
      }
    }
  }
  // findBulkNodes (LBgeometry.h:11-21)
  t.bulk.clear();
  for (int32_t n = 1; n < N; ++n)
    if (is_fluid(n) && t.node_rank[n] == t.rank) t.bulk.push_back(n);
}

}  // namespace lbamd

// ====================================================================== reader
namespace {

int read_vtklb(const std::string& text, const char* path, int lattice, lb_topo& t) {
  Cursor c(text);
  std::string s;
  auto bad = [&](const char* what) {
    return fail_global(LB_ERR_IO, "vtklb %s: %s", path, what);
  };
  // preamble (LBvtk.h:265-290)
  c.rest_of_line();
  if (!c.token(s)) return bad("missing title");
  c.rest_of_line();
  if (!c.token(s) || (s != "ASCII" && s != "BINARY")) return bad("data type must be ASCII or BINARY");
  if (s != "ASCII") return bad("BINARY vtklb is not supported");
  c.rest_of_line();
  // DATASET (LBvtk.h:293-315)
  if (!c.token(s) || s != "DATASET") return bad("expected DATASET");
  if (!c.token(s) || s != "UNSTRUCTURED_LB_GRID") return bad("unknown DATASET type");
  c.rest_of_line();
  // optional NUM_DIMENSIONS, GLOBAL_DIMENSIONS, optional USE_ZERO_GHOST_NODE (LBvtk.h:318-359)
  int nD = 3;
  size_t mark = c.tell();
  c.token(s);
  if (s == "NUM_DIMENSIONS") {
    long v;
    if (!c.integer(v)) return bad("bad NUM_DIMENSIONS");
    nD = static_cast<int>(v);
    c.rest_of_line();
  } else {
    c.seek(mark);
  }
  if (!c.token(s) || s != "GLOBAL_DIMENSIONS") return bad("expected GLOBAL_DIMENSIONS");
  for (int d = 0; d < nD; ++d) {
    long v;
    if (!c.integer(v)) return bad("bad GLOBAL_DIMENSIONS");
    t.global_dims.push_back(static_cast<int32_t>(v));
  }
  c.rest_of_line();
  mark = c.tell();
  c.token(s);
  t.zero_ghost = false;
  if (s == "USE_ZERO_GHOST_NODE") {
    t.zero_ghost = true;
    c.rest_of_line();
  } else {
    c.seek(mark);
  }
  const int Q = lattice == LB_D2Q9 ? 9 : 19;
  const int lnD = lattice == LB_D2Q9 ? 2 : 3;
  if (nD != lnD) return bad("NUM_DIMENSIONS does not match the lattice");
  t.lattice = lattice;
  t.nD = nD;
  t.Q = Q;
  // POINTS (LBvtk.h:362-383)
  long npts;
  if (!c.token(s) || s != "POINTS" || !c.integer(npts) || npts < 0) return bad("expected POINTS n");
  c.token(s);
  c.rest_of_line();
  const int begin = t.zero_ghost ? 1 : 0;
  t.n_nodes = static_cast<int32_t>(npts + begin);
  t.pos.assign(static_cast<size_t>(t.n_nodes) * nD, -1);
  for (long n = 0; n < npts; ++n)
    for (int d = 0; d < nD; ++d) {
      long v;
      if (!c.integer(v)) return bad("bad POINTS entry");
      t.pos[(n + begin) * nD + d] = static_cast<int32_t>(v);
    }
  c.rest_of_line();
  // LATTICE (LBvtk.h:398-440)
  long nq;
  if (!c.token(s) || s != "LATTICE" || !c.integer(nq)) return bad("expected LATTICE");
  c.token(s);
  c.rest_of_line();
  if (nq != Q) return bad("wrong number of basis vectors for the lattice");
  std::vector<int> f2p(Q);
  std::vector<bool> seen(Q, false);
  for (int q = 0; q < Q; ++q) {
    std::vector<int> v(nD);
    for (int d = 0; d < nD; ++d) {
      long x;
      if (!c.integer(x)) return bad("bad LATTICE entry");
      v[d] = static_cast<int>(x);
    }
    f2p[q] = lattice == LB_D2Q9 ? c2q<lbamd::D2Q9>(v) : c2q<lbamd::D3Q19>(v);
    if (f2p[q] < 0) return bad("LATTICE vector not in the lattice");
    seen[f2p[q]] = true;
  }
  for (bool b : seen)
    if (!b) return bad("LATTICE basis is not complete");
  c.rest_of_line();
  // NEIGHBORS (LBvtk.h:443-482), stored in program basis order
  if (!c.token(s) || s != "NEIGHBORS") return bad("expected NEIGHBORS");
  c.rest_of_line();
  t.neighbor.assign(static_cast<size_t>(t.n_nodes) * Q, 0);
  for (long n = 0; n < npts; ++n)
    for (int q = 0; q < Q; ++q) {
      long v;
      if (!c.integer(v)) return bad("bad NEIGHBORS entry");
      if (v < 0 || v >= t.n_nodes) return bad("neighbour index out of range");
      t.neighbor[(n + begin) * Q + f2p[q]] = static_cast<int32_t>(v);
    }
  c.rest_of_line();
  // PARALLEL_COMPUTING + PROCESSOR blocks (LBvtk.h:485-545)
  long rank;
  if (!c.token(s) || s != "PARALLEL_COMPUTING" || !c.integer(rank)) return bad("expected PARALLEL_COMPUTING");
  t.rank = static_cast<int>(rank);
  c.rest_of_line();
  for (;;) {
    mark = c.tell();
    if (!c.token(s) || s != "PROCESSOR") {
      c.seek(mark);
      break;
    }
    long cnt, r;
    if (!c.integer(cnt) || !c.integer(r)) return bad("bad PROCESSOR header");
    c.rest_of_line();
    std::vector<std::pair<int32_t, int32_t>> pairs(cnt);
    for (long i = 0; i < cnt; ++i) {
      long a, b;
      if (!c.integer(a) || !c.integer(b)) return bad("bad PROCESSOR entry");
      pairs[i] = {static_cast<int32_t>(a), static_cast<int32_t>(b)};
    }
    c.rest_of_line();
    t.processors.emplace_back(static_cast<int>(r), std::move(pairs));
  }
  // POINT_DATA + SCALARS blocks (LBvtk.h:548-587)
  long npd;
  if (!c.token(s) || s != "POINT_DATA" || !c.integer(npd)) return bad("expected POINT_DATA");
  c.rest_of_line();
  for (;;) {
    mark = c.tell();
    if (!c.token(s) || s != "SCALARS") {
      c.seek(mark);
      break;
    }
    std::string name, type;
    c.token(name);
    c.token(type);
    c.rest_of_line();
    std::vector<std::string> vals(npd);
    for (long i = 0; i < npd; ++i)
      if (!c.token(vals[i])) return bad("truncated SCALARS block");
    c.rest_of_line();
    t.attributes.emplace(name, std::make_pair(type, std::move(vals)));
  }
  // node ranks (Nodes ctor, LBnodes.h:120-128)
  t.node_rank.assign(t.n_nodes, t.rank);
  t.node_rank[0] = -1;
  for (auto& blk : t.processors)
    for (auto& pr : blk.second)
      if (pr.first >= 0 && pr.first < t.n_nodes) t.node_rank[pr.first] = blk.first;
  // nodetype attribute (LBnodes.h:130-135)
  auto it = t.attributes.find("nodetype");
  if (it == t.attributes.end()) return bad("data attribute 'nodetype' does not exist");
  t.nodetype_attr.assign(t.n_nodes, 0);
  for (long i = 0; i < npts && i < npd; ++i)
    t.nodetype_attr[i + begin] = static_cast<int32_t>(std::strtol(it->second.second[i].c_str(), nullptr, 10));
  lbamd::finish_topology(t);
  return LB_OK;
}

// ====================================================================== builder
struct Padded {
  std::vector<int32_t> dims;  // padded
  std::vector<int64_t> stride;
  int64_t size = 1;
  explicit Padded(const std::vector<int32_t>& d) : dims(d), stride(d.size()) {
    for (int i = static_cast<int>(d.size()) - 1; i >= 0; --i) {
      stride[i] = size;
      size *= d[i];
    }
  }
};

int build_from_geometry(int lattice, int nd, const int32_t* dims, const int32_t* geo,
                        int periodic, int rank, lb_topo& t) {
  const int Q = lattice == LB_D2Q9 ? 9 : 19;
  std::vector<int32_t> pd(nd);
  for (int d = 0; d < nd; ++d) pd[d] = dims[d] + 2;
  Padded P(pd);
  const int64_t M = P.size;
  std::vector<int32_t> g(M, -1), label(M, 0);
  // Interior cells in C order, one run along the last axis at a time: the
  // padded index of the first cell of interior run r (the last axis has
  // stride 1).
  int64_t n_in = 1;
  for (int d = 0; d < nd; ++d) n_in *= dims[d];
  const int32_t nlast = dims[nd - 1];
  const int64_t n_runs = n_in / nlast;
  auto run_start = [&](int64_t r) {
    int64_t flat = 1;
    for (int d = nd - 2; d >= 0; --d) {
      flat += (r % dims[d] + 1) * P.stride[d];
      r /= dims[d];
    }
    return flat;
  };
  // interior copy + per-rank labels in C order (vtklb.py:84-94); the own
  // fluid cells in C order (== label order) come out of the same pass
  // (vtklb.py:184-190), since the rim copies below touch no interior cell
  int max_rank = 0;
  std::vector<int64_t> nodes;
  {
    std::vector<int32_t> counter(1, 0);
    const int32_t* src = geo;
    for (int64_t r = 0; r < n_runs; ++r) {
      const int64_t flat0 = run_start(r);
      for (int32_t k = 0; k < nlast; ++k) {
        const int32_t v = *src++;
        if (v < 0) return fail_global(LB_ERR_ARG, "geometry values must be >= 0");
        const int64_t flat = flat0 + k;
        g[flat] = v;
        if (v > 0) {
          if (v >= static_cast<int32_t>(counter.size())) counter.resize(v + 1, 0);
          label[flat] = ++counter[v];
          max_rank = std::max(max_rank, v);
          if (v == rank) nodes.push_back(flat);
        }
      }
    }
  }
  if (rank < 1 || rank > max_rank) return fail_global(LB_ERR_ARG, "rank %d not present in the geometry", rank);
  // periodic rim copies, axis by axis (vtklb.py:97-114): the layer k = 0 of
  // the axis takes layer n - 2, then layer n - 1 takes layer 1; each layer is
  // a run of stride[ax] cells per index of the axes before it
  for (int ax = 0; ax < nd; ++ax) {
    if (!(periodic & (1 << ax))) continue;
    const int64_t n = pd[ax], inner = P.stride[ax], outer = M / (n * inner);
    for (int pass = 0; pass < 2; ++pass) {
      const int64_t dst_k = pass == 0 ? 0 : n - 1, src_k = pass == 0 ? n - 2 : 1;
      for (int64_t o = 0; o < outer; ++o) {
        const int64_t dst = (o * n + dst_k) * inner, src = (o * n + src_k) * inner;
        std::copy(g.begin() + src, g.begin() + src + inner, g.begin() + dst);
        std::copy(label.begin() + src, label.begin() + src + inner, label.begin() + dst);
      }
    }
  }
  std::vector<int64_t> shift(Q);
  for (int q = 0; q < Q; ++q) {
    shift[q] = 0;
    for (int d = 0; d < nd; ++d) shift[q] += lattice_c(lattice, q, d) * P.stride[d];
  }
  const size_t n_fluid = nodes.size();
  // solid cells reached from own fluid, C order of the padded array (vtklb.py:192-197)
  // (host threads mark with relaxed byte stores; each notes the other ranks it meets)
  std::vector<uint8_t> mark(M, 0);
  std::set<int> nb_ranks;
  {
    std::vector<std::set<int>> seen(lbamd::kMaxHostThreads);
    uint8_t* mk = mark.data();
    lbamd::parallel_ranges(static_cast<int64_t>(n_fluid), [&](int64_t b, int64_t e, int th) {
      std::set<int>& sn = seen[th];
      int32_t last = 0;
      for (int64_t k = b; k < e; ++k)
        for (int q = 0; q < Q; ++q) {
          const int64_t j = nodes[k] + shift[q];
          const int32_t v = g[j];
          if (v == 0) __atomic_store_n(mk + j, uint8_t{1}, __ATOMIC_RELAXED);
          else if (v > 0 && v != rank && v != last) {
            sn.insert(v);
            last = v;
          }
        }
    });
    for (const auto& sn : seen) nb_ranks.insert(sn.begin(), sn.end());
  }
  for (int64_t i = 0; i < M; ++i)
    if (mark[i]) nodes.push_back(i);
  // other-rank cells, one block per neighbouring rank (vtklb.py:199-215);
  // ascending rank order, the writer's set order for small rank numbers
  for (int r : nb_ranks) {
    std::fill(mark.begin(), mark.end(), 0);
    for (size_t k = 0; k < n_fluid; ++k)
      for (int q = 0; q < Q; ++q) {
        const int64_t j = nodes[k] + shift[q];
        if (g[j] == r) mark[j] = 1;
      }
    for (int64_t i = 0; i < M; ++i)
      if (mark[i]) nodes.push_back(i);
  }
  const int32_t npts = static_cast<int32_t>(nodes.size());
  std::vector<int32_t> local(M, 0);
  for (int32_t n = 0; n < npts; ++n) local[nodes[n]] = n + 1;

  t.lattice = lattice;
  t.nD = nd;
  t.Q = Q;
  t.rank = rank - 1;
  t.zero_ghost = true;
  t.global_dims = pd;
  t.n_nodes = npts + 1;
  t.pos.assign(static_cast<size_t>(t.n_nodes) * nd, -1);
  t.neighbor.assign(static_cast<size_t>(t.n_nodes) * Q, 0);
  t.nodetype_attr.assign(t.n_nodes, 0);
  // positions and neighbour index (vtklb.py:23-39), node by node on host threads
  lbamd::parallel_ranges(npts, [&](int64_t b, int64_t e, int) {
    int32_t idx[3];
    for (int64_t n = b; n < e; ++n) {
      const int64_t i = nodes[n];
      for (int d = 0; d < nd; ++d) {
        idx[d] = static_cast<int32_t>((i / P.stride[d]) % pd[d]);
        t.pos[static_cast<size_t>(n + 1) * nd + d] = idx[d] - 1;
      }
      for (int q = 0; q < Q; ++q) {
        bool inside = true;
        for (int d = 0; d < nd; ++d) {
          const int32_t k = idx[d] + lattice_c(lattice, q, d);
          inside = inside && k >= 0 && k < pd[d];
        }
        int32_t v = 0;
        if (inside) {
          const int64_t j = i + shift[q];
          v = local[j];
          if (v == 0 && g[j] == rank) v = label[j];
        }
        t.neighbor[static_cast<size_t>(n + 1) * Q + q] = v;
      }
      t.nodetype_attr[n + 1] = g[i] > 0 ? 1 : 0;
    }
  });
  // PROCESSOR blocks (vtklb.py:271-281)
  for (int r : nb_ranks) {
    std::vector<std::pair<int32_t, int32_t>> pairs;
    for (int32_t n = 0; n < npts; ++n)
      if (g[nodes[n]] == r) pairs.emplace_back(n + 1, label[nodes[n]]);
    t.processors.emplace_back(r - 1, std::move(pairs));
  }
  t.node_rank.assign(t.n_nodes, t.rank);
  t.node_rank[0] = -1;
  for (auto& blk : t.processors)
    for (auto& pr : blk.second) t.node_rank[pr.first] = blk.first;
  lbamd::finish_topology(t);
  return LB_OK;
}

// ====================================================================== sphere pack
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* lb_global_error(void) { return lbamd::global_error().c_str(); }

int lb_topo_read_vtklb(const char* path, int lattice, lb_topo** out) {
  if (!path || !out || (lattice != LB_D2Q9 && lattice != LB_D3Q19))
    return fail_global(LB_ERR_ARG, "lb_topo_read_vtklb: bad argument");
  std::ifstream ifs(path, std::ios::binary);
  if (!ifs) return fail_global(LB_ERR_IO, "could not open file %s", path);
  std::stringstream ss;
  ss << ifs.rdbuf();
  std::string text = ss.str();
  auto t = std::make_unique<lb_topo>();
  int rc = read_vtklb(text, path, lattice, *t);
  if (rc != LB_OK) return rc;
  *out = t.release();
  return LB_OK;
}

int lb_topo_from_geometry(int lattice, int32_t nd, const int32_t* dims, const int32_t* geo,
                          int32_t periodic, int32_t rank, lb_topo** out) {
  if (!dims || !geo || !out || (lattice != LB_D2Q9 && lattice != LB_D3Q19))
    return fail_global(LB_ERR_ARG, "lb_topo_from_geometry: bad argument");
  if (nd != (lattice == LB_D2Q9 ? 2 : 3)) return fail_global(LB_ERR_ARG, "nd does not match lattice");
  for (int d = 0; d < nd; ++d)
    if (dims[d] < 1) return fail_global(LB_ERR_ARG, "dims must be positive");
  try {
    auto t = std::make_unique<lb_topo>();
    int rc = build_from_geometry(lattice, nd, dims, geo, periodic, rank, *t);
    if (rc != LB_OK) return rc;
    *out = t.release();
  } catch (const std::bad_alloc&) {
    return fail_global(LB_ERR_NOMEM, "lb_topo_from_geometry: out of host memory");
  }
  return LB_OK;
}

// ---- binary twin of .vtklb ------------------------------------------------
// Layout (little-endian): "LBVTKB01", int32 {lattice, nD, Q, rank, zero_ghost,
// n_nodes}, int32 global_dims[nD], int32 pos[n_nodes*nD], int32
// neighbor[n_nodes*Q], int32 nodetype_attr[n_nodes], int32 node_rank[n_nodes],
// int32 n_blocks, per block {int32 rank, int32 count, int32 pairs[2*count]},
// int32 n_attr, per attribute {int32 name length, name, int32 type length,
// type, int64 count, double values[count], float values[count]}.
int lb_topo_write_binary(const lb_topo* t, const char* path) {
  if (!t || !path) return fail_global(LB_ERR_ARG, "null argument");
  std::unique_ptr<FILE, int (*)(FILE*)> fh(std::fopen(path, "wb"), std::fclose);
  if (!fh) return fail_global(LB_ERR_IO, "could not open %s for writing", path);
  FILE* f = fh.get();
  auto put = [&](const void* p, size_t bytes) { return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes; };
  auto put32 = [&](int32_t v) { return put(&v, 4); };
  bool ok = put("LBVTKB01", 8);
  const int32_t head[6] = {t->lattice, t->nD, t->Q, t->rank, t->zero_ghost ? 1 : 0, t->n_nodes};
  ok = ok && put(head, sizeof(head));
  ok = ok && put(t->global_dims.data(), 4 * t->global_dims.size());
  ok = ok && put(t->pos.data(), 4 * t->pos.size());
  ok = ok && put(t->neighbor.data(), 4 * t->neighbor.size());
  ok = ok && put(t->nodetype_attr.data(), 4 * t->nodetype_attr.size());
  ok = ok && put(t->node_rank.data(), 4 * t->node_rank.size());
  ok = ok && put32(static_cast<int32_t>(t->processors.size()));
  for (auto& blk : t->processors) {
    ok = ok && put32(blk.first) && put32(static_cast<int32_t>(blk.second.size()));
    for (auto& pr : blk.second) ok = ok && put32(pr.first) && put32(pr.second);
  }
  ok = ok && put32(static_cast<int32_t>(t->attributes.size() + t->attr_values.size()));
  auto put_attr = [&](const std::string& name, const std::string& type, const std::vector<double>& dv,
                      const std::vector<float>& fv) {
    const int64_t n = static_cast<int64_t>(dv.size());
    return put32(static_cast<int32_t>(name.size())) && put(name.data(), name.size()) &&
           put32(static_cast<int32_t>(type.size())) && put(type.data(), type.size()) && put(&n, 8) &&
           put(dv.data(), 8 * dv.size()) && put(fv.data(), 4 * fv.size());
  };
  for (auto& a : t->attributes) {
    const auto& toks = a.second.second;
    std::vector<double> dv(toks.size());
    std::vector<float> fv(toks.size());
    for (size_t i = 0; i < toks.size(); ++i) {
      // `ifs >> double` and `ifs >> float` each round the decimal text once
      dv[i] = std::strtod(toks[i].c_str(), nullptr);
      fv[i] = std::strtof(toks[i].c_str(), nullptr);
    }
    ok = ok && put_attr(a.first, a.second.first, dv, fv);
  }
  for (auto& a : t->attr_values) ok = ok && put_attr(a.first, a.second.type, a.second.d, a.second.f);
  ok = ok && std::fflush(f) == 0;
  return ok ? LB_OK : fail_global(LB_ERR_IO, "write error on %s", path);
}

int lb_topo_read_binary(const char* path, int lattice, lb_topo** out) {
  if (!path || !out || (lattice != LB_D2Q9 && lattice != LB_D3Q19))
    return fail_global(LB_ERR_ARG, "lb_topo_read_binary: bad argument");
  std::unique_ptr<FILE, int (*)(FILE*)> fh(std::fopen(path, "rb"), std::fclose);
  if (!fh) return fail_global(LB_ERR_IO, "could not open file %s", path);
  FILE* f = fh.get();
  auto get = [&](void* p, size_t bytes) { return bytes == 0 || std::fread(p, 1, bytes, f) == bytes; };
  auto bad = [&](const char* what) { return fail_global(LB_ERR_IO, "%s: %s", path, what); };
  char magic[8];
  int32_t head[6];
  if (!get(magic, 8) || std::memcmp(magic, "LBVTKB01", 8) != 0) return bad("not a binary vtklb file");
  if (!get(head, sizeof(head))) return bad("truncated header");
  const int Q = lattice == LB_D2Q9 ? 9 : 19, nD = lattice == LB_D2Q9 ? 2 : 3;
  if (head[0] != lattice || head[1] != nD || head[2] != Q)
    return fail_global(LB_ERR_ARG, "%s holds lattice %d (nD %d, Q %d), %d was asked for", path, head[0], head[1],
                       head[2], lattice);
  if (head[5] < 1) return bad("bad node count");
  try {
    auto t = std::make_unique<lb_topo>();
    t->lattice = lattice;
    t->nD = nD;
    t->Q = Q;
    t->rank = head[3];
    t->zero_ghost = head[4] != 0;
    t->n_nodes = head[5];
    const size_t N = static_cast<size_t>(t->n_nodes);
    t->global_dims.resize(nD);
    t->pos.resize(N * nD);
    t->neighbor.resize(N * Q);
    t->nodetype_attr.resize(N);
    t->node_rank.resize(N);
    if (!get(t->global_dims.data(), 4 * nD) || !get(t->pos.data(), 4 * t->pos.size()) ||
        !get(t->neighbor.data(), 4 * t->neighbor.size()) || !get(t->nodetype_attr.data(), 4 * N) ||
        !get(t->node_rank.data(), 4 * N))
      return bad("truncated arrays");
    for (int32_t v : t->neighbor)
      if (v < 0 || v >= t->n_nodes) return bad("neighbour index out of range");
    int32_t nblk = 0;
    if (!get(&nblk, 4) || nblk < 0) return bad("bad PROCESSOR block count");
    for (int32_t b = 0; b < nblk; ++b) {
      int32_t r = 0, cnt = 0;
      if (!get(&r, 4) || !get(&cnt, 4) || cnt < 0) return bad("bad PROCESSOR block");
      std::vector<std::pair<int32_t, int32_t>> pairs(cnt);
      for (auto& pr : pairs)
        if (!get(&pr.first, 4) || !get(&pr.second, 4)) return bad("truncated PROCESSOR block");
      t->processors.emplace_back(r, std::move(pairs));
    }
    int32_t nattr = 0;
    if (!get(&nattr, 4) || nattr < 0) return bad("bad attribute count");
    for (int32_t a = 0; a < nattr; ++a) {
      int32_t ln = 0, lt = 0;
      int64_t n = 0;
      std::string name, type;
      if (!get(&ln, 4) || ln < 0 || ln > 4096) return bad("bad attribute name");
      name.resize(ln);
      if (!get(&name[0], ln) || !get(&lt, 4) || lt < 0 || lt > 4096) return bad("bad attribute type");
      type.resize(lt);
      if (!get(&type[0], lt) || !get(&n, 8) || n < 0 || n > t->n_nodes) return bad("bad attribute size");
      std::vector<double> dv(n);
      std::vector<float> fv(n);
      if (!get(dv.data(), 8 * n) || !get(fv.data(), 4 * n)) return bad("truncated attribute");
      t->attr_values.emplace(name, lb_topo::AttrValues{type, std::move(dv), std::move(fv)});
    }
    // node types and bulk list exactly as after the text reader
    lbamd::finish_topology(*t);
    *out = t.release();
  } catch (const std::bad_alloc&) {
    return fail_global(LB_ERR_NOMEM, "lb_topo_read_binary: out of host memory");
  }
  return LB_OK;
}

int lb_topo_destroy(lb_topo* t) {
  delete t;
  return LB_OK;
}
int32_t lb_topo_num_nodes(const lb_topo* t) { return t ? t->n_nodes : -1; }
int32_t lb_topo_num_bulk(const lb_topo* t) { return t ? static_cast<int32_t>(t->bulk.size()) : -1; }
int32_t lb_topo_num_dims(const lb_topo* t) { return t ? t->nD : -1; }
int32_t lb_topo_rank(const lb_topo* t) { return t ? t->rank : -1; }

int lb_topo_global_dims(const lb_topo* t, int32_t* dims) {
  if (!t || !dims) return fail_global(LB_ERR_ARG, "null argument");
  std::copy(t->global_dims.begin(), t->global_dims.end(), dims);
  return LB_OK;
}
int lb_topo_neighbors(const lb_topo* t, int32_t* out) {
  if (!t || !out) return fail_global(LB_ERR_ARG, "null argument");
  std::copy(t->neighbor.begin(), t->neighbor.end(), out);
  return LB_OK;
}
int lb_topo_positions(const lb_topo* t, int32_t* out) {
  if (!t || !out) return fail_global(LB_ERR_ARG, "null argument");
  std::copy(t->pos.begin(), t->pos.end(), out);
  return LB_OK;
}
int lb_topo_node_types(const lb_topo* t, int16_t* out) {
  if (!t || !out) return fail_global(LB_ERR_ARG, "null argument");
  std::copy(t->node_type.begin(), t->node_type.end(), out);
  return LB_OK;
}
int lb_topo_node_ranks(const lb_topo* t, int32_t* out) {
  if (!t || !out) return fail_global(LB_ERR_ARG, "null argument");
  std::copy(t->node_rank.begin(), t->node_rank.end(), out);
  return LB_OK;
}
int32_t lb_topo_num_neighbor_ranks(const lb_topo* t) { return t ? static_cast<int32_t>(t->processors.size()) : -1; }

int lb_topo_neighbor_rank(const lb_topo* t, int32_t k, int32_t* rank, int32_t* n_nodes, int32_t* local,
                          int32_t* remote) {
  if (!t || k < 0 || k >= static_cast<int32_t>(t->processors.size()))
    return fail_global(LB_ERR_ARG, "lb_topo_neighbor_rank: no neighbouring rank %d", k);
  const auto& blk = t->processors[static_cast<size_t>(k)];
  if (rank) *rank = blk.first;
  if (n_nodes) *n_nodes = static_cast<int32_t>(blk.second.size());
  for (size_t i = 0; i < blk.second.size(); ++i) {
    if (local) local[i] = blk.second[i].first;
    if (remote) remote[i] = blk.second[i].second;
  }
  return LB_OK;
}

int lb_topo_bulk(const lb_topo* t, int32_t* out) {
  if (!t || !out) return fail_global(LB_ERR_ARG, "null argument");
  std::copy(t->bulk.begin(), t->bulk.end(), out);
  return LB_OK;
}

int lb_topo_attribute(const lb_topo* t, const char* name, int32_t as_float, double* out) {
  if (!t || !name || !out) return fail_global(LB_ERR_ARG, "null argument");
  const int begin = t->zero_ghost ? 1 : 0;
  auto it = t->attributes.find(name);
  if (it == t->attributes.end()) {
    auto jt = t->attr_values.find(name);
    if (jt == t->attr_values.end())
      return fail_global(LB_ERR_IO, "Data attribute with name %s do not exist.", name);
    if (begin) out[0] = 0.0;
    const size_t n = jt->second.d.size();
    for (size_t i = 0; i < n && static_cast<int64_t>(i) + begin < t->n_nodes; ++i)
      out[i + begin] = as_float ? static_cast<double>(jt->second.f[i]) : jt->second.d[i];
    return LB_OK;
  }
  const auto& vals = it->second.second;
  if (begin) out[0] = 0.0;
  for (size_t i = 0; i < vals.size() && static_cast<int64_t>(i) + begin < t->n_nodes; ++i) {
    // `ifs >> float` / `ifs >> double` both round the decimal text directly
    out[i + begin] = as_float ? static_cast<double>(std::strtof(vals[i].c_str(), nullptr))
                              : std::strtod(vals[i].c_str(), nullptr);
  }
  return LB_OK;
}

int lb_topo_write_vtklb(const lb_topo* t, const char* path, int32_t n_attr,
                        const char* const* names, const double* const* values) {
  if (!t || !path || (n_attr > 0 && (!names || !values)))
    return fail_global(LB_ERR_ARG, "null argument");
  if (!t->zero_ghost) return fail_global(LB_ERR_ARG, "writer needs a zero-ghost topology");
  FILE* fh = std::fopen(path, "w");
  if (!fh) return fail_global(LB_ERR_IO, "could not open %s for writing", path);
  const int nD = t->nD, Q = t->Q;
  const int32_t npts = t->n_nodes - 1;
  std::fprintf(fh, "# BADChIMP vtklb Version na\nGeometry file for process %d\nASCII\n", t->rank);
  std::fprintf(fh, "DATASET UNSTRUCTURED_LB_GRID\nNUM_DIMENSIONS %d\nGLOBAL_DIMENSIONS", nD);
  for (int32_t v : t->global_dims) std::fprintf(fh, " %d", v);
  std::fprintf(fh, "\nUSE_ZERO_GHOST_NODE\nPOINTS %d int\n", npts);
  for (int32_t n = 1; n <= npts; ++n) {
    for (int d = 0; d < nD; ++d) std::fprintf(fh, d ? " %d" : "%d", t->pos[static_cast<size_t>(n) * nD + d]);
    std::fputc('\n', fh);
  }
  std::fprintf(fh, "LATTICE %d int\n", Q);
  for (int q = 0; q < Q; ++q) {
    for (int d = 0; d < nD; ++d) std::fprintf(fh, d ? " %d" : "%d", lattice_c(t->lattice, q, d));
    std::fputc('\n', fh);
  }
  std::fprintf(fh, "NEIGHBORS int\n");
  for (int32_t n = 1; n <= npts; ++n) {
    for (int q = 0; q < Q; ++q) std::fprintf(fh, q ? " %d" : "%d", t->neighbor[static_cast<size_t>(n) * Q + q]);
    std::fputc('\n', fh);
  }
  std::fprintf(fh, "PARALLEL_COMPUTING %d\n", t->rank);
  for (auto& blk : t->processors) {
    std::fprintf(fh, "PROCESSOR %zu %d\n", blk.second.size(), blk.first);
    for (auto& pr : blk.second) std::fprintf(fh, "%d %d\n", pr.first, pr.second);
  }
  std::fprintf(fh, "POINT_DATA %d\nSCALARS nodetype int\n", npts);
  for (int32_t n = 1; n <= npts; ++n) std::fprintf(fh, "%d\n", t->nodetype_attr[n]);
  for (int32_t a = 0; a < n_attr; ++a) {
    std::fprintf(fh, "SCALARS %s float\n", names[a]);
    for (int32_t n = 1; n <= npts; ++n) std::fprintf(fh, "%.18e\n", values[a][n]);
  }
  const bool ok = std::ferror(fh) == 0;
  std::fclose(fh);
  return ok ? LB_OK : fail_global(LB_ERR_IO, "write error on %s", path);
}

int lb_sphere_pack(int32_t nd, const int32_t* dims, double radius, double porosity,
                   uint64_t seed, int32_t periodic, int32_t walls, int32_t max_fail,
                   double relax, int32_t* geo, double* info) {
  if (nd < 2 || nd > 3 || !dims || !geo || radius <= 0.0 || porosity < 0.0 || porosity > 1.0 ||
      max_fail < 1 || !(relax > 0.0 && relax < 1.0))
    return fail_global(LB_ERR_ARG, "lb_sphere_pack: bad argument");
  int64_t total = 1;
  std::vector<int64_t> stride(nd);
  for (int d = nd - 1; d >= 0; --d) {
    if (dims[d] < 1) return fail_global(LB_ERR_ARG, "lb_sphere_pack: dims must be positive");
    stride[d] = total;
    total *= dims[d];
  }
  std::fill(geo, geo + total, 1);
  std::vector<bool> wrap(nd);
  for (int d = 0; d < nd; ++d) wrap[d] = (periodic >> d) & 1;
  // walls on the faces of non-periodic axes
  int64_t n_inner = 1;
  for (int d = 0; d < nd; ++d) n_inner *= (walls && !wrap[d]) ? std::max(dims[d] - 2, 0) : dims[d];
  if (walls) {
    for (int64_t i = 0; i < total; ++i)
      for (int d = 0; d < nd; ++d) {
        if (wrap[d]) continue;
        const int64_t k = (i / stride[d]) % dims[d];
        if (k == 0 || k == dims[d] - 1) geo[i] = 0;
      }
  }
  SplitMix rng{seed};
  const double r = radius;
  const int ir = static_cast<int>(std::ceil(r));
  double min_dist = 2.0 * r;
  // cell list with cells of at least 2r
  std::vector<int32_t> ncell(nd);
  int64_t n_cells = 1;
  for (int d = 0; d < nd; ++d) {
    ncell[d] = std::max(1, static_cast<int32_t>(dims[d] / (2.0 * r)));
    n_cells *= ncell[d];
  }
  std::vector<std::vector<int32_t>> cells(n_cells);
  std::vector<double> centres;
  auto cell_coord = [&](const double* c, int d) {
    return std::min(static_cast<int32_t>(c[d] / dims[d] * ncell[d]), ncell[d] - 1);
  };
  auto too_close = [&](const double* c) {
    int32_t k[3] = {0, 0, 0};
    for (int d = 0; d < nd; ++d) k[d] = cell_coord(c, d);
    const int span = nd == 3 ? 27 : 9;
    for (int o = 0; o < span; ++o) {
      int32_t kk[3];
      int oo = o;
      bool ok = true;
      int64_t flat = 0;
      for (int d = 0; d < nd; ++d) {
        int v = k[d] + (oo % 3) - 1;
        oo /= 3;
        if (wrap[d]) {
          v = ((v % ncell[d]) + ncell[d]) % ncell[d];
        } else if (v < 0 || v >= ncell[d]) {
          ok = false;
          break;
        }
        kk[d] = v;
        flat = flat * ncell[d] + v;
      }
      if (!ok) continue;
      (void)kk;
      for (int32_t j : cells[flat]) {
        double d2 = 0.0;
        for (int d = 0; d < nd; ++d) {
          double diff = std::fabs(centres[j * nd + d] - c[d]);
          if (wrap[d]) diff = std::min(diff, dims[d] - diff);
          d2 += diff * diff;
        }
        if (d2 < min_dist * min_dist) return true;
      }
    }
    return false;
  };
  // with small cell counts several offsets map to one cell; deduplicate by
  // checking each candidate against a cell at most once is not needed for
  // correctness (the test is idempotent)
  int64_t fluid = 0;
  for (int64_t i = 0; i < total; ++i) fluid += geo[i] != 0;
  int64_t trials = 0, fails = 0;
  const int64_t max_trials = 200000000;
  double c[3] = {0, 0, 0};
  while (n_inner > 0 && static_cast<double>(fluid) / n_inner > porosity && trials < max_trials) {
    ++trials;
    for (int d = 0; d < nd; ++d) c[d] = rng.uniform() * dims[d];
    if (too_close(c)) {
      if (++fails >= max_fail) {
        min_dist *= relax;
        fails = 0;
      }
      continue;
    }
    fails = 0;
    const int32_t id = static_cast<int32_t>(centres.size() / nd);
    for (int d = 0; d < nd; ++d) centres.push_back(c[d]);
    int64_t flat = 0;
    for (int d = 0; d < nd; ++d) flat = flat * ncell[d] + cell_coord(c, d);
    cells[flat].push_back(id);
    // stamp voxels with centre distance <= r
    int32_t base[3];
    for (int d = 0; d < nd; ++d) base[d] = static_cast<int32_t>(std::floor(c[d]));
    const int w = 2 * ir + 1;
    const int64_t box = nd == 3 ? static_cast<int64_t>(w) * w * w : static_cast<int64_t>(w) * w;
    for (int64_t b = 0; b < box; ++b) {
      int64_t bb = b, idx = 0;
      double d2 = 0.0;
      bool ok = true;
      for (int d = nd - 1; d >= 0; --d) {
        int32_t k = base[d] + static_cast<int32_t>(bb % w) - ir;
        bb /= w;
        const double diff = k - c[d];
        d2 += diff * diff;
        if (wrap[d]) {
          k = ((k % dims[d]) + dims[d]) % dims[d];
        } else if (k < 0 || k >= dims[d]) {
          ok = false;
          break;
        }
        idx += k * stride[d];
      }
      if (!ok || d2 > r * r) continue;
      if (geo[idx] != 0) {
        geo[idx] = 0;
        --fluid;
      }
    }
  }
  if (info) {
    info[0] = n_inner > 0 ? static_cast<double>(fluid) / n_inner : 0.0;
    info[1] = static_cast<double>(centres.size() / nd);
    info[2] = min_dist;
    info[3] = static_cast<double>(trials);
  }
  return LB_OK;
}

}  // extern "C"
