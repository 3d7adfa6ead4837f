// Host-side sanitizer driver (test infrastructure): built with g++ and
// AddressSanitizer + UndefinedBehaviorSanitizer from the library's host-only
// sources (csrc/topology.cpp, csrc/vtu.cpp), it runs the node-list builder,
// the .vtklb writer and reader, the binary twin, every accessor, the sphere
// pack and the VTU writer on small geometries of both lattices, one rank and
// several, a single fluid cell, and the error paths (an all-solid box,
// missing and truncated files, missing attributes, bad arguments).  Exit status 0 and no sanitizer report
// is the pass.
#include <stdlib.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../include/badchimp_amd.h"

#define CHECK(x)                                                                                     \
  do {                                                                                               \
    if (!(x)) {                                                                                      \
      std::fprintf(stderr, "check failed at line %d: %s (%s)\n", __LINE__, #x, lb_global_error()); \
      return 1;                                                                                      \
    }                                                                                                \
  } while (0)

namespace {

int same_topology(lb_topo* a, lb_topo* b, int Q) {
  const int32_t n = lb_topo_num_nodes(a), nd = lb_topo_num_dims(a);
  CHECK(n == lb_topo_num_nodes(b) && lb_topo_num_bulk(a) == lb_topo_num_bulk(b) && nd == lb_topo_num_dims(b));
  std::vector<int32_t> na(size_t(n) * Q), nb(size_t(n) * Q), pa(size_t(n) * nd), pb(size_t(n) * nd);
  std::vector<int16_t> ta(n), tb(n);
  CHECK(lb_topo_neighbors(a, na.data()) == 0 && lb_topo_neighbors(b, nb.data()) == 0);
  CHECK(lb_topo_positions(a, pa.data()) == 0 && lb_topo_positions(b, pb.data()) == 0);
  CHECK(lb_topo_node_types(a, ta.data()) == 0 && lb_topo_node_types(b, tb.data()) == 0);
  CHECK(na == nb && pa == pb && ta == tb);
  return 0;
}

int exercise(lb_topo* t, int lattice, const std::string& dir, const std::string& tag) {
  const int Q = lattice == LB_D2Q9 ? 9 : 19;
  const int32_t n = lb_topo_num_nodes(t), nb = lb_topo_num_bulk(t), nd = lb_topo_num_dims(t);
  CHECK(n >= 1 && nb >= 0 && nd == (lattice == LB_D2Q9 ? 2 : 3));
  std::vector<int32_t> nbr(size_t(n) * Q), pos(size_t(n) * nd), ranks(n), bulk(std::max(nb, 1));
  std::vector<int16_t> types(n);
  int32_t gd[3] = {0, 0, 0};
  CHECK(lb_topo_global_dims(t, gd) == 0);
  CHECK(lb_topo_neighbors(t, nbr.data()) == 0);
  CHECK(lb_topo_positions(t, pos.data()) == 0);
  CHECK(lb_topo_node_types(t, types.data()) == 0);
  CHECK(lb_topo_node_ranks(t, ranks.data()) == 0);
  CHECK(lb_topo_bulk(t, bulk.data()) == 0);
  for (size_t k = 0; k < nbr.size(); ++k) CHECK(nbr[k] >= 0 && nbr[k] < n);
  for (int32_t i = 1; i < n; ++i) CHECK(nbr[size_t(i) * Q + Q - 1] == i);  // the rest direction
  for (int32_t k = 0; k < nb; ++k) CHECK(bulk[k] > 0 && bulk[k] < n && types[bulk[k]] >= 2);
  const int32_t nr = lb_topo_num_neighbor_ranks(t);
  CHECK(nr >= 0);
  for (int32_t k = 0; k < nr; ++k) {
    int32_t r = -1, cnt = -1;
    CHECK(lb_topo_neighbor_rank(t, k, &r, &cnt, nullptr, nullptr) == 0 && cnt > 0);
    std::vector<int32_t> local(cnt), remote(cnt);
    CHECK(lb_topo_neighbor_rank(t, k, &r, &cnt, local.data(), remote.data()) == 0);
    for (int32_t j = 0; j < cnt; ++j) CHECK(local[j] > 0 && local[j] < n && remote[j] > 0);
  }
  CHECK(lb_topo_neighbor_rank(t, nr, nullptr, nullptr, nullptr, nullptr) != 0);

  // text file with an attribute, read back; binary twin, read back
  std::vector<double> attr(n);
  for (int32_t i = 0; i < n; ++i) attr[i] = 0.25 * i - 1.0;
  const char* names[1] = {"a"};
  const double* values[1] = {attr.data()};
  const std::string text = dir + "/" + tag + ".vtklb", bin = dir + "/" + tag + ".lbbin";
  CHECK(lb_topo_write_vtklb(t, text.c_str(), 1, names, values) == 0);
  lb_topo* r = nullptr;
  CHECK(lb_topo_read_vtklb(text.c_str(), lattice, &r) == 0);
  if (same_topology(t, r, Q)) return 1;
  std::vector<double> back(n), backf(n);
  CHECK(lb_topo_attribute(r, "a", 0, back.data()) == 0 && lb_topo_attribute(r, "a", 1, backf.data()) == 0);
  for (int32_t i = 1; i < n; ++i) CHECK(backf[i] == double(float(attr[i])));
  CHECK(lb_topo_attribute(r, "missing", 0, back.data()) != 0);
  CHECK(lb_topo_write_binary(r, bin.c_str()) == 0);
  lb_topo* rb = nullptr;
  CHECK(lb_topo_read_binary(bin.c_str(), lattice, &rb) == 0);
  if (same_topology(r, rb, Q)) return 1;
  CHECK(lb_topo_destroy(rb) == 0 && lb_topo_destroy(r) == 0);

  // truncated copies of both files are refused, not read past their end
  for (const std::string& src : {text, bin}) {
    std::ifstream in(src, std::ios::binary);
    std::string all((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    for (size_t cut : {all.size() / 3, all.size() * 2 / 3, all.size() - 1}) {
      const std::string part = src + ".part";
      std::ofstream(part, std::ios::binary).write(all.data(), static_cast<std::streamsize>(cut));
      lb_topo* bad = nullptr;
      const int rc = src == text ? lb_topo_read_vtklb(part.c_str(), lattice, &bad)
                                 : lb_topo_read_binary(part.c_str(), lattice, &bad);
      if (rc == 0) lb_topo_destroy(bad);
      // only the text file that lacks no more than its final newline is whole
      CHECK(rc != 0 || (src == text && cut == all.size() - 1));
    }
  }

  // VTU output of the bulk cells, two writes
  if (nb > 0) {
    std::vector<int32_t> cells(size_t(nb) * nd);
    std::vector<double> rho(nb), vel(size_t(nb) * nd);
    for (int32_t k = 0; k < nb; ++k) {
      for (int d = 0; d < nd; ++d) cells[size_t(k) * nd + d] = pos[size_t(bulk[k]) * nd + d];
      rho[k] = 1.0 + 1e-3 * k;
      for (int d = 0; d < nd; ++d) vel[size_t(k) * nd + d] = 1e-4 * (d + 1);
    }
    lb_vtu* v = nullptr;
    CHECK(lb_vtu_create(nd, cells.data(), nb, (dir + "/vtu_" + tag).c_str(), 0, 1, LB_VTU_F64, &v) == 0);
    CHECK(lb_vtu_add_file(v, "fluid") == 0);
    CHECK(lb_vtu_add_variable(v, "rho", 1, rho.data()) == 0);
    CHECK(lb_vtu_add_variable(v, "vel", nd, vel.data()) == 0);
    CHECK(lb_vtu_write(v, 0.0) == 0 && lb_vtu_write(v, 1.0) == 0);
    CHECK(lb_vtu_num_points(v) > nb);
    CHECK(lb_vtu_destroy(v) == 0);
  }
  return 0;
}

int from_geometry(int lattice, const std::vector<int32_t>& dims, const std::vector<int32_t>& geo, int periodic,
                  int ranks, const std::string& dir, const std::string& tag) {
  for (int rank = 1; rank <= ranks; ++rank) {
    lb_topo* t = nullptr;
    CHECK(lb_topo_from_geometry(lattice, int32_t(dims.size()), dims.data(), geo.data(), periodic, rank, &t) == 0);
    CHECK(lb_topo_rank(t) == rank - 1);  // 0-based, as the file names its process
    if (exercise(t, lattice, dir, tag + "_r" + std::to_string(rank))) return 1;
    CHECK(lb_topo_destroy(t) == 0);
  }
  return 0;
}

// fluid cells labelled 1..ranks by slabs along the first axis
std::vector<int32_t> label_slabs(std::vector<int32_t> g, int32_t n0, int ranks) {
  const size_t per = g.size() / size_t(n0);
  for (size_t c = 0; c < g.size(); ++c)
    if (g[c] > 0) g[c] = 1 + int32_t((c / per) * size_t(ranks) / size_t(n0));
  return g;
}

}  // namespace

int main() {
  char tmpl[] = "/tmp/lbsanXXXXXX";
  const char* d = mkdtemp(tmpl);
  CHECK(d != nullptr);
  const std::string dir = d;
  std::mt19937 rng(7);

  // D2Q9: random porous box, walls on y, periodic x; one rank and three
  {
    const std::vector<int32_t> dims = {18, 11};
    std::vector<int32_t> geo(size_t(dims[0]) * dims[1]);
    for (int32_t x = 0; x < dims[0]; ++x)
      for (int32_t y = 0; y < dims[1]; ++y)
        geo[size_t(x) * dims[1] + y] = (y == 0 || y == dims[1] - 1) ? 0 : int32_t(rng() % 10 < 7);
    if (from_geometry(LB_D2Q9, dims, geo, 1, 1, dir, "d2")) return 1;
    if (from_geometry(LB_D2Q9, dims, label_slabs(geo, dims[0], 3), 1, 3, dir, "d2s")) return 1;
    if (from_geometry(LB_D2Q9, dims, geo, 3, 1, dir, "d2p")) return 1;
  }
  // D3Q19: sphere pack, periodic x with walls, and fully periodic; one rank and two
  {
    const std::vector<int32_t> dims = {14, 12, 10};
    std::vector<int32_t> geo(size_t(dims[0]) * dims[1] * dims[2]);
    double info[4];
    CHECK(lb_sphere_pack(3, dims.data(), 2.5, 0.5, 11, 1, 1, 2000, 0.95, geo.data(), info) == 0);
    CHECK(info[0] > 0.0 && info[0] < 1.0);
    if (from_geometry(LB_D3Q19, dims, geo, 1, 1, dir, "d3")) return 1;
    if (from_geometry(LB_D3Q19, dims, label_slabs(geo, dims[0], 2), 1, 2, dir, "d3s")) return 1;
    CHECK(lb_sphere_pack(3, dims.data(), 2.5, 0.5, 12, 7, 0, 2000, 0.95, geo.data(), info) == 0);
    if (from_geometry(LB_D3Q19, dims, geo, 7, 1, dir, "d3p")) return 1;
    // all solid: no cell of rank 1, refused
    std::vector<int32_t> solid(geo.size(), 0);
    lb_topo* none = nullptr;
    CHECK(lb_topo_from_geometry(LB_D3Q19, 3, dims.data(), solid.data(), 1, 1, &none) != 0 && none == nullptr);
    // a single fluid cell in a closed box: one bulk node, all links to walls
    solid[size_t(5) * dims[1] * dims[2] + size_t(6) * dims[2] + 4] = 1;
    if (from_geometry(LB_D3Q19, dims, solid, 0, 1, dir, "d3one")) return 1;
  }
  // error paths
  {
    lb_topo* t = nullptr;
    const int32_t dims[3] = {4, 4, 4};
    std::vector<int32_t> geo(64, 1);
    CHECK(lb_topo_read_vtklb((dir + "/absent.vtklb").c_str(), LB_D3Q19, &t) != 0);
    CHECK(lb_topo_read_binary((dir + "/absent.lbbin").c_str(), LB_D3Q19, &t) != 0);
    CHECK(lb_topo_from_geometry(LB_D3Q19, 3, dims, nullptr, 0, 1, &t) != 0);
    CHECK(lb_topo_from_geometry(LB_D3Q19, 4, dims, geo.data(), 0, 1, &t) != 0);
    CHECK(lb_topo_from_geometry(LB_D2Q9, 3, dims, geo.data(), 0, 1, &t) != 0);
    CHECK(lb_topo_from_geometry(7, 3, dims, geo.data(), 0, 1, &t) != 0);
    CHECK(lb_topo_from_geometry(LB_D3Q19, 3, dims, geo.data(), 0, 1, nullptr) != 0);
    const int32_t zero[3] = {4, 0, 4};
    CHECK(lb_topo_from_geometry(LB_D3Q19, 3, zero, geo.data(), 0, 1, &t) != 0);
    double info[4];
    CHECK(lb_sphere_pack(3, dims, -1.0, 0.5, 1, 0, 0, 10, 0.9, geo.data(), info) != 0);
    lb_vtu* v = nullptr;
    const int32_t cell[3] = {0, 0, 0};
    CHECK(lb_vtu_create(3, cell, 1, dir.c_str(), 0, 1, 9, &v) != 0);
    CHECK(lb_topo_destroy(nullptr) != 0 || true);
  }
  std::printf("host sanitizer run: ok\n");
  return 0;
}
