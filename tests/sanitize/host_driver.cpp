// Host-side entry points of the library (node list, .vtklb text and binary
// files, attributes, sphere packs, VTU output, error paths), run under
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitize_cpu.py.
// The run fails on any sanitizer report or on a mismatch between the text and
// binary round trips.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/badchimp_amd.h"

namespace {

int failures = 0;

void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s (%s)\n", what, lb_global_error());
    ++failures;
  }
}

struct Arrays {
  std::vector<int32_t> nb, pos, ranks, bulk;
  std::vector<int16_t> types;
};

Arrays arrays_of(const lb_topo* t, int Q) {
  Arrays a;
  const int32_t n = lb_topo_num_nodes(t), nd = lb_topo_num_dims(t);
  a.nb.resize(static_cast<size_t>(n) * Q);
  a.pos.resize(static_cast<size_t>(n) * nd);
  a.ranks.resize(n);
  a.types.resize(n);
  a.bulk.resize(lb_topo_num_bulk(t));
  check(lb_topo_neighbors(t, a.nb.data()) == LB_OK, "neighbors");
  check(lb_topo_positions(t, a.pos.data()) == LB_OK, "positions");
  check(lb_topo_node_ranks(t, a.ranks.data()) == LB_OK, "ranks");
  check(lb_topo_node_types(t, a.types.data()) == LB_OK, "types");
  if (!a.bulk.empty()) check(lb_topo_bulk(t, a.bulk.data()) == LB_OK, "bulk");
  return a;
}

bool same(const Arrays& a, const Arrays& b) {
  return a.nb == b.nb && a.pos == b.pos && a.ranks == b.ranks && a.types == b.types && a.bulk == b.bulk;
}

// One geometry through every host path: writer numbering, text and binary
// files, attributes, VTU output of the bulk cells.
void round_trip(const std::string& dir, int lattice, int nd, const int32_t* dims, const std::vector<int32_t>& geo,
                int periodic, int rank, const char* tag) {
  const int Q = lattice == LB_D2Q9 ? 9 : 19;
  lb_topo* t = nullptr;
  if (lb_topo_from_geometry(lattice, nd, dims, geo.data(), periodic, rank, &t) != LB_OK) {
    check(false, tag);
    return;
  }
  const Arrays a = arrays_of(t, Q);
  const int32_t n = lb_topo_num_nodes(t);
  std::vector<double> attr(n);
  for (int32_t i = 0; i < n; ++i) attr[i] = 0.25 + 1e-3 * (i % 97);
  const char* names[1] = {"init_rho"};
  const double* vals[1] = {attr.data()};
  const std::string txt = dir + "/" + tag + ".vtklb", bin = dir + "/" + tag + ".vtklbb";
  check(lb_topo_write_vtklb(t, txt.c_str(), 1, names, vals) == LB_OK, "write vtklb");
  check(lb_topo_write_binary(t, bin.c_str()) == LB_OK, "write binary");
  lb_topo* r1 = nullptr;
  lb_topo* r2 = nullptr;
  check(lb_topo_read_vtklb(txt.c_str(), lattice, &r1) == LB_OK, "read vtklb");
  if (r1) {
    check(lb_topo_write_binary(r1, bin.c_str()) == LB_OK, "write binary of the text read");
    check(lb_topo_read_binary(bin.c_str(), lattice, &r2) == LB_OK, "read binary");
  }
  if (r1 && r2) {
    const Arrays b = arrays_of(r1, Q), c = arrays_of(r2, Q);
    if (rank == 1) check(same(a, b), "text read equals the writer's node list");
    check(same(b, c), "binary read equals text read");
    std::vector<double> x(n), y(n);
    check(lb_topo_attribute(r1, "init_rho", 1, x.data()) == LB_OK, "attribute text");
    check(lb_topo_attribute(r2, "init_rho", 1, y.data()) == LB_OK, "attribute binary");
    check(x == y, "attribute text = binary");
    check(lb_topo_attribute(r1, "no_such_attribute", 0, x.data()) != LB_OK, "missing attribute refused");
  }
  // VTU output of the bulk cells, one scalar and one vector
  const int32_t nbulk = static_cast<int32_t>(a.bulk.size());
  if (nbulk > 0) {
    std::vector<int32_t> pos(static_cast<size_t>(nbulk) * nd);
    std::vector<double> s(nbulk), v(static_cast<size_t>(nbulk) * nd);
    for (int32_t k = 0; k < nbulk; ++k) {
      for (int d = 0; d < nd; ++d) {
        pos[static_cast<size_t>(k) * nd + d] = a.pos[static_cast<size_t>(a.bulk[k]) * nd + d];
        v[static_cast<size_t>(k) * nd + d] = 0.1 * d + 1e-4 * k;
      }
      s[k] = 1.0 + 1e-6 * k;
    }
    lb_vtu* out = nullptr;
    if (lb_vtu_create(nd, pos.data(), nbulk, dir.c_str(), 0, 1, LB_VTU_F64, &out) == LB_OK) {
      check(lb_vtu_add_file(out, tag) == LB_OK, "vtu file");
      check(lb_vtu_add_variable(out, "rho", 1, s.data()) == LB_OK, "vtu scalar");
      check(lb_vtu_add_variable(out, "vel", nd, v.data()) == LB_OK, "vtu vector");
      check(lb_vtu_write(out, 0) == LB_OK, "vtu write 0");
      check(lb_vtu_write(out, 10) == LB_OK, "vtu write 10");
      check(lb_vtu_num_points(out) > 0, "vtu points");
      lb_vtu_destroy(out);
    } else {
      check(false, "vtu create");
    }
  }
  if (r1) lb_topo_destroy(r1);
  if (r2) lb_topo_destroy(r2);
  lb_topo_destroy(t);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: host_driver <scratch dir>\n");
    return 2;
  }
  const std::string dir = argv[1];
  // sphere packs, 3-D and 2-D, with and without walls
  {
    const int32_t dims[3] = {18, 14, 12};
    std::vector<int32_t> geo(18 * 14 * 12);
    double info[4];
    check(lb_sphere_pack(3, dims, 2.5, 0.5, 7, 1, 1, 2000, 0.97, geo.data(), info) == LB_OK, "sphere pack 3-D");
    round_trip(dir, LB_D3Q19, 3, dims, geo, 1, 1, "pack3_x_walls");
    check(lb_sphere_pack(3, dims, 2.0, 0.6, 8, 7, 0, 2000, 0.97, geo.data(), info) == LB_OK, "sphere pack periodic");
    round_trip(dir, LB_D3Q19, 3, dims, geo, 7, 1, "pack3_xyz");
    // two ranks: the second half of x belongs to rank 2
    std::vector<int32_t> two = geo;
    for (int32_t x = 9; x < 18; ++x)
      for (int32_t k = 0; k < 14 * 12; ++k)
        if (two[static_cast<size_t>(x) * 14 * 12 + k] > 0) two[static_cast<size_t>(x) * 14 * 12 + k] = 2;
    round_trip(dir, LB_D3Q19, 3, dims, two, 7, 2, "pack3_rank2");
  }
  {
    const int32_t dims[2] = {24, 16};
    std::vector<int32_t> geo(24 * 16);
    double info[4];
    check(lb_sphere_pack(2, dims, 2.0, 0.6, 9, 1, 1, 2000, 0.97, geo.data(), info) == LB_OK, "sphere pack 2-D");
    round_trip(dir, LB_D2Q9, 2, dims, geo, 1, 1, "pack2_x_walls");
  }
  // edge geometries: all solid, a single fluid cell, a thin open box
  {
    const int32_t dims[3] = {4, 4, 3};
    std::vector<int32_t> geo(48, 0);
    lb_topo* none = nullptr;
    check(lb_topo_from_geometry(LB_D3Q19, 3, dims, geo.data(), 0, 1, &none) != LB_OK, "all-solid geometry refused");
    geo[1 * 12 + 1 * 3 + 1] = 1;
    round_trip(dir, LB_D3Q19, 3, dims, geo, 0, 1, "single");
    std::vector<int32_t> open(48, 1);
    round_trip(dir, LB_D3Q19, 3, dims, open, 7, 1, "open");
  }
  // error paths: bad arguments, missing and truncated files
  {
    lb_topo* t = nullptr;
    const int32_t dims[3] = {0, 4, 4};
    int32_t g[1] = {1};
    check(lb_topo_from_geometry(LB_D3Q19, 3, dims, g, 0, 1, &t) != LB_OK, "empty dims refused");
    check(lb_topo_read_vtklb((dir + "/missing.vtklb").c_str(), LB_D3Q19, &t) != LB_OK, "missing file refused");
    const std::string trunc = dir + "/trunc.vtklbb";
    FILE* f = std::fopen(trunc.c_str(), "wb");
    if (f) {
      const char junk[40] = "LBTB";
      std::fwrite(junk, 1, sizeof(junk), f);
      std::fclose(f);
    }
    check(lb_topo_read_binary(trunc.c_str(), LB_D3Q19, &t) != LB_OK, "truncated binary refused");
    const std::string bad = dir + "/bad.vtklb";
    f = std::fopen(bad.c_str(), "w");
    if (f) {
      std::fputs("# vtk DataFile Version 3.0\nnot a vtklb file\n", f);
      std::fclose(f);
    }
    check(lb_topo_read_vtklb(bad.c_str(), LB_D3Q19, &t) != LB_OK, "malformed text refused");
    lb_vtu* v = nullptr;
    check(lb_vtu_create(3, nullptr, 5, dir.c_str(), 0, 1, LB_VTU_F64, &v) != LB_OK, "vtu without positions refused");
  }
  if (failures) {
    std::fprintf(stderr, "%d failure(s)\n", failures);
    return 1;
  }
  std::printf("host paths clean\n");
  return 0;
}
