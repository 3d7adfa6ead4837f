// Test infrastructure (the parity checker, never shipped): runs the
// reference's own initiateLbField (LBinitiatefield.h:33-57, the overload
// with separate population, density and velocity field numbers) on fields
// read from a file and writes the populations it sets.
//
// Input (little endian): int32 lattice (0 D2Q9, 1 D3Q19), int32 n_nodes,
// int32 n_bulk, int32 bulk[n_bulk], int32 n_lb_fields, int32 n_rho_fields,
// int32 n_vel_fields, double rho[n_nodes][n_rho_fields],
// double vel[n_nodes][n_vel_fields][nD], int32 n_calls, then per call
// int32 lbFieldNo, rhoFieldNo, velFieldNo.
// Output file: double f[n_nodes][n_lb_fields][Q] (the LbField layout), zero
// where no call wrote.
//
// Built by oracle/build_ref.sh against the reference's headers.
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <vector>

#include "LBfield.h"
#include "LBinitiatefield.h"
#include "LBlatticetypes.h"

template <typename T>
static bool rd(std::ifstream& in, T* p, size_t n) {
  return static_cast<bool>(in.read(reinterpret_cast<char*>(p), static_cast<std::streamsize>(sizeof(T) * n)));
}

template <typename LT>
static int run(std::ifstream& in, int32_t n_nodes, const char* out_path) {
  int32_t n_bulk = 0, nlb = 0, nrho = 0, nvel = 0;
  if (!rd(in, &n_bulk, 1)) return 2;
  std::vector<int32_t> b(n_bulk);
  if (!rd(in, b.data(), b.size()) || !rd(in, &nlb, 1) || !rd(in, &nrho, 1) || !rd(in, &nvel, 1)) return 2;
  std::vector<double> r(static_cast<size_t>(n_nodes) * nrho), v(static_cast<size_t>(n_nodes) * nvel * LT::nD);
  if (!rd(in, r.data(), r.size()) || !rd(in, v.data(), v.size())) return 2;
  ScalarField rho(nrho, n_nodes);
  VectorField<LT> vel(nvel, n_nodes);
  for (int n = 0; n < n_nodes; ++n) {
    for (int k = 0; k < nrho; ++k) rho(k, n) = r[static_cast<size_t>(n) * nrho + k];
    for (int k = 0; k < nvel; ++k)
      for (int d = 0; d < LT::nD; ++d) vel(k, d, n) = v[(static_cast<size_t>(n) * nvel + k) * LT::nD + d];
  }
  LbField<LT> f(nlb, n_nodes);
  const std::vector<int> bulk(b.begin(), b.end());
  int32_t n_calls = 0;
  if (!rd(in, &n_calls, 1)) return 2;
  for (int32_t c = 0; c < n_calls; ++c) {
    int32_t a[3];
    if (!rd(in, a, 3)) return 2;
    initiateLbField(a[0], a[1], a[2], bulk, rho, vel, f);
  }
  std::vector<double> o(static_cast<size_t>(n_nodes) * nlb * LT::nQ);
  for (int n = 0; n < n_nodes; ++n)
    for (int k = 0; k < nlb; ++k)
      for (int q = 0; q < LT::nQ; ++q) o[(static_cast<size_t>(n) * nlb + k) * LT::nQ + q] = f(k, q, n);
  std::FILE* fo = std::fopen(out_path, "wb");
  if (!fo) return 3;
  const bool ok = std::fwrite(o.data(), sizeof(double), o.size(), fo) == o.size();
  std::fclose(fo);
  return ok ? 0 : 3;
}

int main(int argc, char** argv) {
  int rc = 1;
  if (argc == 3) {
    std::ifstream in(argv[1], std::ios::binary);
    int32_t lattice = -1, n_nodes = 0;
    if (in && rd(in, &lattice, 1) && rd(in, &n_nodes, 1) && n_nodes > 0)
      rc = lattice == 0 ? run<D2Q9>(in, n_nodes, argv[2]) : (lattice == 1 ? run<D3Q19>(in, n_nodes, argv[2]) : 2);
  }
  if (rc != 0) std::fprintf(stderr, "usage: init_harness <input file> <output file>; bad or short input\n");
  return rc;
}
