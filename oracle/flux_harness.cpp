// Test infrastructure (the parity checker, never shipped): runs the
// reference's own global-forcing functions on fields read from a file, so
// that the oracle's flux sums can be pinned against them.
//
//   calcCapNumbForceCartDir (LBglobalforcing.h:35-98): the phase-weighted
//     flux sums  sum_bulk phi_k (sum_q f c)_dir  with phi_k = rho_k / rho,
//     combined as 2 (sCa - (m0fc nu0 + m1fc nu1)) / (m0 nu0 + m1 nu1),
//     where f is the total population and rho(k, n) the phase densities;
//   calcFluxForceCartDir (LBglobalforcing.h:8-33): 2 (fixedFlux - mean fc).
//
// The total field is f0 + f1 (valarray addition) and the phase densities are
// calcRho of each field (LBmacroscopic.h:11-20), the quantities the
// reference's colour-gradient loop keeps (main_TWOPHASE.cpp:246-262).
//
// Input (little endian): int32 lattice (0 D2Q9, 1 D3Q19), int32 n_nodes,
// int32 n_bulk, int32 bulk[n_bulk], double f[n_nodes][2][Q] (the LbField
// layout of two fields), int32 n_cases, then per case int32 dir,
// double sCa, double nu0, double nu1, int32 n_global.
// Output: per case one line "cap flux_tot flux_f0 flux_f1" (%.17g): the
// capillary-number force, and calcFluxForceCartDir with fixedFlux = sCa on
// the total field and on the two fields of f.
//
// Built by oracle/build_ref.sh against the reference's headers.
#include <mpi.h>

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <vector>

#include "LBfield.h"
#include "LBglobalforcing.h"
#include "LBlatticetypes.h"
#include "LBmacroscopic.h"

template <typename T>
static bool rd(std::ifstream& in, T* p, size_t n) {
  return static_cast<bool>(in.read(reinterpret_cast<char*>(p), static_cast<std::streamsize>(sizeof(T) * n)));
}

template <typename LT>
static int run(std::ifstream& in, int32_t n_nodes) {
  int32_t n_bulk = 0;
  if (!rd(in, &n_bulk, 1)) return 2;
  std::vector<int32_t> b(n_bulk);
  if (!rd(in, b.data(), b.size())) return 2;
  std::vector<double> raw(static_cast<size_t>(n_nodes) * 2 * LT::nQ);
  if (!rd(in, raw.data(), raw.size())) return 2;
  LbField<LT> f(2, n_nodes);
  for (int n = 0; n < n_nodes; ++n)
    for (int k = 0; k < 2; ++k)
      for (int q = 0; q < LT::nQ; ++q) f(k, q, n) = raw[(static_cast<size_t>(n) * 2 + k) * LT::nQ + q];
  LbField<LT> ftot(1, n_nodes);
  ScalarField rho(2, n_nodes);
  for (int n = 0; n < n_nodes; ++n) {
    const std::valarray<lbBase_t> t = f(0, n) + f(1, n);
    ftot.set(0, n) = t;
    rho(0, n) = calcRho<LT>(f(0, n));
    rho(1, n) = calcRho<LT>(f(1, n));
  }
  const std::vector<int> bulk(b.begin(), b.end());
  int32_t n_cases = 0;
  if (!rd(in, &n_cases, 1)) return 2;
  for (int32_t c = 0; c < n_cases; ++c) {
    int32_t dir = 0, n_global = 0;
    double sca = 0.0, nu0 = 0.0, nu1 = 0.0;
    if (!rd(in, &dir, 1) || !rd(in, &sca, 1) || !rd(in, &nu0, 1) || !rd(in, &nu1, 1) || !rd(in, &n_global, 1))
      return 2;
    const double cap = calcCapNumbForceCartDir(0, ftot, rho, bulk, dir, sca, nu0, nu1, n_global);
    const double ft = calcFluxForceCartDir(0, ftot, bulk, dir, sca, n_global);
    const double f0 = calcFluxForceCartDir(0, f, bulk, dir, sca, n_global);
    const double f1 = calcFluxForceCartDir(1, f, bulk, dir, sca, n_global);
    std::printf("%.17g %.17g %.17g %.17g\n", cap, ft, f0, f1);
  }
  return 0;
}

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  int rc = 1;
  if (argc == 2) {
    std::ifstream in(argv[1], std::ios::binary);
    int32_t lattice = -1, n_nodes = 0;
    if (in && rd(in, &lattice, 1) && rd(in, &n_nodes, 1) && n_nodes > 0)
      rc = lattice == 0 ? run<D2Q9>(in, n_nodes) : (lattice == 1 ? run<D3Q19>(in, n_nodes) : 2);
  }
  if (rc != 0) std::fprintf(stderr, "usage: flux_harness <input file>; bad or short input\n");
  MPI_Finalize();
  return rc;
}
