// Drop-in check of the C++ shim (lbsolver_gpu.hpp) against the reference's
// own set-up code: a std_case-style driver that builds LBvtk, Grid, Nodes,
// the bulk list and the fields with the reference's lbsolver headers, then
// hands the step loop to the GPU through lbamd::GpuSubdomain.
//
//   std_case_gpu <case.vtklb> <tau> <iterations> <precision 0|1|2> <out.bin> <F_x> <F_y> [<F_z>]
//
// Runs iterations + 1 steps (the loop of std_case/main.cpp:109 runs
// i = 0..max) and writes, for each bulk node in findBulkNodes order, rho and
// the nD velocity components as doubles (the fields the loop stores).
// Built by oracle/build_ref.sh where the reference tree is present.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "LBSOLVER.h"
#include "lbamd/lbsolver_gpu.hpp"

#ifndef LB_LATTICE
#define LB_LATTICE D3Q19
#endif
using LT = LB_LATTICE;

int main(int argc, char** argv) {
  if (argc < 7 + LT::nD - 1) {
    std::fprintf(stderr, "usage: %s case.vtklb tau iterations precision out.bin F...\n", argv[0]);
    return 2;
  }
  const std::string path = argv[1];
  const double tau = std::atof(argv[2]);
  const int iterations = std::atoi(argv[3]);
  const int precision = std::atoi(argv[4]);
  const std::string out = argv[5];
  std::vector<double> force(LT::nD);
  for (int d = 0; d < LT::nD; ++d) force[d] = std::atof(argv[6 + d]);

  // set-up with the reference's classes, as a case main does
  LBvtk<LT> vtk(path);
  Grid<LT> grid(vtk);
  Nodes<LT> nodes(vtk, grid);
  const std::vector<int> bulk = findBulkNodes(nodes);
  ScalarField rho(1, grid.size());
  vtk.toAttribute("init_rho");
  for (int n = vtk.beginNodeNo(); n < vtk.endNodeNo(); ++n) rho(0, n) = vtk.getScalarAttribute<lbBase_t>();
  VectorField<LT> vel(1, grid.size());
  LbField<LT> f(1, grid.size());
  for (int n : bulk)
    for (int q = 0; q < LT::nQ; ++q) f(0, q, n) = LT::w[q] * rho(0, n);

  try {
    lbamd::GpuSubdomain<LT> gpu(grid, nodes, bulk, precision);
    gpu.upload(f);
    gpu.stepBGK(iterations + 1, tau, force);
    gpu.downloadMacro(rho, vel, bulk);
  } catch (const lbamd::GpuError& e) {
    std::fprintf(stderr, "GPU error %d: %s\n", e.code, e.what());
    return 1;
  }
  std::ofstream os(out, std::ios::binary);
  for (int n : bulk) {
    const double r = rho(0, n);
    os.write(reinterpret_cast<const char*>(&r), sizeof(double));
    for (int d = 0; d < LT::nD; ++d) {
      const double v = vel(0, d, n);
      os.write(reinterpret_cast<const char*>(&v), sizeof(double));
    }
  }
  return os ? 0 : 1;
}
