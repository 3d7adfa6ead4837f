// Drop-in check of the C++ shim for the colour-gradient case: the reference's
// set-up classes (LBvtk, Grid, Nodes, findBulkNodes, ScalarField/VectorField,
// initiateLbField) prepare the two-phase fields as a twophase main does, and
// the step loop runs on the GPU through lbamd::GpuSubdomain::stepCG with the
// momentum-controlled force.
//
//   twophase_gpu <case.vtklb> <tau0> <tau1> <sigma> <beta> <momx> <F_y> <F_z> <iterations> <precision> <out.bin>
//
// Runs iterations + 1 steps and writes the last force F_x, then for each bulk
// node in findBulkNodes order rho0, rho1 and the nD velocity components
// (doubles).  LB_LATTICE selects D2Q9 or D3Q19 (F_z is ignored in 2-D).
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
  if (argc != 12) {
    std::fprintf(stderr, "usage: %s case.vtklb tau0 tau1 sigma beta momx Fy Fz iterations precision out.bin\n",
                 argv[0]);
    return 2;
  }
  const std::string path = argv[1];
  const double tau0 = std::atof(argv[2]), tau1 = std::atof(argv[3]);
  const double sigma = std::atof(argv[4]), beta = std::atof(argv[5]), momx = std::atof(argv[6]);
  const std::vector<double> force = {0.0, std::atof(argv[7]), std::atof(argv[8])};
  const int iterations = std::atoi(argv[9]);
  const int precision = std::atoi(argv[10]);
  const std::string out = argv[11];

  LBvtk<LT> vtk(path);
  Grid<LT> grid(vtk);
  Nodes<LT> nodes(vtk, grid);
  const std::vector<int> bulk = findBulkNodes(nodes);
  ScalarField rho(2, grid.size());
  VectorField<LT> vel(1, grid.size());
  // phase densities as the case reads them (single precision attributes), the
  // solid-boundary nodes from the wettability
  const char* names[2] = {"rho0", "rho1"};
  for (int k = 0; k < 2; ++k) {
    vtk.toAttribute(names[k]);
    for (int n = vtk.beginNodeNo(); n < vtk.endNodeNo(); ++n) {
      rho(k, n) = vtk.getScalar<float>();
      for (int d = 0; d < LT::nD; ++d) vel(0, d, n) = 0.0;
    }
  }
  vtk.toAttribute("wettability");
  for (int n = vtk.beginNodeNo(); n < vtk.endNodeNo(); ++n) {
    const float w = vtk.getScalar<float>();
    if (nodes.isSolidBoundary(n)) {
      rho(0, n) = w;
      rho(1, n) = 1 - w;
    }
  }
  LbField<LT> f(2, grid.size());
  initiateLbField(0, 0, 0, bulk, rho, vel, f);
  initiateLbField(1, 1, 0, bulk, rho, vel, f);

  double fx = 0.0;
  try {
    lbamd::GpuSubdomain<LT> gpu(grid, nodes, bulk, precision);
    gpu.upload(f);
    gpu.setWallDensities(rho);
    gpu.stepCG(iterations + 1, tau0, tau1, sigma, beta, momx, force);
    gpu.downloadMacro(rho, vel, bulk);
    fx = gpu.lastForceX();
  } catch (const lbamd::GpuError& e) {
    std::fprintf(stderr, "GPU error %d: %s\n", e.code, e.what());
    return 1;
  }
  std::ofstream os(out, std::ios::binary);
  os.write(reinterpret_cast<const char*>(&fx), sizeof(double));
  for (int n : bulk) {
    double v[2 + LT::nD] = {rho(0, n), rho(1, n)};
    for (int d = 0; d < LT::nD; ++d) v[2 + d] = vel(0, d, n);
    os.write(reinterpret_cast<const char*>(v), sizeof(v));
  }
  return os ? 0 : 1;
}
