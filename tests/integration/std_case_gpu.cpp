// Drop-in check of the C++ shim (lbsolver_gpu.hpp) against the reference's
// own set-up code: a std_case-style driver that builds LBvtk, Grid, Nodes,
// the bulk list and the fields with the reference's lbsolver headers, then
// hands the step loop to the GPU through lbamd::GpuSubdomain.
//
//   std_case_gpu <case.vtklb> <tau | tauSym,tauAnti> <iterations> <precision 0|1|2> <out.bin> <F_x> <F_y> [<F_z>] [<ckpt>]
//
// Runs iterations + 1 steps (the loop of std_case/main.cpp:109 runs
// i = 0..max) and writes, for each bulk node in findBulkNodes order, rho and
// the nD velocity components as doubles (the fields the loop stores).  With
// <ckpt>, the field files are also cross-checked: the library's .lblbf is
// read by the reference's LbField::readFromFile, and the reference's
// writeToFile output is read back by the library; it prints the number of
// bulk entries that differ.
// Built by oracle/build_ref.sh where the reference tree is present.
#include <cstdio>
#include <cstring>
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
  // "tau" (BGK) or "tauSym,tauAnti" (the loop with calcOmegaBGKTRT / calcDeltaOmegaFTRT)
  const double tau = std::atof(argv[2]);
  const char* comma = std::strchr(argv[2], ',');
  const double tauAnti = comma ? std::atof(comma + 1) : 0.0;
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

  const bool ckpt = argc > 6 + LT::nD;
  try {
    lbamd::GpuSubdomain<LT> gpu(grid, nodes, bulk, precision);
    gpu.upload(f);
    if (comma)
      gpu.stepTRT(iterations + 1, tau, tauAnti, force);
    else
      gpu.stepBGK(iterations + 1, tau, force);
    gpu.downloadMacro(rho, vel, bulk);
    if (ckpt) {
      const std::string prefix = argv[6 + LT::nD];
      LbField<LT> now(1, grid.size()), theirs(1, grid.size()), back(1, grid.size());
      gpu.download(now, bulk);
      // library writer -> reference reader
      lbamd::check(lb_write_lbfield(gpu.handle(), (prefix + "_lib").c_str()), gpu.handle());
      theirs.readFromFile(prefix + "_lib");
      // reference writer -> library reader
      now.writeToFile(prefix + "_ref");
      lbamd::check(lb_read_lbfield(gpu.handle(), (prefix + "_ref").c_str()), gpu.handle());
      gpu.download(back, bulk);
      long bad = 0;
      for (int n : bulk)
        for (int q = 0; q < LT::nQ; ++q)
          bad += (theirs(0, q, n) != now(0, q, n)) + (back(0, q, n) != now(0, q, n));
      std::printf("checkpoint mismatches %ld\n", bad);
    }
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
