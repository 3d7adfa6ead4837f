// lbsolver_gpu.hpp -- header-only C++ drop-in over the C ABI (badchimp_amd.h)
// for BADChIMP case mains.
//
// A case main keeps its set-up (LBvtk, Grid, Nodes, findBulkNodes, the
// fields) and replaces the per-node collide/propagate loop, the swap and the
// half-way bounce-back with one call:
//
//   std_case/main.cpp:109-145       ->  gpu.stepBGK(1, tau, bodyForce(0, 0));
//     (with calcOmegaBGKTRT / calcDeltaOmegaFTRT)  gpu.stepTRT(1, tauSym, tauAnti, F);
//   main_TWOPHASE.cpp:238-392       ->  gpu.stepCG(1, tau0, tau1, sigma, beta, momx, Fyz);
//
// The wrapper reads the reference's objects only through their public
// interface -- Grid::neighbor(q, n) / Grid::pos(n, d) (LBgrid.h:187-258),
// Nodes::getType(n) (LBnodes.h:72), LbField::operator()(field, q, n)
// (LBfield.h:298-305), ScalarField / VectorField operator() (LBfield.h:92-100,
// 168-175) -- so it compiles against them unchanged.  Errors throw
// lbamd::GpuError (the C ABI itself never exits).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/badchimp_amd.h"

namespace lbamd {

struct GpuError : std::runtime_error {
  int code;
  GpuError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const lb_ctx* ctx) {
  if (rc != LB_OK) throw GpuError(rc, ctx ? lb_last_error(ctx) : lb_global_error());
}

// Lattice id of a reference lattice struct (D2Q9 / D3Q19: LBd2q9.h, LBd3q19.h).
template <typename DXQY>
constexpr int lattice_id() {
  return DXQY::nQ == 9 ? LB_D2Q9 : LB_D3Q19;
}

// One rank's subdomain on one GPU.  GRID, NODES: the reference's Grid<DXQY>
// and Nodes<DXQY> (or anything with the same neighbor / pos / getType calls).
template <typename DXQY>
class GpuSubdomain {
 public:
  template <typename GRID, typename NODES>
  GpuSubdomain(const GRID& grid, const NODES& nodes, const std::vector<int>& bulkNodes,
               int precision = LB_FP64, int device = 0, bool usePositions = true)
      : nNodes_(grid.size()) {
    check(lb_ctx_create(lattice_id<DXQY>(), precision, device, &ctx_), nullptr);
    std::vector<int32_t> neighbor(static_cast<size_t>(nNodes_) * DXQY::nQ);
    std::vector<int16_t> type(nNodes_);
    std::vector<int32_t> pos(static_cast<size_t>(nNodes_) * DXQY::nD, -1);
    for (int n = 0; n < nNodes_; ++n) {
      for (int q = 0; q < DXQY::nQ; ++q) neighbor[static_cast<size_t>(n) * DXQY::nQ + q] = grid.neighbor(q, n);
      type[n] = static_cast<int16_t>(nodes.getType(n));
      if (n > 0)
        for (int d = 0; d < DXQY::nD; ++d) pos[static_cast<size_t>(n) * DXQY::nD + d] = grid.pos(n, d);
    }
    std::vector<int32_t> bulk(bulkNodes.begin(), bulkNodes.end());
    const int rc = usePositions
                       ? lb_set_topology_grid(ctx_, nNodes_, neighbor.data(), type.data(), bulk.data(),
                                              static_cast<int32_t>(bulk.size()), pos.data(), 0)
                       : lb_set_topology(ctx_, nNodes_, neighbor.data(), type.data(), bulk.data(),
                                         static_cast<int32_t>(bulk.size()), 0);
    check(rc, ctx_);
  }
  GpuSubdomain(const GpuSubdomain&) = delete;
  GpuSubdomain& operator=(const GpuSubdomain&) = delete;
  ~GpuSubdomain() { lb_ctx_destroy(ctx_); }

  lb_ctx* handle() { return ctx_; }
  int layout() const { return lb_layout(ctx_); }

  // LbField<DXQY> with nFields fields -> device (the pre-collision state)
  template <typename LBFIELD>
  void upload(const LBFIELD& f) {
    const int nF = f.num_fields();
    std::vector<double> h(static_cast<size_t>(nNodes_) * nF * DXQY::nQ, 0.0);
    for (int n = 0; n < nNodes_; ++n)
      for (int k = 0; k < nF; ++k)
        for (int q = 0; q < DXQY::nQ; ++q) h[(static_cast<size_t>(n) * nF + k) * DXQY::nQ + q] = f(k, q, n);
    check(lb_upload_f(ctx_, nF, h.data()), ctx_);
  }

  // device -> LbField (bulk nodes only; other entries are left as they are)
  template <typename LBFIELD>
  void download(LBFIELD& f, const std::vector<int>& bulkNodes) {
    const int nF = f.num_fields();
    std::vector<double> h(static_cast<size_t>(nNodes_) * nF * DXQY::nQ, 0.0);
    check(lb_download_f(ctx_, nF, h.data()), ctx_);
    for (int n : bulkNodes)
      for (int k = 0; k < nF; ++k)
        for (int q = 0; q < DXQY::nQ; ++q) f(k, q, n) = h[(static_cast<size_t>(n) * nF + k) * DXQY::nQ + q];
  }

  // rho (ScalarField, 1 or 2 fields) and vel (VectorField, 1 field) of the last step
  template <typename SCALAR, typename VECTOR>
  void downloadMacro(SCALAR& rho, VECTOR& vel, const std::vector<int>& bulkNodes) {
    const int nF = rho.num_fields();
    std::vector<double> r(static_cast<size_t>(nNodes_) * nF), v(static_cast<size_t>(nNodes_) * DXQY::nD);
    check(lb_download_scalar(ctx_, LB_RHO, nF, r.data()), ctx_);
    check(lb_download_vector(ctx_, LB_VEL, 1, v.data()), ctx_);
    for (int n : bulkNodes) {
      for (int k = 0; k < nF; ++k) rho(k, n) = r[static_cast<size_t>(n) * nF + k];
      for (int d = 0; d < DXQY::nD; ++d) vel(0, d, n) = v[static_cast<size_t>(n) * DXQY::nD + d];
    }
  }

  // colour field (ScalarField, 1 field) of the last two-phase step
  template <typename SCALAR>
  void downloadPhi(SCALAR& phi, const std::vector<int>& bulkNodes) {
    std::vector<double> p(nNodes_);
    check(lb_download_scalar(ctx_, LB_PHI, 1, p.data()), ctx_);
    for (int n : bulkNodes) phi(0, n) = p[n];
  }

  // wettability densities of the solid-boundary nodes (ScalarField, 2 fields)
  template <typename SCALAR>
  void setWallDensities(const SCALAR& rho) {
    std::vector<double> r(static_cast<size_t>(nNodes_) * 2);
    for (int n = 0; n < nNodes_; ++n) {
      r[2 * static_cast<size_t>(n)] = rho(0, n);
      r[2 * static_cast<size_t>(n) + 1] = rho(1, n);
    }
    check(lb_set_wall_densities(ctx_, r.data()), ctx_);
  }

  // nsteps of the std_case loop body; force: valarray/vector of nD components
  template <typename VEC>
  void stepBGK(int nsteps, double tau, const VEC& force) {
    double F[3] = {0, 0, 0};
    for (int d = 0; d < DXQY::nD; ++d) F[d] = force[d];
    check(lb_step_bgk(ctx_, nsteps, tau, F), ctx_);
  }

  // the same loop body with calcOmegaBGKTRT / calcDeltaOmegaFTRT (phi = 1)
  template <typename VEC>
  void stepTRT(int nsteps, double tauSym, double tauAnti, const VEC& force) {
    double F[3] = {0, 0, 0};
    for (int d = 0; d < DXQY::nD; ++d) F[d] = force[d];
    check(lb_step_trt(ctx_, nsteps, tauSym, tauAnti, F), ctx_);
  }

  // nsteps of the colour-gradient loop body with the momx flux controller
  // (main_TWOPHASE.cpp:292-308); other force components fixed
  template <typename VEC>
  void stepCG(int nsteps, double tau0, double tau1, double sigma, double beta, double momx,
              const VEC& force) {
    double F[3] = {momx, 0, 0};
    for (int d = 1; d < DXQY::nD; ++d) F[d] = force[d];
    check(lb_step_cg(ctx_, nsteps, tau0, tau1, sigma, beta, LB_FORCE_MOMX, F), ctx_);
  }

  // body force of the last two-phase step (the value main_TWOPHASE writes to force.dat)
  double lastForceX() {
    double F[3] = {0, 0, 0};
    check(lb_last_force(ctx_, F), ctx_);
    return F[0];
  }

 private:
  lb_ctx* ctx_ = nullptr;
  int nNodes_ = 0;
};

}  // namespace lbamd
