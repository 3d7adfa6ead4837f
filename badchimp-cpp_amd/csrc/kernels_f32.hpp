// Launchers of the reduced-precision kernels (shifted populations).
//
// Storage holds g_q = (f_q - w_q) * s in float (FP32S, s = 1) or half
// (FP16S, s = kHalfScale); the arithmetic is float; the flux and
// permeability partial sums are accumulated in double.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "context.hpp"
#include "kernels.hpp"

namespace lbamd {

// Scale of the FP16 storage: deviations f - w of order 1e-7..1e-2 are kept in
// the normal half range (6.1e-5 .. 65504) after multiplying by 2^12.
constexpr float kHalfScale = 4096.0f;

// Two-phase rest state f_k = w rho_k (rho_0 = r0 per compact node, rho_1 = 1 -
// r0) written where the first pull through ad reads it, in the (t, d) form.
hipError_t prestream_lowp(int precision, int lattice, const RowWin& ad, const double* r0, int32_t n, int64_t NP,
                          void* g, hipStream_t s);
hipError_t store_lowp(int precision, int lattice, int nf, const double* f, void* g, int32_t n,
                      int64_t NP, hipStream_t s);
// FP64 populations from the stored form: AoSoA like g, or with rank given,
// the LbField layout over the bulk nodes in ascending node order.
hipError_t load_lowp(int precision, int lattice, int nf, const void* g, double* f, int32_t n,
                     int64_t NP, hipStream_t s, const int32_t* rank = nullptr);
hipError_t gather_lowp(int precision, int lattice, int nf, const void* gin, void* gout,
                       const Addr& ad, int32_t n, int64_t NP, hipStream_t s);
hipError_t bgk_lowp(int precision, int lattice, bool first, bool store, bool reduce,
                    const void* gin, void* gout, const Addr& ad, int32_t n, int64_t NP,
                    const BgkArgs& a, double* rho, double* vel, double* partial, hipStream_t s, bool accum = false);
hipError_t flux_lowp(int precision, int lattice, int nf, bool first, const void* gin,
                     const Addr& ad, int32_t n, int64_t NP, int dir, double* partial,
                     hipStream_t s);
// Batched form in reduced precision (see batch_bgk_f64).
// step0: the global step at the start of the launch sequence, koff: this
// launch's offset from it (the solves' own step counts seed the FP16
// stochastic rounding).
hipError_t batch_bgk_lowp(int precision, int lattice, bool reduce, const BatchSolve* solves, const uint32_t* items,
                          int32_t n_items, int32_t grid, int32_t per_wg, int32_t parity, const uint32_t* step0,
                          int32_t koff, int ns, bool accum, hipStream_t s);
// Both passes of one batched two-phase step (see k_batch_cg_phi_lp).  ns:
// 256-node items per workgroup at a time (1 or 2; workgroups of 256 ns
// threads), as for the single-phase step.
hipError_t batch_cg_lowp(int precision, int lattice, bool reduce, const BatchSolve* solves, const uint32_t* items,
                         int32_t n_items, int32_t grid, int32_t per_wg, int32_t parity, const uint32_t* step0,
                         int32_t koff, int ns, hipStream_t s);
// Two-phase populations are stored as (t, d) = (f0 + f1 - w, f0 - f1).
// rowwin: the row windows of ad with one wall colour, else the link table.
hipError_t cg_phi_lowp(int precision, int lattice, bool first, bool momx, const void* gin, const Addr& ad,
                       bool rowwin, int32_t n, int64_t NP, float* phi, double* partial, hipStream_t s);
hipError_t cg_lowp(int precision, int lattice, bool first, bool store, const void* gin, void* gout, const Addr& ad,
                   bool rowwin, int32_t n, int64_t NP, const CgArgs& a, const float* phi, const WallPhi& wp,
                   const double* dscalar, double* rho, double* vel, hipStream_t s);

}  // namespace lbamd
