// Host-callable launchers of the device kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "addressing.hpp"
#include "context.hpp"

namespace lbamd {

// Address layout of a launch (see addressing.hpp).
struct Addr {
  int layout = 0;  // 0 table, 1 rows, 2 row windows
  TableAddr table{};
  RowAddr rows{};
  RowWin win{};
};

// ---- FP64 (reference arithmetic, -ffp-contract=off) ----------------------
// One BGK + Guo collide-stream step in pull form (std_case/main.cpp:110-145).
// first: buffer holds pre-collision f (identity read); store: write rho/vel;
// reduce: write per-block partial sums of j_x = sum_q f c_x + F_x/2 (accum:
// add them to the partials, summing over the steps of a check interval).
hipError_t bgk_f64(int lattice, bool first, bool store, bool reduce, const double* fin,
                   double* fout, const Addr& ad, int32_t n, int64_t NP, const BgkArgs& a,
                   double* rho, double* vel, double* partial, hipStream_t s, bool accum = false);

// Two-phase pass 1 (+ pass 2 partial sums): phi = (rho0-rho1)/(rho0+rho1)
// and sum_q (f0+f1) c_x per block (main_TWOPHASE.cpp:238-298).  rowwin: the
// row windows of ad (every solid neighbour has the colour wp.uniform), else
// the link table (wall colour per slot).
hipError_t cg_phi_f64(int lattice, bool first, bool momx, const double* fin, const Addr& ad, bool rowwin,
                      int32_t n, int64_t NP, double* phi, double* partial, hipStream_t s);
// Two-phase pass 3 (main_TWOPHASE.cpp:312-374) and bounce-back.
hipError_t cg_f64(int lattice, bool first, bool store, const double* fin, double* fout, const Addr& ad,
                  bool rowwin, int32_t n, int64_t NP, const CgArgs& a, const double* phi, const WallPhi& wp,
                  const double* dscalar, double* rho, double* vel, hipStream_t s);
// Streaming only: fout = pre-collision state of the post-collision buffer fin.
hipError_t gather_f64(int lattice, int nfields, const double* fin, double* fout,
                      const Addr& ad, int32_t n, int64_t NP, hipStream_t s);
// Flux partial sums over the pre-collision state (LBglobalforcing.h:8-33,61-78):
// three per-block partials (sum fc, sum phi0 fc, sum phi1 fc).
hipError_t flux_f64(int lattice, int nfields, bool first, const double* fin, const Addr& ad,
                    int32_t n, int64_t NP, int dir, double* partial, hipStream_t s);
// Equilibrium initialisation on bulk nodes (LBinitiatefield.h:50-56).
// FP64 populations of nf fields (AoSoA) to the LbField layout over the bulk
// nodes in ascending node order (rank[i]: position of compact node i).
hipError_t unstage_f64(int lattice, int nf, const double* g, double* f, int32_t n, int64_t NP,
                       const int32_t* rank, hipStream_t s);
hipError_t init_eq_f64(int lattice, double* f, int nf, int field, int32_t n, int64_t NP,
                       const double* rho, const double* vel, hipStream_t s);

// ---- tagged boundaries (std_one_phase/main.cpp:27-204, 513-575) -------------
// One step with the tag-derived pull table ad, the anti-bounce-back
// direction masks abb, the domain labels/mass-source flags info, the force
// attribute fon and the per-domain scale factors and mass changes.
hipError_t tagged_f64(int lattice, bool first, bool store, const double* fin, double* fout, const TableAddr& ad,
                      int32_t n, int64_t NP, const BgkArgs& a, double rho_w, const uint32_t* info,
                      const double* fon, const uint32_t* abb, const double* scale, const double* mass, double* rho,
                      double* vel, hipStream_t s);
// Block partials of sum (1 - rho) per domain label 1..n_label (partial[(l-1)*nblocks + b]).
hipError_t tag_mass_f64(int lattice, bool first, const double* fin, const TableAddr& ad, int32_t n, int64_t NP,
                        const uint32_t* info, int n_label, double* partial, hipStream_t s);

// ---- batched form -----------------------------------------------------------
// One FP64 step of every solve listed in items (see k_batch_bgk): grid
// workgroups of 256 ns threads (ns = 1 or 2 items at a time), each over a run
// of per_wg items; reduce writes each solve's block partials of j_x (accum:
// adds to them).
hipError_t batch_bgk_f64(int lattice, bool reduce, const BatchSolve* solves, const uint32_t* items,
                         int32_t n_items, int32_t grid, int32_t per_wg, int32_t parity, int ns, bool accum,
                         hipStream_t s);
// Two-phase rest state (phase-0 density r0 per compact node) placed where the
// first pull through ad reads it (FP64 form).
hipError_t prestream_f64(int lattice, const RowWin& ad, const double* r0, int32_t n, int64_t NP, double* f,
                         hipStream_t s);
// Both passes of one FP64 colour-gradient step of every two-phase solve listed
// in items (as batch_bgk_f64); reduce writes each solve's three flux partials.
hipError_t batch_cg_f64(int lattice, bool reduce, const BatchSolve* solves, const uint32_t* items,
                        int32_t n_items, int32_t grid, int32_t per_wg, int32_t parity, int ns, hipStream_t s);
// out[slot] = sum of the solve's block partials, for each slot listed.
hipError_t batch_sum(const BatchSolve* solves, const int32_t* slots, int32_t n_slots, double* out,
                     hipStream_t s);
// Stop tests of the listed solves after an interval of M steps (from the sums
// in out), marking decided solves not live; then advances the step counter
// *step by M.
hipError_t batch_check(BatchSolve* solves, SolveCheck* chk, const int32_t* slots, int32_t n_slots, const double* out,
                       int32_t M, uint32_t* step, hipStream_t s);
// Rest state (f = w) of one FP64 buffer of Q * NP elements.
hipError_t batch_rest_f64(int lattice, double* f, int64_t NP, hipStream_t s);

// ---- shared ---------------------------------------------------------------
// out[c] = sum over blocks of partial[c*nblocks + b], c < ncomp (fixed order).
hipError_t sum_partials(const double* partial, int64_t nblocks, int ncomp, double* out,
                        hipStream_t s);
// dscalar[0] = 2 (momx - sum/n_global) from the partial sums (main_TWOPHASE.cpp:299-308).
hipError_t finish_momx(const double* partial, int64_t nblocks, double momx,
                       int64_t n_global, double* dscalar, hipStream_t s);

// Halo exchange (halo.hip): out[k n + j] = field k (< nf of nf_total) of the
// element elem[j], as the double of the stored value; unpack writes back.
hipError_t halo_pack(int precision, int nf_total, int nf, const void* f, const uint32_t* elem, int64_t n,
                     double* out, hipStream_t s);
hipError_t halo_unpack(int precision, int nf_total, int nf, void* f, const uint32_t* elem, int64_t n,
                       const double* in, hipStream_t s);

}  // namespace lbamd
