/*
 * badchimp_amd.h -- C ABI of the MI355X lattice-Boltzmann engine.
 *
 * One context is one rank's subdomain on one GPU: the labelled sparse node
 * list with its neighbour index (as Grid/Nodes hold it), the population
 * fields (as LbField holds them) and the macroscopic fields (as ScalarField /
 * VectorField hold them), resident in HBM.  A BADChIMP case main keeps its
 * set-up (LBvtk, Grid, Nodes, findBulkNodes) and replaces its per-node loop,
 * LbField::swapData and HalfWayBounceBack::apply with one step call.
 *
 * Conventions
 *   - Every call returns 0 on success or a negative LB_ERR_* code; the message
 *     is available from lb_last_error(ctx) (or lb_global_error() for calls
 *     without a context).  No call exits the process.
 *   - Host arrays are copied in or out; the library keeps no host pointers.
 *   - Node numbers are the reference's: node 0 is the ghost node, nodes
 *     1..n_nodes-1 are the points of the .vtklb file.
 *   - Arrays use the reference layouts:
 *       neighbor  int32[n_nodes*Q]      neighbor[n*Q + q]   (Grid::neigList_, LBgrid.h:119,187-198)
 *       node_type int16[n_nodes]        -1..4               (Nodes::nodeType_, LBnodes.h:47-54)
 *       LbField   double[n_nodes*nF*Q]  f[(nF*Q)*n + Q*field + q]   (LBfield.h:298-305)
 *       Scalar    double[n_nodes*nF]    s[nF*n + field]     (LBfield.h:92-100)
 *       Vector    double[n_nodes*nF*nD] v[nF*nD*n + nD*field + d]   (LBfield.h:168-175)
 *     Only bulk-node entries are read or written by the field calls.
 *   - Calls are synchronous unless named *_async.  A context is not
 *     thread-safe; use one context per rank or per host thread.
 */
#ifndef BADCHIMP_AMD_H
#define BADCHIMP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ constants */
enum lb_lattice { LB_D2Q9 = 0, LB_D3Q19 = 1 };

/* Population storage.  FP64 is the reference's arithmetic (lbBase_t = double,
 * LBglobal.h:10).  FP32S / FP16S store the shifted populations f - w_q in
 * float / half.  The single-phase steps compute in float, the two-phase
 * (colour-gradient) steps in double on the stored values; flux reductions
 * stay in double. */
enum lb_precision { LB_FP64 = 0, LB_FP32S = 1, LB_FP16S = 2 };

enum lb_status {
  LB_OK = 0,
  LB_ERR_ARG = -1,           /* invalid argument */
  LB_ERR_STATE = -2,         /* call out of order (e.g. step before topology) */
  LB_ERR_GHOST0_LINK = -3,   /* a bulk node links to the ghost node 0 */
  LB_ERR_UNSUPPORTED = -4,   /* e.g. a two-phase step on a context with halo nodes */
  LB_ERR_HIP = -5,           /* HIP runtime error */
  LB_ERR_IO = -6,            /* file could not be read or parsed */
  LB_ERR_NOMEM = -7          /* device or host allocation failed */
};

/* lb_set_topology flags */
#define LB_TOPO_ALLOW_GHOST0 1 /* accept fluid links to node 0 and reproduce the
                                  reference's shared-slot result exactly */
#define LB_TOPO_TABLE_ONLY 2   /* keep the per-link pull table even when the
                                  row layout would apply */

/* field selectors */
enum lb_scalar { LB_RHO = 0, LB_PHI = 1, LB_VISC = 2 };
enum lb_vector { LB_VEL = 0 };

/* two-phase body-force modes */
enum lb_force_mode {
  LB_FORCE_FIXED = 0, /* F given */
  LB_FORCE_MOMX = 1   /* F_x = 2 (momx - <sum_q (f0+f1) c_x>), main_TWOPHASE.cpp:292-308 */
};

typedef struct lb_ctx lb_ctx;
typedef struct lb_topo lb_topo;

/* ------------------------------------------------------------ errors */
const char* lb_last_error(const lb_ctx* ctx);
const char* lb_global_error(void);
const char* lb_version(void);

/* ------------------------------------------------------------ context */
/* Replaces the allocation of LbField f, fTmp / ScalarField / VectorField on
 * the host (std_case/main.cpp:65-90, main_TWOPHASE.cpp:138-144). */
int lb_ctx_create(int lattice, int precision, int device, lb_ctx** out);
int lb_ctx_destroy(lb_ctx* ctx);

/* Node list + neighbour index + node types + bulk list.  Replaces what the
 * per-node loop reads from Grid (LBgrid.h:187-198), Nodes (LBnodes.h:72-83),
 * findBulkNodes (LBgeometry.h:11-21) and the link classification of
 * BoundaryHalwWayHelper (LBhalfwayhelperclass.h:110-161).  bulk must be
 * ascending (findBulkNodes order); it is the node order of every sum. */
int lb_set_topology(lb_ctx* ctx, int32_t n_nodes, const int32_t* neighbor,
                    const int16_t* node_type, const int32_t* bulk,
                    int32_t n_bulk, int32_t flags);

/* The same, with the node positions (Grid::pos_, int32[n_nodes*nD],
 * LBgrid.h:120,215-258).  When the bulk nodes are cells of one box (periodic
 * along any axes) the single-phase kernels switch to the row layout: a 4-byte
 * cell coordinate per node and a cached per-row bit mask instead of one
 * 4-byte link per direction.  The layout is used only after every link it
 * yields has been checked against the neighbour index; otherwise the table
 * layout stays (lb_layout_note says why).  Results are identical in both in
 * LB_FP64 and LB_FP32S.  LB_FP16S seeds its stochastic rounding by a node's
 * place in the layout's node order, which the row layout changes when the
 * box has more than 32 rows along y: the two layouts then agree to rounding
 * noise. */
int lb_set_topology_grid(lb_ctx* ctx, int32_t n_nodes, const int32_t* neighbor,
                         const int16_t* node_type, const int32_t* bulk,
                         int32_t n_bulk, const int32_t* pos, int32_t flags);

/* 0 = per-link table, 1 = row layout, -1 = no topology. */
int32_t lb_layout(const lb_ctx* ctx);
const char* lb_layout_note(const lb_ctx* ctx);

/* Total bulk-node count over all ranks, the divisor of the flux mean
 * (main_TWOPHASE.cpp:197-199,301).  Defaults to n_bulk. */
int lb_set_global_bulk_count(lb_ctx* ctx, int64_t n);

/* ---- Halo exchange between ranks (one subdomain per context) ------------
 * A node list split over ranks (lb_topo_from_geometry of a geometry whose
 * fluid cells carry rank + 1, or a rank's .vtklb) holds other-rank fluid
 * nodes (type 4).  lb_set_topology
 * gives each of them a population slot after the bulk nodes, and the bulk
 * nodes gather from those slots through the link table (such a context
 * uses the link-table layout).  Between two steps every rank fills the halo
 * of its neighbours: as communicateLbField (LBmonlatmpi.h:236-300), it sends
 * the post-collision value of every listed (node, direction) and writes
 * what it receives into its halo slots.  The transport (MPI, RCCL, a device
 * copy) is the caller's.  Single-phase steps (BGK, TRT, rheology) read the
 * halo.  The colour-gradient iteration also needs the colour of the halo
 * nodes between its two passes: it runs as lb_step_cg_colour, the colour
 * exchange (lb_halo_pack_colour / lb_halo_unpack_colour below), then
 * lb_step_cg_collide; lb_step_cg refuses a context with halo nodes
 * (LB_ERR_UNSUPPORTED), and so do tagged boundaries and the solves
 * (lb_solve_permeability, lb_solve_permeability_trt, lb_solve_relperm: their
 * stop tests sum over one context; a split domain sums over the ranks).
 *
 * lb_halo_add registers the exchange with one neighbouring rank from the
 * lists MonLatMpi holds after BndMpi::setup (LBbndmpi.h:206-315, the
 * addMpiBnd arguments): nodes_to_send (bulk nodes of this rank), n_dir_send
 * per node and dir_send (the directions, concatenated); nodes_received (this
 * rank's halo nodes of that rank), n_dir_recv and dir_recv.  *index is the
 * exchange's number.  lb_halo_pack writes the n_send values of the nf first
 * fields into send[f * n_send + j] (j in list order), lb_halo_unpack reads
 * recv[f * n_recv + j]; buffers are host or device memory.  The values are
 * the stored ones as double: f in LB_FP64, f - w_q in LB_FP32S and
 * LB_FP16S, so contexts of one precision exchange them exactly.  Both need a
 * post-collision state (after a step; LB_ERR_STATE otherwise), and the
 * first step after an upload reads no halo. */
int lb_halo_add(lb_ctx* ctx, int32_t neighbor_rank, int32_t n_send, const int32_t* nodes_to_send,
                const int32_t* n_dir_send, const int32_t* dir_send, int32_t n_recv,
                const int32_t* nodes_received, const int32_t* n_dir_recv, const int32_t* dir_recv,
                int32_t* index);
int lb_halo_sizes(const lb_ctx* ctx, int32_t index, int32_t* neighbor_rank, int64_t* n_send, int64_t* n_recv);
int lb_halo_pack(lb_ctx* ctx, int32_t index, int32_t nf, double* send);
int lb_halo_unpack(lb_ctx* ctx, int32_t index, int32_t nf, const double* recv);
/* The colour exchange of the colour-gradient iteration,
 * communciateScalarField of cgField (main_TWOPHASE.cpp:287;
 * MonLatMpi::communicateScalarField, LBmonlatmpi.h:181-205):
 * lb_halo_pack_colour writes the colour of the n_send nodes_to_send of
 * exchange `index` (as given to lb_halo_add) into send[j], and
 * lb_halo_unpack_colour writes recv[j] into the colour of its n_recv
 * nodes_received.  One double per node, the stored colour (double in
 * LB_FP64, float in the other forms, so exact).  Both run between
 * lb_step_cg_colour and lb_step_cg_collide of one iteration (LB_ERR_STATE
 * otherwise). */
int lb_halo_pack_colour(lb_ctx* ctx, int32_t index, double* send);
int lb_halo_unpack_colour(lb_ctx* ctx, int32_t index, const double* recv);
int32_t lb_halo_count(const lb_ctx* ctx);      /* exchanges registered */
int32_t lb_halo_num_nodes(const lb_ctx* ctx);  /* halo (other-rank) nodes with a slot */

/* ------------------------------------------------------------ fields */
/* Populations of n_fields fields (1: single phase, 2: two phase) in the
 * LbField layout over n_nodes nodes (LBfield.h:285-305).  upload sets the
 * pre-collision state f_t; download returns it. */
int lb_upload_f(lb_ctx* ctx, int32_t n_fields, const double* f);
int lb_download_f(lb_ctx* ctx, int32_t n_fields, double* f);

/* lb_download_f that allocates no device memory: no counterpart in the
 * reference.  It is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * populations are converted into a fixed 128 MiB staging array that the
 * library keeps in each device's memory until the process ends, then copied
 * out run by run of consecutive bulk node numbers.  Limit: n_bulk * n_fields
 * * Q * 8 bytes <= 128 MiB (a larger context gets LB_ERR_UNSUPPORTED).
 * Results equal lb_download_f's; entries of non-bulk nodes are left as they
 * are, as there.  Every call on one device shares that device's array:
 * calls that overlap on one GPU (from several host threads) race. */
int lb_download_f_noalloc(lb_ctx* ctx, int32_t n_fields, double* f);

/* Equilibrium initialisation f = w rho (1 + 3cu + 4.5(cu^2 - u^2/3)) of field
 * `field` (of n_fields population fields) from ScalarField rho (field
 * rho_field of rho_nfields) and VectorField vel (field 0), over bulk nodes
 * (LBinitiatefield.h:33-57).  A call with a different n_fields, or after a
 * step, starts a new state in which every field is to be initialised. */
int lb_init_equilibrium(lb_ctx* ctx, int32_t n_fields, int32_t field,
                        int32_t rho_nfields, int32_t rho_field,
                        const double* rho, const double* vel);

/* Macroscopic fields stored by the last step call: rho (1 field single
 * phase, 2 fields two phase), phi (colour field, 1 field) and vel (1
 * field).  They are the values of the pre-collision state of the last step,
 * as std_case/main.cpp:119-120 and main_TWOPHASE.cpp:241-245,329 store them. */
int lb_download_scalar(lb_ctx* ctx, int32_t which, int32_t n_fields, double* s);
int lb_download_vector(lb_ctx* ctx, int32_t which, int32_t n_fields, double* v);

/* Densities of the solid-boundary nodes for the two-phase wettability
 * (main_TWOPHASE.cpp:171-178,280-284): ScalarField rho with 2 fields over
 * n_nodes; only solid-boundary (type 1) entries are read. */
int lb_set_wall_densities(lb_ctx* ctx, const double* rho2);

/* Field files in the reference's binary formats, "<prefix>.lblbf" etc.
 * (LbField/ScalarField/VectorField::writeToFile, LBfield.h:102-138,233-276,
 * 378-421): int32 header {nFields, Q, nNodes} / {nFields, nNodes} /
 * {nFields, nD, nNodes}, then the doubles over all nodes in the host layout.
 * The population file holds the pre-collision state (a checkpoint: read it
 * back into a context with the same node list to resume); non-bulk nodes are
 * written as 0. */
int lb_write_lbfield(lb_ctx* ctx, const char* prefix);
int lb_read_lbfield(lb_ctx* ctx, const char* prefix);
int lb_write_scalar(lb_ctx* ctx, int32_t which, int32_t n_fields, const char* prefix);
int lb_write_vector(lb_ctx* ctx, int32_t which, const char* prefix);

/* ------------------------------------------------------------ steps */
/* nsteps iterations of the std_case loop body (std_case/main.cpp:110-145):
 * BGK + Guo forcing, streaming and half-way bounce-back, force uniform. */
int lb_step_bgk(lb_ctx* ctx, int32_t nsteps, double tau, const double* force);

/* Generalized-Newtonian fluid (GeneralizedNewtonian, LBrheology.h:321-476,
 * the case polymer_viscosity_tabulation/main.cpp:105-155): the viscosity
 * table is n rows of (strain-rate square sum_ij E_ij^2, viscosity) with
 * increasing strain-rate squares, extended by one constant row below and
 * above as the class's constructor does.  lb_step_rheology runs the
 * std_case loop with each node's relaxation time tau = nu c2Inv / rho + 1/2
 * looked up from its strain rate (upper_bound, linear interpolation) and
 * the Guo term at that tau.  The node viscosities of the last step are
 * downloaded with lb_download_scalar(ctx, LB_VISC, 1, ...). */
int lb_set_viscosity_table(lb_ctx* ctx, int32_t n, const double* strain_rate_sq,
                           const double* viscosity);
int lb_step_rheology(lb_ctx* ctx, int32_t nsteps, const double* force);

/* Tagged-link boundaries and per-domain mass sources of std_one_phase
 * (main.cpp:27-204, 322-354, 513-612), one rank, FP64 contexts.
 * tags[n]: the node tag (nodetags attribute, Nodes::getTag; tags[0] = -1
 * as Nodes sets it): bits 0-1 phase (0 solid, 1, 2 fluid, 3 pressure
 * ghost), bit 2 fluid-fluid interface, bit 3 fluid-solid interface, bit 4
 * pressure boundary.  The links the reference's boundary lists rewrite
 * become bounce-back (solid-fluid and fluid-fluid) or anti-bounce-back
 * -f + 2 w rho_w (1 + 0.5 (9 cu^2 - 3 u^2)) (pressure).  A link from a bulk
 * node to a non-bulk node that no tag covers, a pressure link to a bulk
 * node and any link to node 0 give LB_ERR_UNSUPPORTED.
 * interior_domains[n]: domain label (0 none); force_on[n]: the force
 * attribute.  lb_step_tagged runs the loop: per-domain mass change
 * sum(1 - rho), mass source q = 1.8 (mass change)/(domain size) on the
 * domain's untagged nodes, body force times force_on, BGK with the Guo and
 * source terms, then the boundary links.  lb_pressure_flux gives the mass
 * fluxes q1, q2 (half the sums of rho u_z over the pressure-boundary nodes
 * of each phase, in node order) of the last step. */
int lb_set_node_tags(lb_ctx* ctx, int32_t n_nodes, const int32_t* neighbor, const int32_t* tags,
                     const int32_t* interior_domains, const double* force_on);
int lb_step_tagged(lb_ctx* ctx, int32_t nsteps, double tau, const double* force, double rho_w);
int lb_pressure_flux(lb_ctx* ctx, double* q);

/* The same loop body with the two-relaxation-time operator in place of
 * calcOmegaBGK / calcDeltaOmegaF: calcOmegaBGKTRT(f, tau_sym, tau_anti, ...)
 * and calcDeltaOmegaFTRT(tau_sym, tau_anti, 1.0, ...) (LBcollision.h:50-75,
 * 214-233).  The even part of each direction pair relaxes with tau_sym (the
 * viscosity, nu = (tau_sym - 1/2)/3), the odd part with tau_anti.  Both > 1/2. */
int lb_step_trt(lb_ctx* ctx, int32_t nsteps, double tau_sym, double tau_anti,
                const double* force);

/* nsteps iterations of the colour-gradient loop body
 * (main_TWOPHASE.cpp:238-392).  force[nD]: in LB_FORCE_MOMX mode force[0]
 * is the target momentum momx and force[1..] the fixed other components. */
int lb_step_cg(lb_ctx* ctx, int32_t nsteps, double tau0, double tau1,
               double sigma, double beta, int32_t force_mode,
               const double* force);

/* One iteration of the same loop body in two calls, for a rank of a split
 * domain (main_TWOPHASE.cpp:238-308 on each rank):
 * lb_step_cg_colour computes cgField = (rho0 - rho1)/(rho0 + rho1) of every
 * bulk node from its gathered populations; with flux_x non-NULL it also
 * writes this rank's sum over its bulk nodes of sum_q (f0 + f1) c_q,x
 * (meanfcX before MPI_Allreduce, main_TWOPHASE.cpp:290-298).
 * The caller then exchanges the colour (lb_halo_pack_colour /
 * lb_halo_unpack_colour) and, in LB_FORCE_MOMX mode, sums flux_x over the
 * ranks.  lb_step_cg_collide runs the collision with the arguments of
 * lb_step_cg; in LB_FORCE_MOMX mode the force is 2 (force[0] -
 * flux_x_total / N), N the global bulk count (lb_set_global_bulk_count),
 * as main_TWOPHASE.cpp:300-307 (flux_x_total is ignored in
 * LB_FORCE_FIXED mode).  It stores rho and u of every iteration, and
 * refuses to run without the colour pass of the same iteration
 * (LB_ERR_STATE).  On a context without halo nodes the two calls give the
 * bits of lb_step_cg(ctx, 1, ...). */
int lb_step_cg_colour(lb_ctx* ctx, double* flux_x);
int lb_step_cg_collide(lb_ctx* ctx, double tau0, double tau1, double sigma,
                       double beta, int32_t force_mode, const double* force,
                       double flux_x_total);

/* Body force used by the last two-phase step (force.dat of main_TWOPHASE). */
int lb_last_force(lb_ctx* ctx, double* force);

/* FP64 device reductions over bulk nodes of the pre-collision state
 * (LBglobalforcing.h:8-33,61-78):
 *   sum_fc      = sum_n sum_q f_q c_q,dir         (f = f0 + f1 for two phase)
 *   sum_phi0_fc = sum_n (rho0/rho) sum_q f_q c_q,dir
 *   sum_phi1_fc = sum_n (rho1/rho) sum_q f_q c_q,dir
 * Any output pointer may be NULL. */
int lb_reduce_flux(lb_ctx* ctx, int32_t dir, double* sum_fc,
                   double* sum_phi0_fc, double* sum_phi1_fc);

/* Single-phase permeability solve on the device: step BGK in check
 * intervals of M = check_every steps until |k_t - k_{t-M}| <= tol |k_t| at two
 * consecutive checks, or max_steps.  k = nu <j_x> / F_x, where <j_x> is
 * sum_n j_x / n_box averaged over the states the interval's steps started
 * from (j = sum_q f c + F/2, nu = (tau - 1/2)/3); the average over an
 * interval keeps the rounding noise of FP16 storage out of the stop test.
 * std_one_phase/main.cpp:603-624 prints the same relative change of its
 * fluxes (it has no stop).  Returns k and the steps taken. */
int lb_solve_permeability(lb_ctx* ctx, double tau, const double* force,
                          int32_t check_every, double tol, int64_t max_steps,
                          double n_box, double* k_out, int64_t* steps_out);
/* The same solve stepping lb_step_trt; nu = (tau_sym - 1/2)/3.  With
 * (tau_sym - 1/2)(tau_anti - 1/2) fixed (3/16 puts a half-way wall exactly
 * half-way for Poiseuille flow) k does not depend on the viscosity. */
int lb_solve_permeability_trt(lb_ctx* ctx, double tau_sym, double tau_anti,
                              const double* force, int32_t check_every,
                              double tol, int64_t max_steps, double n_box,
                              double* k_out, int64_t* steps_out);

/* ------------------------------------------------------------ timing */
int lb_sync(lb_ctx* ctx);
/* Device time (HIP events on the context's stream) summed over the step
 * kernels launched since the last reset, and the number of launches. */
int lb_kernel_time(lb_ctx* ctx, double* ms, int64_t* launches);
int lb_kernel_time_reset(lb_ctx* ctx);
/* Bytes of device memory the context holds. */
int64_t lb_device_bytes(const lb_ctx* ctx);
/* Calibration (no counterpart in the reference): a plain device copy of
 * `bytes` bytes repeated `iters` times; returns read+write GB/s and the time
 * per copy.  Used to scale bandwidth figures and HBM counter readings. */
int lb_bench_copy(int device, int64_t bytes, int32_t iters, double* gbs, double* ms_per_iter);
/* Identity streaming of n_nodes float populations in the AoSoA layout (Q
 * loads and Q stores per node, no gathers, no collision): the bandwidth
 * ceiling of the population layout itself (read + write GB/s). */
int lb_bench_stream(int device, int32_t lattice, int64_t n_nodes, int32_t iters, double* gbs,
                    double* ms_per_iter);

/* ------------------------------------------------------------ topology host side */
/* .vtklb reader (LBvtk<DXQY>, LBvtk.h:221-614) followed by Grid
 * (LBgrid.h:127-148), Nodes::setupNodeType (LBnodes.h:112-209) and
 * findBulkNodes (LBgeometry.h:11-21). */
int lb_topo_read_vtklb(const char* path, int lattice, lb_topo** out);

/* Node list and neighbour index of one rank built natively from a voxel
 * geometry with the reference writer's numbering (PythonScripts/vtklb.py:
 * 6-44,70-124,182-218): geo[x][y][z] (C order, dims[0] slowest) is 0 for
 * solid and the 1-based rank for fluid; periodic is a bit mask (1 = x,
 * 2 = y, 4 = z).  Node types are set as Nodes does for that rank. */
int lb_topo_from_geometry(int lattice, int32_t nd, const int32_t* dims,
                          const int32_t* geo, int32_t periodic, int32_t rank,
                          lb_topo** out);
/* The same node list, neighbour index, positions, node types and bulk list
 * as lb_topo_from_geometry for rank 1, built on GPU `device` (padding and
 * rim copies, fluid/wall flags, a tiled scan for the numbering, neighbour
 * index and node types in kernels).  Single-rank geometries only: values 0
 * (solid) and 1 (fluid); other positive values give LB_ERR_UNSUPPORTED. */
int lb_topo_from_geometry_device(int device, int lattice, int32_t nd, const int32_t* dims,
                                 const int32_t* geo, int32_t periodic, lb_topo** out);
/* Binary twin of a .vtklb file (no counterpart in the reference: a format of
 * this library).  It holds what the text file holds -- global dimensions,
 * positions, neighbour index, nodetype attribute, node ranks, PROCESSOR
 * blocks and the SCALARS attributes (each as double and as float, the two
 * ways the reference reads them) -- as raw little-endian arrays.  Reading it
 * runs the same node typing and bulk list as lb_topo_read_vtklb, so both give
 * the same topology. */
int lb_topo_write_binary(const lb_topo* t, const char* path);
int lb_topo_read_binary(const char* path, int lattice, lb_topo** out);
int lb_topo_destroy(lb_topo* t);
int32_t lb_topo_num_nodes(const lb_topo* t);  /* incl. node 0 */
int32_t lb_topo_num_bulk(const lb_topo* t);
int32_t lb_topo_num_dims(const lb_topo* t);
int32_t lb_topo_rank(const lb_topo* t);      /* the MPI rank of this node list, 0-based (the file's process) */
int lb_topo_global_dims(const lb_topo* t, int32_t* dims);  /* padded dims */
int lb_topo_neighbors(const lb_topo* t, int32_t* neighbor);   /* [n_nodes*Q] */
int lb_topo_positions(const lb_topo* t, int32_t* pos);        /* [n_nodes*nD], node 0 = -1 */
int lb_topo_node_types(const lb_topo* t, int16_t* types);     /* [n_nodes] */
int lb_topo_node_ranks(const lb_topo* t, int32_t* ranks);     /* [n_nodes] */
int lb_topo_bulk(const lb_topo* t, int32_t* bulk);            /* [n_bulk] */
/* PROCESSOR blocks of a node list split over ranks: neighbouring rank k's
 * rank, and its nodes held here as (local node number, node number on that
 * rank) pairs.  Replaces LBvtk::getNumNeigProc / getNeigRank /
 * getNeigNodesNo / getNeigNodesNeigNo (LBvtk.h:152-173).  local / remote may
 * be null (n_nodes alone gives the sizes). */
int32_t lb_topo_num_neighbor_ranks(const lb_topo* t);
int lb_topo_neighbor_rank(const lb_topo* t, int32_t k, int32_t* rank, int32_t* n_nodes, int32_t* local,
                          int32_t* remote);
/* Point attribute of a .vtklb file (SCALARS block), as double over n_nodes
 * (node 0 = 0); as_float rounds each value through float first, the way
 * main_TWOPHASE.cpp:158,166,173 reads rho0/rho1/wettability. */
int lb_topo_attribute(const lb_topo* t, const char* name, int32_t as_float,
                      double* out);
/* Write the topology as a .vtklb file in the reference writer's format
 * (PythonScripts/vtklb.py:222-295), with optional float attributes given
 * over n_nodes (node 0 ignored). */
int lb_topo_write_vtklb(const lb_topo* t, const char* path, int32_t n_attr,
                        const char* const* names, const double* const* values);

/* Seeded sphere pack (geometry generator, not in the reference's C++):
 * spheres of `radius` placed by random sequential addition with centre
 * distance >= 2r, relaxed by `relax` after `max_fail` consecutive failures,
 * until the porosity of the non-wall region is <= porosity.  periodic: bit
 * mask of axes along which spheres wrap; walls != 0 adds one solid layer on
 * the faces of the other axes.  geo: int32[prod(dims)], 1 fluid / 0 solid.
 * info[4] = {porosity, n_spheres, final min centre distance, trials}. */
int lb_sphere_pack(int32_t nd, const int32_t* dims, double radius,
                   double porosity, uint64_t seed, int32_t periodic,
                   int32_t walls, int32_t max_fail, double relax,
                   int32_t* geo, double* info);

/* ------------------------------------------------------------ VTU output */
/* VTU/PVTU files in the reference's format (io/Output.h:25-175,
 * io/VTK.h:316-560, 854-1355), the output std_case and twophase write:
 * one voxel (3D) or pixel (2D) cell centred on each position, shared corner
 * points numbered in order of first use, data appended raw.  lb_vtu_create
 * takes the positions of the cells (Grid::pos(bulkNodes): n_cells*nd
 * ints), the output directory (files go to dir/vtu/RRRR_name_NNNNNNN.vtu and
 * dir/name_NNNNNNN.pvtu), the rank and number of ranks, and the data type
 * of every variable.  A variable is a host array of n_cells*dim values of
 * that type, cell-major; it is read at every lb_vtu_write, like the
 * reference's Output reads its fields (dim 2 is written with a zero third
 * component).  Each rank writes its own piece line of the .pvtu at its
 * slot; rank 0 writes the header and the last rank the footer, so with
 * several ranks the caller orders the calls (rank 0 first). */
enum lb_vtu_dtype { LB_VTU_F64 = 0, LB_VTU_F32 = 1, LB_VTU_I32 = 2 };
typedef struct lb_vtu lb_vtu;
int lb_vtu_create(int32_t nd, const int32_t* pos, int32_t n_cells, const char* dir,
                  int32_t rank, int32_t nprocs, int32_t dtype, lb_vtu** out);
int lb_vtu_add_file(lb_vtu* v, const char* name);                     /* Output::add_file */
int lb_vtu_add_variable(lb_vtu* v, const char* name, int32_t dim, const void* data);
int lb_vtu_write(lb_vtu* v, double time);                             /* Output::write */
int32_t lb_vtu_num_points(const lb_vtu* v);
int lb_vtu_destroy(lb_vtu* v);

/* Relative-permeability solve (colour gradient, fixed body force on the
 * total, LB_FORCE_FIXED): lb_step_cg until both phase permeabilities
 * k_i = nu_i <phi_i sum_q f c_x> / F_x (nu_i = (tau_i - 1/2)/3,
 * LBglobalforcing.h:61-78), taken every check_every steps from the state the
 * last step of the interval starts from, change by at most
 * tol (|k_0| + |k_1|) at two consecutive checks, or max_steps.  The stop of
 * a two-phase solve in a batch; k_phase[2] and the steps taken. */
int lb_solve_relperm(lb_ctx* ctx, double tau0, double tau1, double sigma, double beta,
                     const double* force, int32_t check_every, double tol, int64_t max_steps,
                     double n_box, double* k_phase, int64_t* steps_out);
/* ---- batched and serving form -------------------------------------------
 * No counterpart in the reference: the reference runs one solve per process
 * (std_case/main.cpp).  A batch keeps many independent permeability
 * solves (the std_case step with lb_solve_permeability's stop, or the
 * colour-gradient step of twophase with phase permeabilities)
 * resident in one pre-allocated device pool of pool_bytes and advances all of
 * them with one launch per step.  A solve is admitted when its footprint
 * (populations x 2, coordinates, row-window table, block partials) fits into
 * the pool; it retires when |dk| <= tol |k| between checks every check_every
 * steps, at two consecutive checks (or at max_steps), and frees its memory.  Each solve's per-node
 * arithmetic and reduction order are those of a single context with row
 * windows, so its k and step count are the same alone, in any mix, and
 * through lb_solve_permeability. */
typedef struct lb_batch lb_batch;
int lb_batch_create(int lattice, int precision, int device, int64_t pool_bytes,
                    int32_t check_every, lb_batch** out);
int lb_batch_destroy(lb_batch* b);
const char* lb_batch_last_error(const lb_batch* b);
/* Bytes a solve of this geometry (int32 cells, > 0 fluid, C order over
 * x, y[, z]) takes in a pool (n_fields 1 single-phase, 2 two-phase), and its
 * fluid node count. */
int lb_batch_footprint(int lattice, int precision, int32_t n_fields, int32_t nd,
                       const int32_t* dims, const int32_t* geo, int32_t periodic_mask,
                       int64_t* bytes, int64_t* n_fluid);
/* Admit a solve that starts at rest (f = w, rho = 1): LB_ERR_NOMEM when it
 * does not fit now (retry after others retire).  n_box = prod(dims). */
int lb_batch_add(lb_batch* b, int32_t nd, const int32_t* dims, const int32_t* geo,
                 int32_t periodic_mask, double tau, const double* force, double tol,
                 int64_t max_steps, int64_t tag);
/* Admit a colour-gradient two-phase solve (in the batch's two-phase precision,
 * lb_batch_set_precision): rho0 per
 * geometry cell (phase-0 density, rho1 = 1 - rho0, f_k = w rho_k at rest),
 * the solid-boundary colour from wettability (rho0 = wettability on every
 * solid neighbour, main_TWOPHASE.cpp:171-178,280-284), fixed body force on the
 * total (the LB_FORCE_FIXED mode of lb_step_cg).  It retires when both phase
 * permeabilities k_i = nu_i <phi_i sum_q f c_x> / F_x (nu_i = (tau_i - 1/2)/3,
 * LBglobalforcing.h:61-78) change by at most tol (|k_0| + |k_1|) between
 * checks, at two consecutive checks. */
/* Admit count single-phase solves (geometry k: dims[k*nd..], geos[k]) with
 * the parameters of lb_batch_add, tagged tags[k].  The row layouts are built
 * on up to 16 host threads, then the solves are admitted in order until one
 * does not fit: *n_added says how many were (LB_OK; the rest waits for
 * retirements).  Any other error stops it with that error. */
int lb_batch_add_many(lb_batch* b, int32_t count, int32_t nd, const int32_t* dims,
                      const int32_t* const* geos, int32_t periodic_mask, double tau,
                      const double* force, double tol, int64_t max_steps, const int64_t* tags,
                      int32_t* n_added);
int lb_batch_add_twophase(lb_batch* b, int32_t nd, const int32_t* dims, const int32_t* geo,
                          int32_t periodic_mask, const double* rho0, double wettability,
                          double tau0, double tau1, double sigma, double beta,
                          const double* force, double tol, int64_t max_steps, int64_t tag);
/* Advance every resident solve by up to `intervals` check intervals and wait
 * (lb_batch_launch, then lb_batch_wait).  Each interval ends on the device:
 * the per-solve sums and the stop test, so a solve stops at its own interval
 * (it is skipped by the rest of the sequence) and the host synchronises once
 * per call.  Admissions take effect at the next call. */
int lb_batch_advance(lb_batch* b, int32_t intervals, int32_t* n_finished);
/* The same in two halves: lb_batch_launch enqueues the intervals (HIP graphs
 * of one interval each, replayed) and returns; lb_batch_wait waits for them,
 * retires the solves whose stop test decided and frees their memory.  In
 * between, the host may admit solves (lb_batch_add*): their uploads run on
 * a stream of their own and overlap the intervals in flight. */
int lb_batch_launch(lb_batch* b, int32_t intervals);
int lb_batch_wait(lb_batch* b, int32_t* n_finished);
/* Take up to max_n finished solves: tag, k (two-phase: k_0 + k_1),
 * k_phase[2] (two-phase k_0, k_1), steps, seconds resident, status
 * (0 converged, 1 max_steps reached, 2 k not finite). */
int lb_batch_collect(lb_batch* b, int32_t max_n, int64_t* tags, double* k, double* k_phase,
                     int64_t* steps, double* seconds, int32_t* status, int32_t* n_out);
int lb_batch_status(const lb_batch* b, int32_t* n_resident, int64_t* resident_nodes,
                    int64_t* bytes_used, int64_t* bytes_free, int64_t* largest_free);
/* Precision of the solves of one kind admitted from now on (kind 0
 * single-phase, 1 two-phase; LB_FP64, LB_FP32S or LB_FP16S); both kinds start
 * at the batch's precision.  LB_ERR_STATE while solves of that kind are
 * resident.  A two-phase solve in FP64 has the bits of lb_step_cg (fixed
 * force) on a context with row windows. */
int lb_batch_set_precision(lb_batch* b, int32_t kind, int32_t precision);
/* Launch shape: -1 autotune on live steps after each change of the resident
 * set by 20 % or more (default), 0 one workgroup per 256-node block, k > 0 a
 * persistent launch of k workgroups per CU (a fixed shape uses 256-thread
 * workgroups).  The tuner times nine candidates, workgroups per CU
 * {0, 1, 2, 4, 8} with 256-thread workgroups and {0, 1, 2, 4} with 512-thread
 * workgroups that work on two 256-node blocks at a time, on the live steps
 * of one interval.  No candidate changes a result. */
int lb_batch_set_shape(lb_batch* b, int32_t wg_per_cu);
int lb_batch_shape(const lb_batch* b, int32_t* wg_per_cu);
/* The tuner's choice: workgroups per CU, threads per workgroup, and the
 * resident node count it was tuned at (-1: not tuned yet). */
int lb_batch_tuning(const lb_batch* b, int32_t* wg_per_cu, int32_t* wg_threads, int64_t* tuned_nodes);
/* Device time of the step launches and sums (HIP events), launches, and
 * lattice updates done. */
int lb_batch_kernel_time(const lb_batch* b, double* ms, int64_t* launches, int64_t* lups);

#ifdef __cplusplus
}
#endif

#endif /* BADCHIMP_AMD_H */
