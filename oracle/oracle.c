/*
 * oracle.c -- CPU restatement of the reference's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py load this library, and only as the checker
 * (or as the timed CPU baseline).  The product never calls it.
 *
 * It follows the reference step by step, in its data layout (AoS LbField,
 * LBfield.h:298-305), with push streaming (LBfield.h:350-357), swap
 * (LBfield.h:359) and half-way bounce-back after the push
 * (LBhalfwayhelperclass.h:110-161, LBhalfwaybb.h:37-63).  The moment and
 * collision formulas are written with the reference's explicit term order:
 *   D2Q9  tables and helpers   LBd2q9.h:14-40, 128-177
 *   D3Q19 tables and helpers   LBd3q19.h:14-40, 111-190
 *   calcRho / calcVel          LBmacroscopic.h:11-45
 *   initiateLbField            LBinitiatefield.h:33-57
 *   calcOmegaBGK               LBcollision.h:27-46
 *   calcOmegaBGKTRT            LBcollision.h:50-75
 *   calcDeltaOmegaQ            LBcollision.h:79-98
 *   calcDeltaOmegaF            LBcollision.h:194-213
 *   calcDeltaOmegaFTRT         LBcollision.h:214-233
 *   GeneralizedNewtonian       LBrheology.h:321-476 (polymer_viscosity_tabulation/main.cpp:105-155)
 *   tagged links, mass sources std_one_phase/main.cpp:27-204, 322-354, 513-612
 *   calcDeltaOmegaST / RC      LBcollision2phase.h:7-20, 70-86
 *   grad / vecNorm             LButilities.h:12-22, 62-66
 *   std_case loop              std_case/main.cpp:109-145
 *   twophase loop              twophase/main_TWOPHASE.cpp:236-392
 *   flux sums                  LBglobalforcing.h:8-33, 61-78
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, SSE2 doubles as the
 * reference's x86-64 build uses).
 *
 * Pinned by: the compiled reference's own outputs on the committed
 * geometries (tests/golden/ref_*.npz) -- bit-exact -- and by the closed
 * forms in tests/test_oracle.py (rest state, uniform-force growth, mass
 * conservation, Poiseuille start-up series).  The flux sums are pinned
 * bit-exact against the reference's own calcCapNumbForceCartDir and
 * calcFluxForceCartDir run on two-phase fields by a harness compiled from
 * its headers (tests/golden/forcing_ref.npz, oracle/flux_harness.cpp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SQRT2 1.4142135623730950488
#define DBL_EPS_LB 2.220446049250313080847e-16

/* ------------------------------------------------------------ lattices */
typedef struct {
  int nD, nQ, nDirPairs;
  const int *c;      /* [nQ*nD] */
  const double *w;   /* [nQ] */
  const int *rev;    /* [nQ] */
  const double *B;   /* [nQ] */
  const double *cNorm;
  double w1c2Inv, w2c2Inv;
} lattice_t;

static const int d2q9_c[18] = {1, 0, 1, 1, 0, 1, -1, 1, -1, 0, -1, -1, 0, -1, 1, -1, 0, 0};
static const double d2q9_w[9] = {4.0 / 36.0, 1.0 / 36.0, 4.0 / 36.0, 1.0 / 36.0, 4.0 / 36.0,
                                 1.0 / 36.0, 4.0 / 36.0, 1.0 / 36.0, 16.0 / 36.0};
static const int d2q9_rev[9] = {4, 5, 6, 7, 0, 1, 2, 3, 8};
static const double d2q9_B[9] = {8.0 / 108.0, 5.0 / 108.0, 8.0 / 108.0, 5.0 / 108.0, 8.0 / 108.0,
                                 5.0 / 108.0, 8.0 / 108.0, 5.0 / 108.0, -16.0 / 108.0};
static const double d2q9_cn[9] = {1.0, SQRT2, 1.0, SQRT2, 1.0, SQRT2, 1.0, SQRT2, 0.0};

static const int d3q19_c[57] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 1, 0, 1, -1, 0, 1, 0, 1, 1, 0,
                                -1, 0, 1, 1, 0, 1, -1, -1, 0, 0, 0, -1, 0, 0, 0, -1, -1, -1, 0, -1,
                                1, 0, -1, 0, -1, -1, 0, 1, 0, -1, -1, 0, -1, 1, 0, 0, 0};
#define W1 (2.0 / 36.0)
#define W2 (1.0 / 36.0)
static const double d3q19_w[19] = {W1, W1, W1, W2, W2, W2, W2, W2, W2, W1,
                                   W1, W1, W2, W2, W2, W2, W2, W2, 12.0 / 36.0};
static const int d3q19_rev[19] = {9, 10, 11, 12, 13, 14, 15, 16, 17, 0, 1, 2, 3, 4, 5, 6, 7, 8, 18};
#define B1 (1.0 / 54.0)
#define B2 (2.0 / 54.0)
static const double d3q19_B[19] = {B1, B1, B1, B2, B2, B2, B2, B2, B2, B1,
                                   B1, B1, B2, B2, B2, B2, B2, B2, -12.0 / 54.0};
static const double d3q19_cn[19] = {1.0, 1.0, 1.0, SQRT2, SQRT2, SQRT2, SQRT2, SQRT2, SQRT2, 1.0,
                                    1.0, 1.0, SQRT2, SQRT2, SQRT2, SQRT2, SQRT2, SQRT2, 0.0};

static lattice_t get_lattice(int id) {
  lattice_t L;
  if (id == 0) {
    L.nD = 2; L.nQ = 9; L.nDirPairs = 4;
    L.c = d2q9_c; L.w = d2q9_w; L.rev = d2q9_rev; L.B = d2q9_B; L.cNorm = d2q9_cn;
    L.w1c2Inv = (4.0 / 36.0) * 3.0; L.w2c2Inv = (1.0 / 36.0) * 3.0;
  } else {
    L.nD = 3; L.nQ = 19; L.nDirPairs = 9;
    L.c = d3q19_c; L.w = d3q19_w; L.rev = d3q19_rev; L.B = d3q19_B; L.cNorm = d3q19_cn;
    L.w1c2Inv = W1 * 3.0; L.w2c2Inv = W2 * 3.0;
  }
  return L;
}

static const double c2Inv = 3.0, c4Inv = 9.0, c4Inv0_5 = 0.5 * 9.0;
static const double c2 = 1.0 / 3.0;

/* qSum: left fold from 0.0 (LBd3q19.h:173-180) */
static double qsum(const lattice_t *L, const double *f) {
  double r = 0.0;
  for (int q = 0; q < L->nQ; ++q) r += f[q];
  return r;
}

/* qSumC, written out as the generated headers do (LBd2q9.h:170-177, LBd3q19.h:182-190) */
static void qsumc(const lattice_t *L, const double *d, double *r) {
  if (L->nQ == 9) {
    r[0] = + d[0] + d[1] - d[3] - d[4] - d[5] + d[7];
    r[1] = + d[1] + d[2] + d[3] - d[5] - d[6] - d[7];
  } else {
    r[0] = + d[0] + d[3] + d[4] + d[5] + d[6] - d[9] - d[12] - d[13] - d[14] - d[15];
    r[1] = + d[1] + d[3] - d[4] + d[7] + d[8] - d[10] - d[12] + d[13] - d[16] - d[17];
    r[2] = + d[2] + d[5] - d[6] + d[7] - d[8] - d[11] - d[14] + d[15] - d[16] + d[17];
  }
}

/* cDotAll (LBd2q9.h:128-142, LBd3q19.h:129-153) */
static void cdotall(const lattice_t *L, const double *v, double *r) {
  if (L->nQ == 9) {
    r[0] = +v[0]; r[1] = +v[0] + v[1]; r[2] = +v[1]; r[3] = -v[0] + v[1];
    r[4] = -v[0]; r[5] = -v[0] - v[1]; r[6] = -v[1]; r[7] = +v[0] - v[1]; r[8] = 0.0;
  } else {
    r[0] = +v[0]; r[1] = +v[1]; r[2] = +v[2];
    r[3] = +v[0] + v[1]; r[4] = +v[0] - v[1]; r[5] = +v[0] + v[2]; r[6] = +v[0] - v[2];
    r[7] = +v[1] + v[2]; r[8] = +v[1] - v[2];
    r[9] = -v[0]; r[10] = -v[1]; r[11] = -v[2];
    r[12] = -v[0] - v[1]; r[13] = -v[0] + v[1]; r[14] = -v[0] - v[2]; r[15] = -v[0] + v[2];
    r[16] = -v[1] - v[2]; r[17] = -v[1] + v[2]; r[18] = 0.0;
  }
}

static double dot(const lattice_t *L, const double *a, const double *b) {
  return L->nD == 2 ? a[0] * b[0] + a[1] * b[1] : a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* grad (LBd2q9.h:144-151, LBd3q19.h:155-163) */
static void lattice_grad(const lattice_t *L, const double *s, double *g) {
  const double a = L->w1c2Inv, b = L->w2c2Inv;
  if (L->nQ == 9) {
    g[0] = + a * (+ s[0] - s[4]) + b * (+ s[1] - s[3] - s[5] + s[7]);
    g[1] = + a * (+ s[2] - s[6]) + b * (+ s[1] + s[3] - s[5] - s[7]);
  } else {
    g[0] = + a * (+ s[0] - s[9]) + b * (+ s[3] + s[4] + s[5] + s[6] - s[12] - s[13] - s[14] - s[15]);
    g[1] = + a * (+ s[1] - s[10]) + b * (+ s[3] - s[4] + s[7] + s[8] - s[12] + s[13] - s[16] - s[17]);
    g[2] = + a * (+ s[2] - s[11]) + b * (+ s[5] - s[6] + s[7] - s[8] - s[14] + s[15] - s[16] + s[17]);
  }
}

/* ------------------------------------------------------------ bounce-back */
typedef struct {
  int n;          /* boundary nodes */
  int32_t *node;  /* [n] */
  int32_t *nbeta, *ngamma, *ndelta;  /* [n] */
  int32_t *links; /* [n*nDirPairs]: beta..., gamma..., delta... */
} bb_t;

static int is_fluid(const int16_t *type, int32_t n) { return type[n] > 1; }

/* BoundaryHalwWayHelper ctor (LBhalfwayhelperclass.h:110-161) */
static bb_t bb_setup(const lattice_t *L, const int32_t *nodes, int n, const int32_t *nb,
                     const int16_t *type) {
  bb_t b;
  b.n = n;
  b.node = malloc(sizeof(int32_t) * (n ? n : 1));
  b.nbeta = malloc(sizeof(int32_t) * (n ? n : 1));
  b.ngamma = malloc(sizeof(int32_t) * (n ? n : 1));
  b.ndelta = malloc(sizeof(int32_t) * (n ? n : 1));
  b.links = malloc(sizeof(int32_t) * (n ? n : 1) * L->nDirPairs);
  for (int k = 0; k < n; ++k) {
    const int32_t node = nodes[k];
    int beta[9], gamma[9], delta[9], nbt = 0, ng = 0, nd = 0;
    for (int q = 0; q < L->nDirPairs; ++q) {
      const int32_t a = nb[(int64_t)node * L->nQ + q];
      const int32_t r = nb[(int64_t)node * L->nQ + L->rev[q]];
      if (is_fluid(type, a)) {
        if (is_fluid(type, r)) gamma[ng++] = q;
        else beta[nbt++] = q;
      } else {
        if (is_fluid(type, r)) beta[nbt++] = L->rev[q];
        else delta[nd++] = q;
      }
    }
    b.node[k] = node;
    b.nbeta[k] = nbt; b.ngamma[k] = ng; b.ndelta[k] = nd;
    int32_t *l = b.links + (int64_t)k * L->nDirPairs;
    for (int i = 0; i < nbt; ++i) *l++ = beta[i];
    for (int i = 0; i < ng; ++i) *l++ = gamma[i];
    for (int i = 0; i < nd; ++i) *l++ = delta[i];
  }
  return b;
}

static void bb_free(bb_t *b) {
  free(b->node); free(b->nbeta); free(b->ngamma); free(b->ndelta); free(b->links);
}

/* HalfWayBounceBack::apply (LBhalfwaybb.h:37-63) on field k of nF */
static void bb_apply(const lattice_t *L, const bb_t *b, double *f, int nF, int k, const int32_t *nb) {
  const int Q = L->nQ;
  for (int i = 0; i < b->n; ++i) {
    const int32_t node = b->node[i];
    const int32_t *l = b->links + (int64_t)i * L->nDirPairs;
    for (int m = 0; m < b->nbeta[i]; ++m) {
      const int beta = l[m], br = L->rev[beta];
      f[(int64_t)node * nF * Q + k * Q + beta] = f[(int64_t)nb[(int64_t)node * Q + br] * nF * Q + k * Q + br];
    }
    const int32_t *dl = l + b->nbeta[i] + b->ngamma[i];
    for (int m = 0; m < b->ndelta[i]; ++m) {
      const int d = dl[m], dr = L->rev[d];
      f[(int64_t)node * nF * Q + k * Q + d] = f[(int64_t)nb[(int64_t)node * Q + dr] * nF * Q + k * Q + dr];
      f[(int64_t)node * nF * Q + k * Q + dr] = f[(int64_t)nb[(int64_t)node * Q + d] * nF * Q + k * Q + d];
    }
  }
}

/* ------------------------------------------------------------ std_case */
/* nsteps iterations of std_case/main.cpp:110-145.  f, ftmp: LbField (1 field)
 * over n_nodes; rho, vel: ScalarField / VectorField of the last iteration. */
static int std_case_run(int lattice, int32_t n_nodes, const int32_t *nb, const int16_t *type,
                        const int32_t *bulk, int32_t n_bulk, double *f, double *ftmp, double tau,
                        int trt, double tauAnti, const double *F, int32_t nsteps, double *rho,
                        double *vel, int nthreads) {
  const lattice_t L = get_lattice(lattice);
  const int Q = L.nQ, D = L.nD;
  /* fluid boundary nodes of this rank (LBgeometry.h:48-56) */
  int32_t *fb = malloc(sizeof(int32_t) * (n_nodes ? n_nodes : 1));
  int nfb = 0;
  for (int32_t n = 1; n < n_nodes; ++n)
    if (type[n] == 2) fb[nfb++] = n;
  bb_t bb = bb_setup(&L, fb, nfb, nb, type);
  free(fb);
  const double tau_inv = 1.0 / tau;
  const double tau_factor = (1 - 0.5 / tau);
  /* two-relaxation-time constants (LBcollision.h:62-63, 223-224) */
  const double tauSym_inv = 1.0 / tau, tauAnti_inv = 1.0 / tauAnti;
  const double tauSym_factor = (1 - 0.5 / tau), tauAnti_factor = (1 - 0.5 / tauAnti);
  double cF[19];
  cdotall(&L, F, cF);
  double *a = f, *b = ftmp;
  (void)nthreads;
  for (int32_t it = 0; it < nsteps; ++it) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int32_t k = 0; k < n_bulk; ++k) {
      const int32_t node = bulk[k];
      const double *fn = a + (int64_t)node * Q;
      const double rhoN = qsum(&L, fn);
      double j[3], u[3], cu[19];
      qsumc(&L, fn, j);
      for (int d = 0; d < D; ++d) u[d] = (j[d] + 0.5 * F[d]) / rhoN;
      rho[node] = rhoN;
      for (int d = 0; d < D; ++d) vel[(int64_t)node * D + d] = u[d];
      const double u2 = dot(&L, u, u);
      cdotall(&L, u, cu);
      const double uF = dot(&L, u, F);
      for (int q = 0; q < Q; ++q) {
        double om, dF;
        if (!trt) {
          om = -tau_inv * (fn[q] - rhoN * L.w[q] * (1.0 + c2Inv * cu[q] + c4Inv0_5 * (cu[q] * cu[q] - c2 * u2)));
          dF = L.w[q] * tau_factor * (c2Inv * cF[q] + c4Inv * (cF[q] * cu[q] - c2 * uF));
        } else {
          /* calcOmegaBGKTRT (LBcollision.h:64-71), calcDeltaOmegaFTRT with
           * phi = 1 (LBcollision.h:226-229; w*1.0 == w exactly) */
          const double fqSym = 0.5 * (fn[q] + fn[L.rev[q]]);
          const double fqAnti = 0.5 * (fn[q] - fn[L.rev[q]]);
          om = -tauSym_inv * (fqSym - rhoN * L.w[q] * (1.0 + c4Inv0_5 * (cu[q] * cu[q] - c2 * u2)))
               - tauAnti_inv * (fqAnti - rhoN * L.w[q] * c2Inv * cu[q]);
          dF = L.w[q] * 1.0 * (tauAnti_factor * c2Inv * cF[q] + tauSym_factor * c4Inv * (cF[q] * cu[q] - c2 * uF));
        }
        b[(int64_t)nb[(int64_t)node * Q + q] * Q + q] = fn[q] + om + dF;
      }
    }
    double *t = a; a = b; b = t;
    bb_apply(&L, &bb, a, 1, 0, nb);
  }
  if (a != f) memcpy(f, a, sizeof(double) * (size_t)n_nodes * Q);
  bb_free(&bb);
  return 0;
}

int orc_std_case(int lattice, int32_t n_nodes, const int32_t *nb, const int16_t *type,
                 const int32_t *bulk, int32_t n_bulk, double *f, double *ftmp, double tau,
                 const double *F, int32_t nsteps, double *rho, double *vel, int nthreads) {
  return std_case_run(lattice, n_nodes, nb, type, bulk, n_bulk, f, ftmp, tau, 0, tau, F, nsteps, rho,
                      vel, nthreads);
}

/* std_case with the two-relaxation-time operator: calcOmegaBGK ->
 * calcOmegaBGKTRT(f, tauSym, tauAnti, ...) and calcDeltaOmegaF ->
 * calcDeltaOmegaFTRT(tauSym, tauAnti, 1.0, ...), LBcollision.h:50-75, 214-233. */
int orc_std_case_trt(int lattice, int32_t n_nodes, const int32_t *nb, const int16_t *type,
                     const int32_t *bulk, int32_t n_bulk, double *f, double *ftmp, double tau_sym,
                     double tau_anti, const double *F, int32_t nsteps, double *rho, double *vel,
                     int nthreads) {
  return std_case_run(lattice, n_nodes, nb, type, bulk, n_bulk, f, ftmp, tau_sym, 1, tau_anti, F, nsteps,
                      rho, vel, nthreads);
}

/* ------------------------------------------------------------ std_one_phase */
/* Tagged-link boundaries and per-domain mass sources of
 * std_one_phase/main.cpp, one rank.  Node tags (Nodes::getTag, short, -1 by
 * default; main.cpp:274-281): bits 0-1 phase (0 solid, 1, 2 fluid phases,
 * 3 pressure ghost), bit 2 fluid-fluid interface, bit 3 fluid-solid
 * interface, bit 4 pressure boundary.  Link lists (main.cpp:27-126), each
 * entry {node, q unknown = rev q, neighbour, q known = q}.  f, ftmp: LbField
 * (1 field); label, force_on: interior_domains and force attributes;
 * flux[2]: the mass fluxes q1, q2 of the last iteration (main.cpp:610-624). */
typedef struct {
  int32_t n, *e; /* [n][4] */
} links_t;

static void links_add(links_t *l, int32_t a, int32_t b, int32_t c, int32_t d) {
  l->e = realloc(l->e, sizeof(int32_t) * 4 * (size_t)(l->n + 1));
  l->e[4 * l->n] = a;
  l->e[4 * l->n + 1] = b;
  l->e[4 * l->n + 2] = c;
  l->e[4 * l->n + 3] = d;
  l->n++;
}

static double cdotref(const lattice_t *L, int q, const double *v) {
  /* c(q,0)*v[0] + c(q,1)*v[1] (+ c(q,2)*v[2]), LBd3q19.h:124-127 */
  double r = L->c[q * L->nD] * v[0] + L->c[q * L->nD + 1] * v[1];
  if (L->nD == 3) r = r + L->c[q * L->nD + 2] * v[2];
  return r;
}

int orc_one_phase(int lattice, int32_t n_nodes, const int32_t *nb, const int32_t *bulk, int32_t n_bulk,
                  const int32_t *tags, const int32_t *label, const double *force_on, const double *normals,
                  double *f, double *ftmp, double tau, const double *F, double rho_w, int32_t nsteps,
                  double *rho, double *vel, double *flux) {
  const lattice_t L = get_lattice(lattice);
  const int Q = L.nQ, D = L.nD;
  links_t sf = {0, NULL}, pf = {0, NULL}, ff = {0, NULL};
  for (int32_t n = 1; n < n_nodes; ++n) {
    const int t = (short)tags[n];
    for (int q = 0; q < Q - 1; ++q) {
      const int32_t nn = nb[(int64_t)n * Q + q];
      const int tn = (short)tags[nn];
      if (((t >> 3) & 1) && (tn & 3) == 0) links_add(&sf, n, L.rev[q], nn, q);
      if (((t >> 4) & 1) && (tn & 3) == 3) links_add(&pf, n, L.rev[q], nn, q);
      if (((t >> 2) & 1) && (t & 3) == 1 && (tn & 3) == 2) links_add(&ff, n, L.rev[q], nn, q);
    }
  }
  /* mass-source set-up (main.cpp:322-354) */
  int lmax = 0;
  for (int32_t n = 1; n < n_nodes; ++n)
    if (label[n] > lmax) lmax = label[n];
  double *scale = calloc((size_t)lmax + 1, sizeof(double));
  double *mc = calloc((size_t)lmax + 1, sizeof(double));
  double *add = calloc((size_t)n_nodes, sizeof(double));
  for (int32_t n = 1; n < n_nodes; ++n)
    if (label[n] > 0 && (short)tags[n] < 3) {
      scale[label[n]] += 1;
      add[n] = 1.0;
    }
  int rc = 0;
  for (int i = 1; i <= lmax; ++i) {
    if (scale[i] == 0.0) rc = -1;
    else scale[i] = 1.0 / scale[i];
  }
  double *a = f, *b = ftmp;
  for (int32_t it = 0; it < nsteps && rc == 0; ++it) {
    for (int i = 0; i <= lmax; ++i) mc[i] = 0.0;
    for (int32_t k = 0; k < n_bulk; ++k) {
      const int32_t node = bulk[k];
      mc[label[node]] += 1.0 - qsum(&L, a + (int64_t)node * Q);
    }
    for (int32_t k = 0; k < n_bulk; ++k) {
      const int32_t node = bulk[k];
      const double *fn = a + (int64_t)node * Q;
      double rhoN = qsum(&L, fn);
      const int lb = label[node];
      const double qm = 0.9 * 2 * scale[lb] * mc[lb] * add[node];
      rhoN += 0.5 * qm;
      double Fn[3] = {0, 0, 0}, j[3], u[3], cu[19], cF[19];
      for (int d = 0; d < D; ++d) Fn[d] = F[d] * force_on[node];
      qsumc(&L, fn, j);
      for (int d = 0; d < D; ++d) u[d] = (j[d] + 0.5 * Fn[d]) / rhoN;
      rho[node] = rhoN;
      for (int d = 0; d < D; ++d) vel[(int64_t)node * D + d] = u[d];
      const double u2 = dot(&L, u, u);
      cdotall(&L, u, cu);
      const double uF = dot(&L, u, Fn);
      cdotall(&L, Fn, cF);
      const double tau_inv = 1.0 / tau, tau_factor = (1 - 0.5 / tau);
      for (int q = 0; q < Q; ++q) {
        const double om = -tau_inv * (fn[q] - rhoN * L.w[q] * (1.0 + c2Inv * cu[q] + c4Inv0_5 * (cu[q] * cu[q] - c2 * u2)));
        const double dF = L.w[q] * tau_factor * (c2Inv * cF[q] + c4Inv * (cF[q] * cu[q] - c2 * uF));
        const double dQ = tau_factor * qm * L.w[q] * (1.0 + c2Inv * cu[q] + c4Inv0_5 * (cu[q] * cu[q] - c2 * u2));
        b[(int64_t)nb[(int64_t)node * Q + q] * Q + q] = fn[q] + om + dF + dQ;
      }
    }
    double *t = a; a = b; b = t;
    /* applySolidFluidBoundary (main.cpp:138-153) */
    for (int32_t k = 0; k < sf.n; ++k) {
      const int32_t *e = sf.e + 4 * k;
      a[(int64_t)e[0] * Q + e[1]] = a[(int64_t)e[2] * Q + e[3]];
    }
    /* applyPressureFluidBoundary (main.cpp:156-175) */
    for (int32_t k = 0; k < pf.n; ++k) {
      const int32_t *e = pf.e + 4 * k;
      const double *v = vel + (int64_t)e[0] * D;
      const double u2 = dot(&L, v, v);
      const double cu = cdotref(&L, e[3], v);
      const double w = L.w[e[3]];
      a[(int64_t)e[0] * Q + e[1]] = -a[(int64_t)e[2] * Q + e[3]] + 2 * w * rho_w * (1 + 0.5 * (c4Inv * cu * cu - c2Inv * u2));
    }
    /* applyFluidFluidBoundary (main.cpp:178-204) */
    for (int32_t k = 0; k < ff.n; ++k) {
      const int32_t *e = ff.e + 4 * k;
      const int32_t n1 = e[0], qu = e[1], n2 = e[2], qk = e[3];
      const double f1 = a[(int64_t)n1 * Q + qu];
      double uw[3] = {0, 0, 0}, nv[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) {
        uw[d] = 0.5 * (vel[(int64_t)n1 * D + d] + vel[(int64_t)n2 * D + d]);
        nv[d] = 0.5 * (normals[(int64_t)n1 * 3 + d] + normals[(int64_t)n2 * 3 + d]);
      }
      const double length = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
      for (int d = 0; d < D; ++d) nv[d] /= (length + 1e-14);
      const double cn = cdotref(&L, qu, nv);
      const double cu = cdotref(&L, qu, uw);
      const double un = dot(&L, uw, nv);
      const double w = L.w[qu];
      const int isbnd1 = 1 - (((short)tags[n1] >> 3) & 1);
      const int isbnd2 = 1 - (((short)tags[n2] >> 3) & 1);
      const double dfu = 0 * 6 * w * (cu - un * cn) * isbnd1 * isbnd2;
      a[(int64_t)n1 * Q + qu] = a[(int64_t)n2 * Q + qk] + dfu;
      a[(int64_t)n2 * Q + qk] = f1 - dfu;
    }
    /* mass fluxes over the pressure-boundary nodes (main.cpp:600-612) */
    if (flux && it == nsteps - 1 && D == 3) {
      double m[2] = {0, 0};
      for (int32_t n = 1; n < n_nodes; ++n) {
        if ((((short)tags[n] >> 4) & 1) == 0) continue;
        const int ph = ((short)tags[n] & 3) - 1;
        if (ph != 0 && ph != 1) {
          rc = -2;
          break;
        }
        m[ph] += vel[(int64_t)n * 3 + 2] * rho[n];
      }
      flux[0] = 0.5 * m[0];
      flux[1] = 0.5 * m[1];
    }
  }
  if (a != f) memcpy(f, a, sizeof(double) * (size_t)n_nodes * Q);
  free(sf.e);
  free(pf.e);
  free(ff.e);
  free(scale);
  free(mc);
  free(add);
  return rc;
}

/* ------------------------------------------------------------ rheology */
/* GeneralizedNewtonian (LBrheology.h:321-476): table with a constant
 * extension below and above (:370-384), viscosity by linear interpolation
 * of the strain-rate square, located with upper_bound over the inner
 * entries (:459-462). */
typedef struct {
  int n;          /* table_length + 2 */
  double *sr, *nu;
} visc_table_t;

static visc_table_t table_setup(int32_t m, const double *sr, const double *nu) {
  visc_table_t t;
  t.n = m + 2;
  t.sr = malloc(sizeof(double) * t.n);
  t.nu = malloc(sizeof(double) * t.n);
  for (int i = 0; i < m; ++i) {
    t.sr[i + 1] = sr[i];
    t.nu[i + 1] = nu[i];
  }
  t.sr[0] = t.sr[1] - 1;
  t.nu[0] = t.nu[1];
  t.sr[m + 1] = t.sr[m] + 1;
  t.nu[m + 1] = t.nu[m];
  return t;
}

/* std::upper_bound(sr + 1, sr + n - 1, v) - (sr + 1), libstdc++'s loop */
static int upper_index(const visc_table_t *t, double v) {
  int first = 1, len = t->n - 2;
  while (len > 0) {
    const int half = len >> 1, mid = first + half;
    if (v < t->sr[mid]) {
      len = half;
    } else {
      first = mid + 1;
      len = len - half - 1;
    }
  }
  return first - 1;
}

/* nsteps iterations of polymer_viscosity_tabulation/main.cpp:105-155: the
 * std_case loop with GeneralizedNewtonian::omegaBGK (source 0) and the Guo
 * term at the node's relaxation time; visc: the node viscosities of the
 * last iteration. */
int orc_rheology(int lattice, int32_t n_nodes, const int32_t *nb, const int16_t *type, const int32_t *bulk,
                 int32_t n_bulk, double *f, double *ftmp, int32_t n_table, const double *sr, const double *nu,
                 const double *F, int32_t nsteps, double *rho, double *vel, double *visc, int nthreads) {
  const lattice_t L = get_lattice(lattice);
  const int Q = L.nQ, D = L.nD;
  int32_t *fb = malloc(sizeof(int32_t) * (n_nodes ? n_nodes : 1));
  int nfb = 0;
  for (int32_t n = 1; n < n_nodes; ++n)
    if (type[n] == 2) fb[nfb++] = n;
  bb_t bb = bb_setup(&L, fb, nfb, nb, type);
  free(fb);
  visc_table_t tab = table_setup(n_table, sr, nu);
  double cF[19];
  cdotall(&L, F, cF);
  double *a = f, *b = ftmp;
  (void)nthreads;
  for (int32_t it = 0; it < nsteps; ++it) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int32_t k = 0; k < n_bulk; ++k) {
      const int32_t node = bulk[k];
      const double *fn = a + (int64_t)node * Q;
      const double rhoN = qsum(&L, fn);
      double j[3], u[3], cu[19], feq[19], S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      qsumc(&L, fn, j);
      for (int d = 0; d < D; ++d) u[d] = (j[d] + 0.5 * F[d]) / rhoN;
      rho[node] = rhoN;
      for (int d = 0; d < D; ++d) vel[(int64_t)node * D + d] = u[d];
      const double u2 = dot(&L, u, u);
      cdotall(&L, u, cu);
      /* LBrheology.h:437-447 */
      for (int q = 0; q < Q; ++q) {
        feq[q] = L.w[q] * rhoN * (1 + c2Inv * cu[q] + c4Inv0_5 * (cu[q] * cu[q] - c2 * u2));
        for (int i = 0; i < D; ++i) {
          const int ci = L.c[q * D + i];
          S[i + D * i] += (fn[q] - feq[q]) * (ci * ci - c2);
          for (int jj = i + 1; jj < D; ++jj) S[i + D * jj] += (fn[q] - feq[q]) * ci * L.c[q * D + jj];
        }
      }
      /* LBrheology.h:449-457, source = 0 */
      const double source = 0.0;
      double sq = 0;
      for (int i = 0; i < D; ++i) {
        S[i + D * i] += (u[i] * F[i]);
        S[i + D * i] += 0.5 * u2 * source;
        sq += S[i + D * i] * S[i + D * i];
        for (int jj = i + 1; jj < D; ++jj) {
          S[i + D * jj] += 0.5 * (u[i] * F[jj] + u[jj] * F[i]);
          S[i + D * jj] += 0.5 * u[i] * u[jj] * source;
          sq += 2 * S[i + D * jj] * S[i + D * jj];
        }
      }
      /* LBrheology.h:459-465 */
      const int i0 = upper_index(&tab, sq);
      const double nuN = (tab.nu[i0 + 1] - tab.nu[i0]) * (sq - tab.sr[i0]) / (tab.sr[i0 + 1] - tab.sr[i0]) + tab.nu[i0];
      const double tau = nuN * c2Inv / rhoN + 0.5;
      const double tau_inv = 1.0 / tau;
      visc[node] = nuN;
      /* calcDeltaOmegaF at the node's tau (main.cpp:130-134) */
      const double uF = dot(&L, u, F);
      const double tau_factor = (1 - 0.5 / tau);
      for (int q = 0; q < Q; ++q) {
        const double om = -tau_inv * (fn[q] - feq[q]);
        const double dF = L.w[q] * tau_factor * (c2Inv * cF[q] + c4Inv * (cF[q] * cu[q] - c2 * uF));
        b[(int64_t)nb[(int64_t)node * Q + q] * Q + q] = fn[q] + om + dF;
      }
    }
    double *t = a; a = b; b = t;
    bb_apply(&L, &bb, a, 1, 0, nb);
  }
  if (a != f) memcpy(f, a, sizeof(double) * (size_t)n_nodes * Q);
  bb_free(&bb);
  free(tab.sr);
  free(tab.nu);
  return 0;
}

/* ------------------------------------------------------------ twophase */
/* nsteps iterations of main_TWOPHASE.cpp:238-392.
 *   f, ftmp : LbField with 2 fields over n_nodes
 *   rho     : ScalarField with 2 fields; solid-boundary entries hold the
 *             wettability densities on entry (main_TWOPHASE.cpp:171-178)
 *   cg      : ScalarField (colour field) over n_nodes, zero-initialised
 *   force_mode 0: F fixed; 1: F[0] = momx (flux controller :292-308)
 *   force_out[nD] : body force of the last iteration */
int orc_twophase(int lattice, int32_t n_nodes, const int32_t *nb, const int16_t *type,
                 const int32_t *bulk, int32_t n_bulk, double *f, double *ftmp, double *rho,
                 double *vel, double *cg, double tau0, double tau1, double sigma, double beta,
                 int force_mode, const double *Fin, int64_t n_global, int32_t nsteps,
                 double *force_out) {
  const lattice_t L = get_lattice(lattice);
  const int Q = L.nQ, D = L.nD, nF = 2;
  bb_t bb = bb_setup(&L, bulk, n_bulk, nb, type);  /* built over all bulk nodes (:81) */
  const double nu0Inv = 1.0 / (c2 * (tau0 - 0.5));
  const double nu1Inv = 1.0 / (c2 * (tau1 - 0.5));
  double F[3] = {Fin[0], Fin[1], D == 3 ? Fin[2] : 0.0};
  const double momx = Fin[0];
  const double q0Node = 0.0, q1Node = 0.0;  /* Q(0,n) = Q(1,n) = 0 (:132-136) */
  double *a = f, *b = ftmp;
  for (int32_t it = 0; it < nsteps; ++it) {
    /* pass 1 */
    for (int32_t k = 0; k < n_bulk; ++k) {
      const int32_t n = bulk[k];
      const double r0 = rho[(int64_t)n * 2 + 0] = qsum(&L, a + (int64_t)n * nF * Q);
      const double r1 = rho[(int64_t)n * 2 + 1] = qsum(&L, a + (int64_t)n * nF * Q + Q);
      cg[n] = (r0 - r1) / (r0 + r1);
    }
    for (int32_t n = 1; n < n_nodes; ++n)
      if (type[n] == 1) {
        const double r0 = rho[(int64_t)n * 2 + 0], r1 = rho[(int64_t)n * 2 + 1];
        cg[n] = (r0 - r1) / (r0 + r1);
      }
    /* pass 2 */
    if (force_mode == 1) {
      double meanfcX = 0.0;
      for (int32_t k = 0; k < n_bulk; ++k) {
        const int32_t n = bulk[k];
        double ft[19], sfc[3];
        for (int q = 0; q < Q; ++q) ft[q] = a[(int64_t)n * nF * Q + q] + a[(int64_t)n * nF * Q + Q + q];
        qsumc(&L, ft, sfc);
        meanfcX += sfc[0];
      }
      double meanG = meanfcX;
      meanG /= (double)n_global;
      F[0] = 2 * (momx - meanG);
    }
    /* pass 3 */
    double cF[19];
    cdotall(&L, F, cF);
    for (int32_t k = 0; k < n_bulk; ++k) {
      const int32_t n = bulk[k];
      const double *f0 = a + (int64_t)n * nF * Q, *f1 = f0 + Q;
      double ft[19], j[3], u[3], cu[19];
      for (int q = 0; q < Q; ++q) ft[q] = f0[q] + f1[q];
      double r0 = rho[(int64_t)n * 2 + 0];
      double r1 = rho[(int64_t)n * 2 + 1];
      const double rhoN = r0 + r1;
      qsumc(&L, ft, j);
      for (int d = 0; d < D; ++d) u[d] = (j[d] + 0.5 * F[d]) / rhoN;
      for (int d = 0; d < D; ++d) vel[(int64_t)n * D + d] = u[d];
      rho[(int64_t)n * 2 + 0] = r0 += 0.5 * q0Node;
      rho[(int64_t)n * 2 + 1] = r1 += 0.5 * q1Node;
      const double tau = c2Inv * rhoN / (r0 * nu0Inv + r1 * nu1Inv) + 0.5;
      const double uu = dot(&L, u, u);
      cdotall(&L, u, cu);
      const double tau_inv = 1.0 / tau;
      const double tau_factor = (1 - 0.5 / tau);
      const double uF = dot(&L, u, F);
      /* grad of cg over the neighbours */
      double s[19], g[3], cg_n[19];
      for (int q = 0; q < Q; ++q) s[q] = cg[nb[(int64_t)n * Q + q]];
      lattice_grad(&L, s, g);
      const double CGNorm = sqrt(dot(&L, g, g));
      const double inv = 1.0 / (CGNorm + (CGNorm < DBL_EPS_LB));
      for (int d = 0; d < D; ++d) g[d] *= inv;
      cdotall(&L, g, cg_n);
      const double AF0_5 = 1.125 * CGNorm * sigma / tau;
      const double rhoFacBeta = beta * r0 * r1 / rhoN;
      const double c0 = (r0 / rhoN), c1 = (r1 / rhoN);
      for (int q = 0; q < Q; ++q) {
        const double om = -tau_inv * (ft[q] - rhoN * L.w[q] * (1.0 + c2Inv * cu[q] + c4Inv0_5 * (cu[q] * cu[q] - c2 * uu)));
        const double dF = L.w[q] * tau_factor * (c2Inv * cF[q] + c4Inv * (cF[q] * cu[q] - c2 * uF));
        const double dQ0 = tau_factor * q0Node * L.w[q] * (1.0 + c2Inv * cu[q] + c4Inv0_5 * (cu[q] * cu[q] - c2 * uu));
        const double dQ1 = tau_factor * q1Node * L.w[q] * (1.0 + c2Inv * cu[q] + c4Inv0_5 * (cu[q] * cu[q] - c2 * uu));
        double st, rc;
        if (q < Q - 1) {
          st = AF0_5 * (L.w[q] * cg_n[q] * cg_n[q] - L.B[q]);
          rc = rhoFacBeta * L.w[q] * cg_n[q] / L.cNorm[q];
        } else {
          st = -AF0_5 * L.B[q];
          rc = 0.0;
        }
        const int32_t to = nb[(int64_t)n * Q + q];
        b[(int64_t)to * nF * Q + q] = c0 * (ft[q] + om + dF + st) + rc + dQ0;
        b[(int64_t)to * nF * Q + Q + q] = c1 * (ft[q] + om + dF + st) - rc + dQ1;
      }
    }
    double *t = a; a = b; b = t;
    bb_apply(&L, &bb, a, nF, 0, nb);
    bb_apply(&L, &bb, a, nF, 1, nb);
  }
  if (a != f) memcpy(f, a, sizeof(double) * (size_t)n_nodes * nF * Q);
  if (force_out) for (int d = 0; d < D; ++d) force_out[d] = F[d];
  bb_free(&bb);
  return 0;
}

/* ------------------------------------------------------------ equilibrium init */
/* initiateLbField (LBinitiatefield.h:33-57, the overload with separate field
 * numbers): f(lb_field, q, n) = w_q rho(rho_field, n) (1 + c2Inv cu_q +
 * c4Inv0_5 (cu_q^2 - c2 uu)) on the bulk nodes, with cu = cDotAll(u) and
 * uu = dot(u, u) of velocity field vel_field.  f: LbField with n_lb fields;
 * rho: ScalarField with n_rho fields; vel: VectorField with n_vel fields. */
int orc_init_eq(int lattice, const int32_t *bulk, int32_t n_bulk, double *f, int n_lb, int lb_field,
                const double *rho, int n_rho, int rho_field, const double *vel, int n_vel, int vel_field) {
  const lattice_t L = get_lattice(lattice);
  const int Q = L.nQ, D = L.nD;
  for (int32_t k = 0; k < n_bulk; ++k) {
    const int64_t n = bulk[k];
    const double *u = vel + (n * n_vel + vel_field) * D;
    double cu[19];
    cdotall(&L, u, cu);
    const double uu = dot(&L, u, u);
    const double r = rho[n * n_rho + rho_field];
    double *fn = f + (n * n_lb + lb_field) * Q;
    for (int q = 0; q < Q; ++q) fn[q] = L.w[q] * r * (1.0 + c2Inv * cu[q] + c4Inv0_5 * (cu[q] * cu[q] - c2 * uu));
  }
  return 0;
}

/* ------------------------------------------------------------ flux sums */
/* LBglobalforcing.h:23-26 (sum of f c over bulk, field sum for nF = 2) and
 * :61-78 (phi-weighted sums of calcCapNumbForceCartDir, with f the total
 * f0 + f1 and phi_k = rho_k / rho from the two fields); sequential in bulk
 * order.  out[3]. */
int orc_flux(int lattice, int32_t n_nodes, const int32_t *bulk, int32_t n_bulk, const double *f,
             int nF, int dir, double *out) {
  const lattice_t L = get_lattice(lattice);
  const int Q = L.nQ;
  (void)n_nodes;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int32_t k = 0; k < n_bulk; ++k) {
    const int32_t n = bulk[k];
    const double *fn = f + (int64_t)n * nF * Q;
    double ft[19], fc[3];
    for (int q = 0; q < Q; ++q) ft[q] = nF == 2 ? fn[q] + fn[Q + q] : fn[q];
    qsumc(&L, ft, fc);
    s0 += fc[dir];
    if (nF == 2) {
      const double r0 = qsum(&L, fn), r1 = qsum(&L, fn + Q);
      const double rt = r0 + r1;
      s1 += (r0 / rt) * fc[dir];
      s2 += (r1 / rt) * fc[dir];
    }
  }
  out[0] = s0; out[1] = s1; out[2] = s2;
  return 0;
}

/* ------------------------------------------------------------ node types */
/* Nodes::setupNodeType (LBnodes.h:156-209); attr = 'nodetype' attribute
 * (0 solid, nonzero fluid), rank[] per node, my_rank. */
int orc_node_types(int lattice, int32_t n_nodes, const int32_t *nb, const int32_t *attr,
                   const int32_t *rank, int32_t my_rank, int16_t *type) {
  const lattice_t L = get_lattice(lattice);
  const int Q = L.nQ;
  type[0] = -1;
  for (int32_t n = 1; n < n_nodes; ++n) type[n] = attr[n] == 0 ? 0 : 3;
  for (int32_t n = 1; n < n_nodes; ++n) {
    if (type[n] < 2) {
      int fl = 0;
      for (int q = 0; q < Q; ++q) fl |= type[nb[(int64_t)n * Q + q]] > 1;
      type[n] = fl ? 1 : 0;
    } else {
      if (rank[n] != my_rank) {
        type[n] = 4;
      } else {
        int so = 0;
        for (int q = 0; q < Q; ++q) so |= type[nb[(int64_t)n * Q + q]] < 2;
        type[n] = so ? 2 : 3;
      }
    }
  }
  return 0;
}
