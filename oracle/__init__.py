"""Oracle: CPU restatement of the reference's hot path (test infrastructure only).

Only tests/, tests/golden/ scripts, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this package, and only as the checker.

  oracle.vtklb_oracle   numpy restatement of PythonScripts/vtklb.py (node list)
  oracle.oracle.c       C restatement of the std_case / twophase loops
  oracle.run_*          ctypes wrappers of liboracle.so (built by oracle/Makefile)
  oracle.ref_run        timing of the compiled reference (oracle/_ref) under MPI
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None

Q_OF = {0: 9, 1: 19}
ND_OF = {0: 2, 1: 3}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_std_case.restype = C.c_int
        L.orc_std_case.argtypes = [C.c_int, C.c_int32, _i32p, _i16p, _i32p, C.c_int32, _f64p, _f64p,
                                   C.c_double, _f64p, C.c_int32, _f64p, _f64p, C.c_int]
        L.orc_std_case_trt.restype = C.c_int
        L.orc_std_case_trt.argtypes = [C.c_int, C.c_int32, _i32p, _i16p, _i32p, C.c_int32, _f64p, _f64p,
                                       C.c_double, C.c_double, _f64p, C.c_int32, _f64p, _f64p, C.c_int]
        L.orc_one_phase.restype = C.c_int
        L.orc_one_phase.argtypes = [C.c_int, C.c_int32, _i32p, _i32p, C.c_int32, _i32p, _i32p, _f64p, _f64p,
                                    _f64p, _f64p, C.c_double, _f64p, C.c_double, C.c_int32, _f64p, _f64p, _f64p]
        L.orc_rheology.restype = C.c_int
        L.orc_rheology.argtypes = [C.c_int, C.c_int32, _i32p, _i16p, _i32p, C.c_int32, _f64p, _f64p, C.c_int32,
                                   _f64p, _f64p, _f64p, C.c_int32, _f64p, _f64p, _f64p, C.c_int]
        L.orc_twophase.restype = C.c_int
        L.orc_twophase.argtypes = [C.c_int, C.c_int32, _i32p, _i16p, _i32p, C.c_int32, _f64p, _f64p,
                                   _f64p, _f64p, _f64p, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_int, _f64p, C.c_int64, C.c_int32, _f64p]
        L.orc_flux.restype = C.c_int
        L.orc_flux.argtypes = [C.c_int, C.c_int32, _i32p, C.c_int32, _f64p, C.c_int, C.c_int, _f64p]
        L.orc_init_eq.restype = C.c_int
        L.orc_init_eq.argtypes = [C.c_int, _i32p, C.c_int32, _f64p, C.c_int, C.c_int, _f64p, C.c_int, C.c_int,
                                  _f64p, C.c_int, C.c_int]
        L.orc_node_types.restype = C.c_int
        L.orc_node_types.argtypes = [C.c_int, C.c_int32, _i32p, _i32p, _i32p, C.c_int32, _i16p]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def run_std_case(lattice, neighbor, node_type, bulk, f, tau, force, nsteps, nthreads=1):
    """nsteps iterations of the std_case body.  f: LbField (n_nodes, Q), updated
    copy returned with the rho / vel of the last iteration.  tau: the BGK
    relaxation time, or a pair (tau_sym, tau_anti) for the two-relaxation-time
    operator (calcOmegaBGKTRT / calcDeltaOmegaFTRT with phi = 1)."""
    n_nodes = node_type.shape[0]
    Q, D = Q_OF[lattice], ND_OF[lattice]
    f = _c(f, np.float64).reshape(-1).copy()
    ftmp = f.copy()
    rho = np.zeros(n_nodes)
    vel = np.zeros(n_nodes * D)
    F = np.zeros(3)
    F[:D] = force[:D]
    bulk = _c(bulk, np.int32)
    args = (lattice, n_nodes, _c(neighbor, np.int32).ravel(), _c(node_type, np.int16),
            bulk if bulk.size else np.zeros(1, np.int32), bulk.size, f, ftmp)
    tail = (F, int(nsteps), rho, vel, int(nthreads))
    if np.ndim(tau):
        lib().orc_std_case_trt(*args, float(tau[0]), float(tau[1]), *tail)
    else:
        lib().orc_std_case(*args, float(tau), *tail)
    return f.reshape(n_nodes, Q), rho, vel.reshape(n_nodes, D)


def run_one_phase(lattice, neighbor, bulk, tags, label, force_on, normals, f, tau, force, rho_w, nsteps):
    """nsteps iterations of std_one_phase/main.cpp (one rank): tagged-link
    bounce-back, anti-bounce-back and fluid-fluid swaps, per-domain mass
    sources.  tags, label: int per node (tags[0] = -1 as Nodes sets it);
    force_on: per node; normals: (n_nodes, 3).  Returns f, rho, vel and the
    mass fluxes (q1, q2) of the last iteration."""
    n_nodes = len(tags)
    Q, D = Q_OF[lattice], ND_OF[lattice]
    f = _c(f, np.float64).reshape(-1).copy()
    ftmp = np.zeros_like(f)
    rho = np.zeros(n_nodes)
    vel = np.zeros(n_nodes * D)
    flux = np.zeros(2)
    F = np.zeros(3)
    F[:D] = force[:D]
    bulk = _c(bulk, np.int32)
    rc = lib().orc_one_phase(lattice, n_nodes, _c(neighbor, np.int32).ravel(), bulk if bulk.size else
                             np.zeros(1, np.int32), bulk.size, _c(tags, np.int32), _c(label, np.int32),
                             _c(force_on, np.float64), _c(normals, np.float64).ravel(), f, ftmp, float(tau), F,
                             float(rho_w), int(nsteps), rho, vel, flux)
    if rc != 0:
        raise ValueError("orc_one_phase: %s" % ("an interior domain has no interior nodes" if rc == -1
                                                else "a pressure-boundary node is not fluid"))
    return f.reshape(n_nodes, Q), rho, vel.reshape(n_nodes, D), flux


def run_rheology(lattice, neighbor, node_type, bulk, f, table, force, nsteps, nthreads=1):
    """nsteps iterations of the generalized-Newtonian loop
    (polymer_viscosity_tabulation/main.cpp).  table: (m, 2) rows of
    (strain-rate square, viscosity).  Returns f, rho, vel and the node
    viscosities of the last iteration."""
    n_nodes = node_type.shape[0]
    Q, D = Q_OF[lattice], ND_OF[lattice]
    f = _c(f, np.float64).reshape(-1).copy()
    ftmp = f.copy()
    rho, visc = np.zeros(n_nodes), np.zeros(n_nodes)
    vel = np.zeros(n_nodes * D)
    F = np.zeros(3)
    F[:D] = force[:D]
    bulk = _c(bulk, np.int32)
    table = _c(table, np.float64)
    lib().orc_rheology(lattice, n_nodes, _c(neighbor, np.int32).ravel(), _c(node_type, np.int16),
                       bulk if bulk.size else np.zeros(1, np.int32), bulk.size, f, ftmp, table.shape[0],
                       _c(table[:, 0], np.float64), _c(table[:, 1], np.float64), F, int(nsteps), rho, vel, visc,
                       int(nthreads))
    return f.reshape(n_nodes, Q), rho, vel.reshape(n_nodes, D), visc


def run_twophase(lattice, neighbor, node_type, bulk, f, rho, tau0, tau1, sigma, beta,
                 force_mode, force, nsteps, n_global=None):
    """nsteps iterations of the twophase body.  f: LbField (n_nodes, 2, Q);
    rho: ScalarField (n_nodes, 2) carrying the wall densities.  Returns
    (f, rho, vel, cg, force_of_last_iteration)."""
    n_nodes = node_type.shape[0]
    Q, D = Q_OF[lattice], ND_OF[lattice]
    f = _c(f, np.float64).reshape(-1).copy()
    ftmp = f.copy()
    rho = _c(rho, np.float64).reshape(-1).copy()
    vel = np.zeros(n_nodes * D)
    cg = np.zeros(n_nodes)
    F = np.zeros(3)
    F[: len(force)] = force
    Fo = np.zeros(3)
    bulk = _c(bulk, np.int32)
    lib().orc_twophase(lattice, n_nodes, _c(neighbor, np.int32).ravel(), _c(node_type, np.int16),
                       bulk, bulk.size, f, ftmp, rho, vel, cg, float(tau0), float(tau1),
                       float(sigma), float(beta), int(force_mode), F,
                       int(n_global if n_global is not None else bulk.size), int(nsteps), Fo)
    return (f.reshape(n_nodes, 2, Q), rho.reshape(n_nodes, 2), vel.reshape(n_nodes, D), cg,
            Fo[:D])


def flux(lattice, bulk, f, n_fields, direction):
    f = _c(f, np.float64).ravel()
    out = np.zeros(3)
    bulk = _c(bulk, np.int32)
    n_nodes = f.size // (n_fields * Q_OF[lattice])
    lib().orc_flux(lattice, n_nodes, bulk, bulk.size, f, n_fields, direction, out)
    return out


def init_equilibrium(lattice, bulk, f, lb_field, rho, rho_field, vel, vel_field):
    """initiateLbField (LBinitiatefield.h:33-57) in place on f (n_nodes,
    n_lb, Q); rho (n_nodes, n_rho), vel (n_nodes, n_vel, nD)."""
    assert f.dtype == np.float64 and f.flags.c_contiguous
    rho = _c(rho, np.float64)
    vel = _c(vel, np.float64)
    n_nodes = f.shape[0]
    bulk = _c(bulk, np.int32)
    lib().orc_init_eq(lattice, bulk, bulk.size, f.reshape(-1), f.shape[1], int(lb_field), rho.reshape(-1),
                      rho.size // n_nodes, int(rho_field), vel.reshape(-1), vel.size // (n_nodes * ND_OF[lattice]),
                      int(vel_field))
    return f


def node_types(lattice, neighbor, nodetype_attr, ranks, my_rank):
    n_nodes = nodetype_attr.shape[0]
    out = np.zeros(n_nodes, dtype=np.int16)
    lib().orc_node_types(lattice, n_nodes, _c(neighbor, np.int32).ravel(), _c(nodetype_attr, np.int32),
                         _c(ranks, np.int32), int(my_rank), out)
    return out


def bulk_nodes(node_type, ranks, my_rank):
    """findBulkNodes (LBgeometry.h:11-21): fluid nodes of this rank, ascending."""
    n = np.arange(node_type.shape[0])
    return n[(n >= 1) & (node_type > 1) & (ranks == my_rank)].astype(np.int32)


def rest_state(lattice, n_nodes, bulk, rho):
    """f = w rho on bulk nodes (std_case/main.cpp:92-96), zeros elsewhere."""
    w = WEIGHTS[lattice]
    f = np.zeros((n_nodes, Q_OF[lattice]))
    f[bulk] = w[None, :] * np.asarray(rho)[bulk, None]
    return f


WEIGHTS = {
    0: np.array([4.0 / 36.0, 1.0 / 36.0] * 4 + [16.0 / 36.0]),
    1: np.array([2.0 / 36.0] * 3 + [1.0 / 36.0] * 6 + [2.0 / 36.0] * 3 + [1.0 / 36.0] * 6
                + [12.0 / 36.0]),
}
