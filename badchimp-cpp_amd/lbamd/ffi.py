"""ctypes binding of the C ABI in include/badchimp_amd.h.

The shared library is badchimp-cpp_amd/lib/libbadchimp_amd.so, built by
badchimp-cpp_amd/build.py.  Loading fails loudly when it is missing: there
is no fallback implementation.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LBAMD_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("LBAMD_LIB") or os.path.join(os.path.dirname(HERE), "lib", "libbadchimp_amd.so")

D2Q9, D3Q19 = 0, 1
FP64, FP32S, FP16S = 0, 1, 2
RHO, PHI, VISC = 0, 1, 2
VEL = 0
FORCE_FIXED, FORCE_MOMX = 0, 1
TOPO_ALLOW_GHOST0 = 1
TOPO_TABLE_ONLY = 2
Q_OF = {D2Q9: 9, D3Q19: 19}
ND_OF = {D2Q9: 2, D3Q19: 3}
LATTICE_NAME = {"D2Q9": D2Q9, "D3Q19": D3Q19}
PERIODIC_BITS = {"x": 1, "y": 2, "z": 4}

_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class LBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("[%d] %s" % (code, msg))
        self.code = code


SIGNATURES = {
    "lb_last_error": (C.c_char_p, [_vp]),
    "lb_global_error": (C.c_char_p, []),
    "lb_version": (C.c_char_p, []),
    "lb_ctx_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "lb_ctx_destroy": (C.c_int, [_vp]),
    "lb_set_topology": (C.c_int, [_vp, C.c_int32, _i32p, _i16p, _i32p, C.c_int32, C.c_int32]),
    "lb_set_topology_grid": (C.c_int, [_vp, C.c_int32, _i32p, _i16p, _i32p, C.c_int32, _i32p,
                                       C.c_int32]),
    "lb_layout": (C.c_int32, [_vp]),
    "lb_layout_note": (C.c_char_p, [_vp]),
    "lb_set_global_bulk_count": (C.c_int, [_vp, C.c_int64]),
    "lb_halo_add": (C.c_int, [_vp, C.c_int32, C.c_int32, _i32p, _i32p, _i32p, C.c_int32, _i32p, _i32p, _i32p,
                              C.POINTER(C.c_int32)]),
    "lb_halo_sizes": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int64)]),
    "lb_halo_pack": (C.c_int, [_vp, C.c_int32, C.c_int32, _f64p]),
    "lb_halo_unpack": (C.c_int, [_vp, C.c_int32, C.c_int32, _f64p]),
    "lb_halo_pack_colour": (C.c_int, [_vp, C.c_int32, _f64p]),
    "lb_halo_unpack_colour": (C.c_int, [_vp, C.c_int32, _f64p]),
    "lb_halo_count": (C.c_int32, [_vp]),
    "lb_halo_num_nodes": (C.c_int32, [_vp]),
    "lb_upload_f": (C.c_int, [_vp, C.c_int32, _f64p]),
    "lb_download_f": (C.c_int, [_vp, C.c_int32, _f64p]),
    "lb_download_f_noalloc": (C.c_int, [_vp, C.c_int32, _f64p]),
    "lb_init_equilibrium": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _f64p, _f64p]),
    "lb_download_scalar": (C.c_int, [_vp, C.c_int32, C.c_int32, _f64p]),
    "lb_download_vector": (C.c_int, [_vp, C.c_int32, C.c_int32, _f64p]),
    "lb_set_wall_densities": (C.c_int, [_vp, _f64p]),
    "lb_step_bgk": (C.c_int, [_vp, C.c_int32, C.c_double, _f64p]),
    "lb_step_trt": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double, _f64p]),
    "lb_set_viscosity_table": (C.c_int, [_vp, C.c_int32, _f64p, _f64p]),
    "lb_set_node_tags": (C.c_int, [_vp, C.c_int32, _i32p, _i32p, _i32p, _f64p]),
    "lb_step_tagged": (C.c_int, [_vp, C.c_int32, C.c_double, _f64p, C.c_double]),
    "lb_pressure_flux": (C.c_int, [_vp, _f64p]),
    "lb_step_rheology": (C.c_int, [_vp, C.c_int32, _f64p]),
    "lb_step_cg": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                             C.c_int32, _f64p]),
    "lb_step_cg_colour": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "lb_step_cg_collide": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32, _f64p,
                                     C.c_double]),
    "lb_last_force": (C.c_int, [_vp, _f64p]),
    "lb_reduce_flux": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    "lb_solve_permeability": (C.c_int, [_vp, C.c_double, _f64p, C.c_int32, C.c_double, C.c_int64,
                                        C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "lb_solve_relperm": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, _f64p, C.c_int32,
                                   C.c_double, C.c_int64, C.c_double, _f64p, C.POINTER(C.c_int64)]),
    "lb_solve_permeability_trt": (C.c_int, [_vp, C.c_double, C.c_double, _f64p, C.c_int32, C.c_double,
                                            C.c_int64, C.c_double, C.POINTER(C.c_double),
                                            C.POINTER(C.c_int64)]),
    "lb_sync": (C.c_int, [_vp]),
    "lb_kernel_time": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "lb_kernel_time_reset": (C.c_int, [_vp]),
    "lb_device_bytes": (C.c_int64, [_vp]),
    "lb_bench_copy": (C.c_int, [C.c_int, C.c_int64, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "lb_bench_stream": (C.c_int, [C.c_int, C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]),
    "lb_topo_read_vtklb": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_vp)]),
    "lb_topo_from_geometry": (C.c_int, [C.c_int, C.c_int32, _i32p, _i32p, C.c_int32, C.c_int32,
                                        C.POINTER(_vp)]),
    "lb_topo_from_geometry_device": (C.c_int, [C.c_int, C.c_int, C.c_int32, _i32p, _i32p, C.c_int32,
                                               C.POINTER(_vp)]),
    "lb_topo_write_binary": (C.c_int, [_vp, C.c_char_p]),
    "lb_topo_read_binary": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_vp)]),
    "lb_topo_destroy": (C.c_int, [_vp]),
    "lb_vtu_create": (C.c_int, [C.c_int32, _i32p, C.c_int32, C.c_char_p, C.c_int32, C.c_int32, C.c_int32,
                                C.POINTER(_vp)]),
    "lb_vtu_add_file": (C.c_int, [_vp, C.c_char_p]),
    "lb_vtu_add_variable": (C.c_int, [_vp, C.c_char_p, C.c_int32, _vp]),
    "lb_vtu_write": (C.c_int, [_vp, C.c_double]),
    "lb_vtu_num_points": (C.c_int32, [_vp]),
    "lb_vtu_destroy": (C.c_int, [_vp]),
    "lb_topo_num_nodes": (C.c_int32, [_vp]),
    "lb_topo_num_bulk": (C.c_int32, [_vp]),
    "lb_topo_num_dims": (C.c_int32, [_vp]),
    "lb_topo_rank": (C.c_int32, [_vp]),
    "lb_topo_global_dims": (C.c_int, [_vp, _i32p]),
    "lb_topo_neighbors": (C.c_int, [_vp, _i32p]),
    "lb_topo_positions": (C.c_int, [_vp, _i32p]),
    "lb_topo_node_types": (C.c_int, [_vp, _i16p]),
    "lb_topo_node_ranks": (C.c_int, [_vp, _i32p]),
    "lb_topo_num_neighbor_ranks": (C.c_int32, [_vp]),
    "lb_topo_neighbor_rank": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _vp, _vp]),
    "lb_topo_bulk": (C.c_int, [_vp, _i32p]),
    "lb_topo_attribute": (C.c_int, [_vp, C.c_char_p, C.c_int32, _f64p]),
    "lb_topo_write_vtklb": (C.c_int, [_vp, C.c_char_p, C.c_int32, C.POINTER(C.c_char_p),
                                      C.POINTER(C.POINTER(C.c_double))]),
    "lb_sphere_pack": (C.c_int, [C.c_int32, _i32p, C.c_double, C.c_double, C.c_uint64, C.c_int32,
                                 C.c_int32, C.c_int32, C.c_double, _i32p, _f64p]),
    "lb_write_lbfield": (C.c_int, [_vp, C.c_char_p]),
    "lb_read_lbfield": (C.c_int, [_vp, C.c_char_p]),
    "lb_write_scalar": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_char_p]),
    "lb_write_vector": (C.c_int, [_vp, C.c_int32, C.c_char_p]),
    "lb_batch_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int32, C.POINTER(_vp)]),
    "lb_batch_destroy": (C.c_int, [_vp]),
    "lb_batch_last_error": (C.c_char_p, [_vp]),
    "lb_batch_footprint": (C.c_int, [C.c_int, C.c_int, C.c_int32, C.c_int32, _i32p, _i32p, C.c_int32,
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "lb_batch_add_many": (C.c_int, [_vp, C.c_int32, C.c_int32, _i32p, C.c_void_p, C.c_int32, C.c_double, _f64p,
                                    C.c_double, C.c_int64, _i64p, C.POINTER(C.c_int32)]),
    "lb_batch_add_twophase": (C.c_int, [_vp, C.c_int32, _i32p, _i32p, C.c_int32, _f64p, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, _f64p, C.c_double, C.c_int64,
                                        C.c_int64]),
    "lb_batch_add": (C.c_int, [_vp, C.c_int32, _i32p, _i32p, C.c_int32, C.c_double, _f64p, C.c_double,
                               C.c_int64, C.c_int64]),
    "lb_batch_advance": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_int32)]),
    "lb_batch_launch": (C.c_int, [_vp, C.c_int32]),
    "lb_batch_wait": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "lb_batch_collect": (C.c_int, [_vp, C.c_int32, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                   _f64p, _f64p, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                   _f64p, _i32p, C.POINTER(C.c_int32)]),
    "lb_batch_status": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "lb_batch_set_precision": (C.c_int, [_vp, C.c_int32, C.c_int32]),
    "lb_batch_set_shape": (C.c_int, [_vp, C.c_int32]),
    "lb_batch_shape": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "lb_batch_tuning": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "lb_batch_kernel_time": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
}


def lib():
    """Load the native library (once)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("native library missing: %s (run badchimp-cpp_amd/build.py)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check_global(rc):
    if rc != 0:
        raise LBError(rc, lib().lb_global_error().decode())


def periodic_mask(periodic):
    m = 0
    for ch in periodic.lower():
        m |= PERIODIC_BITS.get(ch, 0)
    return m


def sphere_pack(shape, radius, porosity, seed, periodic="x", walls=True, max_fail=2000,
                relax=0.97):
    """Seeded sphere pack (see lb_sphere_pack).  Returns (geo int32 array, info dict)."""
    dims = np.ascontiguousarray(shape, dtype=np.int32)
    geo = np.empty(int(np.prod(dims)), dtype=np.int32)
    info = np.zeros(4)
    _check_global(lib().lb_sphere_pack(len(dims), dims, float(radius), float(porosity),
                                       int(seed) & 0xFFFFFFFFFFFFFFFF, periodic_mask(periodic),
                                       1 if walls else 0, int(max_fail), float(relax), geo, info))
    d = {"porosity": float(info[0]), "n_spheres": int(info[1]),
         "min_center_distance": float(info[2]), "trials": int(info[3]), "radius": float(radius),
         "overlap_allowed": bool(info[2] < 2 * radius - 1e-12)}
    return geo.reshape(tuple(int(x) for x in dims)), d


def bench_copy(nbytes, iters=20, device=0):
    """Plain device copy: (read+write GB/s, ms per copy)."""
    g, m = C.c_double(), C.c_double()
    _check_global(lib().lb_bench_copy(device, int(nbytes), int(iters), C.byref(g), C.byref(m)))
    return g.value, m.value


def bench_stream(n_nodes, lattice=D3Q19, iters=20, device=0):
    """Identity AoSoA streaming of float populations: (read+write GB/s, ms)."""
    g, m = C.c_double(), C.c_double()
    _check_global(lib().lb_bench_stream(device, lattice, int(n_nodes), int(iters), C.byref(g), C.byref(m)))
    return g.value, m.value


class Topology:
    """Node list of one rank (lb_topo)."""

    def __init__(self, handle, lattice):
        self._h = handle
        self.lattice = lattice
        self.Q = Q_OF[lattice]
        self.nD = ND_OF[lattice]

    @classmethod
    def from_vtklb(cls, path, lattice):
        h = _vp()
        _check_global(lib().lb_topo_read_vtklb(path.encode(), lattice, C.byref(h)))
        return cls(h, lattice)

    @classmethod
    def from_geometry(cls, geo, lattice, periodic="", rank=1):
        geo = np.ascontiguousarray(geo, dtype=np.int32)
        dims = np.ascontiguousarray(geo.shape, dtype=np.int32)
        h = _vp()
        _check_global(lib().lb_topo_from_geometry(lattice, geo.ndim, dims, geo.ravel(),
                                                  periodic_mask(periodic), rank, C.byref(h)))
        return cls(h, lattice)

    @classmethod
    def from_geometry_device(cls, geo, lattice, periodic="", device=0):
        """The node list of from_geometry (rank 1), built on the GPU."""
        geo = np.ascontiguousarray(geo, dtype=np.int32)
        dims = np.ascontiguousarray(geo.shape, dtype=np.int32)
        h = _vp()
        _check_global(lib().lb_topo_from_geometry_device(device, lattice, geo.ndim, dims, geo.ravel(),
                                                         periodic_mask(periodic), C.byref(h)))
        return cls(h, lattice)

    @classmethod
    def from_binary(cls, path, lattice):
        """Read the binary twin of a .vtklb file (write_binary)."""
        h = _vp()
        _check_global(lib().lb_topo_read_binary(os.fsencode(path), lattice, C.byref(h)))
        return cls(h, lattice)

    def write_binary(self, path):
        _check_global(lib().lb_topo_write_binary(self._h, os.fsencode(path)))

    def close(self):
        if self._h:
            lib().lb_topo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_nodes(self):
        return lib().lb_topo_num_nodes(self._h)

    @property
    def n_bulk(self):
        return lib().lb_topo_num_bulk(self._h)

    @property
    def rank(self):
        return lib().lb_topo_rank(self._h)

    def global_dims(self):
        out = np.zeros(self.nD, dtype=np.int32)
        _check_global(lib().lb_topo_global_dims(self._h, out))
        return out

    def neighbors(self):
        out = np.zeros(self.n_nodes * self.Q, dtype=np.int32)
        _check_global(lib().lb_topo_neighbors(self._h, out))
        return out.reshape(self.n_nodes, self.Q)

    def positions(self):
        out = np.zeros(self.n_nodes * self.nD, dtype=np.int32)
        _check_global(lib().lb_topo_positions(self._h, out))
        return out.reshape(self.n_nodes, self.nD)

    def node_types(self):
        out = np.zeros(self.n_nodes, dtype=np.int16)
        _check_global(lib().lb_topo_node_types(self._h, out))
        return out

    def node_ranks(self):
        out = np.zeros(self.n_nodes, dtype=np.int32)
        _check_global(lib().lb_topo_node_ranks(self._h, out))
        return out

    def neighbor_ranks(self):
        """PROCESSOR blocks: [(rank, local node numbers, node numbers on that
        rank)] (LBvtk::getNeigRank / getNeigNodesNo / getNeigNodesNeigNo)."""
        out = []
        for k in range(lib().lb_topo_num_neighbor_ranks(self._h)):
            r, n = C.c_int32(), C.c_int32()
            _check_global(lib().lb_topo_neighbor_rank(self._h, k, C.byref(r), C.byref(n), None, None))
            loc = np.zeros(max(n.value, 1), dtype=np.int32)
            rem = np.zeros(max(n.value, 1), dtype=np.int32)
            _check_global(lib().lb_topo_neighbor_rank(self._h, k, None, None, loc.ctypes.data_as(C.c_void_p),
                                                      rem.ctypes.data_as(C.c_void_p)))
            out.append((r.value, loc[: n.value], rem[: n.value]))
        return out

    def bulk(self):
        out = np.zeros(max(self.n_bulk, 1), dtype=np.int32)
        _check_global(lib().lb_topo_bulk(self._h, out))
        return out[: self.n_bulk]

    def attribute(self, name, as_float=False):
        out = np.zeros(self.n_nodes)
        _check_global(lib().lb_topo_attribute(self._h, name.encode(), 1 if as_float else 0, out))
        return out

    def write_vtklb(self, path, attributes=()):
        names = (C.c_char_p * max(len(attributes), 1))()
        vals = (C.POINTER(C.c_double) * max(len(attributes), 1))()
        keep = []
        for k, (name, v) in enumerate(attributes):
            arr = np.ascontiguousarray(v, dtype=np.float64)
            assert arr.shape == (self.n_nodes,)
            keep.append(arr)
            names[k] = name.encode()
            vals[k] = arr.ctypes.data_as(C.POINTER(C.c_double))
        _check_global(lib().lb_topo_write_vtklb(self._h, path.encode(), len(attributes), names, vals))


class Context:
    """One rank's subdomain on one GPU (lb_ctx)."""

    def __init__(self, lattice, precision=FP64, device=0):
        h = _vp()
        _check_global(lib().lb_ctx_create(lattice, precision, device, C.byref(h)))
        self._h = h
        self.lattice = lattice
        self.precision = precision
        self.Q = Q_OF[lattice]
        self.nD = ND_OF[lattice]
        self.n_nodes = 0
        self._halo_node_counts = []  # per exchange: (nodes to send, nodes received)

    def _check(self, rc):
        if rc != 0:
            raise LBError(rc, lib().lb_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().lb_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # topology
    def set_topology(self, neighbor, node_type, bulk, allow_ghost0=False):
        neighbor = np.ascontiguousarray(neighbor, dtype=np.int32)
        node_type = np.ascontiguousarray(node_type, dtype=np.int16)
        bulk = np.ascontiguousarray(bulk, dtype=np.int32)
        n_nodes = node_type.shape[0]
        assert neighbor.size == n_nodes * self.Q
        self._check(lib().lb_set_topology(self._h, n_nodes, neighbor.ravel(), node_type,
                                          bulk if bulk.size else np.zeros(1, np.int32),
                                          bulk.size, TOPO_ALLOW_GHOST0 if allow_ghost0 else 0))
        self.n_nodes = n_nodes
        self._halo_node_counts = []  # a new topology drops the exchanges

    def set_topology_grid(self, neighbor, node_type, bulk, pos, allow_ghost0=False, table_only=False):
        neighbor = np.ascontiguousarray(neighbor, dtype=np.int32)
        node_type = np.ascontiguousarray(node_type, dtype=np.int16)
        bulk = np.ascontiguousarray(bulk, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.int32)
        n_nodes = node_type.shape[0]
        assert neighbor.size == n_nodes * self.Q and pos.size == n_nodes * self.nD
        flags = (TOPO_ALLOW_GHOST0 if allow_ghost0 else 0) | (TOPO_TABLE_ONLY if table_only else 0)
        self._check(lib().lb_set_topology_grid(self._h, n_nodes, neighbor.ravel(), node_type,
                                               bulk if bulk.size else np.zeros(1, np.int32),
                                               bulk.size, pos.ravel(), flags))
        self.n_nodes = n_nodes
        self._halo_node_counts = []

    def set_topology_from(self, topo, allow_ghost0=False, table_only=False):
        self.set_topology_grid(topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions(),
                               allow_ghost0, table_only)

    @property
    def layout(self):
        return lib().lb_layout(self._h)

    @property
    def layout_note(self):
        return lib().lb_layout_note(self._h).decode()

    def set_global_bulk_count(self, n):
        self._check(lib().lb_set_global_bulk_count(self._h, int(n)))

    # halo exchange between ranks
    @property
    def halo_nodes(self):
        return lib().lb_halo_num_nodes(self._h)

    def halo_add(self, neighbor_rank, nodes_to_send, n_dir_send, dir_send, nodes_received, n_dir_recv, dir_recv):
        """Register the exchange with one neighbouring rank (the MonLatMpi
        lists); returns its index."""
        a = [np.ascontiguousarray(x, dtype=np.int32) for x in (nodes_to_send, n_dir_send, dir_send,
                                                                nodes_received, n_dir_recv, dir_recv)]
        a = [x if x.size else np.zeros(1, dtype=np.int32) for x in a]
        k = C.c_int32()
        self._check(lib().lb_halo_add(self._h, int(neighbor_rank), len(nodes_to_send), a[0], a[1], a[2],
                                      len(nodes_received), a[3], a[4], a[5], C.byref(k)))
        assert k.value == len(self._halo_node_counts)
        self._halo_node_counts.append((len(nodes_to_send), len(nodes_received)))
        return k.value

    def halo_sizes(self, k):
        r, ns, nr = C.c_int32(), C.c_int64(), C.c_int64()
        self._check(lib().lb_halo_sizes(self._h, int(k), C.byref(r), C.byref(ns), C.byref(nr)))
        return r.value, ns.value, nr.value

    def halo_pack(self, k, nf=1):
        """The values this rank sends to exchange k: array (nf, n_send)."""
        _, ns, _ = self.halo_sizes(k)
        out = np.zeros(max(nf * ns, 1))
        self._check(lib().lb_halo_pack(self._h, int(k), int(nf), out))
        return out[: nf * ns].reshape(nf, ns)

    def halo_unpack(self, k, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        nf = v.shape[0] if v.ndim == 2 else 1
        flat = v.ravel() if v.size else np.zeros(1)
        self._check(lib().lb_halo_unpack(self._h, int(k), int(nf), flat))

    def halo_pack_colour(self, k):
        """The colour of the nodes this rank sends to exchange k (one per node)."""
        if not 0 <= k < len(self._halo_node_counts):
            raise LBError(-1, "no halo exchange %d" % k)
        n = self._halo_node_counts[k][0]
        out = np.zeros(max(n, 1))
        self._check(lib().lb_halo_pack_colour(self._h, int(k), out))
        return out[:n]

    def halo_unpack_colour(self, k, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._check(lib().lb_halo_unpack_colour(self._h, int(k), v if v.size else np.zeros(1)))

    # fields
    def upload_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        nf = f.size // (self.n_nodes * self.Q)
        self._check(lib().lb_upload_f(self._h, nf, f.ravel()))

    def download_f(self, nf=1, out=None):
        if out is None:
            out = np.zeros(self.n_nodes * nf * self.Q)
        self._check(lib().lb_download_f(self._h, nf, out))
        return out.reshape(self.n_nodes, nf, self.Q)

    def download_f_noalloc(self, nf=1, out=None):
        """lb_download_f through the library's fixed staging memory (no device
        allocation; at most 128 MiB of FP64 populations)."""
        if out is None:
            out = np.zeros(self.n_nodes * nf * self.Q)
        self._check(lib().lb_download_f_noalloc(self._h, nf, out))
        return out.reshape(self.n_nodes, nf, self.Q)

    def init_equilibrium(self, n_fields, field, rho, vel, rho_field=0):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        nfr = rho.size // self.n_nodes
        self._check(lib().lb_init_equilibrium(self._h, n_fields, field, nfr, rho_field, rho.ravel(),
                                              vel.ravel()))

    def download_scalar(self, which=RHO, nf=1):
        out = np.zeros(self.n_nodes * nf)
        self._check(lib().lb_download_scalar(self._h, which, nf, out))
        return out.reshape(self.n_nodes, nf) if nf > 1 else out

    def download_vel(self):
        out = np.zeros(self.n_nodes * self.nD)
        self._check(lib().lb_download_vector(self._h, VEL, 1, out))
        return out.reshape(self.n_nodes, self.nD)

    def set_wall_densities(self, rho2):
        self._check(lib().lb_set_wall_densities(self._h, np.ascontiguousarray(rho2, dtype=np.float64).ravel()))

    # steps
    def step_bgk(self, nsteps, tau, force):
        F = np.zeros(3)
        F[: self.nD] = force[: self.nD]
        self._check(lib().lb_step_bgk(self._h, int(nsteps), float(tau), F))

    def step_trt(self, nsteps, tau_sym, tau_anti, force):
        """std_case steps with the two-relaxation-time operator
        (calcOmegaBGKTRT / calcDeltaOmegaFTRT, LBcollision.h:50-75, 214-233)."""
        F = np.zeros(3)
        F[: self.nD] = force[: self.nD]
        self._check(lib().lb_step_trt(self._h, int(nsteps), float(tau_sym), float(tau_anti), F))

    def set_viscosity_table(self, strain_rate_sq, viscosity):
        """GeneralizedNewtonian's table (LBrheology.h:358-388)."""
        sr = np.ascontiguousarray(strain_rate_sq, dtype=np.float64)
        nu = np.ascontiguousarray(viscosity, dtype=np.float64)
        if sr.shape != nu.shape or sr.ndim != 1:
            raise ValueError("strain-rate squares and viscosities must be 1-D arrays of one length")
        self._check(lib().lb_set_viscosity_table(self._h, sr.size, sr, nu))

    def step_rheology(self, nsteps, force):
        """std_case steps with the generalized-Newtonian relaxation time
        (polymer_viscosity_tabulation/main.cpp)."""
        F = np.zeros(3)
        F[: self.nD] = force[: self.nD]
        self._check(lib().lb_step_rheology(self._h, int(nsteps), F))

    def set_node_tags(self, neighbors, tags, interior_domains, force_on):
        """std_one_phase's node tags, interior domains and force attribute
        over all nodes (tags[0] = -1)."""
        nb = np.ascontiguousarray(neighbors, dtype=np.int32)
        self._check(lib().lb_set_node_tags(self._h, nb.shape[0], nb.ravel(), np.ascontiguousarray(tags, np.int32),
                                           np.ascontiguousarray(interior_domains, np.int32),
                                           np.ascontiguousarray(force_on, np.float64)))

    def step_tagged(self, nsteps, tau, force, rho_w=1.0):
        F = np.zeros(3)
        F[: self.nD] = force[: self.nD]
        self._check(lib().lb_step_tagged(self._h, int(nsteps), float(tau), F, float(rho_w)))

    def pressure_flux(self):
        out = np.zeros(2)
        self._check(lib().lb_pressure_flux(self._h, out))
        return out

    def step_cg(self, nsteps, tau0, tau1, sigma, beta, force_mode, force):
        F = np.zeros(3)
        F[: len(force)] = force
        self._check(lib().lb_step_cg(self._h, int(nsteps), float(tau0), float(tau1), float(sigma),
                                     float(beta), int(force_mode), F))

    def step_cg_colour(self, flux=False):
        """The colour pass of one colour-gradient iteration (lb_step_cg_colour);
        with flux=True returns this rank's sum of sum_q (f0 + f1) c_x over its
        bulk nodes."""
        s = C.c_double()
        self._check(lib().lb_step_cg_colour(self._h, C.byref(s) if flux else None))
        return s.value if flux else None

    def step_cg_collide(self, tau0, tau1, sigma, beta, force_mode, force, flux_x_total=0.0):
        """The collision pass of the iteration begun by step_cg_colour; in
        FORCE_MOMX mode flux_x_total is the sum of every rank's flux."""
        F = np.zeros(3)
        F[: len(force)] = force
        self._check(lib().lb_step_cg_collide(self._h, float(tau0), float(tau1), float(sigma), float(beta),
                                             int(force_mode), F, float(flux_x_total)))

    def last_force(self):
        out = np.zeros(3)
        self._check(lib().lb_last_force(self._h, out))
        return out[: self.nD]

    def reduce_flux(self, direction=0):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self._check(lib().lb_reduce_flux(self._h, direction, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def solve_permeability(self, tau, force, check_every, tol, max_steps, n_box):
        F = np.zeros(3)
        F[: self.nD] = force[: self.nD]
        k, s = C.c_double(), C.c_int64()
        self._check(lib().lb_solve_permeability(self._h, float(tau), F, int(check_every), float(tol),
                                                int(max_steps), float(n_box), C.byref(k), C.byref(s)))
        return k.value, s.value

    def solve_relperm(self, tau0, tau1, sigma, beta, force, check_every, tol, max_steps, n_box):
        """lb_solve_relperm: ((k0, k1), steps)."""
        F = np.zeros(3)
        F[: self.nD] = force[: self.nD]
        k, s = np.zeros(2), C.c_int64()
        self._check(lib().lb_solve_relperm(self._h, float(tau0), float(tau1), float(sigma), float(beta), F,
                                           int(check_every), float(tol), int(max_steps), float(n_box), k,
                                           C.byref(s)))
        return (float(k[0]), float(k[1])), s.value

    def solve_permeability_trt(self, tau_sym, tau_anti, force, check_every, tol, max_steps, n_box):
        F = np.zeros(3)
        F[: self.nD] = force[: self.nD]
        k, s = C.c_double(), C.c_int64()
        self._check(lib().lb_solve_permeability_trt(self._h, float(tau_sym), float(tau_anti), F,
                                                    int(check_every), float(tol), int(max_steps), float(n_box),
                                                    C.byref(k), C.byref(s)))
        return k.value, s.value

    def write_lbfield(self, prefix):
        """Population checkpoint in the reference's .lblbf format."""
        self._check(lib().lb_write_lbfield(self._h, os.fsencode(prefix)))

    def read_lbfield(self, prefix):
        self._check(lib().lb_read_lbfield(self._h, os.fsencode(prefix)))

    def write_scalar(self, prefix, which=RHO, nf=1):
        self._check(lib().lb_write_scalar(self._h, which, nf, os.fsencode(prefix)))

    def write_vector(self, prefix, which=VEL):
        self._check(lib().lb_write_vector(self._h, which, os.fsencode(prefix)))

    def sync(self):
        self._check(lib().lb_sync(self._h))

    def kernel_time(self):
        ms, n = C.c_double(), C.c_int64()
        self._check(lib().lb_kernel_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def kernel_time_reset(self):
        self._check(lib().lb_kernel_time_reset(self._h))

    def device_bytes(self):
        return lib().lb_device_bytes(self._h)


VTU_DTYPES = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.int32): 2}


class VtuOutput:
    """VTU/PVTU writer in the reference's format (io/Output.h, io/VTK.h).

    pos: (n_cells, nd) cell positions (Grid::pos of the output nodes).
    Variables are arrays of `dtype` with shape (n_cells,) or (n_cells, nd);
    the writer reads them at every write(), so update them in place."""

    def __init__(self, pos, directory, rank=0, nprocs=1, dtype=np.float64):
        pos = np.ascontiguousarray(pos, dtype=np.int32)
        self.n_cells, self.nd = pos.shape
        self.dtype = np.dtype(dtype)
        self._vars = []
        self._h = _vp()
        _check_global(lib().lb_vtu_create(self.nd, pos.ravel() if pos.size else np.zeros(1, np.int32),
                                          self.n_cells, os.fsencode(directory), rank, nprocs,
                                          VTU_DTYPES[self.dtype], C.byref(self._h)))

    def add_file(self, name):
        _check_global(lib().lb_vtu_add_file(self._h, name.encode()))

    def add_variable(self, name, array):
        if array.dtype != self.dtype or not array.flags.c_contiguous or array.shape[0] != self.n_cells:
            raise ValueError("variable %r must be a C-contiguous %s array over %d cells" % (name, self.dtype,
                                                                                          self.n_cells))
        dim = 1 if array.ndim == 1 else array.shape[1]
        self._vars.append(array)  # the writer reads it at every write
        _check_global(lib().lb_vtu_add_variable(self._h, name.encode(), dim, array.ctypes.data_as(_vp)))

    def write(self, time=0.0):
        _check_global(lib().lb_vtu_write(self._h, float(time)))

    @property
    def n_points(self):
        return lib().lb_vtu_num_points(self._h)

    def close(self):
        if self._h:
            lib().lb_vtu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_footprint(lattice, precision, geo, periodic="x", n_fields=1):
    """(bytes, fluid nodes) a solve of this geometry takes in a Batch pool."""
    g = np.ascontiguousarray(geo, dtype=np.int32)
    dims = np.ascontiguousarray(g.shape, dtype=np.int32)
    nbytes, nf = C.c_int64(), C.c_int64()
    _check_global(lib().lb_batch_footprint(lattice, precision, int(n_fields), len(dims), dims, g.ravel(),
                                           periodic_mask(periodic), C.byref(nbytes), C.byref(nf)))
    return nbytes.value, nf.value


class Batch:
    """Many permeability solves resident in one device pool (lb_batch)."""

    def __init__(self, lattice, precision=FP32S, pool_bytes=1 << 30, check_every=100, device=0):
        h = _vp()
        _check_global(lib().lb_batch_create(lattice, precision, device, int(pool_bytes), int(check_every),
                                            C.byref(h)))
        self._h = h
        self.lattice = lattice
        self.precision = precision
        self.prec = [precision, precision]  # per kind: single-phase, two-phase
        self.nD = ND_OF[lattice]
        self.check_every = int(check_every)

    def _check(self, rc):
        if rc != 0:
            raise LBError(rc, lib().lb_batch_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().lb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, geo, periodic, tau, force, tol, max_steps, tag):
        """Admit a solve; False when it does not fit into the pool now."""
        g = np.ascontiguousarray(geo, dtype=np.int32)
        dims = np.ascontiguousarray(g.shape, dtype=np.int32)
        F = np.zeros(3)
        F[: self.nD] = np.asarray(force, dtype=float)[: self.nD]
        rc = lib().lb_batch_add(self._h, len(dims), dims, g.ravel(), periodic_mask(periodic), float(tau), F,
                                float(tol), int(max_steps), int(tag))
        if rc == -7:
            return False
        self._check(rc)
        return True

    def add_many(self, geos, periodic, tau, force, tol, max_steps, tags):
        """Admit many solves (row layouts built on host threads); returns how
        many, in order, fitted into the pool now."""
        gs = [np.ascontiguousarray(g, dtype=np.int32) for g in geos]
        if not gs:
            return 0
        dims = np.ascontiguousarray([g.shape for g in gs], dtype=np.int32)
        ptrs = (C.c_void_p * len(gs))(*[g.ctypes.data for g in gs])
        F = np.zeros(3)
        F[: self.nD] = np.asarray(force, dtype=float)[: self.nD]
        tg = np.ascontiguousarray(tags, dtype=np.int64)
        got = C.c_int32()
        self._check(lib().lb_batch_add_many(self._h, len(gs), dims.shape[1], dims.ravel(), ptrs,
                                            periodic_mask(periodic), float(tau), F, float(tol), int(max_steps),
                                            tg, C.byref(got)))
        return got.value

    def add_twophase(self, geo, periodic, rho0, wettability, tau0, tau1, sigma, beta, force, tol, max_steps, tag):
        """Admit a colour-gradient solve (rho0 per geometry cell); False when
        it does not fit into the pool now."""
        g = np.ascontiguousarray(geo, dtype=np.int32)
        r = np.ascontiguousarray(rho0, dtype=np.float64)
        if r.shape != g.shape:
            raise ValueError("rho0 must have the geometry's shape")
        dims = np.ascontiguousarray(g.shape, dtype=np.int32)
        F = np.zeros(3)
        F[: self.nD] = np.asarray(force, dtype=float)[: self.nD]
        rc = lib().lb_batch_add_twophase(self._h, len(dims), dims, g.ravel(), periodic_mask(periodic), r.ravel(),
                                         float(wettability), float(tau0), float(tau1), float(sigma), float(beta),
                                         F, float(tol), int(max_steps), int(tag))
        if rc == -7:
            return False
        self._check(rc)
        return True

    def advance(self, intervals=1):
        """Run up to `intervals` check intervals and wait; returns how many
        solves finished."""
        n = C.c_int32()
        self._check(lib().lb_batch_advance(self._h, int(intervals), C.byref(n)))
        return n.value

    def launch(self, intervals=1):
        """Enqueue `intervals` check intervals and return at once."""
        self._check(lib().lb_batch_launch(self._h, int(intervals)))

    def wait(self):
        """Wait for the launched intervals; returns how many solves finished."""
        n = C.c_int32()
        self._check(lib().lb_batch_wait(self._h, C.byref(n)))
        return n.value

    def collect(self, max_n=1 << 16):
        tags = np.zeros(max_n, dtype=np.int64)
        k = np.zeros(max_n)
        kp = np.zeros(2 * max_n)
        steps = np.zeros(max_n, dtype=np.int64)
        sec = np.zeros(max_n)
        st = np.zeros(max_n, dtype=np.int32)
        n = C.c_int32()
        self._check(lib().lb_batch_collect(self._h, int(max_n), tags, k, kp, steps, sec, st, C.byref(n)))
        m = n.value
        return [{"tag": int(tags[i]), "k": float(k[i]), "k_phase": (float(kp[2 * i]), float(kp[2 * i + 1])),
                 "steps": int(steps[i]), "seconds": float(sec[i]), "status": int(st[i])} for i in range(m)]

    def status(self):
        a, b, c, d, e = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        self._check(lib().lb_batch_status(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d), C.byref(e)))
        return {"resident": a.value, "resident_nodes": b.value, "bytes_used": c.value, "bytes_free": d.value,
                "largest_free": e.value}

    def set_precision(self, kind, precision):
        """Precision of the solves of one kind admitted from now on: kind
        "perm" (single-phase) or "relperm" (two-phase)."""
        k = {"perm": 0, "relperm": 1}.get(kind, kind)
        self._check(lib().lb_batch_set_precision(self._h, int(k), int(precision)))
        self.prec[int(k)] = int(precision)

    def set_shape(self, wg_per_cu):
        self._check(lib().lb_batch_set_shape(self._h, int(wg_per_cu)))

    def shape(self):
        v = C.c_int32()
        self._check(lib().lb_batch_shape(self._h, C.byref(v)))
        return v.value

    def tuning(self):
        """The launch parameters the tuner chose: workgroups per CU, threads
        per workgroup, and the resident nodes they were tuned at."""
        a, t, n = C.c_int32(), C.c_int32(), C.c_int64()
        self._check(lib().lb_batch_tuning(self._h, C.byref(a), C.byref(t), C.byref(n)))
        return {"wg_per_cu": a.value, "wg_threads": t.value, "tuned_nodes": n.value}

    def kernel_time(self):
        ms, n, l = C.c_double(), C.c_int64(), C.c_int64()
        self._check(lib().lb_batch_kernel_time(self._h, C.byref(ms), C.byref(n), C.byref(l)))
        return ms.value, n.value, l.value
