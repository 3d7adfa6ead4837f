"""Edge cases of the node list: no bulk node, a single fluid cell, one-cell
wide channels, D2Q9 in the reduced-precision forms, a box beyond the row
layouts' coordinate range."""
import numpy as np
import pytest

import lbamd
import oracle
from helpers import OracleTopo

pytestmark = pytest.mark.gpu


def test_empty_bulk_list(gpu_available):
    geo = np.zeros((6, 5, 4), dtype=np.int64)
    T = OracleTopo(geo, "D3Q19", "x")
    assert T.bulk.size == 0
    for precision in (lbamd.FP64, lbamd.FP32S):
        ctx = lbamd.Context(lbamd.D3Q19, precision)
        ctx.set_topology(T.neighbors, T.node_types, T.bulk)
        ctx.upload_f(np.zeros((T.n_nodes, 19)))
        ctx.step_bgk(3, 1.0, np.array([1e-6, 0.0, 0.0]))
        ctx.step_bgk(300, 1.0, np.array([1e-6, 0.0, 0.0]))  # a run long enough for the graph replay
        assert ctx.reduce_flux(0) == (0.0, 0.0, 0.0)
        k, steps = ctx.solve_permeability(1.0, np.array([1e-6, 0.0, 0.0]), 10, 1e-6, 100, float(geo.size))
        # k = 0 at every check: the test holds from the second check on and
        # stops at the third (two consecutive checks)
        assert k == 0.0 and steps == 30
        ctx.close()


@pytest.mark.parametrize("lattice,shape,periodic", [("D3Q19", (5, 5, 5), ""), ("D2Q9", (4, 3), "")])
def test_single_fluid_cell(gpu_available, lattice, shape, periodic):
    """One fluid cell enclosed by solids: every link bounces back; FP64 equals
    the oracle bit for bit and the population is conserved."""
    geo = np.zeros(shape, dtype=np.int64)
    geo[tuple(s // 2 for s in shape)] = 1
    T = OracleTopo(geo, lattice, periodic)
    lat = 0 if lattice == "D2Q9" else 1
    Q = 9 if lattice == "D2Q9" else 19
    assert T.bulk.size == 1
    rng = np.random.default_rng(4)
    f0 = oracle.rest_state(lat, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    f0[T.bulk] *= 1.0 + 1e-3 * rng.standard_normal((1, Q))
    F = np.array([1e-5, -2e-6, 3e-6])[: 2 if lat == 0 else 3]
    fo, ro, vo = oracle.run_std_case(lat, T.neighbors, T.node_types, T.bulk, f0, 0.9, F, 25)
    ctx = lbamd.Context(lbamd.D2Q9 if lat == 0 else lbamd.D3Q19, lbamd.FP64)
    ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
    ctx.upload_f(f0)
    ctx.step_bgk(25, 0.9, F)
    f = ctx.download_f(1)
    ctx.close()
    assert np.array_equal(f.reshape(T.n_nodes, Q)[T.bulk], fo.reshape(T.n_nodes, Q)[T.bulk])


@pytest.mark.parametrize("precision,vtol", [(lbamd.FP64, 0.0), (lbamd.FP32S, 1e-4), (lbamd.FP16S, 5e-3)])
def test_d2q9_thin_channel(gpu_available, precision, vtol):
    """A channel three cells wide (D2Q9), every form against the oracle."""
    geo = np.ones((16, 5), dtype=np.int64)
    geo[:, 0] = 0
    geo[:, -1] = 0
    T = OracleTopo(geo, "D2Q9", "x")
    f0 = oracle.rest_state(0, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    F = np.array([1e-5, 0.0])
    fo, ro, vo = oracle.run_std_case(0, T.neighbors, T.node_types, T.bulk, f0, 0.8, F, 200)
    ctx = lbamd.Context(lbamd.D2Q9, precision)
    ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
    ctx.upload_f(f0)
    ctx.step_bgk(200, 0.8, F)
    v = ctx.download_vel()
    ctx.close()
    b = T.bulk
    err = np.abs(v[b] - vo[b]).max() / np.abs(vo[b]).max()
    print("vel err %.3e" % err)
    assert err <= vtol


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S])
def test_box_beyond_packed_coordinates_uses_link_table(gpu_available, precision):
    """A box longer than the row layouts' packed coordinates allow (x > 1024
    cells) is refused by them with the reason in the layout note; the link
    table carries it, FP64 bit-exact against the oracle, FP32S within 1e-4."""
    geo = np.ones((1100, 5, 3), dtype=np.int64)
    geo[:, 0] = 0
    geo[:, -1] = 0
    geo[::7, 2, 1] = 0  # obstacles along the channel
    T = OracleTopo(geo, "D3Q19", "xz")
    rng = np.random.default_rng(9)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, 1.0 + 0.01 * rng.random(T.n_nodes))
    F = np.array([1e-5, 0.0, 2e-6])
    fo, ro, vo = oracle.run_std_case(1, T.neighbors, T.node_types, T.bulk, f0, 0.8, F, 30)
    ctx = lbamd.Context(lbamd.D3Q19, precision)
    ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
    assert ctx.layout == 0 and "too large" in ctx.layout_note, (ctx.layout, ctx.layout_note)
    ctx.upload_f(f0)
    ctx.step_bgk(30, 0.8, F)
    r, v = ctx.download_scalar(lbamd.RHO), ctx.download_vel()
    ctx.close()
    b = T.bulk
    if precision == lbamd.FP64:
        assert np.array_equal(r[b], ro[b]) and np.array_equal(v[b], vo[b])
    else:
        assert np.abs(v[b] - vo[b]).max() / np.abs(vo[b]).max() < 1e-4
