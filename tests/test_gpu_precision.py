"""Reduced-precision forms (shifted populations f - w_q stored in FP32 or FP16,
float arithmetic, FP64 reductions) against the FP64 oracle.

Tolerances (stated per quantity, norm-wise: max |a - b| / max |b|):
  FP32S  velocity 1e-4, density deviation 1e-3 (absolute 1e-9), permeability 1e-4
  FP16S  velocity 2e-2, permeability 2e-2
and, for every layout, the same precision gives bit-identical results
(the per-node arithmetic does not depend on how sources are addressed).
"""
import numpy as np
import pytest

import lbamd
import oracle
from helpers import OracleTopo

pytestmark = pytest.mark.gpu


def pack(shape, radius, porosity, seed, periodic="x"):
    geo, _ = lbamd.sphere_pack(shape, radius, porosity, seed, periodic=periodic, walls=True)
    return geo.astype(np.int64), OracleTopo(geo.astype(np.int64), "D3Q19", periodic)


def run(T, precision, f0, tau, force, nsteps, layout):
    ctx = lbamd.Context(T.lattice, precision)
    if layout == "table":
        ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    else:
        ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
        assert ctx.layout == 2, ctx.layout_note
    ctx.upload_f(f0)
    ctx.step_bgk(nsteps, tau, np.asarray(force, dtype=float))
    out = ctx.download_scalar(lbamd.RHO), ctx.download_vel(), ctx.download_f(1)[:, 0]
    ctx.close()
    return out


def normerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("precision,vtol", [(lbamd.FP32S, 1e-4), (lbamd.FP16S, 2e-2)])
def test_reduced_precision_vs_oracle(gpu_available, precision, vtol):
    geo, T = pack((28, 24, 20), 4.0, 0.5, 21)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    force = np.array([2e-5, 0.0, 0.0])
    steps = 300
    fo, ro, vo = oracle.run_std_case(1, T.neighbors, T.node_types, T.bulk, f0, 0.9, force, steps)
    r, v, f = run(T, precision, f0, 0.9, force, steps, "rows")
    b = T.bulk
    assert normerr(v[b], vo[b]) < vtol
    if precision == lbamd.FP32S:
        assert np.abs(r[b] - ro[b]).max() < max(1e-3 * np.abs(ro[b] - 1).max(), 1e-9)


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S, lbamd.FP16S])
def test_layouts_give_identical_results(gpu_available, precision):
    geo, T = pack((24, 20, 18), 3.0, 0.5, 5)
    rng = np.random.default_rng(2)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    f0[T.bulk] += 1e-4 * rng.standard_normal((T.bulk.size, 19))
    a = run(T, precision, f0, 0.8, [1e-5, 2e-6, 0.0], 40, "table")
    b = run(T, precision, f0, 0.8, [1e-5, 2e-6, 0.0], 40, "rows")
    for x, y in zip(a, b):
        assert np.array_equal(x[T.bulk], y[T.bulk])


@pytest.mark.parametrize("precision,ktol", [(lbamd.FP32S, 1e-4), (lbamd.FP16S, 2e-2)])
def test_permeability_error(gpu_available, precision, ktol):
    geo, T = pack((48, 40, 40), 5.0, 0.45, 9)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    ks = {}
    for p in (lbamd.FP64, precision):
        ctx = lbamd.Context(lbamd.D3Q19, p)
        ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
        ctx.upload_f(f0)
        ks[p] = ctx.solve_permeability(1.0, np.array([1e-6, 0, 0]), 100, 1e-7, 40000, float(geo.size))
        ctx.close()
    k64, s64 = ks[lbamd.FP64]
    k, s = ks[precision]
    assert k64 > 0 and s64 < 40000
    assert abs(k - k64) / k64 < ktol


def test_permeability_matches_oracle_flux(gpu_available):
    """k from the device reduction equals nu * sum_bulk(j_x) / (N_box F_x)
    computed from the oracle's state after the same number of steps."""
    geo, T = pack((20, 16, 16), 3.0, 0.5, 4)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    F = np.array([1e-6, 0.0, 0.0])
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP64)
    ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
    ctx.upload_f(f0)
    k, steps = ctx.solve_permeability(1.0, F, 50, 1e-30, 200, float(geo.size))
    ctx.close()
    assert steps == 200
    # j_x is summed over the state the last step started from (after 199 steps)
    fo, ro, vo = oracle.run_std_case(1, T.neighbors, T.node_types, T.bulk, f0, 1.0, F, 200)
    jx = (ro[T.bulk] * vo[T.bulk, 0]).sum()
    k_ref = (0.5 / 3.0) * jx / geo.size / F[0]
    assert abs(k - k_ref) <= 1e-12 * abs(k_ref)
