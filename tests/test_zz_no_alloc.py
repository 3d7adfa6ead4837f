"""lb_download_f_noalloc (the download that allocates no device memory)
against lb_download_f on the same context state.

Bar: bit-identical outputs (array_equal over every entry, non-bulk entries
left as the caller filled them by both), for FP64, FP32S and FP16S, one and
two population fields, in the row layout and in the link-table layout with
the node numbers shuffled (bulk nodes then form many runs of consecutive
numbers).  Two calls in a row share the device's staging array, the second
on a smaller context; a context above the staging limit is refused.
"""
import numpy as np
import pytest

import lbamd
import oracle
from helpers import OracleTopo

pytestmark = pytest.mark.gpu

LIMIT = 128 << 20


def _shuffled(T, seed):
    """The same node list with node numbers 1..n-1 permuted."""
    rng = np.random.default_rng(seed)
    n = T.n_nodes
    perm = np.concatenate([[0], 1 + rng.permutation(n - 1)]).astype(np.int32)  # old -> new
    inv = np.empty_like(perm)
    inv[perm] = np.arange(n, dtype=np.int32)
    nb = perm[T.neighbors[inv]].astype(np.int32)
    types = T.node_types[inv].copy()
    bulk = np.sort(perm[T.bulk]).astype(np.int32)
    return nb, types, bulk, inv


def _state(ctx, T_nodes, bulk, nf, Q, seed, lattice):
    rng = np.random.default_rng(seed)
    w = oracle.WEIGHTS[lattice]
    f = np.zeros((T_nodes, nf, Q))
    f[bulk] = w[None, None, :] * (0.5 if nf == 2 else 1.0)
    f[bulk] += 1e-3 * rng.standard_normal((bulk.size, nf, Q))
    ctx.upload_f(f)


def _compare(ctx, n_nodes, nf, Q, sentinel_seed):
    rng = np.random.default_rng(sentinel_seed)
    fill = rng.standard_normal(n_nodes * nf * Q)
    a = ctx.download_f(nf, out=fill.copy())
    b = ctx.download_f_noalloc(nf, out=fill.copy())
    assert np.array_equal(a, b)
    return a


def _run(T, precision, nf, table_shuffle, seed):
    lattice = T.lattice
    Q = 19 if lattice == lbamd.D3Q19 else 9
    D = 3 if lattice == lbamd.D3Q19 else 2
    ctx = lbamd.Context(lattice, precision)
    if table_shuffle:
        nb, types, bulk, _ = _shuffled(T, seed)
        ctx.set_topology(nb, types, bulk)
        assert ctx.layout == 0
    else:
        nb, types, bulk = T.neighbors, T.node_types, T.bulk
        ctx.set_topology_grid(nb, types, bulk, T.positions)
        assert ctx.layout in (1, 2)
    _state(ctx, T.n_nodes, bulk, nf, Q, seed, lattice)
    # the pre-collision state as uploaded
    f0 = _compare(ctx, T.n_nodes, nf, Q, seed + 1)
    force = np.array([1e-5, -2e-6, 3e-6])[:D]
    if nf == 1:
        ctx.step_bgk(7, 0.8, force)
    else:
        ctx.set_wall_densities(np.tile([1.0, 0.0], (T.n_nodes, 1)).ravel())
        ctx.step_cg(5, 1.0, 0.9, 0.01, 0.7, lbamd.FORCE_FIXED, force)
    # after steps: the state comes through the pull gather of the last call
    f1 = _compare(ctx, T.n_nodes, nf, Q, seed + 2)
    ctx.close()
    assert not np.array_equal(f0, f1)


def _geo(shape, seed):
    geo, _ = lbamd.sphere_pack(shape, 2.5, 0.55, seed, periodic="x", walls=True)
    return geo.astype(np.int64)


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S, lbamd.FP16S])
@pytest.mark.parametrize("nf", [1, 2])
def test_noalloc_download_equals_download(gpu_available, precision, nf):
    big = OracleTopo(_geo((26, 22, 20), 5), "D3Q19", "x")
    small = OracleTopo(_geo((12, 10, 9), 6), "D3Q19", "x")
    assert small.bulk.size < big.bulk.size
    print("noalloc precision %d nf %d: %d then %d bulk nodes" % (precision, nf, big.bulk.size, small.bulk.size),
          flush=True)
    for T, shuffle, seed in ((big, False, 11), (small, False, 12), (big, True, 13), (small, True, 14)):
        _run(T, precision, nf, shuffle, seed)


def test_noalloc_download_d2q9(gpu_available):
    geo = np.ones((40, 24), dtype=np.int64)
    geo[:, 0] = geo[:, -1] = 0
    T = OracleTopo(geo, "D2Q9", "x")
    for precision in (lbamd.FP64, lbamd.FP32S, lbamd.FP16S):
        _run(T, precision, 1, False, 21)
        _run(T, precision, 1, True, 22)


def test_noalloc_download_refuses_above_limit(gpu_available):
    """1 000 000 bulk nodes x 19 directions x 8 B = 152 MB > 128 MiB."""
    n = 100
    topo = lbamd.Topology.from_geometry(np.ones((n, n, n), np.int32), lbamd.D3Q19, "xyz")
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    assert bulk.size * 19 * 8 > LIMIT
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP32S)
    ctx.set_topology_grid(nb, types, bulk, pos)
    f = np.zeros((topo.n_nodes, 1, 19))
    f[bulk] = oracle.WEIGHTS[lbamd.D3Q19][None, None, :]
    ctx.upload_f(f)
    with pytest.raises(lbamd.LBError) as e:
        ctx.download_f_noalloc(1)
    assert "128 MiB" in str(e.value)
    # the context is still usable, and the regular download works
    got = ctx.download_f(1)
    assert np.array_equal(got[bulk], f[bulk])
    ctx.close()
    topo.close()
