"""A domain split over ranks, one context per rank on one GPU, with the halo
exchanged between steps (lb_halo_pack / lb_halo_unpack over the lists of
BndMpi::setup, LBbndmpi.h:206-315), against the whole domain in one context.

Bar: bit for bit (populations after the last exchange, and the stored rho
and u of the last step) in LB_FP64 and LB_FP32S, for BGK and TRT, D3Q19 and
D2Q9, two and three ranks over the periodic axis.  The per-node arithmetic
is the same in both runs; only where a gather finds a neighbour's value
differs (a halo slot filled by the exchange instead of the neighbour's own
slot).  The whole-domain context is pinned to the compiled reference
elsewhere (test_gpu_parity.py).  The split run is also pinned directly: the
compiled reference std_case run on 2 and 3 MPI ranks (its own halo exchange;
fixtures from tests/golden/make_mpi_fixtures.py) gives every rank's rho and
u bit for bit.
"""
import numpy as np
import pytest

import lbamd
from helpers import load_ref, ref_names
from lbamd.halo import exchange, exchange_lists

pytestmark = pytest.mark.gpu


def split_geometry(lattice, shape, n_ranks, seed):
    if lattice == lbamd.D3Q19:
        geo, _ = lbamd.sphere_pack(shape, 3.0, 0.45, seed, periodic="x", walls=True)
    else:
        rng = np.random.default_rng(seed)
        geo = (rng.random(shape) > 0.3).astype(np.int32)
        geo[:, 0] = 0
        geo[:, -1] = 0
    fluid = (np.asarray(geo) > 0).astype(np.int32)
    g = fluid.copy()
    cut = np.linspace(0, shape[0], n_ranks + 1).astype(int)
    for r in range(n_ranks):
        sl = g[cut[r]:cut[r + 1]]
        sl[sl > 0] = r + 1
    return fluid, g


def initial(topo, cell_f):
    """LbField over the node list: the per-cell populations at each node's
    position (solid and halo nodes get them too; they are not read)."""
    pos = topo.positions()
    Q = topo.Q
    f = np.zeros((topo.n_nodes, Q))
    ok = np.all(pos >= 0, axis=1)
    idx = tuple(np.mod(pos[ok], cell_f.shape[:-1]).T)
    f[ok] = cell_f[idx]
    return f


def run_whole(lattice, fluid, precision, cell_f, nsteps, op):
    t = lbamd.Topology.from_geometry(fluid, lattice, "x")
    ctx = lbamd.Context(lattice, precision)
    ctx.set_topology(t.neighbors(), t.node_types(), t.bulk())
    ctx.upload_f(initial(t, cell_f))
    op(ctx, nsteps)
    out = (t.positions()[t.bulk()], ctx.download_f(1)[t.bulk()], ctx.download_scalar(lbamd.RHO)[t.bulk()],
           ctx.download_vel()[t.bulk()])
    ctx.close()
    t.close()
    return out


def run_split(lattice, g, n_ranks, precision, cell_f, nsteps, op):
    topos = [lbamd.Topology.from_geometry(g, lattice, "x", rank=r + 1) for r in range(n_ranks)]
    lists = exchange_lists(topos)
    ctxs, reg = [], []
    for r, t in enumerate(topos):
        ctx = lbamd.Context(lattice, precision)
        ctx.set_topology_grid(t.neighbors(), t.node_types(), t.bulk(), t.positions())
        assert ctx.layout == 0 and ctx.halo_nodes > 0, ctx.layout_note
        ctx.upload_f(initial(t, cell_f))
        reg.append({x[0]: ctx.halo_add(*x) for x in lists[r]})
        ctxs.append(ctx)
    for _ in range(nsteps):
        for ctx in ctxs:
            op(ctx, 1)
        exchange(ctxs, reg)
    out = []
    for t, ctx in zip(topos, ctxs):
        b = t.bulk()
        out.append((t.positions()[b], ctx.download_f(1)[b], ctx.download_scalar(lbamd.RHO)[b],
                    ctx.download_vel()[b]))
        ctx.close()
        t.close()
    return out


def bgk(ctx, n):
    ctx.step_bgk(n, 0.87, [2e-5, -1e-5, 5e-6][: ctx.nD])


def trt(ctx, n):
    ctx.step_trt(n, 1.0, 0.875, [2e-5, 0.0, 3e-6][: ctx.nD])


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S])
@pytest.mark.parametrize("lattice,shape,n_ranks,op", [
    (lbamd.D3Q19, (24, 16, 12), 2, bgk),
    (lbamd.D3Q19, (30, 14, 12), 3, bgk),
    (lbamd.D3Q19, (24, 16, 12), 2, trt),
    (lbamd.D2Q9, (40, 18), 2, bgk),
])
def test_split_equals_whole(gpu_available, precision, lattice, shape, n_ranks, op):
    fluid, g = split_geometry(lattice, shape, n_ranks, 5)
    Q = 19 if lattice == lbamd.D3Q19 else 9
    rng = np.random.default_rng(11)
    w = {19: np.array([2.0] * 3 + [1.0] * 6 + [2.0] * 3 + [1.0] * 6 + [12.0]) / 36.0,
         9: np.array([4.0, 1.0] * 4 + [16.0]) / 36.0}[Q]
    cell_f = w * (1.0 + 1e-3 * rng.standard_normal(shape + (Q,)))
    nsteps = 25
    wp, wf, wr, wv = run_whole(lattice, fluid, precision, cell_f, nsteps, op)
    where = {tuple(p): k for k, p in enumerate(wp)}
    parts = run_split(lattice, g, n_ranks, precision, cell_f, nsteps, op)
    assert sum(len(p[0]) for p in parts) == len(wp)
    for pos, f, rho, vel in parts:
        k = np.array([where[tuple(p)] for p in pos])
        assert np.array_equal(f, wf[k])
        assert np.array_equal(rho, wr[k])
        assert np.array_equal(vel, wv[k])


@pytest.mark.parametrize("name", ref_names("mpi_"))
def test_split_vs_reference_mpi_run(gpu_available, name):
    """The compiled reference std_case run on R MPI ranks (its own halo
    exchange, tests/golden/make_mpi_fixtures.py) against R contexts on one
    GPU exchanging through lb_halo_pack / lb_halo_unpack: every rank's rho
    and u of the last iteration, bit for bit."""
    d = load_ref(name)
    p = d["params"]
    lattice = {"D2Q9": lbamd.D2Q9, "D3Q19": lbamd.D3Q19}[p["lattice"]]
    ranks = p["ranks"]
    topos = [lbamd.Topology.from_geometry(d["geo"], lattice, p["periodic"], rank=r + 1) for r in range(ranks)]
    lists = exchange_lists(topos)
    W = {lbamd.D3Q19: np.array([2.0] * 3 + [1.0] * 6 + [2.0] * 3 + [1.0] * 6 + [12.0]) / 36.0,
         lbamd.D2Q9: np.array([4.0, 1.0] * 4 + [16.0]) / 36.0}[lattice]
    ctxs, reg = [], []
    for r, t in enumerate(topos):
        ctx = lbamd.Context(lattice, lbamd.FP64)
        ctx.set_topology(t.neighbors(), t.node_types(), t.bulk())
        f0 = np.zeros((t.n_nodes, W.size))
        f0[t.bulk()] = W  # std_case/main.cpp:67-96 with init_rho = 1
        ctx.upload_f(f0)
        reg.append({x[0]: ctx.halo_add(*x) for x in lists[r]})
        ctxs.append(ctx)
    force = np.asarray(p["force"], dtype=float)
    for _ in range(p["steps"] + 1):  # the last iteration's stored fields (std_case/main.cpp:109-155)
        for ctx in ctxs:
            ctx.step_bgk(1, p["tau"], force)
        exchange(ctxs, reg)
    for r, (t, ctx) in enumerate(zip(topos, ctxs)):
        b = t.bulk()
        rho, vel = ctx.download_scalar(lbamd.RHO)[b], ctx.download_vel()[b]
        assert np.array_equal(rho, d["rho_%d" % r]), (name, r)
        assert np.array_equal(vel, d["vel_%d" % r][:, : vel.shape[1]]), (name, r)
        ctx.close()
        t.close()


def test_halo_refusals(gpu_available):
    fluid, g = split_geometry(lbamd.D3Q19, (16, 10, 8), 2, 3)
    t = lbamd.Topology.from_geometry(g, lbamd.D3Q19, "x", rank=1)
    lists = exchange_lists([t, lbamd.Topology.from_geometry(g, lbamd.D3Q19, "x", rank=2)])
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP64)
    ctx.set_topology(t.neighbors(), t.node_types(), t.bulk())
    s, ns, nds, ds, nr, ndr, dr = lists[0][0]
    # a receive node that is not a halo node, a send node that is not bulk
    with pytest.raises(lbamd.LBError) as e:
        ctx.halo_add(s, ns, nds, ds, ns, nds, ds)
    assert e.value.code == -1
    with pytest.raises(lbamd.LBError) as e:
        ctx.halo_add(s, nr, ndr, dr, nr, ndr, dr)
    assert e.value.code == -1
    k = ctx.halo_add(s, ns, nds, ds, nr, ndr, dr)
    assert ctx.halo_sizes(k) == (s, int(nds.sum()), int(ndr.sum()))
    W = np.array([2.0] * 3 + [1.0] * 6 + [2.0] * 3 + [1.0] * 6 + [12.0]) / 36.0
    f0 = np.tile(W, (t.n_nodes, 1))
    ctx.upload_f(f0)
    # the halo carries post-collision values: nothing to exchange before a step
    with pytest.raises(lbamd.LBError) as e:
        ctx.halo_pack(k)
    assert e.value.code == -2
    ctx.step_bgk(1, 1.0, [1e-6, 0.0, 0.0])
    assert ctx.halo_pack(k).shape == (1, int(nds.sum()))
    # colour-gradient steps need the halo's colour too: refused
    ctx.upload_f(np.stack([f0, 0.0 * f0], 1))
    with pytest.raises(lbamd.LBError) as e:
        ctx.step_cg(1, 1.0, 1.0, 0.01, 0.7, lbamd.FORCE_FIXED, [1e-6, 0.0, 0.0])
    assert e.value.code == -4
    ctx.close()
