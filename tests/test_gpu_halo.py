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
from helpers import OracleTopo, load_ref, ref_names, tp_start
from lbamd.halo import exchange, exchange_lists, step_twophase
from oracle.ref_run import rank_blocks

pytestmark = pytest.mark.gpu


def split_geometry(lattice, shape, periodic, blocks, seed):
    if lattice == lbamd.D3Q19:
        geo, _ = lbamd.sphere_pack(shape, 3.0, 0.45, seed, periodic=periodic, walls=True)
    else:
        rng = np.random.default_rng(seed)
        geo = (rng.random(shape) > 0.3).astype(np.int32)
        if "y" not in periodic:
            geo[:, 0] = 0
            geo[:, -1] = 0
    fluid = (np.asarray(geo) > 0).astype(np.int32)
    return fluid, rank_blocks(fluid, blocks)


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


def run_whole(lattice, fluid, periodic, precision, cell_f, nsteps, op):
    t = lbamd.Topology.from_geometry(fluid, lattice, periodic)
    ctx = lbamd.Context(lattice, precision)
    ctx.set_topology(t.neighbors(), t.node_types(), t.bulk())
    ctx.upload_f(initial(t, cell_f))
    op(ctx, nsteps)
    out = (t.positions()[t.bulk()], ctx.download_f(1)[t.bulk()], ctx.download_scalar(lbamd.RHO)[t.bulk()],
           ctx.download_vel()[t.bulk()])
    ctx.close()
    t.close()
    return out


def run_split(lattice, g, periodic, n_ranks, precision, cell_f, nsteps, op):
    topos = [lbamd.Topology.from_geometry(g, lattice, periodic, rank=r + 1) for r in range(n_ranks)]
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
@pytest.mark.parametrize("lattice,shape,periodic,blocks,op", [
    (lbamd.D3Q19, (24, 16, 12), "x", (2,), bgk),
    (lbamd.D3Q19, (30, 14, 12), "x", (3,), bgk),
    (lbamd.D3Q19, (24, 16, 12), "x", (2,), trt),
    (lbamd.D2Q9, (40, 18), "x", (2,), bgk),
    (lbamd.D3Q19, (20, 16, 10), "xy", (2, 2), bgk),
    (lbamd.D3Q19, (18, 18, 12), "xyz", (3, 2), trt),
    (lbamd.D2Q9, (24, 20), "xy", (2, 2), bgk),
])
def test_split_equals_whole(gpu_available, precision, lattice, shape, periodic, blocks, op):
    """Slabs, and blocks whose ranks also meet along edges and corners."""
    fluid, g = split_geometry(lattice, shape, periodic, blocks, 5)
    n_ranks = int(np.prod(blocks))
    Q = 19 if lattice == lbamd.D3Q19 else 9
    rng = np.random.default_rng(11)
    w = {19: np.array([2.0] * 3 + [1.0] * 6 + [2.0] * 3 + [1.0] * 6 + [12.0]) / 36.0,
         9: np.array([4.0, 1.0] * 4 + [16.0]) / 36.0}[Q]
    cell_f = w * (1.0 + 1e-3 * rng.standard_normal(shape + (Q,)))
    nsteps = 25
    wp, wf, wr, wv = run_whole(lattice, fluid, periodic, precision, cell_f, nsteps, op)
    where = {tuple(p): k for k, p in enumerate(wp)}
    parts = run_split(lattice, g, periodic, n_ranks, precision, cell_f, nsteps, op)
    assert sum(len(p[0]) for p in parts) == len(wp)
    for pos, f, rho, vel in parts:
        k = np.array([where[tuple(p)] for p in pos])
        assert np.array_equal(f, wf[k])
        assert np.array_equal(rho, wr[k])
        assert np.array_equal(vel, wv[k])


@pytest.mark.parametrize("name", [n for n in ref_names("mpi_") if not n.startswith("mpi_tp_")])
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
    fluid, g = split_geometry(lbamd.D3Q19, (16, 10, 8), "x", (2,), 3)
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
    # a solve's stop test sums over one context: refused on a rank of a split domain
    with pytest.raises(lbamd.LBError) as e:
        ctx.solve_permeability(1.0, np.array([1e-6, 0.0, 0.0]), 10, 1e-6, 100, 1000.0)
    assert e.value.code == -4
    ctx.close()


# ---- the colour-gradient iteration over a split domain

def tp_contexts(ranked, lattice, periodic, ranks, precision, rho0, wett):
    """One context per rank (link table), the two-phase start state and
    wall colours of tp_start, the exchanges registered."""
    topos = [lbamd.Topology.from_geometry(ranked, lattice, periodic, rank=r + 1) for r in range(ranks)]
    lists = exchange_lists(topos)
    ctxs, reg, where = [], [], []
    d = {"rho0_attr": rho0, "wett_attr": wett}
    for r, t in enumerate(topos):
        T = OracleTopo(ranked, lattice, periodic, rank=r + 1)
        assert np.array_equal(T.neighbors, t.neighbors())
        f0, rho = tp_start(d, T)
        ctx = lbamd.Context(lattice, precision)
        ctx.set_topology(T.neighbors, T.node_types, T.bulk)
        assert ctx.halo_nodes > 0
        ctx.upload_f(f0)
        ctx.set_wall_densities(rho)
        reg.append({x[0]: ctx.halo_add(*x) for x in lists[r]})
        ctxs.append(ctx)
        where.append(T)
        t.close()
    n_global = sum(T.bulk.size for T in where)
    for ctx in ctxs:
        ctx.set_global_bulk_count(n_global)
    return ctxs, reg, where


def tp_whole(fluid, lattice, periodic, precision, rho0, wett):
    T = OracleTopo(fluid, lattice, periodic)
    f0, rho = tp_start({"rho0_attr": rho0, "wett_attr": wett}, T)
    ctx = lbamd.Context(lattice, precision)
    ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    ctx.upload_f(f0)
    ctx.set_wall_densities(rho)
    return ctx, T


def tp_fields(ctx, T):
    b = T.bulk
    return (T.positions[b], ctx.download_f(2)[b], ctx.download_scalar(lbamd.RHO, 2)[b], ctx.download_vel()[b],
            ctx.download_scalar(lbamd.PHI)[b])


def tp_geometry(lattice, shape, periodic, blocks, seed):
    if lattice == lbamd.D3Q19:
        geo, _ = lbamd.sphere_pack(shape, 3.0, 0.5, seed, periodic=periodic, walls=periodic != "xyz")
        fluid = (np.asarray(geo) > 0).astype(np.int64)
    else:
        rng = np.random.default_rng(seed)
        fluid = (rng.random(shape) > 0.3).astype(np.int64)
        fluid[:, 0] = 0
        fluid[:, -1] = 0
    rng = np.random.default_rng(seed + 1)
    rho0 = rng.choice([0.0, 0.3, 1.0], size=shape)
    wett = np.full(shape, 0.8)
    return fluid, rank_blocks(fluid, blocks).astype(np.int64), rho0, wett


TP = dict(tau0=0.8, tau1=1.2, sigma=0.01, beta=0.7)


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S])
@pytest.mark.parametrize("lattice,shape,periodic,blocks,mode", [
    (lbamd.D3Q19, (20, 12, 10), "xyz", (2,), lbamd.FORCE_FIXED),
    (lbamd.D3Q19, (24, 12, 10), "x", (3,), lbamd.FORCE_FIXED),
    (lbamd.D2Q9, (40, 18), "x", (2,), lbamd.FORCE_FIXED),
    (lbamd.D3Q19, (20, 12, 10), "xyz", (2,), lbamd.FORCE_MOMX),
    (lbamd.D3Q19, (20, 16, 10), "xyz", (2, 2), lbamd.FORCE_FIXED),
    (lbamd.D2Q9, (24, 20), "x", (2, 2), lbamd.FORCE_FIXED),
])
def test_twophase_split_equals_whole(gpu_available, precision, lattice, shape, periodic, blocks, mode):
    """The colour-gradient loop on R contexts (colour pass, colour exchange,
    collision pass, population exchange per iteration) against the whole
    domain in one context: bit for bit with a fixed force (the per-node
    arithmetic is the same; only where a gather or the gradient finds a
    neighbour's value differs); with the momentum-controlled force to 1e-12
    of each field's scale (the flux is summed per rank, then over the
    ranks)."""
    ranks = int(np.prod(blocks))
    fluid, ranked, rho0, wett = tp_geometry(lattice, shape, periodic, blocks, 7)
    force = [1e-4 if mode == lbamd.FORCE_MOMX else 3e-6, 1e-6, -2e-6][: 3 if lattice == lbamd.D3Q19 else 2]
    nsteps = 30
    ctx, T = tp_whole(fluid, lattice, periodic, precision, rho0, wett)
    ctx.step_cg(nsteps, force_mode=mode, force=force, **TP)
    wp, wf, wr, wv, wphi = tp_fields(ctx, T)
    wF = ctx.last_force()
    ctx.close()
    ctxs, reg, topos = tp_contexts(ranked, lattice, periodic, ranks, precision, rho0, wett)
    for _ in range(nsteps):
        step_twophase(ctxs, reg, TP["tau0"], TP["tau1"], TP["sigma"], TP["beta"], mode, force)
    where = {tuple(p): k for k, p in enumerate(wp)}
    seen = 0
    for ctx, Tr in zip(ctxs, topos):
        pos, f, r, v, phi = tp_fields(ctx, Tr)
        k = np.array([where[tuple(p)] for p in pos])
        seen += k.size
        if mode == lbamd.FORCE_FIXED:
            assert np.array_equal(f, wf[k])
            assert np.array_equal(r, wr[k])
            assert np.array_equal(v, wv[k])
            assert np.array_equal(phi, wphi[k])
            assert np.array_equal(ctx.last_force(), wF)
        else:
            for got, ref in ((f, wf[k]), (r, wr[k]), (v, wv[k]), (phi, wphi[k])):
                assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
            assert abs(ctx.last_force()[0] - wF[0]) <= 1e-12 * abs(wF[0])
        ctx.close()
    assert seen == len(wp)


@pytest.mark.parametrize("name", ref_names("mpi_tp_"))
def test_twophase_split_vs_reference_mpi_run(gpu_available, name):
    """The compiled reference twophase run on R MPI ranks (its own exchanges
    of populations and colour, its MPI_Allreduce of the flux;
    tests/golden/make_mpi_fixtures.py) against R contexts on one GPU: every
    rank's rho0, rho1 and u of the last iteration, and the body force, to
    1e-12 of their scale (the flux sums run in another order)."""
    d = load_ref(name)
    p = d["params"]
    ranked = d["geo"].astype(np.int64)
    lattice = {"D2Q9": lbamd.D2Q9, "D3Q19": lbamd.D3Q19}[p["lattice"]]
    ctxs, reg, topos = tp_contexts(ranked, lattice, p["periodic"], p["ranks"], lbamd.FP64, d["rho0_attr"],
                                   d["wett_attr"])
    force = [p["momx"]] + list(p["force"][1:])
    for _ in range(p["steps"] + 1):
        step_twophase(ctxs, reg, p["tau0"], p["tau1"], p["sigma"], p["beta"], lbamd.FORCE_MOMX, force)
    for r, (ctx, T) in enumerate(zip(ctxs, topos)):
        _, _, rho, vel, _ = tp_fields(ctx, T)
        for got, ref in ((rho[:, 0], d["rho0_%d" % r]), (rho[:, 1], d["rho1_%d" % r]),
                         (vel, d["vel_%d" % r][:, : vel.shape[1]])):
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (name, r)
        F = ctx.last_force()[0]
        assert abs(F - d["force_x"][-1, 1]) <= 1e-12 * abs(d["force_x"][-1, 1])
        ctx.close()


@pytest.mark.parametrize("precision,layout,mode", [
    (lbamd.FP64, "table", lbamd.FORCE_FIXED), (lbamd.FP64, "table", lbamd.FORCE_MOMX),
    (lbamd.FP64, "grid", lbamd.FORCE_MOMX), (lbamd.FP32S, "grid", lbamd.FORCE_MOMX),
    (lbamd.FP16S, "grid", lbamd.FORCE_FIXED)])
def test_twophase_colour_then_collide_equals_step_cg(gpu_available, precision, layout, mode):
    """On one context without a halo, lb_step_cg_colour + lb_step_cg_collide
    per iteration (the flux handed back unchanged) give the bits of
    lb_step_cg."""
    fluid, _, rho0, wett = tp_geometry(lbamd.D3Q19, (18, 14, 12), "x", (1,), 3)
    T = OracleTopo(fluid, lbamd.D3Q19, "x")
    f0, rho = tp_start({"rho0_attr": rho0, "wett_attr": wett}, T)
    force = [1e-4 if mode == lbamd.FORCE_MOMX else 3e-6, 1e-6, -2e-6]
    out = []
    for split in (False, True):
        ctx = lbamd.Context(lbamd.D3Q19, precision)
        if layout == "table":
            ctx.set_topology(T.neighbors, T.node_types, T.bulk)
        else:
            ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
        ctx.upload_f(f0)
        ctx.set_wall_densities(rho)
        for _ in range(12):
            if split:
                s = ctx.step_cg_colour(flux=mode == lbamd.FORCE_MOMX)
                ctx.step_cg_collide(force_mode=mode, force=force, flux_x_total=s or 0.0, **TP)
            else:
                ctx.step_cg(1, force_mode=mode, force=force, **TP)
        out.append(tp_fields(ctx, T) + (ctx.last_force(),))
        ctx.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_twophase_split_refusals(gpu_available):
    fluid, ranked, rho0, wett = tp_geometry(lbamd.D3Q19, (16, 10, 8), "x", (2,), 5)
    ctxs, reg, _ = tp_contexts(ranked, lbamd.D3Q19, "x", 2, lbamd.FP64, rho0, wett)
    ctx, k = ctxs[0], reg[0][1]
    # lb_step_cg cannot exchange the colour between its passes
    with pytest.raises(lbamd.LBError) as e:
        ctx.step_cg(1, force_mode=lbamd.FORCE_FIXED, force=[1e-6, 0.0, 0.0], **TP)
    assert e.value.code == -4
    # the collision and the colour exchange need the colour pass of the iteration
    with pytest.raises(lbamd.LBError) as e:
        ctx.step_cg_collide(force_mode=lbamd.FORCE_FIXED, force=[1e-6, 0.0, 0.0], **TP)
    assert e.value.code == -2
    with pytest.raises(lbamd.LBError) as e:
        ctx.halo_pack_colour(k)
    assert e.value.code == -2
    ctx.step_cg_colour()
    assert ctx.halo_pack_colour(k).shape == (ctx._halo_node_counts[k][0],)
    ctx.step_cg_collide(force_mode=lbamd.FORCE_FIXED, force=[1e-6, 0.0, 0.0], **TP)
    with pytest.raises(lbamd.LBError) as e:
        ctx.solve_relperm(TP["tau0"], TP["tau1"], TP["sigma"], TP["beta"], [1e-6, 0.0, 0.0], 10, 1e-6, 100, 1000.0)
    assert e.value.code == -4
    # one colour pass per iteration: the next collision needs a new one
    with pytest.raises(lbamd.LBError) as e:
        ctx.step_cg_collide(force_mode=lbamd.FORCE_FIXED, force=[1e-6, 0.0, 0.0], **TP)
    assert e.value.code == -2
    for c in ctxs:
        c.close()
