"""Halo exchange lists of a domain split over ranks (lbamd.halo.exchange_lists,
the lists BndMpi::setup builds, LBbndmpi.h:206-315).  For every pair of
ranks, what one rank sends is what the other receives: the k-th value sent
is the value of the same cell (up to the periodic wrap of the halo) and the
same direction as the k-th value received, and every direction listed
points from the halo node into the receiving rank."""
import numpy as np
import pytest

import lbamd
import oracle
from helpers import OracleTopo, load_ref, ref_names, tp_start
from lbamd.halo import exchange_lists
from oracle.ref_run import rank_blocks


def split(shape, blocks, periodic, seed, lattice):
    if lattice == lbamd.D3Q19:
        geo, _ = lbamd.sphere_pack(shape, 3.0, 0.45, seed, periodic=periodic, walls=True)
    else:
        rng = np.random.default_rng(seed)
        geo = (rng.random(shape) > 0.3).astype(np.int32)
        if periodic == "x":
            geo[:, 0] = 0
            geo[:, -1] = 0
    return rank_blocks(np.asarray(geo), blocks)


@pytest.mark.parametrize("lattice,shape,periodic,blocks", [
    (lbamd.D3Q19, (24, 16, 12), "x", (2,)), (lbamd.D3Q19, (30, 14, 12), "x", (3,)), (lbamd.D2Q9, (40, 18), "x", (2,)),
    (lbamd.D3Q19, (20, 16, 10), "xy", (2, 2)), (lbamd.D3Q19, (18, 18, 12), "xyz", (3, 2)),
    (lbamd.D2Q9, (24, 20), "xy", (2, 2))])
def test_exchange_lists_pair_up(lattice, shape, periodic, blocks):
    """Slabs and blocks (ranks that share only an edge or a corner exchange
    the diagonal directions alone)."""
    g = split(shape, blocks, periodic, 5, lattice)
    n_ranks = int(np.prod(blocks))
    topos = [lbamd.Topology.from_geometry(g, lattice, periodic, rank=r + 1) for r in range(n_ranks)]
    lists = exchange_lists(topos)
    pos = [t.positions() for t in topos]
    ranks = [t.node_ranks() for t in topos]
    nbs = [t.neighbors() for t in topos]
    n_pairs = 0
    for r in range(n_ranks):
        assert lists[r], "rank %d has no neighbour" % r
        for (s, ns, nds, ds, nr, ndr, dr) in lists[r]:
            # r's receive side: halo nodes of rank s, directions into rank r
            assert np.all(ranks[r][nr] == s)
            recv_nodes = np.repeat(nr, ndr)
            assert np.all(ranks[r][nbs[r][recv_nodes, dr]] == r)
            # r's send side: its own bulk nodes
            assert np.all(np.isin(ns, topos[r].bulk()))
            # s receives exactly what r sends, value by value: the same cell
            # up to the periodic wrap, the same direction
            other = [x for x in lists[s] if x[0] == r]
            assert len(other) == 1
            _, _, _, _, nr2, ndr2, dr2 = other[0]
            a, b = np.repeat(ns, nds), np.repeat(nr2, ndr2)
            assert len(a) == len(b) == len(ds) and np.array_equal(ds, dr2)
            d = pos[r][a] - pos[s][b]
            for ax in range(len(shape)):
                if "xyz"[ax] in periodic:
                    assert np.all(d[:, ax] % shape[ax] == 0)
                else:
                    assert np.all(d[:, ax] == 0)
            n_pairs += 1
    assert n_pairs == sum(len(x) for x in lists)
    if len(blocks) == 2:
        # the block of rank 1 touches the diagonal block along an edge only
        assert any(x[0] == blocks[1] + 1 for x in lists[0])


@pytest.mark.parametrize("name", ref_names("mpi_tp_"))
def test_reference_mpi_twophase_equals_oracle_whole_domain(name):
    """The fixtures of the split two-phase run (the compiled reference's
    twophase on R MPI ranks, exchanging populations, the colour field and
    the flux sum) against the oracle on the undivided domain: every rank's
    rho0, rho1 and u at the same cells, and the body force, to 1e-12 of
    their scale (the flux of the force controller is summed per rank, then
    over the ranks)."""
    d = load_ref(name)
    p = d["params"]
    ranked = d["geo"].astype(np.int64)
    T = OracleTopo((ranked > 0).astype(np.int64), p["lattice"], p["periodic"])
    f0, rho = tp_start(d, T)
    _, r, v, _, F = oracle.run_twophase(T.lattice, T.neighbors, T.node_types, T.bulk, f0, rho, p["tau0"],
                                        p["tau1"], p["sigma"], p["beta"], 1, [p["momx"]] + p["force"][1:],
                                        p["steps"] + 1)
    where = {tuple(x): n for n, x in enumerate(T.positions)}
    seen = 0
    for rank in range(p["ranks"]):
        Tr = OracleTopo(ranked, p["lattice"], p["periodic"], rank=rank + 1)
        at = np.array([where[tuple(x)] for x in Tr.positions[Tr.bulk]])
        seen += at.size
        for got, ref in ((r[at, 0], d["rho0_%d" % rank]), (r[at, 1], d["rho1_%d" % rank]),
                         (v[at], d["vel_%d" % rank][:, : v.shape[1]])):
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (name, rank)
    assert seen == T.bulk.size
    assert abs(F[0] - d["force_x"][-1, 1]) <= 1e-12 * abs(d["force_x"][-1, 1])


@pytest.mark.parametrize("name", [n for n in ref_names("mpi_") if not n.startswith("mpi_tp_")])
def test_reference_mpi_std_case_equals_oracle_whole_domain(name):
    """The compiled reference's std_case on R MPI ranks against the oracle on
    the undivided domain: with no global sum in the loop, every rank's rho
    and u equal the whole domain's at the same cells bit for bit."""
    d = load_ref(name)
    p = d["params"]
    ranked = d["geo"].astype(np.int64)
    T = OracleTopo((ranked > 0).astype(np.int64), p["lattice"], p["periodic"])
    W = oracle.WEIGHTS[T.lattice]
    f0 = np.zeros((T.n_nodes, W.size))
    f0[T.bulk] = W  # init_rho = 1
    _, r, v = oracle.run_std_case(T.lattice, T.neighbors, T.node_types, T.bulk, f0, p["tau"],
                                  np.asarray(p["force"], dtype=float), p["steps"] + 1)
    where = {tuple(x): n for n, x in enumerate(T.positions)}
    seen = 0
    for rank in range(p["ranks"]):
        Tr = OracleTopo(ranked, p["lattice"], p["periodic"], rank=rank + 1)
        at = np.array([where[tuple(x)] for x in Tr.positions[Tr.bulk]])
        seen += at.size
        assert np.array_equal(r[at], d["rho_%d" % rank]), (name, rank)
        assert np.array_equal(v[at], d["vel_%d" % rank][:, : v.shape[1]]), (name, rank)
    assert seen == T.bulk.size
