"""Halo exchange lists for a domain split over ranks, as the reference's
BndMpi::setup builds them (LBbndmpi.h:206-315) when every rank is known in
one process (for example several contexts on one GPU, or a test).

Per rank r and each neighbouring rank s, MonLatMpi holds:
* nodes_received: r's halo nodes owned by s (LBvtk::getNeigNodesNo), and per
  node the directions q with neighbour(q, node) on rank r (makeDirList,
  LBbndmpi.h:177-203), concatenated;
* nodes_to_send: the same exchange seen from s, i.e. s's halo nodes owned by r
  in r's own numbering (s's getNeigNodesNeigNo) with s's direction lists.
A rank packs its post-collision values of (node, q) in send order, the
neighbour unpacks them in its receive order: the k-th value of one is the
k-th value of the other.
"""
import numpy as np

from .ffi import FORCE_MOMX


def dir_lists(topo, rank, ghosts):
    """makeDirList: per halo node, the directions (rest excluded) whose
    neighbour belongs to `rank`; returns (n_dir per node, directions)."""
    nb = topo.neighbors()
    ranks = topo.node_ranks()
    qn = nb.shape[1] - 1  # DXQY::nQNonZero_
    ndir, dirs = [], []
    for g in ghosts:
        qs = [q for q in range(qn) if ranks[nb[g, q]] == rank]
        ndir.append(len(qs))
        dirs.extend(qs)
    return np.array(ndir, dtype=np.int32), np.array(dirs, dtype=np.int32)


def exchange_lists(topos):
    """topos[r]: the node list of rank r (0-based, as its PROCESSOR blocks
    number the ranks).  Returns, per rank, a list of (neighbour rank,
    nodes_to_send, n_dir_send, dir_send, nodes_received, n_dir_recv,
    dir_recv), the arguments of Context.halo_add."""
    seen = {}
    for r, t in enumerate(topos):
        for s, local, remote in t.neighbor_ranks():
            ndir, dirs = dir_lists(t, r, local)
            seen[(r, s)] = (np.asarray(local, np.int32), np.asarray(remote, np.int32), ndir, dirs)
    out = []
    for r in range(len(topos)):
        lists = []
        for (a, s), (local, _remote, ndir, dirs) in sorted(seen.items()):
            if a != r:
                continue
            if (s, r) not in seen:
                raise ValueError("rank %d lists rank %d as a neighbour, but not the other way round" % (r, s))
            _l2, remote2, ndir2, dirs2 = seen[(s, r)]
            lists.append((s, remote2, ndir2, dirs2, local, ndir, dirs))
        out.append(lists)
    return out


def exchange(contexts, registered, nf=1):
    """One halo exchange between contexts of one process: every rank packs
    for each neighbour and the neighbour unpacks (the first nf population
    fields).  registered[r] maps a neighbour rank to the exchange index of
    contexts[r] (Context.halo_add)."""
    packed = {}
    for r, ctx in enumerate(contexts):
        for s, k in registered[r].items():
            packed[(r, s)] = ctx.halo_pack(k, nf)
    for r, ctx in enumerate(contexts):
        for s, k in registered[r].items():
            ctx.halo_unpack(k, packed[(s, r)])


def exchange_colour(contexts, registered):
    """The colour exchange of a colour-gradient iteration
    (communciateScalarField of cgField, main_TWOPHASE.cpp:287)."""
    packed = {}
    for r, ctx in enumerate(contexts):
        for s, k in registered[r].items():
            packed[(r, s)] = ctx.halo_pack_colour(k)
    for r, ctx in enumerate(contexts):
        for s, k in registered[r].items():
            ctx.halo_unpack_colour(k, packed[(s, r)])


def step_twophase(contexts, registered, tau0, tau1, sigma, beta, force_mode, force):
    """One iteration of the colour-gradient loop over a split domain, every
    rank in this process (main_TWOPHASE.cpp:238-387 on each rank): the colour
    pass, the colour exchange, the flux summed over the ranks in rank order
    (MPI_Allreduce), the collision pass, the exchange of both population
    fields.  Every context needs lb_set_global_bulk_count for FORCE_MOMX."""
    momx = force_mode == FORCE_MOMX
    fluxes = [ctx.step_cg_colour(flux=momx) for ctx in contexts]
    exchange_colour(contexts, registered)
    total = 0.0
    if momx:
        for s in fluxes:
            total += s
    for ctx in contexts:
        ctx.step_cg_collide(tau0, tau1, sigma, beta, force_mode, force, total)
    exchange(contexts, registered, nf=2)
