"""Batched form: many permeability solves resident in one pool.

Bars: every solve's k and step count are bit-identical to the single-context
path (lb_solve_permeability with row windows, same precision), alone in a
batch and in any mix, whatever launch shape the batch uses; admission follows
the pool's free memory.  FP64 batch results therefore also carry the
single path's parity with the reference loop.
"""
import numpy as np
import pytest

import lbamd

pytestmark = pytest.mark.gpu

TAU, FORCE = 1.0, np.array([1e-6, 0.0, 0.0])


def pack(shape, radius, porosity, seed):
    geo, _ = lbamd.sphere_pack(shape, radius, porosity, seed, periodic="x", walls=True)
    return geo


def single(geo, precision, check_every, tol, max_steps):
    topo = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    ctx = lbamd.Context(lbamd.D3Q19, precision)
    ctx.set_topology_grid(nb, types, bulk, pos)
    assert ctx.layout == 2, ctx.layout_note
    W = np.array([2.0] * 3 + [1.0] * 6 + [2.0] * 3 + [1.0] * 6 + [12.0]) / 36.0
    f0 = np.zeros((topo.n_nodes, 19))
    f0[bulk] = W[None, :]
    ctx.upload_f(f0)
    k, steps = ctx.solve_permeability(TAU, FORCE, check_every, tol, max_steps, float(geo.size))
    ctx.close()
    return k, steps


def run_batch(geos, precision, check_every, tol, max_steps, shape=-1, pool=1 << 28, late=None):
    """Admit geos (tag = index); `late` maps an interval number to indices that
    are admitted only then.  Returns {tag: (k, steps)}."""
    b = lbamd.Batch(lbamd.D3Q19, precision, pool, check_every)
    b.set_shape(shape)
    late = late or {}
    held = {i for v in late.values() for i in v}
    for i, g in enumerate(geos):
        if i not in held:
            assert b.add(g, "x", TAU, FORCE, tol, max_steps, i)
    out = {}
    interval = 0
    while len(out) < len(geos):
        for i in late.get(interval, []):
            assert b.add(geos[i], "x", TAU, FORCE, tol, max_steps, i)
        b.advance(1)
        interval += 1
        for r in b.collect():
            out[r["tag"]] = (r["k"], r["steps"], r["status"])
        assert interval < 10 * max_steps // check_every
    b.close()
    return out


@pytest.fixture(scope="module")
def geos():
    return [pack((24, 20, 20), 3.0, 0.45, 1), pack((32, 28, 24), 4.0, 0.40, 2), pack((20, 24, 36), 3.0, 0.5, 3)]


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S, lbamd.FP16S])
def test_batch_equals_single_context(gpu_available, geos, precision):
    got = run_batch(geos, precision, 50, 1e-6, 20000)
    for i, g in enumerate(geos):
        k, steps = single(g, precision, 50, 1e-6, 20000)
        assert got[i][2] == 0
        assert got[i][0] == k and got[i][1] == steps, (i, got[i], k, steps)


@pytest.mark.parametrize("precision", [lbamd.FP32S, lbamd.FP16S])
def test_batch_alone_equals_mix_and_shapes(gpu_available, geos, precision):
    """FP16S also checks the stochastic rounding's step count: it is the
    solve's own, whenever the solve was admitted."""
    extra = [pack((28, 28, 28), 3.5, 0.42, 9), pack((16, 16, 16), 2.5, 0.5, 10)]
    alone = run_batch([geos[0]], precision, 50, 1e-6, 20000)[0]
    # in a mix, admitted later than the others, under each launch shape
    for shape in (-1, 0, 1, 3):
        mix = run_batch(extra + [geos[0], geos[1]], precision, 50, 1e-6, 20000, shape=shape, late={3: [2]})
        assert mix[2] == alone, (shape, mix[2], alone)


def test_admission_by_footprint(gpu_available, geos):
    sizes = [lbamd.batch_footprint(lbamd.D3Q19, lbamd.FP32S, g, "x")[0] for g in geos]
    pool = sizes[0] + sizes[1] + 1024
    b = lbamd.Batch(lbamd.D3Q19, lbamd.FP32S, pool, 50)
    assert b.add(geos[0], "x", TAU, FORCE, 1e-6, 20000, 0)
    assert b.add(geos[1], "x", TAU, FORCE, 1e-6, 20000, 1)
    st = b.status()
    assert st["resident"] == 2 and st["bytes_used"] == sizes[0] + sizes[1]
    assert not b.add(geos[2], "x", TAU, FORCE, 1e-6, 20000, 2)
    done = []
    while not done:
        b.advance(1)
        done = b.collect()
    # a retired solve frees its region; the waiting one fits if the freed
    # memory suffices
    freed = sum(sizes[r["tag"]] for r in done)
    assert b.status()["bytes_used"] == sizes[0] + sizes[1] - freed
    if freed + 1024 >= sizes[2]:
        assert b.add(geos[2], "x", TAU, FORCE, 1e-6, 20000, 2)
    b.close()


def test_add_many_equals_one_by_one(gpu_available, geos):
    """lb_batch_add_many (row layouts built on host threads) admits the same
    solves as lb_batch_add one by one: same results, same footprint; it
    stops at the first solve that does not fit and reports how many went in."""
    extra = [pack((28, 28, 28), 3.5, 0.42, 9), pack((16, 16, 16), 2.5, 0.5, 10)]
    gs = geos + extra
    ref = run_batch(gs, lbamd.FP32S, 50, 1e-6, 20000)
    b = lbamd.Batch(lbamd.D3Q19, lbamd.FP32S, 1 << 28, 50)
    assert b.add_many(gs, "x", TAU, FORCE, 1e-6, 20000, list(range(len(gs)))) == len(gs)
    out = {}
    while len(out) < len(gs):
        b.advance(1)
        for r in b.collect():
            out[r["tag"]] = (r["k"], r["steps"], r["status"])
    b.close()
    assert out == ref
    sizes = [lbamd.batch_footprint(lbamd.D3Q19, lbamd.FP32S, g, "x")[0] for g in gs]
    b = lbamd.Batch(lbamd.D3Q19, lbamd.FP32S, sizes[0] + sizes[1] + 1024, 50)
    assert b.add_many(gs, "x", TAU, FORCE, 1e-6, 20000, list(range(len(gs)))) == 2
    assert b.status()["bytes_used"] == sizes[0] + sizes[1]
    b.close()


def test_fp16s_batch_error(gpu_available, geos):
    ref = run_batch(geos, lbamd.FP64, 50, 1e-6, 20000)
    got = run_batch(geos, lbamd.FP16S, 50, 1e-6, 20000)
    for i in ref:
        print("FP16S batch k err %.3e" % (abs(got[i][0] - ref[i][0]) / ref[i][0]))
        assert abs(got[i][0] - ref[i][0]) / ref[i][0] < 1e-4


def tp_single(geo, rho0, wett, precision, tau0, tau1, sigma, beta, steps):
    """Phase permeabilities of the single-context two-phase path (fixed force)
    over the state after steps - 1 steps, as the batch reports them."""
    topo = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    n = topo.n_nodes
    r0 = np.zeros(n)
    r0[bulk] = rho0[pos[bulk, 0], pos[bulk, 1], pos[bulk, 2]]
    r1 = 1.0 - r0
    sb = types == 1
    r0[sb], r1[sb] = wett, 1.0 - wett
    W = np.array([2.0] * 3 + [1.0] * 6 + [2.0] * 3 + [1.0] * 6 + [12.0]) / 36.0
    f0 = np.zeros((n, 2, 19))
    f0[bulk, 0] = W[None] * r0[bulk, None]
    f0[bulk, 1] = W[None] * r1[bulk, None]
    ctx = lbamd.Context(lbamd.D3Q19, precision)
    ctx.set_topology_grid(nb, types, bulk, pos)
    ctx.upload_f(f0)
    ctx.set_wall_densities(np.stack([r0, r1], 1))
    ctx.step_cg(steps - 1, tau0, tau1, sigma, beta, lbamd.FORCE_FIXED, FORCE)
    s0, s1, s2 = ctx.reduce_flux(0)
    ctx.close()
    nbox = float(geo.size)
    return (tau0 - 0.5) / 3.0 * (s1 / nbox) / FORCE[0], (tau1 - 0.5) / 3.0 * (s2 / nbox) / FORCE[0]


@pytest.mark.parametrize("precision", [lbamd.FP32S, lbamd.FP16S])
def test_batch_twophase_equals_single_context(gpu_available, geos, precision):
    """Two-phase solves in a batch (mixed with single-phase ones, admitted at
    different times) give the phase permeabilities of the single-context
    two-phase path bit for bit."""
    rng = np.random.default_rng(5)
    params = (1.0, 0.8, 0.01, 0.7)
    steps = 200
    rho0s = [np.where(g > 0, (rng.random(g.shape) < 0.5).astype(float), 0.0) for g in geos[:2]]
    b = lbamd.Batch(lbamd.D3Q19, precision, 1 << 28, steps)
    assert b.add(geos[2], "x", TAU, FORCE, 1e-6, 20000, 99)
    assert b.add_twophase(geos[0], "x", rho0s[0], 1.0, *params, FORCE, 1e-12, steps, 0)
    b.advance(1)
    assert b.add_twophase(geos[1], "x", rho0s[1], 0.7, *params, FORCE, 1e-12, steps, 1)
    got = {}
    while len(got) < 3:
        b.advance(1)
        for r in b.collect():
            got[r["tag"]] = r
    b.close()
    for i, wett in ((0, 1.0), (1, 0.7)):
        k0, k1 = tp_single(geos[i], rho0s[i], wett, precision, *params, steps)
        assert got[i]["status"] == 1 and got[i]["steps"] == steps
        assert got[i]["k_phase"] == (k0, k1), (i, got[i]["k_phase"], (k0, k1))
    # the single-phase solve is unaffected by its two-phase neighbours
    alone = run_batch([geos[2]], precision, steps, 1e-6, 20000)[0]
    assert (got[99]["k"], got[99]["steps"]) == alone[:2]


def test_batch_d2q9_equals_single_context(gpu_available):
    """The batch runs D2Q9 solves too: channels of different widths, each
    bit-identical to its single-context solve."""
    from lbamd.geometry import channel_2d
    W9 = np.array([4.0, 1.0, 4.0, 1.0, 4.0, 1.0, 4.0, 1.0, 16.0]) / 36.0
    F = np.array([1e-6, 0.0])
    geos = [channel_2d(32, 18), channel_2d(48, 26), channel_2d(20, 12)]
    b = lbamd.Batch(lbamd.D2Q9, lbamd.FP32S, 1 << 24, 50)
    for i, g in enumerate(geos):
        assert b.add(g, "x", 1.0, F, 1e-6, 50000, i)
    got = {}
    while len(got) < len(geos):
        b.advance(1)
        for r in b.collect():
            got[r["tag"]] = (r["k"], r["steps"])
    b.close()
    for i, g in enumerate(geos):
        topo = lbamd.Topology.from_geometry(g, lbamd.D2Q9, "x")
        nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
        ctx = lbamd.Context(lbamd.D2Q9, lbamd.FP32S)
        ctx.set_topology_grid(nb, types, bulk, pos)
        assert ctx.layout == 2, ctx.layout_note
        f0 = np.zeros((topo.n_nodes, 9))
        f0[bulk] = W9[None, :]
        ctx.upload_f(f0)
        k, steps = ctx.solve_permeability(1.0, F, 50, 1e-6, 50000, float(g.size))
        ctx.close()
        assert got[i] == (k, steps), (i, got[i], k, steps)


def test_batch_rejections(gpu_available, geos):
    """Geometries the row windows cannot address and bad arguments are
    refused with a message; the precision of a kind cannot change while
    solves of that kind are resident."""
    b = lbamd.Batch(lbamd.D3Q19, lbamd.FP64, 1 << 26, 50)
    rho0 = (geos[0] > 0).astype(float)
    assert b.add_twophase(geos[0], "x", rho0, 1.0, 1.0, 1.0, 0.01, 0.7, FORCE, 1e-4, 1000, 0)
    with pytest.raises(lbamd.LBError) as e:
        b.set_precision("relperm", lbamd.FP32S)
    assert e.value.code == -2
    b.set_precision("perm", lbamd.FP16S)  # no single-phase solve resident
    b.close()
    b = lbamd.Batch(lbamd.D3Q19, lbamd.FP64, 1 << 26, 50)
    with pytest.raises(lbamd.LBError) as e:
        b.add(geos[0], "xyz", TAU, FORCE, 1e-6, 1000, 1)  # periodic along the row axis
    assert e.value.code == -4
    with pytest.raises(lbamd.LBError) as e:
        b.add(geos[0], "x", 0.4, FORCE, 1e-6, 1000, 2)  # tau <= 1/2
    assert e.value.code == -1
    assert b.status()["resident"] == 0
    b.close()


def test_serving_results_equal_alone(gpu_available, geos):
    """Continuous batching: a stream of permeability and relative-permeability
    requests arriving over time through lbamd.serve gives, for every request,
    the bits of the same request served alone."""
    from lbamd.serve import Request, Server, run_stream
    extra = [pack((26, 22, 24), 3.0, 0.45, 21), pack((18, 20, 16), 2.5, 0.5, 22)]
    allg = geos + extra

    def requests():
        out = []
        for i, g in enumerate(allg):
            t = 0.002 * i
            if i % 2:
                rho0 = np.where(g > 0, (np.random.default_rng(i).random(g.shape) < 0.5).astype(float), 0.0)
                out.append(Request(i, g, "x", force=FORCE, tol=1e-4, max_steps=1500, t_arrival=t, kind="relperm",
                                   rho0=rho0, wett=1.0, tau0=1.0, tau1=0.9, sigma=0.01, beta=0.7))
            else:
                out.append(Request(i, g, "x", tau=TAU, force=FORCE, tol=1e-6, max_steps=20000, t_arrival=t))
        return out

    srv = Server(1 << 28, precision=lbamd.FP32S, check_every=50)
    mix, _ = run_stream(srv, requests())
    srv.close()
    mix = {d["tag"]: d for d in mix}
    assert sorted(mix) == list(range(len(allg)))
    for r in requests():
        solo = Server(1 << 28, precision=lbamd.FP32S, check_every=50)
        r.t_arrival = 0.0
        (d,), _ = run_stream(solo, [r])
        solo.close()
        m = mix[r.tag]
        assert (d["k"], d["k_phase"], d["steps"], d["status"]) == (m["k"], m["k_phase"], m["steps"], m["status"]), r.tag
