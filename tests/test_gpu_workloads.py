"""The batched ensemble and the serving stream at their own sizes.

Bars:
* Batched ensemble (512 sphere-pack permeability solves, cubic sizes
  log-uniform in 32..128): a size-stratified draw of 64 of its samples
  admitted into one batch gives, per solve, k within 1e-4 (relative) of the
  plain FP64 single-context solve in FP32S and in FP16S; on a subset the
  batch's k and step count equal the single-context solve of the same
  precision bit for bit.
* Serving stream (sizes 48..160, 70 % permeability, 30 % relative
  permeability): six of its requests, both kinds, served together in FP32S
  give permeabilities within 1e-4 of the FP64 single-context solve; the
  relative-permeability requests (served in FP32S storage, computed in FP64)
  equal the FP32S single-context solve (lb_solve_relperm) bit for bit, phase
  permeabilities and step counts, and their phase permeabilities are within
  1e-4 of k0 + k1 of the FP64 single-context solve.
* A two-phase solve in an FP64 batch equals lb_solve_relperm bit for bit,
  alone and beside single-phase solves in a reduced precision.
"""
import numpy as np
import pytest

import lbamd
from lbamd.workloads import TWO_PHASE, ensemble, request_mix

pytestmark = pytest.mark.gpu

W19 = np.array([2.0] * 3 + [1.0] * 6 + [2.0] * 3 + [1.0] * 6 + [12.0]) / 36.0
F = np.array([1e-6, 0.0, 0.0])


def single_perm(geo, precision, tol, max_steps, check_every=100):
    topo = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    ctx = lbamd.Context(lbamd.D3Q19, precision)
    ctx.set_topology_grid(nb, types, bulk, pos)
    assert ctx.layout == 2, ctx.layout_note
    f0 = np.zeros((topo.n_nodes, 19))
    f0[bulk] = W19[None, :]
    ctx.upload_f(f0)
    k, steps = ctx.solve_permeability(1.0, F, check_every, tol, max_steps, float(geo.size))
    ctx.close()
    topo.close()
    return k, steps


def single_relperm(geo, rho0, precision, tol, max_steps, tp=TWO_PHASE, check_every=100):
    """lb_solve_relperm from f_k = w rho_k at rest, solid-boundary colour from
    the wettability."""
    topo = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    n = topo.n_nodes
    r0 = np.zeros(n)
    r0[bulk] = rho0[pos[bulk, 0], pos[bulk, 1], pos[bulk, 2]]
    r1 = 1.0 - r0
    sb = types == 1
    r0[sb], r1[sb] = tp["wett"], 1.0 - tp["wett"]
    f0 = np.zeros((n, 2, 19))
    f0[bulk, 0] = W19[None] * r0[bulk, None]
    f0[bulk, 1] = W19[None] * r1[bulk, None]
    ctx = lbamd.Context(lbamd.D3Q19, precision)
    ctx.set_topology_grid(nb, types, bulk, pos)
    assert ctx.layout == 2, ctx.layout_note
    ctx.upload_f(f0)
    ctx.set_wall_densities(np.stack([r0, r1], 1))
    kp, steps = ctx.solve_relperm(tp["tau0"], tp["tau1"], tp["sigma"], tp["beta"], F, check_every, tol, max_steps,
                                  float(geo.size))
    ctx.close()
    topo.close()
    return kp, steps


def drain(b, n):
    out = {}
    while len(out) < n:
        b.advance(1)
        for r in b.collect():
            out[r["tag"]] = r
    return out


def test_twophase_fp64_batch_equals_single(gpu_available):
    geos = [lbamd.sphere_pack(s, r, p, seed, periodic="x", walls=True)[0]
            for s, r, p, seed in (((28, 24, 22), 3.0, 0.45, 31), ((22, 26, 30), 3.5, 0.5, 32))]
    rho0s = [np.where(g > 0, (np.random.default_rng(40 + i).random(g.shape) < 0.4 + 0.2 * i).astype(float), 0.0)
             for i, g in enumerate(geos)]
    tol, cap = 1e-5, 2500
    ref = [single_relperm(g, r, lbamd.FP64, tol, cap) for g, r in zip(geos, rho0s)]
    tp = TWO_PHASE
    for mixed in (False, True):
        b = lbamd.Batch(lbamd.D3Q19, lbamd.FP32S if mixed else lbamd.FP64, 1 << 28, 100)
        b.set_precision("relperm", lbamd.FP64)
        for i, (g, r) in enumerate(zip(geos, rho0s)):
            assert b.add_twophase(g, "x", r, tp["wett"], tp["tau0"], tp["tau1"], tp["sigma"], tp["beta"], F, tol,
                                  cap, i)
        if mixed:
            assert b.add(geos[0], "x", 1.0, F, 1e-6, 4000, 9)
        got = drain(b, 3 if mixed else 2)
        b.close()
        for i in range(2):
            assert got[i]["k_phase"] == ref[i][0] and got[i]["steps"] == ref[i][1], (mixed, i, got[i], ref[i])


def test_ensemble_stratified_draw(gpu_available):
    n_all = 512
    L = np.rint(np.exp(np.random.default_rng(2024).uniform(np.log(32), np.log(128), n_all))).astype(int)
    order = np.argsort(L, kind="stable")
    idx = order[4::8]  # 64 samples, every 8th by size
    geos, par = ensemble(n_all, indices=idx)
    assert len(geos) == 64 and par["L"].min() < 40 and par["L"].max() > 118
    print("ensemble draw: 64 samples, L %d..%d, %.1f M cells" % (par["L"].min(), par["L"].max(),
                                                                  sum(g.size for g in geos) / 1e6), flush=True)
    tol, cap = 1e-6, 400000
    res = {}
    for prec in (lbamd.FP32S, lbamd.FP16S):
        pool = int(sum(lbamd.batch_footprint(lbamd.D3Q19, prec, g, "x")[0] for g in geos) * 1.02) + (16 << 20)
        b = lbamd.Batch(lbamd.D3Q19, prec, pool, 100)
        assert b.add_many(geos, "x", 1.0, F, tol, cap, list(range(len(geos)))) == len(geos)
        res[prec] = drain(b, len(geos))
        b.close()
        assert all(r["status"] == 0 for r in res[prec].values())
    err = {lbamd.FP32S: [], lbamd.FP16S: []}
    for j, g in enumerate(geos):
        k64, _ = single_perm(g, lbamd.FP64, tol, cap)
        for prec in err:
            err[prec].append(abs(res[prec][j]["k"] - k64) / abs(k64))
    print("k error vs FP64: FP32S max %.2e, FP16S max %.2e" % (max(err[lbamd.FP32S]), max(err[lbamd.FP16S])),
          flush=True)
    assert max(err[lbamd.FP32S]) <= 1e-4
    assert max(err[lbamd.FP16S]) <= 1e-4
    # batch = single context, bit for bit, on the smallest, a middle and the largest sample
    for j in (0, 31, 63):
        for prec in (lbamd.FP32S, lbamd.FP16S):
            k, steps = single_perm(geos[j], prec, tol, cap)
            assert (res[prec][j]["k"], res[prec][j]["steps"]) == (k, steps), (j, prec)


def test_serving_requests_at_stream_sizes(gpu_available):
    from lbamd.serve import Server, run_stream
    idx = [18, 10, 26, 8, 31, 0]  # sizes 61, 80, 159 (permeability), 54, 110, 154 (relative permeability)
    cap_two = 6000  # bounds the test's time; the two-phase flows are still moving then

    def reqs():
        rs = request_mix(48, 7, indices=idx)
        for r in rs:
            if r.kind == "relperm":
                r.max_steps = cap_two
        return rs

    kinds = {r.tag: r.kind for r in reqs()}
    assert sorted(kinds.values()) == ["perm"] * 3 + ["relperm"] * 3
    srv = Server(12 << 30, precision=lbamd.FP32S, check_every=100)
    done, _ = run_stream(srv, reqs())
    srv.close()
    got = {d["tag"]: d for d in done}
    assert sorted(got) == sorted(idx)
    for r in reqs():
        d = got[r.tag]
        print("request %d (%s, %d^3): %d steps" % (r.tag, r.kind, r.geo.shape[0], d["steps"]), flush=True)
        if r.kind == "perm":
            assert d["status"] == 0
            k64, _ = single_perm(r.geo, lbamd.FP64, r.tol, r.max_steps)
            assert abs(d["k"] - k64) <= 1e-4 * abs(k64), (r.tag, d["k"], k64)
        else:
            kp, steps = single_relperm(r.geo, r.rho0, lbamd.FP32S, r.tol, r.max_steps)
            assert d["k_phase"] == kp and d["steps"] == steps, (r.tag, d, kp, steps)
            k64, steps64 = single_relperm(r.geo, r.rho0, lbamd.FP64, r.tol, r.max_steps)
            assert steps64 == steps, (r.tag, steps64, steps)
            err = max(abs(kp[0] - k64[0]), abs(kp[1] - k64[1])) / (abs(k64[0]) + abs(k64[1]))
            print("  phase permeabilities FP32S %s FP64 %s: error %.2e of k0 + k1" % (kp, k64, err), flush=True)
            assert err <= 1e-4, (r.tag, kp, k64)
