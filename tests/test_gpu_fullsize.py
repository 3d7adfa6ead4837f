"""The benchmark's full sizes, checked directly against the oracle for a few
steps (it finishes them in seconds) and through properties that do not
depend on size.

  256^3 sphere pack (D3Q19, radius 8, porosity 0.40, 6.6 M bulk nodes):
    * FP64 bit-exact against the oracle after 3 steps from a perturbed state
      (populations, density, velocity);
    * the link-table and row-window layouts give identical bits after 20
      steps (FP64, FP32S);
    * total mass (sum over bulk nodes and directions) conserved over 200
      FP64 steps to 1e-12 relative: streaming, half-way bounce-back, the
      collision and the Guo term each conserve it;
    * FP32S and FP16S stepped in chunks equal stepped at once, bit for bit;
    * FP32S velocity within 1e-4 of FP64 (norm-wise) after 200 steps,
      FP16S within 2e-2: its error is rounding noise, and the maximum over
      6.6 M nodes sits higher than over the small packs of the other tests
      (1.2e-2 measured).
  192^3 two-phase pack (porosity 0.38, saturation 0.5): FP64 fixed-force
    colour gradient bit-exact against the oracle after 3 steps.
  128^3 (the ensemble's largest size): a batched solve equals its
    single-context solve (k and step count), FP32S.
"""
import os

import numpy as np
import pytest

import lbamd
import oracle

pytestmark = pytest.mark.gpu

TAU, F3 = 1.0, np.array([1e-6, 0.0, 0.0])
W19 = oracle.WEIGHTS[1]
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def pack256():
    geo, _ = lbamd.sphere_pack((256,) * 3, 8.0, 0.40, 1, periodic="x", walls=True)
    T = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    out = {"geo": geo, "nb": T.neighbors(), "types": T.node_types(), "bulk": T.bulk(), "pos": T.positions(),
           "n": T.n_nodes}
    rng = np.random.default_rng(11)
    f0 = np.zeros((T.n_nodes, 19))
    b = out["bulk"]
    f0[b] = W19[None] * (1.0 + 0.01 * rng.random(b.size))[:, None]
    f0[b] += 1e-4 * rng.standard_normal((b.size, 19))
    out["f0"] = f0
    return out


def context(P, precision, layout="rows"):
    ctx = lbamd.Context(lbamd.D3Q19, precision)
    if layout == "rows":
        ctx.set_topology_grid(P["nb"], P["types"], P["bulk"], P["pos"])
        assert ctx.layout == 2, ctx.layout_note
    else:
        ctx.set_topology(P["nb"], P["types"], P["bulk"])
    ctx.upload_f(P["f0"])
    return ctx


def run(P, precision, chunks, layout="rows", force=F3):
    ctx = context(P, precision, layout)
    for n in chunks:
        ctx.step_bgk(n, TAU, force)
    out = ctx.download_f(1)[:, 0], ctx.download_scalar(lbamd.RHO), ctx.download_vel()
    ctx.close()
    return out


def test_256_fp64_bitexact_vs_oracle(gpu_available, pack256):
    P = pack256
    force = np.array([1e-6, 2e-7, -1e-7])
    print("256^3: %d bulk nodes, oracle on %d threads" % (P["bulk"].size, THREADS), flush=True)
    fo, ro, vo = oracle.run_std_case(1, P["nb"], P["types"], P["bulk"], P["f0"], TAU, force, 3, nthreads=THREADS)
    f, r, v = run(P, lbamd.FP64, [1, 2], force=force)
    b = P["bulk"]
    assert np.array_equal(f[b], fo[b])
    assert np.array_equal(r[b], ro[b])
    assert np.array_equal(v[b], vo[b])


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S])
def test_256_layouts_identical(gpu_available, pack256, precision):
    b = pack256["bulk"]
    a = run(pack256, precision, [20], layout="table")
    c = run(pack256, precision, [20], layout="rows")
    for x, y in zip(a, c):
        assert np.array_equal(x[b], y[b])


def test_256_mass_conserved(gpu_available, pack256):
    b = pack256["bulk"]
    m0 = pack256["f0"][b].sum()
    f, _, _ = run(pack256, lbamd.FP64, [200])
    m = f[b].sum()
    print("mass %.17g -> %.17g (rel %.2e)" % (m0, m, abs(m - m0) / m0))
    assert abs(m - m0) <= 1e-12 * m0


@pytest.mark.parametrize("precision", [lbamd.FP32S, lbamd.FP16S])
def test_256_chunking_identical(gpu_available, pack256, precision):
    b = pack256["bulk"]
    a = run(pack256, precision, [100])
    c = run(pack256, precision, [1, 36, 63])
    for x, y in zip(a, c):
        assert np.array_equal(x[b], y[b])


def test_256_reduced_precision_vs_fp64(gpu_available, pack256):
    b = pack256["bulk"]
    _, _, v64 = run(pack256, lbamd.FP64, [200])
    scale = np.abs(v64[b]).max()
    for precision, tol in ((lbamd.FP32S, 1e-4), (lbamd.FP16S, 2e-2)):
        _, _, v = run(pack256, precision, [200])
        err = np.abs(v[b] - v64[b]).max() / scale
        print("precision %d: velocity error %.3e" % (precision, err))
        assert err < tol


def test_192_twophase_fp64_bitexact_vs_oracle(gpu_available):
    geo, _ = lbamd.sphere_pack((192,) * 3, 8.0, 0.38, 3, periodic="x", walls=True)
    T = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    nb, types, bulk, pos, n = T.neighbors(), T.node_types(), T.bulk(), T.positions(), T.n_nodes
    rng = np.random.default_rng(4)
    r0 = np.zeros(n)
    r0[bulk] = (rng.random(bulk.size) < 0.5).astype(float)
    r1 = np.zeros(n)
    r1[bulk] = 1.0 - r0[bulk]
    sb = types == 1
    r0[sb], r1[sb] = 1.0, 0.0
    rho = np.stack([r0, r1], 1)
    f0 = np.zeros((n, 2, 19))
    f0[bulk, 0] = W19[None] * r0[bulk, None]
    f0[bulk, 1] = W19[None] * r1[bulk, None]
    tau0, tau1, sigma, beta, force = 1.0, 1.0, 0.01, 0.7, [1e-6, 0.0, 0.0]
    print("192^3 two-phase: %d bulk nodes" % bulk.size, flush=True)
    fo, ro, vo, cgo, _ = oracle.run_twophase(1, nb, types, bulk, f0, rho, tau0, tau1, sigma, beta, 0, force, 3)
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP64)
    ctx.set_topology_grid(nb, types, bulk, pos)
    ctx.upload_f(f0)
    ctx.set_wall_densities(rho)
    ctx.step_cg(3, tau0, tau1, sigma, beta, lbamd.FORCE_FIXED, force)
    f, r, v, phi = ctx.download_f(2), ctx.download_scalar(lbamd.RHO, 2), ctx.download_vel(), \
        ctx.download_scalar(lbamd.PHI)
    ctx.close()
    assert np.array_equal(f[bulk], fo[bulk])
    assert np.array_equal(r[bulk], ro[bulk])
    assert np.array_equal(v[bulk], vo[bulk])
    assert np.array_equal(phi[bulk], cgo[bulk])


def test_128_batched_equals_single(gpu_available):
    big, _ = lbamd.sphere_pack((128,) * 3, 8.0, 0.40, 5, periodic="x", walls=True)
    small, _ = lbamd.sphere_pack((40, 36, 32), 4.0, 0.45, 6, periodic="x", walls=True)
    T = lbamd.Topology.from_geometry(big, lbamd.D3Q19, "x")
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP32S)
    ctx.set_topology_grid(T.neighbors(), T.node_types(), T.bulk(), T.positions())
    f0 = np.zeros((T.n_nodes, 19))
    f0[T.bulk()] = W19[None]
    ctx.upload_f(f0)
    k, steps = ctx.solve_permeability(TAU, F3, 50, 1e-6, 40000, float(big.size))
    ctx.close()
    b = lbamd.Batch(lbamd.D3Q19, lbamd.FP32S, 1 << 30, 50)
    assert b.add(small, "x", TAU, F3, 1e-6, 40000, 1)
    assert b.add(big, "x", TAU, F3, 1e-6, 40000, 0)
    got = {}
    while len(got) < 2:
        b.advance(1)
        for r in b.collect():
            got[r["tag"]] = r
    b.close()
    assert got[0]["status"] == 0 and steps < 40000
    assert (got[0]["k"], got[0]["steps"]) == (k, steps)
