"""Contexts whose population buffers exceed 4 GiB, against the oracle.

Byte offsets into such a buffer pass 2^32, so these runs check that no
address is formed in 32 bits on any path they take (element indices stay
32-bit: lb_set_topology refuses a context where two fields of Q
directions would not fit).

  * single phase, FP64, one field of 31 M bulk nodes (1.10 x 4 GiB): the
    row layout over 64-bit addresses (the row windows need byte offsets
    below 4 GiB and step aside); bit-exact against the oracle after 3
    steps from a perturbed state: populations, density, velocity;
  * two phase, FP64, two fields of 15.5 M bulk nodes (1.09 x 4 GiB; one
    field fits below 4 GiB, so the context holds the row windows): the
    colour gradient steps aside to the link table over 64-bit addresses,
    the flux reduction over both fields to the rows; the fixed-force
    colour gradient bit-exact against the oracle after 2 steps
    (populations, densities, velocity, colour), the total and per-phase
    fluxes within 1e-12 of the oracle's (another summation order).

The geometry is a sphere pack of 128 x 96 x 256 cells repeated along the
row axis (16 and 8 times); periodic along all three axes, or, for the
row windows, along the other two with walls closing the row axis.
"""
import os

import numpy as np
import pytest

import lbamd
import oracle

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
W19 = oracle.WEIGHTS[1]


def big_pack(repeat, periodic="xyz"):
    geo, _ = lbamd.sphere_pack((128, 96, 256), 6.0, 0.62, 4, periodic=periodic, walls=periodic != "xyz")
    big = np.tile(np.asarray(geo), (1, 1, repeat))
    T = lbamd.Topology.from_geometry(big, lbamd.D3Q19, periodic)
    P = {"nb": T.neighbors(), "types": T.node_types(), "bulk": T.bulk(), "pos": T.positions(), "n": T.n_nodes}
    T.close()
    return P


def test_single_phase_above_4gib_vs_oracle(gpu_available):
    P = big_pack(16)
    b = P["bulk"]
    assert 19 * 8 * b.size > 2**32
    print("single phase: %d bulk nodes, %.2f x 4 GiB per buffer" % (b.size, 19 * 8 * b.size / 2**32), flush=True)
    rng = np.random.default_rng(3)
    f0 = np.zeros((P["n"], 19))
    f0[b] = W19[None] * (1.0 + 0.01 * rng.random(b.size))[:, None]
    f0[b] += 1e-4 * rng.standard_normal((b.size, 19))
    force = np.array([1e-6, 2e-7, -1e-7])
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP64)
    ctx.set_topology_grid(P["nb"], P["types"], b, P["pos"])
    assert ctx.layout == 1 and "4 GiB" in ctx.layout_note, ctx.layout_note
    ctx.upload_f(f0)
    ctx.step_bgk(1, 1.0, force)
    ctx.step_bgk(2, 1.0, force)
    f = ctx.download_f(1)[b, 0]
    r = ctx.download_scalar(lbamd.RHO)[b]
    v = ctx.download_vel()[b]
    ctx.close()
    print("device done; oracle on %d threads" % THREADS, flush=True)
    fo, ro, vo = oracle.run_std_case(1, P["nb"], P["types"], b, f0, 1.0, force, 3, nthreads=THREADS)
    assert np.array_equal(f, fo[b])
    assert np.array_equal(r, ro[b])
    assert np.array_equal(v, vo[b])


def test_two_phase_above_4gib_vs_oracle(gpu_available):
    P = big_pack(8, periodic="xy")
    b = P["bulk"]
    assert 2 * 19 * 8 * b.size > 2**32 > 19 * 8 * b.size
    print("two phase: %d bulk nodes, %.2f x 4 GiB per buffer" % (b.size, 2 * 19 * 8 * b.size / 2**32), flush=True)
    rng = np.random.default_rng(5)
    # phases drawn per node; solid-boundary nodes carry wettability 1
    rho = np.zeros((P["n"], 2))
    r0 = (rng.random(b.size) > 0.5).astype(float)
    rho[b, 0] = r0
    rho[b, 1] = 1.0 - r0
    sb = P["types"] == 1
    rho[sb, 0] = 1.0
    f0 = np.zeros((P["n"], 2, 19))
    f0[b, 0] = W19[None] * r0[:, None]
    f0[b, 1] = W19[None] * (1.0 - r0)[:, None]
    tau0, tau1, sigma, beta = 0.8, 1.2, 0.01, 0.7
    force = [3e-6, 1e-6, -2e-6]
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP64)
    ctx.set_topology_grid(P["nb"], P["types"], b, P["pos"])
    assert ctx.layout == 2, ctx.layout_note
    ctx.upload_f(f0)
    ctx.set_wall_densities(rho)
    ctx.step_cg(2, tau0, tau1, sigma, beta, lbamd.FORCE_FIXED, force)
    f = ctx.download_f(2)[b]
    r = ctx.download_scalar(lbamd.RHO, 2)[b]
    v = ctx.download_vel()[b]
    phi = ctx.download_scalar(lbamd.PHI)[b]
    flux = [np.array(ctx.reduce_flux(d)) for d in range(3)]
    ctx.close()
    print("device done; oracle", flush=True)
    fo, ro, vo, cgo, _ = oracle.run_twophase(1, P["nb"], P["types"], b, f0, rho, tau0, tau1, sigma, beta, 0, force, 2)
    assert np.array_equal(f, fo[b])
    assert np.array_equal(r, ro[b])
    assert np.array_equal(v, vo[b])
    assert np.array_equal(phi, cgo[b])
    for d in range(3):
        ref = oracle.flux(1, b, fo, 2, d)
        assert np.all(np.abs(flux[d] - ref) <= 1e-12 * np.abs(ref).max()), (d, flux[d], ref)
