"""Reduced-precision forms (shifted populations f - w_q stored in FP32 or FP16,
float arithmetic for the single-phase step, FP64 arithmetic for the two-phase
step, FP64 reductions) against the FP64 oracle.

Tolerances (stated per quantity, norm-wise: max |a - b| / max |b|):
  FP32S  velocity 1e-4, density deviation 1e-3 (absolute 1e-9), permeability 1e-4
  FP16S  velocity 5e-3, permeability 1e-4 (stores round stochastically, so
         slow relaxations are not cut short by round-to-nearest)
two-phase (phase field absolute, velocity norm-wise, fluxes relative):
  FP32S  phi 1e-4, velocity 1e-4, fluxes 1e-5
  FP16S  phi 2e-3, velocity 2e-2, fluxes 2e-3
two-phase FP32S over 20 000 steps against the FP64 path: phase
permeabilities within 1e-4 of k0 + k1 at every 1000-step sample,
and, for every layout, the same precision gives bit-identical results
(the per-node arithmetic does not depend on how sources are addressed).
FP16S stores round stochastically: increments below half a unit of a
stored half are kept on average (test_fp16s_small_increments_are_not_lost).
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


@pytest.mark.parametrize("precision,vtol", [(lbamd.FP32S, 1e-4), (lbamd.FP16S, 5e-3)])
def test_reduced_precision_vs_oracle(gpu_available, precision, vtol):
    geo, T = pack((28, 24, 20), 4.0, 0.5, 21)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    force = np.array([2e-5, 0.0, 0.0])
    steps = 300
    fo, ro, vo = oracle.run_std_case(1, T.neighbors, T.node_types, T.bulk, f0, 0.9, force, steps)
    r, v, f = run(T, precision, f0, 0.9, force, steps, "rows")
    b = T.bulk
    print("vel err %.3e" % normerr(v[b], vo[b]))
    assert normerr(v[b], vo[b]) < vtol
    if precision == lbamd.FP32S:
        assert np.abs(r[b] - ro[b]).max() < max(1e-3 * np.abs(ro[b] - 1).max(), 1e-9)


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S, lbamd.FP16S])
def test_layouts_give_identical_results(gpu_available, precision):
    # 20 rows along y fit one 32-row tile, so the row layout keeps the bulk
    # order and FP16S, which seeds its rounding by a node's place in that
    # order, gives the same bits too
    geo, T = pack((24, 20, 18), 3.0, 0.5, 5)
    rng = np.random.default_rng(2)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    f0[T.bulk] += 1e-4 * rng.standard_normal((T.bulk.size, 19))
    a = run(T, precision, f0, 0.8, [1e-5, 2e-6, 0.0], 40, "table")
    b = run(T, precision, f0, 0.8, [1e-5, 2e-6, 0.0], 40, "rows")
    for x, y in zip(a, b):
        assert np.array_equal(x[T.bulk], y[T.bulk])


@pytest.mark.parametrize("precision,ktol", [(lbamd.FP32S, 1e-4), (lbamd.FP16S, 1e-4)])
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
    print("k err %.3e steps %d / %d" % (abs(k - k64) / k64, s, s64))
    assert abs(k - k64) / k64 < ktol


def test_permeability_matches_oracle_flux(gpu_available):
    """k from the device reduction equals nu <j_x> / (N_box F_x), <j_x> the
    average over the last check interval of sum_bulk(j_x), computed from the
    oracle's states."""
    geo, T = pack((20, 16, 16), 3.0, 0.5, 4)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    F = np.array([1e-6, 0.0, 0.0])
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP64)
    ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
    ctx.upload_f(f0)
    k, steps = ctx.solve_permeability(1.0, F, 50, 1e-30, 200, float(geo.size))
    ctx.close()
    assert steps == 200
    # j_x averaged over the states the steps of the last interval of 50
    # started from (after 150 .. 199 steps)
    jx = oracle_interval_jx(T, f0, 1.0, F, 150, 50)
    k_ref = (0.5 / 3.0) * jx / geo.size / F[0]
    assert abs(k - k_ref) <= 1e-12 * abs(k_ref)


def oracle_interval_jx(T, f0, tau, F, start, count):
    """Mean over `count` consecutive states, from the state after `start`
    steps on, of sum_bulk(rho u_x) = sum_bulk(sum_q f c_x + F_x/2)."""
    f, _, _ = oracle.run_std_case(1, T.neighbors, T.node_types, T.bulk, f0, tau, F, start)
    total = 0.0
    for _ in range(count):
        f, ro, vo = oracle.run_std_case(1, T.neighbors, T.node_types, T.bulk, f, tau, F, 1)
        total += (ro[T.bulk] * vo[T.bulk, 0]).sum()
    return total / count


def tp_case(shape, seed, sat=0.5, wett=1.0):
    """Sphere pack with phases drawn per pore node (Bernoulli(sat)), solid
    boundary wettability `wett` (rho0 = wett, rho1 = 1 - wett on type-1 nodes)."""
    geo, T = pack(shape, 3.0, 0.5, seed)
    rng = np.random.default_rng(seed)
    r0 = np.zeros(T.n_nodes)
    r0[T.bulk] = (rng.random(T.bulk.size) < sat).astype(float)
    r1 = np.zeros(T.n_nodes)
    r1[T.bulk] = 1.0 - r0[T.bulk]
    sb = T.node_types == 1
    r0[sb] = wett
    r1[sb] = 1.0 - wett
    W = oracle.WEIGHTS[1]
    f0 = np.zeros((T.n_nodes, 2, 19))
    f0[T.bulk, 0] = W[None] * r0[T.bulk, None]
    f0[T.bulk, 1] = W[None] * r1[T.bulk, None]
    return T, f0, np.stack([r0, r1], 1)


@pytest.mark.parametrize("precision,ptol,vtol,qtol", [(lbamd.FP32S, 1e-4, 1e-4, 1e-5),
                                                      (lbamd.FP16S, 2e-3, 2e-2, 2e-3)])
def test_twophase_reduced_precision_vs_oracle(gpu_available, precision, ptol, vtol, qtol):
    """Colour-gradient steps on (f0 + f1 - w, f0 - f1) in reduced precision
    against the FP64 oracle: phase field (absolute), velocity (norm-wise) and
    the total and per-phase fluxes (relative)."""
    T, f0, rho = tp_case((24, 20, 18), 13)
    tau0, tau1, sigma, beta, force, steps = 0.8, 1.2, 0.01, 0.7, [1e-5, 0.0, 0.0], 200
    fo, ro, vo, cgo, Fo = oracle.run_twophase(1, T.neighbors, T.node_types, T.bulk, f0, rho, tau0, tau1,
                                              sigma, beta, 0, force, steps)
    ctx = lbamd.Context(T.lattice, precision)
    ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    ctx.upload_f(f0)
    ctx.set_wall_densities(rho)
    ctx.step_cg(steps, tau0, tau1, sigma, beta, lbamd.FORCE_FIXED, force)
    phi, v = ctx.download_scalar(lbamd.PHI), ctx.download_vel()
    q = np.array(ctx.reduce_flux(0))
    ctx.close()
    qo = oracle.flux(1, T.bulk, fo, 2, 0)
    b = T.bulk
    ep, ev, eq = np.abs(phi[b] - cgo[b]).max(), normerr(v[b], vo[b]), np.abs(q - qo).max() / np.abs(qo).max()
    print("phi %.3e  vel %.3e  flux %.3e" % (ep, ev, eq))
    assert ep < ptol and ev < vtol and eq < qtol


@pytest.mark.parametrize("L,radius,porosity,seed,sat", [(64, 5.0, 0.40, 101, 0.5), (96, 6.0, 0.40, 202, 0.3)])
def test_twophase_fp32s_tracks_fp64(gpu_available, L, radius, porosity, seed, sat):
    """FP32 storage with FP64 two-phase arithmetic follows the FP64 run over a
    long transient: phase permeabilities (phase-weighted flux sums, nu = 1/6)
    within 1e-4 of k0 + k1 at every 1000-step sample of 20 000 steps, with
    phases drawn per pore cell (Bernoulli(sat)) and wall colour 1."""
    geo, _ = lbamd.sphere_pack((L,) * 3, radius, porosity, seed, periodic="x", walls=True)
    rng = np.random.default_rng(seed + 1)
    rho0 = np.where(geo > 0, (rng.random(geo.shape) < sat).astype(float), 0.0)
    topo = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    r0 = np.zeros(topo.n_nodes)
    r0[bulk] = rho0[pos[bulk, 0], pos[bulk, 1], pos[bulk, 2]]
    r1 = 1.0 - r0
    sb = types == 1
    r0[sb], r1[sb] = 1.0, 0.0
    W = oracle.WEIGHTS[1]
    f0 = np.zeros((topo.n_nodes, 2, 19))
    f0[bulk, 0] = W[None] * r0[bulk, None]
    f0[bulk, 1] = W[None] * r1[bulk, None]
    force, nbox = [1e-6, 0.0, 0.0], float(geo.size)
    hist = {}
    for precision in (lbamd.FP64, lbamd.FP32S):
        ctx = lbamd.Context(lbamd.D3Q19, precision)
        ctx.set_topology_grid(nb, types, bulk, pos)
        ctx.upload_f(f0)
        ctx.set_wall_densities(np.stack([r0, r1], 1))
        h = []
        for _ in range(20):
            ctx.step_cg(1000, 1.0, 1.0, 0.01, 0.7, lbamd.FORCE_FIXED, force)
            _, s1, s2 = ctx.reduce_flux(0)
            h.append((s1 / (6.0 * nbox * force[0]), s2 / (6.0 * nbox * force[0])))
        ctx.close()
        hist[precision] = np.array(h)
    topo.close()
    a, b = hist[lbamd.FP64], hist[lbamd.FP32S]
    err = np.abs(b - a).max(1) / np.abs(a).sum(1)
    print("L=%d: k0+k1 %.4f -> %.4f, max error %.2e of k0 + k1" % (L, a[0].sum(), a[-1].sum(), err.max()))
    assert err.max() <= 1e-4, err


@pytest.mark.parametrize("precision", [lbamd.FP32S, lbamd.FP16S])
def test_twophase_roundtrip_and_chunking(gpu_available, precision):
    """Upload/download of (f0, f1) through the (t, d) storage is exact to the
    storage precision, and stepping in chunks equals stepping at once."""
    T, f0, rho = tp_case((16, 14, 12), 3)
    ctx = lbamd.Context(T.lattice, precision)
    ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    ctx.upload_f(f0)
    back = ctx.download_f(2)
    tol = 1e-7 if precision == lbamd.FP32S else 2e-4
    assert np.abs(back[T.bulk] - f0[T.bulk]).max() < tol
    ctx.set_wall_densities(rho)
    outs = []
    for chunks in ([30], [1, 9, 20]):
        ctx.upload_f(f0)
        for n in chunks:
            ctx.step_cg(n, 1.0, 1.0, 0.01, 0.7, lbamd.FORCE_MOMX, [1e-5, 0.0, 0.0])
        outs.append((ctx.download_f(2), ctx.download_scalar(lbamd.PHI), ctx.last_force()))
    ctx.close()
    assert np.array_equal(outs[0][0][T.bulk], outs[1][0][T.bulk])
    assert np.array_equal(outs[0][1][T.bulk], outs[1][1][T.bulk])
    assert np.array_equal(outs[0][2], outs[1][2])


@pytest.mark.parametrize("precision", [lbamd.FP32S, lbamd.FP16S])
def test_bgk_chunking_is_exact(gpu_available, precision):
    """Reduced-precision single-phase steps give the same bits however the
    steps are split into calls (the stored-moment variant of the kernel
    rounds like the plain one)."""
    geo, T = pack((20, 18, 16), 3.0, 0.5, 8)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    outs = []
    for chunks in ([40], [1, 19, 20]):
        ctx = lbamd.Context(T.lattice, precision)
        ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
        ctx.upload_f(f0)
        for n in chunks:
            ctx.step_bgk(n, 0.9, np.array([2e-5, 1e-6, 0.0]))
        outs.append((ctx.download_f(1), ctx.download_vel()))
        ctx.close()
    assert np.array_equal(outs[0][0][T.bulk], outs[1][0][T.bulk])
    assert np.array_equal(outs[0][1][T.bulk], outs[1][1][T.bulk])


@pytest.mark.parametrize("force,steps", [(5e-9, 3000), (2e-8, 3000), (1e-7, 2000)])
def test_fp16s_small_increments_are_not_lost(gpu_available, force, steps):
    """A channel driven by a force so small that the per-step change of a
    stored half is below half a unit of its last place: FP16S stores round
    stochastically, so the mean velocity follows FP64 to 1e-4 (measured
    1e-6 to 6e-6).  Round to nearest loses those increments: a build that
    rounds to nearest was off by 1.3e-2 to 2.8e-2 on these cases."""
    geo = np.ones((16, 14, 8), dtype=np.int64)
    geo[:, 0] = 0
    geo[:, -1] = 0
    T = OracleTopo(geo, "D3Q19", "xz")
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    mean = {}
    for p in (lbamd.FP64, lbamd.FP16S):
        ctx = lbamd.Context(lbamd.D3Q19, p)
        ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
        ctx.upload_f(f0)
        ctx.step_bgk(steps, 1.0, np.array([force, 0.0, 0.0]))
        mean[p] = ctx.download_vel()[T.bulk, 0].mean()
        ctx.close()
    err = abs(mean[lbamd.FP16S] - mean[lbamd.FP64]) / abs(mean[lbamd.FP64])
    print("mean u_x error %.3e" % err)
    assert err < 1e-4
