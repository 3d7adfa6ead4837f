"""GPU parity: the HIP path (through the C ABI) against the oracle and the
reference's own outputs.

Bar: FP64 bit-exact (array_equal) for populations, rho, vel and phi, except
where a sum runs over nodes (the two-phase flux controller and the flux
reductions), where the tolerance is 1e-12 relative.
"""
import numpy as np
import pytest

import lbamd
import oracle
from helpers import OracleTopo, float_round, load_ref, ref_names, std_ref_names, tp_start

pytestmark = pytest.mark.gpu


def gpu_std_case(T, f0, tau, force, nsteps, precision=lbamd.FP64, allow_ghost0=False, chunks=None,
                 layout="table"):
    ctx = lbamd.Context(T.lattice, precision)
    if layout == "rows":
        ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions, allow_ghost0=allow_ghost0)
        # a row layout applies unless the node list links to the ghost node:
        # row windows (2) when the row axis is not periodic, else rows (1)
        if allow_ghost0:
            assert ctx.layout == 0
        else:
            assert ctx.layout == (1 if T.geometry.periodic.endswith("z") or
                                  (T.lattice == 0 and "y" in T.geometry.periodic) else 2), ctx.layout_note
    else:
        ctx.set_topology(T.neighbors, T.node_types, T.bulk, allow_ghost0=allow_ghost0)
        assert ctx.layout == 0
    ctx.upload_f(f0)
    for n in (chunks or [nsteps]):
        ctx.step_bgk(n, tau, np.asarray(force, dtype=float))
    rho = ctx.download_scalar(lbamd.RHO)
    vel = ctx.download_vel()
    f = ctx.download_f(1)[:, 0, :]
    ctx.close()
    return f, rho, vel


def rest(T, rho_nodes):
    return oracle.rest_state(T.lattice, T.n_nodes, T.bulk, rho_nodes)


@pytest.mark.parametrize("layout", ["table", "rows"])
@pytest.mark.parametrize("name", std_ref_names())
def test_std_case_vs_reference_fixture(gpu_available, name, layout):
    d = load_ref(name)
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0 = rest(T, T.attribute(d["init_rho"]))
    ghost0 = bool((T.neighbors[T.bulk] == 0).any())
    f, rho, vel = gpu_std_case(T, f0, p["tau"], p["force"], p["steps"] + 1, allow_ghost0=ghost0, layout=layout)
    D = vel.shape[1]
    if "rho" in d:
        assert np.array_equal(rho[T.bulk], d["rho"])
        assert np.array_equal(vel[T.bulk], d["vel"][:, :D])
    else:
        ny = d["geo"].shape[1] - 2
        nx = d["geo"].shape[0]
        assert np.array_equal(rho[T.bulk].reshape(nx, ny)[0], d["rho_profile"])
        assert np.array_equal(vel[T.bulk, 0].reshape(nx, ny)[0], d["velx_profile"])


@pytest.mark.parametrize("layout", ["table", "rows"])
@pytest.mark.parametrize("lattice,shape,periodic,seed", [
    ("D2Q9", (40, 23), "xy", 1),
    ("D2Q9", (33, 17), "x", 2),
    ("D3Q19", (17, 12, 9), "xyz", 3),
    ("D3Q19", (20, 16, 13), "x", 4),
    ("D3Q19", (40, 9, 70), "xz", 5),
])
def test_std_case_vs_oracle_random(gpu_available, lattice, shape, periodic, seed, layout):
    rng = np.random.default_rng(seed)
    geo = (rng.random(shape) > 0.3).astype(np.int64)
    # walls on the faces of non-periodic axes keep the ghost node out
    if "y" not in periodic:
        geo[:, 0] = 0
        geo[:, -1] = 0
    if len(shape) == 3 and "z" not in periodic:
        geo[:, :, 0] = 0
        geo[:, :, -1] = 0
    T = OracleTopo(geo, lattice, periodic)
    rho0 = np.ones(T.n_nodes) + 0.01 * rng.random(T.n_nodes)
    f0 = rest(T, rho0)
    f0[T.bulk] += 1e-4 * rng.standard_normal((T.bulk.size, f0.shape[1]))
    force = [2e-5, -1e-5, 5e-6][: len(shape)]
    nsteps = 37
    fo, ro, vo = oracle.run_std_case(T.lattice, T.neighbors, T.node_types, T.bulk, f0, 0.83, np.array(force), nsteps)
    fg, rg, vg = gpu_std_case(T, f0, 0.83, force, nsteps, chunks=[1, 10, 26], layout=layout)
    assert np.array_equal(fg[T.bulk], fo[T.bulk])
    assert np.array_equal(rg[T.bulk], ro[T.bulk])
    assert np.array_equal(vg[T.bulk], vo[T.bulk])


def test_ghost0_rejected_without_flag(gpu_available):
    d = load_ref("d3q19_open_4x4x3")
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    ctx = lbamd.Context(T.lattice)
    with pytest.raises(lbamd.LBError) as e:
        ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    assert e.value.code == -3
    ctx.close()


def tp_initial(T, d):
    return tp_start(d, T)


def gpu_twophase(T, f0, rho, p, mode, force, nsteps, chunks=None, layout="table", precision=lbamd.FP64):
    ctx = lbamd.Context(T.lattice, precision)
    if layout == "table":
        ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    else:
        ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
    ctx.upload_f(f0)
    ctx.set_wall_densities(rho)
    for n in (chunks or [nsteps]):
        ctx.step_cg(n, p["tau0"], p["tau1"], p["sigma"], p["beta"], mode, force)
    out = (ctx.download_f(2), ctx.download_scalar(lbamd.RHO, 2), ctx.download_vel(),
           ctx.download_scalar(lbamd.PHI), ctx.last_force())
    ctx.close()
    return out


@pytest.mark.parametrize("name", [n for n in ref_names("tp_")])
def test_twophase_vs_reference_fixture(gpu_available, name):
    d = load_ref(name)
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0, rho = tp_initial(T, d)
    force = [p["momx"]] + list(p["force"][1:])
    f, r, v, phi, F = gpu_twophase(T, f0, rho, p, lbamd.FORCE_MOMX, force, p["steps"] + 1)
    b = T.bulk
    # the flux controller sums over nodes in a different order: 1e-12 relative
    for got, ref in ((r[b, 0], d["rho0"]), (r[b, 1], d["rho1"]), (v[b], d["vel"][:, : v.shape[1]])):
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-12 * scale
    assert abs(F[0] - d["force_x"][-1, 1]) <= 1e-12 * abs(d["force_x"][-1, 1])


@pytest.mark.parametrize("name,layout", [("tp_pack_16", "table"), ("tp_walls_12x14x10", "table"),
                                         ("tp_walls_12x14x10", "grid")])
def test_twophase_fixed_force_bitexact(gpu_available, name, layout):
    """Fixed-force colour gradient bit for bit against the oracle, with the
    link table and (uniform wall colour) with the row windows."""
    d = load_ref(name)
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0, rho = tp_initial(T, d)
    force = [3e-6, 1e-6, -2e-6]
    fo, ro, vo, cgo, Fo = oracle.run_twophase(1, T.neighbors, T.node_types, T.bulk, f0, rho, p["tau0"], p["tau1"],
                                              p["sigma"], p["beta"], 0, force, 25)
    f, r, v, phi, F = gpu_twophase(T, f0, rho, p, lbamd.FORCE_FIXED, force, 25, chunks=[1, 24], layout=layout)
    b = T.bulk
    assert np.array_equal(f[b], fo[b])
    assert np.array_equal(r[b], ro[b])
    assert np.array_equal(v[b], vo[b])
    assert np.array_equal(phi[b], cgo[b])


@pytest.mark.parametrize("precision,layout", [(lbamd.FP64, "table"), (lbamd.FP64, "grid"),
                                              (lbamd.FP32S, "grid"), (lbamd.FP16S, "grid")])
def test_twophase_d2q9_vs_oracle(gpu_available, precision, layout):
    """The colour-gradient step on D2Q9 (random 33 x 17 geometry, periodic x,
    walls on y; phases drawn per node; wall colour 1): FP64 bit for bit
    against the oracle, both layouts, in chunks; FP32S within phi 1e-4
    (absolute) and velocity 1e-4 (norm-wise), FP16S within 2e-3 and 2e-2."""
    rng = np.random.default_rng(7)
    geo = (rng.random((33, 17)) > 0.3).astype(np.int64)
    geo[:, 0] = 0
    geo[:, -1] = 0
    T = OracleTopo(geo, "D2Q9", "x")
    r0 = np.zeros(T.n_nodes)
    r0[T.bulk] = (rng.random(T.bulk.size) < 0.5).astype(float)
    r1 = np.zeros(T.n_nodes)
    r1[T.bulk] = 1.0 - r0[T.bulk]
    sb = T.node_types == 1
    r0[sb], r1[sb] = 1.0, 0.0
    W = oracle.WEIGHTS[T.lattice]
    f0 = np.zeros((T.n_nodes, 2, 9))
    f0[T.bulk, 0] = W[None] * r0[T.bulk, None]
    f0[T.bulk, 1] = W[None] * r1[T.bulk, None]
    rho = np.stack([r0, r1], 1)
    p = {"tau0": 0.9, "tau1": 1.3, "sigma": 0.01, "beta": 0.7}
    force = [4e-6, -1e-6]
    fo, ro, vo, cgo, _ = oracle.run_twophase(T.lattice, T.neighbors, T.node_types, T.bulk, f0, rho, p["tau0"],
                                             p["tau1"], p["sigma"], p["beta"], 0, force, 30)
    f, r, v, phi, _ = gpu_twophase(T, f0, rho, p, lbamd.FORCE_FIXED, force, 30, chunks=[1, 29], layout=layout,
                                   precision=precision)
    b = T.bulk
    if precision == lbamd.FP64:
        assert np.array_equal(f[b], fo[b])
        assert np.array_equal(r[b], ro[b])
        assert np.array_equal(v[b], vo[b])
        assert np.array_equal(phi[b], cgo[b])
    else:
        ptol, vtol = (1e-4, 1e-4) if precision == lbamd.FP32S else (2e-3, 2e-2)
        ep = np.abs(phi[b] - cgo[b]).max()
        ev = np.abs(v[b] - vo[b]).max() / np.abs(vo[b]).max()
        print("phi %.3e  vel %.3e" % (ep, ev))
        assert ep < ptol and ev < vtol


def test_flux_reduction(gpu_available):
    d = load_ref("tp_pack_16")
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0, rho = tp_initial(T, d)
    force = [3e-6, 1e-6, -2e-6]
    fo, *_ = oracle.run_twophase(1, T.neighbors, T.node_types, T.bulk, f0, rho, p["tau0"], p["tau1"],
                                 p["sigma"], p["beta"], 0, force, 12)
    ctx = lbamd.Context(T.lattice)
    ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    ctx.upload_f(f0)
    ctx.set_wall_densities(rho)
    ctx.step_cg(12, p["tau0"], p["tau1"], p["sigma"], p["beta"], lbamd.FORCE_FIXED, force)
    for direction in range(3):
        got = np.array(ctx.reduce_flux(direction))
        ref = oracle.flux(1, T.bulk, fo, 2, direction)
        assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref).max())
    ctx.close()


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S, lbamd.FP16S])
def test_twophase_layouts_identical(gpu_available, precision):
    """The row-window two-phase path (uniform wall colour) gives the same bits
    as the link-table path, in the momentum-controlled mode too."""
    d = load_ref("tp_walls_12x14x10")
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0, rho = tp_initial(T, d)
    outs = []
    for layout in ("table", "grid"):
        outs.append(gpu_twophase(T, f0, rho, p, lbamd.FORCE_MOMX, [p["momx"], 0.0, 0.0], 30, chunks=[1, 29],
                                 layout=layout, precision=precision))
    b = T.bulk
    for x, y in zip(outs[0][:4], outs[1][:4]):
        assert np.array_equal(x[b], y[b])
    assert np.array_equal(outs[0][4], outs[1][4])
