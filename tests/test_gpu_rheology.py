"""Generalized-Newtonian step on the device (lb_set_viscosity_table,
lb_step_rheology) against the compiled reference's
polymer_viscosity_tabulation case and the oracle.

Bar: FP64 bit-exact (array_equal) for populations, rho, vel and the node
viscosities.  Reduced precision norm-wise against the FP64 oracle: FP32S
velocity 1e-4 and viscosity 1e-5, FP16S velocity 5e-3.
"""
import os
import sys

import numpy as np
import pytest

import lbamd
import oracle
from helpers import GOLDEN, OracleTopo, load_ref, ref_names

sys.path.insert(0, GOLDEN)
from make_ref_fixtures import carreau_table  # noqa: E402

pytestmark = pytest.mark.gpu


def gpu_rheo(T, f0, table, force, chunks, precision=lbamd.FP64, layout="table"):
    ctx = lbamd.Context(T.lattice, precision)
    if layout == "rows":
        ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
        assert ctx.layout != 0, ctx.layout_note
    else:
        ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    ctx.upload_f(f0)
    ctx.set_viscosity_table(table[:, 0], table[:, 1])
    for n in chunks:
        ctx.step_rheology(n, np.asarray(force, dtype=float))
    out = (ctx.download_f(1)[:, 0, :], ctx.download_scalar(lbamd.RHO), ctx.download_vel(),
           ctx.download_scalar(lbamd.VISC))
    ctx.close()
    return out


@pytest.mark.parametrize("layout", ["table", "rows"])
@pytest.mark.parametrize("name", ref_names("rheo_"))
def test_rheology_vs_reference_fixture(gpu_available, name, layout):
    d = load_ref(name)
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0 = oracle.rest_state(T.lattice, T.n_nodes, T.bulk, T.attribute(d["init_rho"]))
    f, rho, vel, visc = gpu_rheo(T, f0, d["table"], p["force"], [p["steps"] + 1], layout=layout)
    b = T.bulk
    assert np.array_equal(rho[b], d["rho"])
    assert np.array_equal(vel[b], d["vel"][:, :vel.shape[1]])
    assert np.array_equal(visc[b], d["viscosity"])


@pytest.mark.parametrize("layout", ["table", "rows"])
@pytest.mark.parametrize("lattice,shape,periodic,seed,lam2", [
    ("D2Q9", (40, 23), "xy", 1, 1e9),
    ("D3Q19", (17, 12, 9), "xyz", 3, 1e9),
    ("D3Q19", (20, 16, 13), "x", 4, 1e11),
])
def test_rheology_vs_oracle_random(gpu_available, lattice, shape, periodic, seed, lam2, layout):
    rng = np.random.default_rng(seed)
    geo = (rng.random(shape) > 0.3).astype(np.int64)
    if "y" not in periodic:
        geo[:, 0] = geo[:, -1] = 0
    if len(shape) == 3 and "z" not in periodic:
        geo[:, :, 0] = geo[:, :, -1] = 0
    T = OracleTopo(geo, lattice, periodic)
    f0 = oracle.rest_state(T.lattice, T.n_nodes, T.bulk, 1.0 + 0.01 * rng.random(T.n_nodes))
    f0[T.bulk] += 1e-4 * rng.standard_normal((T.bulk.size, f0.shape[1]))
    force = [2e-5, -1e-5, 5e-6][: len(shape)]
    _, table = carreau_table(lam2=lam2)
    fo, ro, vo, wo = oracle.run_rheology(T.lattice, T.neighbors, T.node_types, T.bulk, f0, table, np.array(force), 37)
    fg, rg, vg, wg = gpu_rheo(T, f0, table, force, [1, 10, 26], layout=layout)
    b = T.bulk
    assert np.array_equal(fg[b], fo[b])
    assert np.array_equal(rg[b], ro[b])
    assert np.array_equal(vg[b], vo[b])
    assert np.array_equal(wg[b], wo[b])
    assert np.unique(wo[b]).size > 10  # the table is exercised, not one row


@pytest.mark.parametrize("precision,vtol,wtol", [(lbamd.FP32S, 1e-4, 1e-5), (lbamd.FP16S, 5e-3, None)])
def test_rheology_reduced_precision_vs_oracle(gpu_available, precision, vtol, wtol):
    geo, _ = lbamd.sphere_pack((28, 24, 20), 4.0, 0.5, 21, periodic="x", walls=True)
    T = OracleTopo(geo.astype(np.int64), "D3Q19", "x")
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    force = np.array([5e-5, 0.0, 0.0])
    _, table = carreau_table(lam2=1e9)
    fo, ro, vo, wo = oracle.run_rheology(1, T.neighbors, T.node_types, T.bulk, f0, table, force, 300)
    f, r, v, w = gpu_rheo(T, f0, table, force, [300], precision=precision, layout="rows")
    b = T.bulk
    print("vel err %.3e visc err %.3e" % (np.abs(v[b] - vo[b]).max() / np.abs(vo[b]).max(),
                                          np.abs(w[b] - wo[b]).max() / np.abs(wo[b]).max()))
    assert np.abs(v[b] - vo[b]).max() / np.abs(vo[b]).max() < vtol
    if wtol is not None:
        assert np.abs(w[b] - wo[b]).max() / np.abs(wo[b]).max() < wtol


def test_rheology_state_and_argument_errors(gpu_available):
    T = OracleTopo(np.ones((6, 5), dtype=np.int64), "D2Q9", "xy")
    ctx = lbamd.Context(T.lattice, lbamd.FP64)
    ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    ctx.upload_f(oracle.rest_state(0, T.n_nodes, T.bulk, np.ones(T.n_nodes)))
    with pytest.raises(lbamd.LBError) as e:  # no table yet
        ctx.step_rheology(1, np.zeros(2))
    assert e.value.code == -2
    for sr, nu in [([1e-8, 1e-9], [0.1, 0.1]), ([1e-8, np.nan], [0.1, 0.1]), ([], [])]:
        with pytest.raises(lbamd.LBError) as e:
            ctx.set_viscosity_table(np.array(sr, dtype=float), np.array(nu, dtype=float))
        assert e.value.code == -1
    ctx.step_bgk(2, 1.0, np.zeros(2))
    with pytest.raises(lbamd.LBError) as e:  # no rheology step has stored viscosities
        ctx.download_scalar(lbamd.VISC)
    assert e.value.code == -2
    ctx.set_viscosity_table(np.array([1e-10]), np.array([1.0 / 6.0]))  # one row: constant viscosity
    ctx.step_rheology(3, np.array([1e-6, 0.0]))
    assert np.all(ctx.download_scalar(lbamd.VISC)[T.bulk] == 1.0 / 6.0)
    ctx.close()
