"""Tagged-link boundaries and per-domain mass sources on the device
(lb_set_node_tags, lb_step_tagged, lb_pressure_flux) against the compiled
reference's std_one_phase and the oracle.

Bar: FP64 bit-exact (array_equal) for rho and vel when no interior domain
carries a mass source; with sources the per-domain mass change is a sum over
nodes whose order differs from the reference's sequential one, so the bar is
1e-12 relative (as for the two-phase flux controller).  Mass fluxes: bit-exact
against the oracle without sources, 1e-12 with; against the reference's
printed 6 digits, 1e-5.
"""
import sys

import numpy as np
import pytest

import lbamd
import oracle
from helpers import GOLDEN, OracleTopo, load_ref, ref_names

sys.path.insert(0, GOLDEN)
from make_ref_fixtures import one_phase_geometry  # noqa: E402

pytestmark = pytest.mark.gpu


def inputs(T, tags_cells, label_cells, fon_cells):
    tags = np.round(T.attribute(tags_cells)).astype(np.int32)
    tags[0] = -1
    label = np.round(T.attribute(label_cells)).astype(np.int32)
    fon = T.attribute(fon_cells)
    return tags, label, fon


def gpu_tagged(T, f0, tags, label, fon, tau, force, chunks, layout="table"):
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP64)
    if layout == "rows":
        ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
    else:
        ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    ctx.upload_f(f0)
    ctx.set_node_tags(T.neighbors, tags, label, fon)
    for n in chunks:
        ctx.step_tagged(n, tau, np.asarray(force, dtype=float), 1.0)
    out = ctx.download_scalar(lbamd.RHO), ctx.download_vel(), ctx.pressure_flux()
    ctx.close()
    return out


def close(a, b, rel):
    return np.abs(a - b).max() <= rel * np.abs(b).max()


@pytest.mark.parametrize("layout", ["table", "rows"])
@pytest.mark.parametrize("name", ref_names("op_"))
def test_tagged_vs_reference_fixture(gpu_available, name, layout):
    d = load_ref(name)
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), "D3Q19", p["periodic"])
    tags, label, fon = inputs(T, d["tags"], d["label"], d["force_on"])
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    rho, vel, flux = gpu_tagged(T, f0, tags, label, fon, p["tau"], p["force"], [p["steps"] + 1], layout)
    b = T.bulk
    if label.max() == 0:
        assert np.array_equal(rho[b], d["rho"])
        assert np.array_equal(vel[b], d["vel"])
    else:
        assert close(rho[b] - 1, d["rho"] - 1, 1e-9) and close(vel[b], d["vel"], 1e-12)
    assert np.allclose(flux, d["flux"], rtol=1e-5, atol=1e-21)


@pytest.mark.parametrize("shape,seed,labels,chunks", [
    ((10, 9, 14), 21, False, [1, 9, 30]),
    ((12, 8, 11), 22, True, [40]),
    ((8, 12, 16), 23, True, [3, 1, 36]),
])
def test_tagged_vs_oracle(gpu_available, shape, seed, labels, chunks):
    geo, tags_c, label_c, fon_c = one_phase_geometry(shape, seed, labels)
    T = OracleTopo(geo, "D3Q19", "xy")
    tags, label, fon = inputs(T, tags_c, label_c, fon_c)
    rng = np.random.default_rng(seed)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, 1.0 + 0.002 * rng.random(T.n_nodes))
    tau, force = 0.9, [3e-6, -1e-6, 2e-5]
    fo, ro, vo, qo = oracle.run_one_phase(1, T.neighbors, T.bulk, tags, label, fon, np.zeros((T.n_nodes, 3)), f0,
                                          tau, np.array(force), 1.0, sum(chunks))
    rg, vg, qg = gpu_tagged(T, f0, tags, label, fon, tau, force, chunks, "rows")
    b = T.bulk
    if labels:
        assert close(rg[b] - 1, ro[b] - 1, 1e-9) and close(vg[b], vo[b], 1e-12)
        assert close(qg, qo, 1e-12)
    else:
        assert np.array_equal(rg[b], ro[b]) and np.array_equal(vg[b], vo[b])
        assert np.array_equal(qg, qo)


def test_tagged_rejections(gpu_available):
    geo, tags_c, label_c, fon_c = one_phase_geometry((8, 8, 10), 3, True)
    T = OracleTopo(geo, "D3Q19", "xy")
    tags, label, fon = inputs(T, tags_c, label_c, fon_c)
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP64)
    ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    ctx.upload_f(oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes)))
    with pytest.raises(lbamd.LBError) as e:  # no tags yet
        ctx.step_tagged(1, 1.0, np.zeros(3))
    assert e.value.code == -2
    bad = tags.copy()
    wall_side = next(n for n in T.bulk if bad[n] & 8)
    bad[wall_side] &= ~8  # a fluid node next to solid without its fluid-solid bit
    with pytest.raises(lbamd.LBError) as e:
        ctx.set_node_tags(T.neighbors, bad, label, fon)
    assert e.value.code == -4 and "stale" in str(e.value)
    lab = label.copy()
    lab[lab == 2] = 3  # domain 2 left without nodes
    with pytest.raises(lbamd.LBError) as e:
        ctx.set_node_tags(T.neighbors, tags, lab, fon)
    assert e.value.code == -1 and "Interior domain 2" in str(e.value)
    ctx.set_node_tags(T.neighbors, tags, label, fon)
    ctx.step_tagged(2, 1.0, np.array([0.0, 0.0, 1e-5]))
    assert np.isfinite(ctx.pressure_flux()).all()
    ctx.close()
    c32 = lbamd.Context(lbamd.D3Q19, lbamd.FP32S)
    c32.set_topology(T.neighbors, T.node_types, T.bulk)
    with pytest.raises(lbamd.LBError) as e:
        c32.set_node_tags(T.neighbors, tags, label, fon)
    assert e.value.code == -4
    c32.close()
