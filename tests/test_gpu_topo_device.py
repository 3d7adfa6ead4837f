"""Node list built on the GPU (lb_topo_from_geometry_device) against the
host builder, which is itself pinned to files written by the reference's
vtklb.py (tests/test_topology.py): every array identical, integer work."""
import os

import numpy as np
import pytest

import lbamd
from helpers import GOLDEN

pytestmark = pytest.mark.gpu

VT = os.path.join(GOLDEN, "vtklb")
ARRAYS = ("neighbors", "positions", "node_types", "bulk", "node_ranks", "global_dims")


def same(geo, lattice, periodic):
    host = lbamd.Topology.from_geometry(geo, lattice, periodic)
    dev = lbamd.Topology.from_geometry_device(geo, lattice, periodic)
    assert dev.n_nodes == host.n_nodes and dev.n_bulk == host.n_bulk and dev.rank == host.rank
    for m in ARRAYS:
        assert np.array_equal(getattr(host, m)(), getattr(dev, m)()), m
    return host.n_nodes


@pytest.mark.parametrize("name,lattice,periodic", [
    ("d2q9_channel_8x6_px", lbamd.D2Q9, "x"),
    ("d2q9_random_7x5_pxy", lbamd.D2Q9, "xy"),
    ("d3q19_open_4x4x3", lbamd.D3Q19, ""),
    ("d3q19_pack_6x5x4_px_walls", lbamd.D3Q19, "x"),
    ("d3q19_pack_6x5x4_pxyz", lbamd.D3Q19, "xyz"),
])
def test_device_builder_on_reference_geometries(gpu_available, name, lattice, periodic):
    same(np.load(os.path.join(VT, name + "_geo.npy")), lattice, periodic)


@pytest.mark.parametrize("lattice,shape", [(lbamd.D2Q9, (37, 23)), (lbamd.D3Q19, (19, 14, 11))])
@pytest.mark.parametrize("periodic", ["", "x", "y", "xy", "xz", "yz", "xyz"])
def test_device_builder_random(gpu_available, lattice, shape, periodic):
    if lattice == lbamd.D2Q9 and "z" in periodic:
        pytest.skip("no z axis in 2D")
    rng = np.random.default_rng(len(periodic) * 7 + len(shape))
    geo = (rng.random(shape) > 0.35).astype(np.int32)
    same(geo, lattice, periodic)


@pytest.mark.parametrize("lattice,shape,periodic", [
    (lbamd.D2Q9, (5, 1), "y"),          # one layer, its own periodic copy
    (lbamd.D2Q9, (1, 1), "xy"),
    (lbamd.D3Q19, (3, 1, 2), "xyz"),
    (lbamd.D3Q19, (9, 8, 7), "xyz"),    # all fluid: no wall nodes
    (lbamd.D3Q19, (9, 8, 7), ""),       # all fluid, closed: the rim is outside
])
def test_device_builder_thin_and_full(gpu_available, lattice, shape, periodic):
    same(np.ones(shape, dtype=np.int32), lattice, periodic)


def test_device_builder_single_fluid_cell_and_packs(gpu_available):
    g = np.zeros((6, 7, 5), dtype=np.int32)
    g[2, 3, 1] = 1
    assert same(g, lbamd.D3Q19, "") == 1 + 1 + 18
    for seed, per in [(1, "x"), (2, "xyz"), (3, "yz")]:
        geo, _ = lbamd.sphere_pack((48, 40, 36), 5.0, 0.45, seed, periodic=per, walls=True)
        same(geo, lbamd.D3Q19, per)
    geo, _ = lbamd.sphere_pack((96, 96, 96), 8.0, 0.40, 4, periodic="x", walls=True)
    same(geo, lbamd.D3Q19, "x")


def test_device_builder_rejections(gpu_available):
    geo = np.load(os.path.join(VT, "d3q19_2rank_5x6x4_pxyz_geo.npy"))
    with pytest.raises(lbamd.LBError) as e:  # two ranks: the host builder's job
        lbamd.Topology.from_geometry_device(geo, lbamd.D3Q19, "xyz")
    assert e.value.code == -4
    with pytest.raises(lbamd.LBError) as e:
        lbamd.Topology.from_geometry_device(np.zeros((4, 4, 4), np.int32), lbamd.D3Q19, "")
    assert e.value.code == -1
    g = np.ones((4, 4, 4), np.int32)
    g[1, 1, 1] = -2
    with pytest.raises(lbamd.LBError) as e:
        lbamd.Topology.from_geometry_device(g, lbamd.D3Q19, "")
    assert e.value.code == -1
    # the library still works after the refusals
    same(np.ones((4, 5, 6), np.int32), lbamd.D3Q19, "x")
