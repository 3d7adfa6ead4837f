"""Node list and neighbour index: the oracle writer, the native builder and the
native .vtklb reader against files written by the reference writer
(tests/golden/vtklb, from PythonScripts/vtklb.py via make_vtklb_fixtures.py),
plus geometric invariants on larger random geometries."""
import os
import sys

import numpy as np
import pytest

import lbamd
import oracle
from helpers import GOLDEN, OracleTopo
from oracle.vtklb_oracle import BASIS, Geometry

sys.path.insert(0, os.path.join(GOLDEN))
import make_vtklb_fixtures as mk  # noqa: E402

VT = os.path.join(GOLDEN, "vtklb")
CASES = list(mk.cases())


def fixture_text(name, rank):
    with open(os.path.join(VT, "%s%d.vtklb" % (name, rank - 1))) as fh:
        return fh.read()


def stored_attrs(name, attrs):
    return [(k, np.load(os.path.join(VT, "%s_attr_%s.npy" % (name, k)))) for k in sorted(attrs)]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_writer_matches_reference_files(case):
    name, _, lattice, periodic, attrs = case
    geo = np.load(os.path.join(VT, name + "_geo.npy"))
    g = Geometry(geo, lattice, periodic)
    for r in range(1, g.n_ranks + 1):
        assert g.vtklb_text(r, stored_attrs(name, attrs)) == fixture_text(name, r)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_native_builder_and_reader_match_reference_files(case):
    name, _, lattice, periodic, attrs = case
    geo = np.load(os.path.join(VT, name + "_geo.npy"))
    L = lbamd.LATTICE_NAME[lattice]
    for r in range(1, int(geo.max()) + 1):
        built = lbamd.Topology.from_geometry(geo, L, periodic, rank=r)
        read = lbamd.Topology.from_vtklb(os.path.join(VT, "%s%d.vtklb" % (name, r - 1)), L)
        for m in ("neighbors", "positions", "node_types", "bulk", "node_ranks", "global_dims"):
            assert np.array_equal(getattr(built, m)(), getattr(read, m)()), m
        assert built.rank == read.rank == r - 1
        # node types against the oracle's restatement of Nodes::setupNodeType
        T = OracleTopo(geo, lattice, periodic, rank=r)
        assert np.array_equal(built.node_types(), T.node_types)
        assert np.array_equal(built.bulk(), T.bulk)
        assert np.array_equal(built.neighbors(), T.neighbors)


@pytest.mark.parametrize("case", [c for c in CASES if all(
    not np.issubdtype(np.asarray(v).dtype, np.integer) for v in c[4].values())], ids=lambda c: c[0])
def test_native_writer_matches_reference_files(case, tmp_path):
    name, _, lattice, periodic, attrs = case
    geo = np.load(os.path.join(VT, name + "_geo.npy"))
    L = lbamd.LATTICE_NAME[lattice]
    for r in range(1, int(geo.max()) + 1):
        t = lbamd.Topology.from_geometry(geo, L, periodic, rank=r)
        T = OracleTopo(geo, lattice, periodic, rank=r)
        vals = [(k, T.attribute(v)) for k, v in stored_attrs(name, attrs)]
        out = str(tmp_path / "w.vtklb")
        t.write_vtklb(out, vals)
        assert open(out).read() == fixture_text(name, r)


def test_reader_attributes_and_float_rounding():
    path = os.path.join(VT, "d3q19_pack_6x5x4_pxyz0.vtklb")
    t = lbamd.Topology.from_vtklb(path, lbamd.D3Q19)
    T = OracleTopo(np.load(os.path.join(VT, "d3q19_pack_6x5x4_pxyz_geo.npy")), "D3Q19", "xyz")
    r0 = np.load(os.path.join(VT, "d3q19_pack_6x5x4_pxyz_attr_rho0.npy"))
    assert np.array_equal(t.attribute("rho0"), T.attribute(r0))
    src = t.attribute("source")
    assert np.all(src == 0.0)
    with pytest.raises(lbamd.LBError):
        t.attribute("does_not_exist")


def test_reader_maps_a_permuted_file_basis(tmp_path):
    """LATTICE may list the basis in any order; NEIGHBORS follow that order
    and the reader maps them to the program basis (LBvtk.h:418-430, 473-482)."""
    name = "d3q19_pack_6x5x4_px_walls"
    text = fixture_text(name, 1).splitlines(keepends=True)
    perm = np.random.default_rng(5).permutation(19)
    i_lat = next(i for i, l in enumerate(text) if l.startswith("LATTICE"))
    i_nb = next(i for i, l in enumerate(text) if l.startswith("NEIGHBORS"))
    i_par = next(i for i, l in enumerate(text) if l.startswith("PARALLEL_COMPUTING"))
    basis = [text[i_lat + 1 + q] for q in range(19)]
    out = text[: i_lat + 1] + [basis[p] for p in perm] + text[i_lat + 20: i_nb + 1]
    for line in text[i_nb + 1: i_par]:
        v = line.split()
        out.append(" ".join(v[p] for p in perm) + "\n")
    out += text[i_par:]
    p = tmp_path / "perm.vtklb"
    p.write_text("".join(out))
    a = lbamd.Topology.from_vtklb(str(p), lbamd.D3Q19)
    b = lbamd.Topology.from_vtklb(os.path.join(VT, name + "0.vtklb"), lbamd.D3Q19)
    assert np.array_equal(a.neighbors(), b.neighbors())


def test_reader_errors(tmp_path):
    with pytest.raises(lbamd.LBError) as e:
        lbamd.Topology.from_vtklb(str(tmp_path / "missing.vtklb"), lbamd.D3Q19)
    assert e.value.code == -6
    with pytest.raises(lbamd.LBError):  # wrong number of basis vectors for the lattice
        lbamd.Topology.from_vtklb(os.path.join(VT, "d2q9_channel_8x6_px0.vtklb"), lbamd.D3Q19)
    bad = tmp_path / "bad.vtklb"
    bad.write_text(fixture_text("d2q9_channel_8x6_px", 1).replace("SCALARS nodetype int", "SCALARS other int"))
    with pytest.raises(lbamd.LBError):
        lbamd.Topology.from_vtklb(str(bad), lbamd.D2Q9)
    with pytest.raises(lbamd.LBError):
        lbamd.Topology.from_geometry(np.ones((4, 4), dtype=np.int32), lbamd.D2Q9, "", rank=2)


@pytest.mark.parametrize("lattice,shape,periodic", [
    ("D2Q9", (31, 17), "xy"), ("D2Q9", (20, 9), "x"), ("D3Q19", (14, 11, 9), "xyz"),
    ("D3Q19", (12, 10, 7), "xz"), ("D3Q19", (9, 8, 6), ""),
])
def test_neighbour_index_invariants(lattice, shape, periodic):
    """pos(neighbor(q,n)) = pos(n) + c_q (periodic where the axis is), rest = self,
    links off the padded array are 0, and fluid links are symmetric."""
    rng = np.random.default_rng(len(shape) * 100 + shape[0])
    geo = (rng.random(shape) > 0.35).astype(np.int32)
    L = lbamd.LATTICE_NAME[lattice]
    t = lbamd.Topology.from_geometry(geo, L, periodic)
    nb, pos, typ = t.neighbors(), t.positions(), t.node_types()
    c = BASIS[lattice]
    nd = len(shape)
    dims = np.array(shape)
    for n in range(1, t.n_nodes):
        for q in range(c.shape[0]):
            m = nb[n, q]
            target = pos[n] + c[q]
            if q == c.shape[0] - 1:
                assert m == n
                continue
            # a fluid neighbour across a periodic face is reached through the rim copy
            wrapped = target.copy()
            for d in range(nd):
                if "xyz"[d] in periodic:
                    wrapped[d] %= dims[d]
            if m == 0:
                continue
            if typ[m] > 1:
                assert np.array_equal(pos[m], wrapped)
                if typ[n] > 1:
                    qr = [i for i in range(c.shape[0]) if np.array_equal(c[i], -c[q])][0]
                    assert nb[m, qr] == n
            else:
                # solid neighbours (incl. rim copies) sit exactly at pos + c_q
                assert np.array_equal(pos[m], target)
    # bulk = fluid nodes in ascending order and C order of positions
    bulk = t.bulk()
    assert np.all(np.diff(bulk) > 0)
    keys = np.ravel_multi_index(tuple(pos[bulk].T), shape)
    assert np.all(np.diff(keys) > 0)


def test_sphere_pack_is_seeded_and_reaches_porosity():
    g1, i1 = lbamd.sphere_pack((40, 40, 40), 4.0, 0.45, 17, periodic="x", walls=True)
    g2, i2 = lbamd.sphere_pack((40, 40, 40), 4.0, 0.45, 17, periodic="x", walls=True)
    g3, _ = lbamd.sphere_pack((40, 40, 40), 4.0, 0.45, 18, periodic="x", walls=True)
    assert np.array_equal(g1, g2) and not np.array_equal(g1, g3)
    assert i1 == i2
    inner = g1[:, 1:-1, 1:-1]
    assert abs(inner.mean() - i1["porosity"]) < 1e-12
    assert i1["porosity"] <= 0.45 and i1["porosity"] > 0.40
    assert np.all(g1[:, 0, :] == 0) and np.all(g1[:, :, -1] == 0)
    # a loose target is reached without overlap
    _, i4 = lbamd.sphere_pack((40, 40, 40), 4.0, 0.9, 3, periodic="xyz", walls=False)
    assert not i4["overlap_allowed"]


def test_oracle_node_types_known_case():
    # D2Q9 channel 4 x 4 periodic in x: rows y=1,2 fluid next to the walls -> type 2,
    # wall rows -> type 1 (solid with fluid neighbours)
    geo = np.ones((4, 4), dtype=np.int64)
    geo[:, 0] = 0
    geo[:, -1] = 0
    T = OracleTopo(geo, "D2Q9", "x")
    attr = np.zeros(T.n_nodes, dtype=np.int32)
    attr[1:] = T.rank_topology.nodetype
    assert T.node_types[0] == -1
    assert np.all(T.node_types[T.bulk] == 2)
    assert T.bulk.size == 8
    solids = np.nonzero(attr == 0)[0][1:]
    assert np.all(T.node_types[solids] == 1)
    assert np.array_equal(oracle.bulk_nodes(T.node_types, T.ranks, 0), np.arange(1, 9))


TOPO_ARRAYS = ("neighbors", "positions", "node_types", "bulk", "node_ranks", "global_dims")


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_binary_twin_round_trip_of_reference_files(case, tmp_path):
    """A .vtklb written by the reference writer, read as text, written as the
    binary twin and read back: the same topology, and every attribute gives
    the same doubles and the same float roundings as from the text."""
    name, _, lattice, periodic, attrs = case
    geo = np.load(os.path.join(VT, name + "_geo.npy"))
    L = lbamd.LATTICE_NAME[lattice]
    for r in range(1, int(geo.max()) + 1):
        text = lbamd.Topology.from_vtklb(os.path.join(VT, "%s%d.vtklb" % (name, r - 1)), L)
        path = str(tmp_path / "t.vtklbb")
        text.write_binary(path)
        binr = lbamd.Topology.from_binary(path, L)
        for m in TOPO_ARRAYS:
            assert np.array_equal(getattr(text, m)(), getattr(binr, m)()), m
        assert text.rank == binr.rank
        for k in ["nodetype"] + sorted(attrs):
            for as_float in (False, True):
                assert np.array_equal(text.attribute(k, as_float), binr.attribute(k, as_float)), (k, as_float)
        # and a second generation (binary -> binary) is byte-identical
        path2 = str(tmp_path / "t2.vtklbb")
        binr.write_binary(path2)
        assert open(path, "rb").read() == open(path2, "rb").read()


def test_binary_twin_of_built_topology_and_rejections(tmp_path):
    geo, _ = lbamd.sphere_pack((20, 18, 16), 3.0, 0.5, 3, periodic="xz", walls=True)
    t = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "xz")
    path = str(tmp_path / "p.vtklbb")
    t.write_binary(path)
    b = lbamd.Topology.from_binary(path, lbamd.D3Q19)
    for m in TOPO_ARRAYS:
        assert np.array_equal(getattr(t, m)(), getattr(b, m)()), m
    with pytest.raises(lbamd.LBError):  # wrong lattice
        lbamd.Topology.from_binary(path, lbamd.D2Q9)
    raw = open(path, "rb").read()
    bad = str(tmp_path / "bad.vtklbb")
    open(bad, "wb").write(raw[: len(raw) // 2])
    with pytest.raises(lbamd.LBError):  # truncated
        lbamd.Topology.from_binary(bad, lbamd.D3Q19)
    open(bad, "wb").write(b"XXXXXXXX" + raw[8:])
    with pytest.raises(lbamd.LBError):  # not a binary twin
        lbamd.Topology.from_binary(bad, lbamd.D3Q19)


@pytest.mark.parametrize("lattice,shape,periodic", [
    ("D3Q19", (144, 64, 64), "xz"), ("D2Q9", (960, 640), "xy"),
])
def test_native_builder_large_multirank_matches_oracle(lattice, shape, periodic):
    """Geometries large enough that the native builder's passes run on host
    threads (over 64 k fluid nodes per rank), three ranks in blocks along the
    first two axes so that every rank meets two others: every rank's node
    list, positions, owner ranks, node types and bulk list equal the oracle's
    restatement of the reference writer."""
    rng = np.random.default_rng(shape[0])
    fluid = rng.random(shape) > 0.3
    x = np.arange(shape[0])[:, None] * 3 // shape[0]
    y = np.arange(shape[1])[None, :] * 2 // shape[1]
    owner = np.where(x < 2, x + 1, 3) + 0 * y
    owner = np.where((x == 1) & (y == 1), 3, owner)
    if len(shape) == 3:
        owner = owner[:, :, None]
    geo = np.where(fluid, owner, 0).astype(np.int32)
    L = lbamd.LATTICE_NAME[lattice]
    for r in (1, 2, 3):
        built = lbamd.Topology.from_geometry(geo, L, periodic, rank=r)
        T = OracleTopo(geo, lattice, periodic, rank=r)
        assert (T.node_types > 0).sum() > 65536
        assert np.array_equal(built.neighbors(), T.neighbors)
        assert np.array_equal(built.positions(), T.positions)
        assert np.array_equal(built.node_ranks(), T.ranks)
        assert np.array_equal(built.node_types(), T.node_types)
        assert np.array_equal(built.bulk(), T.bulk)
