"""VTU/PVTU writer (lb_vtu_*) against files written by the reference's own
Output/VTK classes (tests/golden/vtu/, kept by make_ref_fixtures.py from
compiled reference runs): the .vtu byte for byte, the .pvtu byte for byte
apart from its creation-time comment.  The cell data are the fixture's
fields (the reference's FP64 results), the positions those of the bulk
nodes of the same node list."""
import os

import numpy as np
import pytest

import lbamd
from helpers import GOLDEN, OracleTopo, load_ref

VTU = os.path.join(GOLDEN, "vtu")


def bulk_positions(d):
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    return T.positions[T.bulk]


def without_created(text):
    return b"\n".join(l for l in text.split(b"\n") if not l.startswith(b"<!-- Created "))


def check_files(tmp, base, name):
    got = open(os.path.join(tmp, "vtu", "0000_%s_0000000.vtu" % base), "rb").read()
    assert got == open(os.path.join(VTU, name + ".vtu"), "rb").read()
    gp = open(os.path.join(tmp, "%s_0000000.pvtu" % base), "rb").read()
    rp = open(os.path.join(VTU, name + ".pvtu"), "rb").read()
    # the reference's file is its second write (_0000001); the piece line names it
    rp = rp.replace(b"_0000001.vtu", b"_0000000.vtu")
    assert without_created(gp) == without_created(rp)
    assert gp.count(b"<!-- Created ") == 1


@pytest.mark.parametrize("name", ["d2q9_random_7x5", "d3q19_pack_6x5x4_pxyz", "trt_d3q19_pack_6x5x4_px_walls"])
def test_vtu_matches_reference_std_case_output(name, tmp_path):
    d = load_ref(name)
    pos = bulk_positions(d)
    nd = pos.shape[1]
    out = lbamd.VtuOutput(pos, str(tmp_path))
    out.add_file("lb_run")
    rho = np.ascontiguousarray(d["rho"])
    vel = np.ascontiguousarray(d["vel"][:, :nd])
    out.add_variable("rho", rho)
    out.add_variable("vel", vel)
    out.write(d["params"]["steps"])
    check_files(str(tmp_path), "lb_run", name)


def test_vtu_matches_reference_twophase_output(tmp_path):
    d = load_ref("tp_pack_6x5x4")
    out = lbamd.VtuOutput(bulk_positions(d), str(tmp_path))
    out.add_file("fluid")
    arrays = [np.ascontiguousarray(d[k]) for k in ("rho0", "rho1", "vel")]
    for k, a in zip(("rho0", "rho1", "vel"), arrays):
        out.add_variable(k, a)
    out.write(d["params"]["steps"])
    check_files(str(tmp_path), "fluid", "tp_pack_6x5x4")


def test_vtu_reads_variables_at_write_time_and_counts_writes(tmp_path):
    d = load_ref("d3q19_pack_6x5x4_pxyz")
    pos = bulk_positions(d)
    rho = np.zeros(pos.shape[0])
    vel = np.zeros((pos.shape[0], 3))
    out = lbamd.VtuOutput(pos, str(tmp_path))
    out.add_file("lb_run")
    out.add_variable("rho", rho)
    out.add_variable("vel", vel)
    out.write(0)
    rho[:] = d["rho"]
    vel[:] = d["vel"]
    out.write(d["params"]["steps"])
    got = open(os.path.join(str(tmp_path), "vtu", "0000_lb_run_0000001.vtu"), "rb").read()
    assert got == open(os.path.join(VTU, "d3q19_pack_6x5x4_pxyz.vtu"), "rb").read()
    assert os.path.exists(os.path.join(str(tmp_path), "lb_run_0000001.pvtu"))


def test_vtu_mesh_shares_corner_points():
    # a 3x2 block of cells: (3+1)(2+1) corner points in 2D, (3+1)(2+1)(1+1) in 3D
    pos2 = np.array([[x, y] for x in range(3) for y in range(2)])
    assert lbamd.VtuOutput(pos2, "/tmp/lbamd_vtu_mesh").n_points == 12
    pos3 = np.array([[x, y, 0] for x in range(3) for y in range(2)])
    assert lbamd.VtuOutput(pos3, "/tmp/lbamd_vtu_mesh").n_points == 24


def test_vtu_rejections(tmp_path):
    pos = np.zeros((4, 3), np.int32)
    out = lbamd.VtuOutput(pos, str(tmp_path))
    with pytest.raises(lbamd.LBError):  # no file yet
        out.add_variable("rho", np.zeros(4))
    out.add_file("x")
    with pytest.raises(ValueError):
        out.add_variable("rho", np.zeros(4, np.float32))
    with pytest.raises(lbamd.LBError):  # dim must be 1 or nd
        out.add_variable("v2", np.zeros((4, 2)))
