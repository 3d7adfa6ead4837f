"""Field files in the reference's binary formats (.lblbf/.lbsca/.lbvec,
LBfield.h:102-138,233-276,378-421) and checkpoint/resume through them."""
import os
import tempfile

import numpy as np
import pytest

import lbamd
import oracle
from helpers import OracleTopo

pytestmark = pytest.mark.gpu


def setup(precision, seed=3):
    geo, _ = lbamd.sphere_pack((20, 18, 16), 3.0, 0.5, seed, periodic="x", walls=True)
    T = OracleTopo(geo.astype(np.int64), "D3Q19", "x")
    ctx = lbamd.Context(lbamd.D3Q19, precision)
    ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    return T, ctx, f0


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S])
def test_checkpoint_resume(gpu_available, precision):
    """Stepping 30 + 20 equals stepping 30, writing a checkpoint, reading it
    into a fresh context and stepping 20 (bit for bit in FP64; in FP32S the
    FP64 file round-trips the float storage to within 1e-15)."""
    F = np.array([2e-5, 1e-6, 0.0])
    T, a, f0 = setup(precision)
    a.upload_f(f0)
    a.step_bgk(30, 0.8, F)
    with tempfile.TemporaryDirectory() as tmp:
        ck = os.path.join(tmp, "state")
        a.write_lbfield(ck)
        hdr = np.fromfile(ck + ".lblbf", dtype=np.int32, count=3)
        assert list(hdr) == [1, 19, T.n_nodes]
        _, b, _ = setup(precision)
        b.read_lbfield(ck)
    a.step_bgk(20, 0.8, F)
    b.step_bgk(20, 0.8, F)
    fa, fb = a.download_f(1), b.download_f(1)
    a.close()
    b.close()
    if precision == lbamd.FP64:
        assert np.array_equal(fa[T.bulk], fb[T.bulk])
    else:
        assert np.abs(fa[T.bulk] - fb[T.bulk]).max() < 1e-15


def test_file_layouts(gpu_available):
    T, ctx, f0 = setup(lbamd.FP64)
    ctx.upload_f(f0)
    ctx.step_bgk(10, 1.0, np.array([1e-5, 0.0, 0.0]))
    f = ctx.download_f(1)
    rho, vel = ctx.download_scalar(lbamd.RHO), ctx.download_vel()
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "x")
        ctx.write_lbfield(p)
        ctx.write_scalar(p, lbamd.RHO, 1)
        ctx.write_vector(p, lbamd.VEL)
        lbf = np.fromfile(p + ".lblbf", dtype=np.uint8)
        sca = np.fromfile(p + ".lbsca", dtype=np.uint8)
        vec = np.fromfile(p + ".lbvec", dtype=np.uint8)
    ctx.close()
    n = T.n_nodes
    assert list(lbf[:12].view(np.int32)) == [1, 19, n]
    assert list(sca[:8].view(np.int32)) == [1, n]
    assert list(vec[:12].view(np.int32)) == [1, 3, n]
    b = T.bulk
    assert np.array_equal(lbf[12:].view(np.float64).reshape(n, 19)[b], f.reshape(n, 19)[b])
    assert np.array_equal(sca[8:].view(np.float64)[b], rho[b])
    assert np.array_equal(vec[12:].view(np.float64).reshape(n, 3)[b], vel[b])


def test_read_rejects_other_node_lists(gpu_available):
    T, a, f0 = setup(lbamd.FP64, seed=3)
    a.upload_f(f0)
    with tempfile.TemporaryDirectory() as tmp:
        ck = os.path.join(tmp, "s")
        a.write_lbfield(ck)
        geo, _ = lbamd.sphere_pack((20, 18, 17), 3.0, 0.5, 4, periodic="x", walls=True)
        U = OracleTopo(geo.astype(np.int64), "D3Q19", "x")
        b = lbamd.Context(lbamd.D3Q19, lbamd.FP64)
        b.set_topology(U.neighbors, U.node_types, U.bulk)
        with pytest.raises(lbamd.LBError) as e:
            b.read_lbfield(ck)
        assert e.value.code == -1
        with pytest.raises(lbamd.LBError) as e:
            b.read_lbfield(os.path.join(tmp, "missing"))
        assert e.value.code == -6
        b.close()
    a.close()


@pytest.mark.parametrize("name,twophase", [("d3q19_pack_6x5x4_pxyz", False), ("d2q9_random_7x5", False),
                                           ("tp_pack_6x5x4", True)])
def test_vtu_of_device_state_matches_reference_file(gpu_available, name, twophase, tmp_path):
    """The reference case run on the device (FP64), its fields downloaded and
    written with lb_vtu_*: the same bytes as the .vtu the compiled reference
    wrote for that case (tests/golden/vtu/)."""
    from helpers import GOLDEN, float_round, load_ref
    d = load_ref(name)
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    ctx = lbamd.Context(T.lattice, lbamd.FP64)
    ctx.set_topology(T.neighbors, T.node_types, T.bulk)
    b = T.bulk
    if twophase:
        r0 = float_round(T.attribute(d["rho0_attr"]))
        r1 = float_round(T.attribute(1.0 - d["rho0_attr"]))
        w = float_round(T.attribute(d["wett_attr"]))
        sb = T.node_types == 1
        r0[sb], r1[sb] = w[sb], 1 - w[sb]
        W = oracle.WEIGHTS[1]
        f0 = np.zeros((T.n_nodes, 2, 19))
        f0[b, 0] = W[None] * r0[b, None]
        f0[b, 1] = W[None] * r1[b, None]
        ctx.upload_f(f0)
        ctx.set_wall_densities(np.stack([r0, r1], 1))
        ctx.step_cg(p["steps"] + 1, p["tau0"], p["tau1"], p["sigma"], p["beta"], lbamd.FORCE_MOMX,
                    [p["momx"]] + p["force"][1:])
    else:
        f0 = oracle.rest_state(T.lattice, T.n_nodes, b, T.attribute(d["init_rho"]))
        ctx.upload_f(f0)
        ctx.step_bgk(p["steps"] + 1, p["tau"], np.asarray(p["force"], dtype=float))
    vel = np.ascontiguousarray(ctx.download_vel()[b])
    out = lbamd.VtuOutput(T.positions[b], str(tmp_path))
    if twophase:
        out.add_file("fluid")
        rr = ctx.download_scalar(lbamd.RHO, 2)
        fields = [np.ascontiguousarray(rr[b, k]) for k in range(2)]
        out.add_variable("rho0", fields[0])
        out.add_variable("rho1", fields[1])
        base = "fluid"
    else:
        out.add_file("lb_run")
        rho = np.ascontiguousarray(ctx.download_scalar(lbamd.RHO)[b])
        out.add_variable("rho", rho)
        base = "lb_run"
    out.add_variable("vel", vel)
    out.write(p["steps"])
    ctx.close()
    got = open(os.path.join(str(tmp_path), "vtu", "0000_%s_0000000.vtu" % base), "rb").read()
    ref = open(os.path.join(GOLDEN, "vtu", name + ".vtu"), "rb").read()
    if twophase:
        # the momentum-controlled force sums over nodes in another order:
        # header and mesh identical, the fields agree to 1e-12
        import sys
        sys.path.insert(0, GOLDEN)
        from make_ref_fixtures import parse_vtu
        head = ref.index(b"<AppendedData")
        assert got[:head] == ref[:head] and len(got) == len(ref)
        fg = os.path.join(str(tmp_path), "g.vtu")
        fr = os.path.join(GOLDEN, "vtu", name + ".vtu")
        a, r = parse_vtu(fg if open(fg, "wb").write(got) else fg), parse_vtu(fr)
        for k in ("points", "connectivity", "offsets", "types"):
            assert np.array_equal(a[k], r[k]), k
        for k in ("rho0", "rho1", "vel"):
            assert np.abs(a[k] - r[k]).max() <= 1e-12 * np.abs(r[k]).max(), k
    else:
        assert got == ref
