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
