"""Long runs of single-phase steps (at least 256 in one lb_step_bgk,
lb_step_trt or lb_step_rheology call, FP64 and FP32S) replay a captured graph
of 128 steps, with the first step after an upload, the steps left over and
the last one launched on their own.  The results equal the same steps made
one call at a time, bit for bit: populations, density, velocity (and node
viscosity), from a fresh upload and from a post-collision state, for step
counts that leave 0, 1 and many steps over.  FP16S runs, which launch every
step on its own, are checked the same way."""
import numpy as np
import pytest

import lbamd
import oracle
from helpers import OracleTopo

pytestmark = pytest.mark.gpu


def case(lattice):
    rng = np.random.default_rng(21)
    if lattice == lbamd.D2Q9:
        geo = (rng.random((40, 24)) > 0.25).astype(np.int64)
        geo[:, 0] = 0
        geo[:, -1] = 0
        T = OracleTopo(geo, "D2Q9", "x")
    else:
        geo, _ = lbamd.sphere_pack((20, 16, 14), 3.0, 0.5, 3, periodic="x", walls=True)
        T = OracleTopo(np.asarray(geo).astype(np.int64), "D3Q19", "x")
    W = oracle.WEIGHTS[T.lattice]
    f0 = np.zeros((T.n_nodes, W.size))
    f0[T.bulk] = W * (1.0 + 1e-3 * rng.standard_normal((T.bulk.size, W.size)))
    return T, f0


def stepper(op, nD):
    force = [2e-5, -1e-5, 3e-6][:nD]
    if op == "bgk":
        return lambda ctx, n: ctx.step_bgk(n, 0.8, force)
    if op == "trt":
        return lambda ctx, n: ctx.step_trt(n, 0.9, 0.75, force)
    return lambda ctx, n: ctx.step_rheology(n, force)


def run(T, f0, precision, op, chunks):
    ctx = lbamd.Context(T.lattice, precision)
    ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
    ctx.upload_f(f0)
    if op == "rheology":
        sr = np.linspace(0.0, 1e-6, 20)
        ctx.set_viscosity_table(sr, 0.2 - 0.05 * np.sqrt(sr / sr[-1]))
    step = stepper(op, ctx.nD)
    for n in chunks:
        step(ctx, n)
    out = [ctx.download_f(1), ctx.download_scalar(lbamd.RHO), ctx.download_vel()]
    if op == "rheology":
        out.append(ctx.download_scalar(lbamd.VISC))
    ctx.close()
    return out


@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S, lbamd.FP16S])
@pytest.mark.parametrize("lattice,op", [(lbamd.D2Q9, "bgk"), (lbamd.D3Q19, "bgk"), (lbamd.D3Q19, "trt"),
                                        (lbamd.D2Q9, "rheology")])
@pytest.mark.parametrize("n,pre", [(256, 0), (258, 0), (385, 1), (300, 1)])
def test_long_run_equals_single_steps(gpu_available, precision, lattice, op, n, pre):
    T, f0 = case(lattice)
    head = [1] * pre
    once = run(T, f0, precision, op, head + [n])
    single = run(T, f0, precision, op, head + [1] * n)
    b = T.bulk
    for x, y in zip(once, single):
        assert np.array_equal(x[b], y[b])
