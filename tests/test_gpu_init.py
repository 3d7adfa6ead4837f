"""GPU parity of the equilibrium initialisation (lb_init_equilibrium,
k_init_eq) against the reference's own initiateLbField
(LBinitiatefield.h:33-57), whose outputs the compiled harness wrote to
tests/golden/init_ref.npz.

Bar: FP64 bit-exact on every population; FP32S and FP16S within the
rounding of their storage, value by value.  The reduced forms store the
shifted value t = f - w (one field), or t = f0 + f1 - w and d = f0 - f1 (two
fields), in float, or in half after a scale of 4096.  Rounding to nearest
moves each stored value by at most half a unit in its last place, so a
population read back as t + w, or as (t + d)/2 and (t - d)/2, is off by at
most half a unit of t, or a quarter of a unit of t plus a quarter of a unit
of d (the float step of the double to half conversion adds half a float unit).
A run started from the initialised state equals the oracle's run from the
reference's state, bit for bit in FP64.
"""
import importlib.util
import os

import numpy as np
import pytest

import lbamd
import oracle
from helpers import GOLDEN



def _mod():
    spec = importlib.util.spec_from_file_location("make_init_fixture", os.path.join(GOLDEN, "make_init_fixture.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


CASES = [c[0] for c in _mod().cases()]


def _case(name):
    mod = _mod()
    ref = np.load(os.path.join(GOLDEN, "init_ref.npz"))
    for c in mod.cases():
        if c[0] == name:
            cname, lattice, geo, periodic, n_lb, n_rho, calls, seed = c
            T, rho, vel = mod.inputs(cname, lattice, geo, periodic, n_lb, n_rho, seed)
            return T, rho, vel[:, 0], n_lb, calls, ref[name + "/f"]
    raise KeyError(name)


def _ctx(T, precision):
    ctx = lbamd.Context(T.lattice, precision)
    ctx.set_topology_grid(T.neighbors, T.node_types, T.bulk, T.positions)
    return ctx


HALF_SCALE = 4096.0


def _ulp(x, precision):
    """Unit in the last place of the stored form of each value of x."""
    if precision == lbamd.FP32S:
        return np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)
    return np.spacing(np.abs(x * HALF_SCALE).astype(np.float16)).astype(np.float64) / HALF_SCALE


def _round_bound(state, lattice, precision):
    """Read-back error of each population of `state` (nodes, fields, Q) that
    one store in the reduced form allows: round to nearest of each stored
    value, plus the float step of a half."""
    w = oracle.WEIGHTS[lattice]
    f32 = lambda x: np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)  # noqa: E731
    if state.shape[1] == 1:
        t = state[:, 0, :] - w
        b = 0.5 * _ulp(t, precision) + (0.5 * f32(t) if precision == lbamd.FP16S else 0.0)
        return b[:, None, :]
    t = state[:, 0, :] + state[:, 1, :] - w
    d = state[:, 0, :] - state[:, 1, :]
    b = 0.25 * (_ulp(t, precision) + _ulp(d, precision))
    if precision == lbamd.FP16S:
        b = b + 0.25 * (f32(t) + f32(d))
    return np.stack([b, b], 1)


def storage_bound(f_ref, calls, lattice, precision):
    """Largest read-back error after the init calls.  Every call loads the
    stored state, sets one field exactly and stores the state again, so the
    roundings of all calls add up.  Before its call a field holds what the
    zeroed storage reads back as: w/2 for each of two fields (t = 0, d = 0),
    w for one field."""
    n_lb = f_ref.shape[1]
    w = oracle.WEIGHTS[lattice]
    state = np.broadcast_to(w / n_lb, f_ref.shape).copy()
    total = np.zeros(f_ref.shape)
    seen = set()
    for lb, _, _ in calls:
        assert lb not in seen  # each call sets a different field
        seen.add(lb)
        state[:, lb, :] = f_ref[:, lb, :]
        total += _round_bound(state, lattice, precision)
    return total


@pytest.mark.parametrize("precision", [lbamd.FP32S, lbamd.FP16S])
def test_storage_bound_is_not_a_unit_loose(precision):
    """A planted mistake, on the host: a conversion that rounds t and d a
    whole unit the same way (truncation instead of round to nearest) moves
    the read-back f0 = (t + d)/2 by half a unit of each, twice what the bound
    allows for the worst case of correct rounding."""
    _, _, _, n_lb, _, f_ref = _case("d3q19_two_fields")
    b = storage_bound(f_ref, [(0, 0, 0)], lbamd.D3Q19, precision)
    one = f_ref.copy()
    one[:, 1, :] = oracle.WEIGHTS[lbamd.D3Q19] / 2
    w = oracle.WEIGHTS[lbamd.D3Q19]
    t = one[:, 0, :] + one[:, 1, :] - w
    d = one[:, 0, :] - one[:, 1, :]
    moved = 0.5 * (_ulp(t, precision) + _ulp(d, precision))
    assert np.all(moved > 1.5 * b[:, 0, :])


def _emulated_init(f_ref, calls, lattice, precision):
    """The init calls on the stored forms, rounded by numpy (round to nearest
    of float32, and of float16 after a float32 step and the 4096 scale)."""
    w = oracle.WEIGHTS[lattice]
    n = f_ref.shape[0]
    S = np.float32 if precision == lbamd.FP32S else np.float16
    g = np.zeros((n, 19 if lattice == lbamd.D3Q19 else 9), S)
    d = g.copy()

    def put(x):
        x = x.astype(np.float32)
        return x if S is np.float32 else (x * np.float32(HALF_SCALE)).astype(np.float16)

    def get(x):
        x = x.astype(np.float32)
        return (x if S is np.float32 else x * np.float32(1.0 / HALF_SCALE)).astype(np.float64)

    for lb, _, _ in calls:
        t = get(g) + w
        f = [0.5 * (t + get(d)), 0.5 * (t - get(d))]
        f[lb] = f_ref[:, lb, :]
        g, d = put((f[0] + f[1]) - w), put(f[0] - f[1])
    t = get(g) + w
    return np.stack([0.5 * (t + get(d)), 0.5 * (t - get(d))], 1)


@pytest.mark.parametrize("precision", [lbamd.FP32S, lbamd.FP16S])
def test_storage_bound_covers_emulated_rounding(precision):
    """Two init calls on two fields, emulated on the host with round to
    nearest: the error stays within the bound and uses most of it."""
    _, _, _, n_lb, calls, f_ref = _case("d3q19_two_fields")
    got = _emulated_init(f_ref, calls, lbamd.D3Q19, precision)
    ratio = np.abs(got - f_ref) / storage_bound(f_ref, calls, lbamd.D3Q19, precision)
    assert ratio.max() <= 1.0 and ratio.max() > 0.5


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("precision", [lbamd.FP64, lbamd.FP32S, lbamd.FP16S])
def test_init_equilibrium_vs_reference(gpu_available, name, precision):
    T, rho, vel, n_lb, calls, f_ref = _case(name)
    print("init case %s precision %d" % (name, precision), flush=True)
    ctx = _ctx(T, precision)
    for lb, rf, vf in calls:
        assert vf == 0
        ctx.init_equilibrium(n_lb, lb, rho, vel, rho_field=rf)
    f = ctx.download_f(n_lb)
    ctx.close()
    b = T.bulk
    if precision == lbamd.FP64:
        assert np.array_equal(f[b], f_ref[b])
    else:
        bound = storage_bound(f_ref, calls, T.lattice, precision)
        assert np.all(np.abs(f[b] - f_ref[b]) <= bound[b] * (1 + 1e-9))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["d3q19_walls", "d2q9_random"])
def test_steps_from_initialised_state_match_oracle(gpu_available, name):
    T, rho, vel, n_lb, calls, f_ref = _case(name)
    ctx = _ctx(T, lbamd.FP64)
    ctx.init_equilibrium(1, 0, rho, vel)
    D = vel.shape[1]
    force = np.array([2e-6, -1e-6, 5e-7])[:D]
    ctx.step_bgk(12, 0.9, force)
    f = ctx.download_f(1)[:, 0, :]
    r = ctx.download_scalar(lbamd.RHO)
    ctx.close()
    fo, ro, _ = oracle.run_std_case(T.lattice, T.neighbors, T.node_types, T.bulk, f_ref[:, 0].copy(), 0.9, force, 12)
    b = T.bulk
    assert np.array_equal(f[b], fo[b])
    assert np.array_equal(r[b], ro[b])
