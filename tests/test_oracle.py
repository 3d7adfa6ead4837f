"""Pinning the oracle (CPU restatement of the reference loops).

1. Against the compiled reference's own outputs (tests/golden/ref_*.npz,
   written by make_ref_fixtures.py from oracle/_ref builds): bit-exact.
2. Against what the mathematics fixes: rest state, Guo momentum growth in a
   periodic box, mass conservation, the Poiseuille start-up series, phase
   symmetry and the one-phase limit of the colour-gradient model.
"""
import math
import os

import numpy as np
import pytest

import oracle
from helpers import OracleTopo, float_round, load_ref, ref_names, std_ref_names, tp_start


def ref_std_run(name):
    d = load_ref(name)
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0 = oracle.rest_state(T.lattice, T.n_nodes, T.bulk, T.attribute(d["init_rho"]))
    f, rho, vel = oracle.run_std_case(T.lattice, T.neighbors, T.node_types, T.bulk, f0, p["tau"],
                                      np.array(p["force"]), p["steps"] + 1)
    return d, T, f, rho, vel


@pytest.mark.parametrize("name", std_ref_names())
def test_std_case_bitexact_vs_compiled_reference(name):
    d, T, f, rho, vel = ref_std_run(name)
    D = vel.shape[1]
    if "rho" in d:
        assert np.array_equal(rho[T.bulk], d["rho"])
        assert np.array_equal(vel[T.bulk], d["vel"][:, :D])
    else:
        nx, ny = d["geo"].shape[0], d["geo"].shape[1] - 2
        r = rho[T.bulk].reshape(nx, ny)
        u = vel[T.bulk, 0].reshape(nx, ny)
        assert np.array_equal(r[0], d["rho_profile"])
        assert np.array_equal(u[0], d["velx_profile"])
        # x-invariance (bit-identical columns), as in the reference run
        assert np.ptp(r, axis=0).max() == d["x_spread"][0] == 0.0
        assert np.ptp(u, axis=0).max() == d["x_spread"][1] == 0.0


@pytest.mark.parametrize("name", ref_names("trt_"))
def test_trt_bitexact_vs_compiled_reference(name):
    """std_case with calcOmegaBGKTRT / calcDeltaOmegaFTRT (phi = 1), built from
    the reference's own main.cpp by oracle/build_ref.sh."""
    d, T, f, rho, vel = ref_std_run(name)
    assert len(d["params"]["tau"]) == 2
    D = vel.shape[1]
    assert np.array_equal(rho[T.bulk], d["rho"])
    assert np.array_equal(vel[T.bulk], d["vel"][:, :D])


@pytest.mark.parametrize("lattice,shape,periodic", [("D2Q9", (9, 7), "xy"), ("D3Q19", (6, 5, 7), "x")])
def test_trt_with_equal_times_is_bgk(lattice, shape, periodic):
    """tau_sym = tau_anti: the even and odd relaxations recombine into BGK and
    both Guo terms into calcDeltaOmegaF (equal up to rounding)."""
    rng = np.random.default_rng(8)
    geo = (rng.random(shape) > 0.25).astype(np.int64)
    T = OracleTopo(geo, lattice, periodic)
    f0 = oracle.rest_state(T.lattice, T.n_nodes, T.bulk, 1.0 + 0.01 * rng.random(T.n_nodes))
    F = np.array([2e-5, -1e-5, 5e-6])
    fb, rb, vb = oracle.run_std_case(T.lattice, T.neighbors, T.node_types, T.bulk, f0, 0.7, F, 40)
    ft, rt, vt = oracle.run_std_case(T.lattice, T.neighbors, T.node_types, T.bulk, f0, (0.7, 0.7), F, 40)
    b = T.bulk
    # populations are O(w); j = sum_q f c cancels to O(1e-5), so both carry
    # absolute rounding of a few 1e-16
    assert np.abs(ft[b] - fb[b]).max() <= 1e-14
    assert np.abs(vt[b] - vb[b]).max() <= 1e-14
    # and a distinct tau_anti changes the flow
    fa, ra, va = oracle.run_std_case(T.lattice, T.neighbors, T.node_types, T.bulk, f0, (0.7, 1.1), F, 40)
    assert np.abs(va[b] - vb[b]).max() > 1e-3 * np.abs(vb[b]).max()


@pytest.mark.parametrize("tau_sym,tau_anti", [(1.0, 0.875), (0.875, 1.0), (0.6, 0.5 + 0.1875 / 0.1)])
def test_trt_magic_parameter_gives_exact_poiseuille(tau_sym, tau_anti):
    """(tau_sym - 1/2)(tau_anti - 1/2) = 3/16: the steady channel flow with
    half-way bounce-back is the parabola u = F/(2 nu) (y+1/2)(H-1/2-y) with
    nu = (tau_sym - 1/2)/3, walls exactly half-way (BGK at tau = 1 misses it
    by 5e-3).  A swap of the two times changes nu and fails this."""
    H = 8
    geo = np.ones((4, H + 2), dtype=np.int64)
    geo[:, 0] = geo[:, -1] = 0
    T = OracleTopo(geo, "D2Q9", "x")
    f0 = oracle.rest_state(0, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    F = 1e-5
    f, rho, vel = oracle.run_std_case(0, T.neighbors, T.node_types, T.bulk, f0, (tau_sym, tau_anti),
                                      np.array([F, 0.0]), 20000)
    y = T.positions[T.bulk, 1] - 1
    nu = (tau_sym - 0.5) / 3
    exact = F / (2 * nu) * (y + 0.5) * (H - 0.5 - y)
    assert np.abs(vel[T.bulk, 0] - exact).max() <= 1e-10 * exact.max()
    assert np.abs(vel[T.bulk, 1]).max() <= 1e-12 * exact.max()


@pytest.mark.parametrize("name", ref_names("rheo_"))
def test_rheology_bitexact_vs_compiled_reference(name):
    """polymer_viscosity_tabulation (GeneralizedNewtonian) as compiled from the
    reference by oracle/build_ref.sh: rho, vel and the node viscosities."""
    d = load_ref(name)
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0 = oracle.rest_state(T.lattice, T.n_nodes, T.bulk, T.attribute(d["init_rho"]))
    f, rho, vel, visc = oracle.run_rheology(T.lattice, T.neighbors, T.node_types, T.bulk, f0, d["table"],
                                            np.array(p["force"]), p["steps"] + 1)
    b = T.bulk
    assert np.array_equal(rho[b], d["rho"])
    assert np.array_equal(vel[b], d["vel"][:, :vel.shape[1]])
    assert np.array_equal(visc[b], d["viscosity"])


def test_rheology_constant_table_is_bgk():
    """A table of one viscosity nu: tau = 3 nu / rho + 1/2 at every node, BGK
    with tau = 3 nu + 1/2 up to the density's departure from 1."""
    rng = np.random.default_rng(4)
    geo = (rng.random((14, 11, 9)) > 0.3).astype(np.int64)
    T = OracleTopo(geo, "D3Q19", "xyz")
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    F = np.array([2e-5, 0.0, 1e-5])
    nu = 0.1
    table = np.array([[1e-12, nu], [1e-6, nu]])
    fr, rr, vr, visc = oracle.run_rheology(1, T.neighbors, T.node_types, T.bulk, f0, table, F, 60)
    fb, rb, vb = oracle.run_std_case(1, T.neighbors, T.node_types, T.bulk, f0, 3 * nu + 0.5, F, 60)
    b = T.bulk
    assert np.all(visc[b] == nu)
    drho = np.abs(rb[b] - 1).max()
    assert 0 < drho < 1e-3
    assert np.abs(vr[b] - vb[b]).max() <= 10 * drho * np.abs(vb[b]).max()


def test_rheology_channel_stress_balance():
    """Steady generalized-Newtonian channel flow: the node viscosities
    satisfy mu(y) du/dy = F (y_c - y), the momentum balance, to the
    discretisation error; a shear-thinning fluid is least viscous at the
    walls.  The table is the Carreau curve of the fixtures."""
    import sys
    from helpers import GOLDEN
    sys.path.insert(0, GOLDEN)
    from make_ref_fixtures import carreau_table
    _, table = carreau_table()
    H = 22
    geo = np.ones((4, H + 2), dtype=np.int64)
    geo[:, 0] = geo[:, -1] = 0
    T = OracleTopo(geo, "D2Q9", "x")
    f0 = oracle.rest_state(0, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    F = 2e-5
    f, rho, vel, visc = oracle.run_rheology(0, T.neighbors, T.node_types, T.bulk, f0, table, np.array([F, 0.0]),
                                            40000)
    y = T.positions[T.bulk, 1]
    col = T.positions[T.bulk, 0] == 0
    yy, u, mu = y[col], vel[T.bulk, 0][col], visc[T.bulk][col]
    s = np.argsort(yy)
    yy, u, mu = yy[s], u[s], mu[s]
    # fluid rows y = 1..H, half-way walls at 1/2 and H + 1/2; interior faces
    # between rows k and k+1: stress mu du/dy against F (y_c - y_face)
    yc = (H + 1) / 2.0
    yface = 0.5 * (yy[1:] + yy[:-1])
    stress = 0.5 * (mu[1:] + mu[:-1]) * np.diff(u)
    exact = F * (yc - yface)
    # the face viscosity is the mean of its two nodes: a few per cent where mu
    # changes fastest (the Carreau knee near the centre)
    assert np.abs(stress - exact).max() <= 0.05 * np.abs(exact).max()
    assert mu[H // 2] > mu[0] and mu[H // 2] > mu[-1]
    assert np.allclose(mu, mu[::-1], rtol=0, atol=1e-12)


def one_phase_inputs(d, T):
    """Per-node tags (node 0: -1, as Nodes sets it), domain labels and force
    attribute of a std_one_phase fixture."""
    tags = np.round(T.attribute(d["tags"])).astype(np.int32)
    tags[0] = -1
    label = np.round(T.attribute(d["label"])).astype(np.int32)
    fon = T.attribute(d["force_on"])
    return tags, label, fon


@pytest.mark.parametrize("name", ref_names("op_"))
def test_one_phase_bitexact_vs_compiled_reference(name):
    """std_one_phase (tagged bounce-back, anti-bounce-back, fluid-fluid swaps,
    per-domain mass sources) as compiled from the reference: rho and vel bit
    for bit; the mass fluxes to the 6 digits the case prints."""
    d = load_ref(name)
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), "D3Q19", p["periodic"])
    tags, label, fon = one_phase_inputs(d, T)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    f, rho, vel, flux = oracle.run_one_phase(1, T.neighbors, T.bulk, tags, label, fon, np.zeros((T.n_nodes, 3)), f0,
                                             p["tau"], np.array(p["force"]), p["rho_w"], p["steps"] + 1)
    b = T.bulk
    assert np.array_equal(rho[b], d["rho"])
    assert np.array_equal(vel[b], d["vel"])
    assert np.allclose(flux, d["flux"], rtol=1e-5, atol=1e-21)


def test_one_phase_without_tags_conserves_and_bounces():
    """Closed box, every fluid node next to solid tagged fluid-solid and no
    sources: the tagged bounce-back is the half-way bounce-back, so the loop
    equals std_case's bit for bit; and a pressure-free, source-free run
    conserves mass."""
    rng = np.random.default_rng(12)
    geo = (rng.random((9, 8, 7)) > 0.25).astype(np.int64)
    geo[:, :, 0] = geo[:, :, -1] = 0
    T = OracleTopo(geo, "D3Q19", "xy")
    tags = np.zeros(T.n_nodes, np.int32)
    tags[0] = -1
    fluid = T.node_types > 1
    tags[fluid] = 1
    nbt = T.node_types[T.neighbors]
    tags[fluid & (nbt < 2).any(axis=1)] |= 8
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    F = np.array([1e-5, 2e-6, 0.0])
    f1, r1, v1, _ = oracle.run_one_phase(1, T.neighbors, T.bulk, tags, np.zeros(T.n_nodes, np.int32),
                                         np.ones(T.n_nodes), np.zeros((T.n_nodes, 3)), f0, 0.7, F, 1.0, 40)
    f2, r2, v2 = oracle.run_std_case(1, T.neighbors, T.node_types, T.bulk, f0, 0.7, F, 40)
    b = T.bulk
    assert np.array_equal(f1[b], f2[b]) and np.array_equal(v1[b], v2[b])
    assert abs(f1[b].sum() - f0[b].sum()) <= 1e-12 * f0[b].sum()


def test_one_phase_mass_source_restores_domain_mass():
    """A closed domain that starts with a mass deficit: the source
    0.9*2*(mass change)/(domain size) returns its mass towards rho = 1."""
    geo = np.ones((6, 6, 8), dtype=np.int64)
    geo[:, :, 0] = geo[:, :, -1] = 0
    T = OracleTopo(geo, "D3Q19", "xy")
    tags = np.zeros(T.n_nodes, np.int32)
    tags[0] = -1
    fluid = T.node_types > 1
    tags[fluid] = 2
    nbt = T.node_types[T.neighbors]
    tags[fluid & (nbt < 2).any(axis=1)] |= 8
    label = np.where(fluid, 1, 0).astype(np.int32)
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, np.full(T.n_nodes, 0.98))
    b = T.bulk
    m0 = f0[b].sum()
    f, rho, vel, _ = oracle.run_one_phase(1, T.neighbors, T.bulk, tags, label, np.zeros(T.n_nodes),
                                          np.zeros((T.n_nodes, 3)), f0, 1.0, np.zeros(3), 1.0, 60)
    deficit0 = b.size - m0
    deficit = b.size - f[b].sum()
    assert 0 < deficit < 1e-3 * deficit0


@pytest.mark.parametrize("name", ref_names("tp_"))
def test_twophase_bitexact_vs_compiled_reference(name):
    d = load_ref(name)
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0, rho = tp_start(d, T)
    f, r, v, cg, F = oracle.run_twophase(T.lattice, T.neighbors, T.node_types, T.bulk, f0, rho, p["tau0"],
                                         p["tau1"], p["sigma"], p["beta"], 1, [p["momx"]] + p["force"][1:],
                                         p["steps"] + 1)
    assert np.array_equal(r[T.bulk, 0], d["rho0"])
    assert np.array_equal(r[T.bulk, 1], d["rho1"])
    assert np.array_equal(v[T.bulk], d["vel"][:, : v.shape[1]])
    assert F[0] == d["force_x"][-1, 1]


def periodic_box(lattice, shape):
    return OracleTopo(np.ones(shape, dtype=np.int64), lattice, "xyz"[: len(shape)])


@pytest.mark.parametrize("lattice,shape", [("D2Q9", (7, 5)), ("D3Q19", (5, 4, 3))])
def test_rest_state_is_exact(lattice, shape):
    T = periodic_box(lattice, shape)
    f0 = oracle.rest_state(T.lattice, T.n_nodes, T.bulk, np.full(T.n_nodes, 1.25))
    f, rho, vel = oracle.run_std_case(T.lattice, T.neighbors, T.node_types, T.bulk, f0, 0.77,
                                      np.zeros(3), 25)
    # sum_q w_q rho and sum_q w_q c_q round to a few ulp: the rest state stays
    # put to that accuracy and does not drift
    assert np.abs(f[T.bulk] - f0[T.bulk]).max() <= 4e-16
    assert np.abs(vel[T.bulk]).max() <= 1e-16


@pytest.mark.parametrize("lattice,shape", [("D2Q9", (6, 5)), ("D3Q19", (4, 5, 3))])
def test_guo_forcing_momentum_growth(lattice, shape):
    """Uniform force in a periodic box: after t steps j = t F and the stored
    velocity of step t is (t + 1/2) F / rho."""
    T = periodic_box(lattice, shape)
    D = len(shape)
    F = np.array([2e-5, -1e-5, 3e-6])[:D]
    f0 = oracle.rest_state(T.lattice, T.n_nodes, T.bulk, np.ones(T.n_nodes))
    for t in (1, 10, 40):
        f, rho, vel = oracle.run_std_case(T.lattice, T.neighbors, T.node_types, T.bulk, f0, 0.9, F, t + 1)
        assert np.allclose(vel[T.bulk], (t + 0.5) * F[None, :], rtol=1e-12, atol=0)
        assert np.allclose(rho[T.bulk], 1.0, rtol=0, atol=1e-14)


def test_mass_conservation_with_walls():
    rng = np.random.default_rng(3)
    geo = (rng.random((12, 10, 8)) > 0.3).astype(np.int64)
    geo[:, 0] = geo[:, -1] = 0
    geo[:, :, 0] = geo[:, :, -1] = 0
    T = OracleTopo(geo, "D3Q19", "x")
    f0 = oracle.rest_state(1, T.n_nodes, T.bulk, 1.0 + 0.01 * rng.random(T.n_nodes))
    m0 = f0[T.bulk].sum()
    f, rho, vel = oracle.run_std_case(1, T.neighbors, T.node_types, T.bulk, f0, 0.7,
                                      np.array([1e-4, 2e-5, 0.0]), 60)
    assert abs(f[T.bulk].sum() - m0) <= 1e-12 * m0


def poiseuille_startup(y, t, F, nu, H, nterms=2001):
    """u(y,t) of a channel started from rest: steady parabola minus the odd
    sine series (yh = y - 1/2 measured from the half-way wall)."""
    yh = y - 0.5
    u = F / (2 * nu) * yh * (H - yh)
    for n in range(1, nterms, 2):
        u -= (4 * F * H ** 2 / (nu * math.pi ** 3 * n ** 3)) * np.sin(n * math.pi * yh / H) * \
            math.exp(-(n * math.pi) ** 2 * nu * t / H ** 2)
    return u


def test_channel_matches_poiseuille_startup_series():
    """C1: D2Q9 256x128 channel, tau=1, F=1e-6, 5000 steps, from the reference
    run; half-way walls at y=1/2 and y=126.5 (H=126), nu=1/6.  BGK + half-way
    bounce-back carries a tau-dependent slip, so the bar is 1e-3 of the peak."""
    d = load_ref("d2q9_channel_256x128")
    p = d["params"]
    y = np.arange(1, 127)
    u = poiseuille_startup(y, p["steps"] + 0.5, p["force"][0], 1.0 / 6.0, 126.0)
    err = np.abs(d["velx_profile"] - u).max() / u.max()
    assert err < 1e-3
    # the flow is far from steady at 5000 steps (about 39% of the steady peak)
    steady = p["force"][0] / (2 / 6.0) * 63 * 63
    assert 0.3 < d["velx_profile"].max() / steady < 0.5


def test_twophase_phase_swap_symmetry():
    d = load_ref("tp_pack_16")
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0, rho = tp_start(d, T)
    args = (p["tau0"], p["tau0"], p["sigma"], p["beta"], 0, [1e-6, 0.0, 0.0], 30)
    fa, ra, va, ca, _ = oracle.run_twophase(1, T.neighbors, T.node_types, T.bulk, f0, rho, *args)
    fb, rb, vb, cb, _ = oracle.run_twophase(1, T.neighbors, T.node_types, T.bulk, f0[:, ::-1].copy(),
                                            rho[:, ::-1].copy(), *args)
    b = T.bulk
    assert np.allclose(fa[b, 0], fb[b, 1], rtol=0, atol=1e-14)
    assert np.allclose(fa[b, 1], fb[b, 0], rtol=0, atol=1e-14)
    assert np.allclose(ca[b], -cb[b], rtol=0, atol=1e-12)
    assert np.abs(va[b] - vb[b]).max() <= 1e-12 * np.abs(va[b]).max()


def test_twophase_one_phase_limit_is_bgk():
    """rho1 = 0 everywhere and fully wetting walls: phi = 1, no colour
    gradient, no recolouring; the model is BGK + Guo with tau = tau0."""
    geo = OracleTopo(np.ones((6, 5, 4), dtype=np.int64), "D3Q19", "xyz")
    rng = np.random.default_rng(1)
    T = geo
    r0 = np.ones(T.n_nodes) + 0.001 * rng.random(T.n_nodes)
    f0 = np.zeros((T.n_nodes, 2, 19))
    W = oracle.WEIGHTS[1]
    f0[T.bulk, 0] = W[None] * r0[T.bulk, None]
    rho = np.stack([r0, np.zeros(T.n_nodes)], 1)
    F = [1e-5, 0.0, 2e-6]
    f2, *_ = oracle.run_twophase(1, T.neighbors, T.node_types, T.bulk, f0, rho, 0.8, 1.3, 0.01, 0.7, 0, F, 20)
    f1, _, _ = oracle.run_std_case(1, T.neighbors, T.node_types, T.bulk, f0[:, 0], 0.8, np.array(F), 20)
    assert np.allclose(f2[T.bulk, 0], f1[T.bulk], rtol=0, atol=1e-15)
    assert np.all(f2[T.bulk, 1] == 0.0)


def test_flux_sums_match_numpy():
    d = load_ref("tp_pack_6x5x4")
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0, rho = tp_start(d, T)
    f, *_ = oracle.run_twophase(1, T.neighbors, T.node_types, T.bulk, f0, rho, 1.0, 1.0, 0.01, 0.7, 0,
                                [1e-5, 0, 0], 5)
    from oracle.vtklb_oracle import BASIS
    c = BASIS["D3Q19"][:, 0]
    ft = f[T.bulk, 0] + f[T.bulk, 1]
    fc = ft @ c
    r0, r1 = f[T.bulk, 0].sum(1), f[T.bulk, 1].sum(1)
    s = oracle.flux(1, T.bulk, f, 2, 0)
    assert np.isclose(s[0], fc.sum(), rtol=1e-13, atol=1e-18)
    assert np.isclose(s[1], (r0 / (r0 + r1) * fc).sum(), rtol=1e-13, atol=1e-18)
    assert np.isclose(s[2], (r1 / (r0 + r1) * fc).sum(), rtol=1e-13, atol=1e-18)


def _forcing_fields():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_flux_fixture", os.path.join(os.path.dirname(__file__), "golden", "make_flux_fixture.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _seq(v):
    """Left-fold sum in order (the reference's accumulation loops)."""
    s = 0.0
    for x in v:
        s += float(x)
    return s


def test_flux_sums_pinned_to_reference_global_forcing():
    """oracle.flux against the reference's own calcCapNumbForceCartDir and
    calcFluxForceCartDir (LBglobalforcing.h:8-98), run by the compiled harness
    on the populations of every two-phase fixture and on random fields
    (forcing_ref.npz, tests/golden/make_flux_fixture.py).  With nu1 = 0 the
    capillary-number force is -2 <phi0 fc> / <phi0>, with nu0 = 0 it is
    -2 <phi1 fc> / <phi1>: each phase-weighted sum is pinned on its own, in
    each direction.  The flux forces pin sum fc of the total and of each
    field."""
    mod = _forcing_fields()
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "forcing_ref.npz"))
    seen = 0
    for name, T, f in mod.fields():
        assert str(ref[name + "/sha256"]) == mod.digest(f, T.bulk), "regenerated input differs for " + name
        b = T.bulk
        r0 = np.array([_seq(f[n, 0]) for n in b])
        r1 = np.array([_seq(f[n, 1]) for n in b])
        m0 = _seq(r0 / (r0 + r1))
        m1 = _seq(r1 / (r0 + r1))
        for (d, sca, nu0, nu1, ng), res in zip(ref[name + "/cases"], ref[name + "/result"]):
            d, ng = int(d), int(ng)
            s = oracle.flux(1, b, f, 2, d)
            s0a = oracle.flux(1, b, f[:, 0].copy(), 1, d)[0]
            s1a = oracle.flux(1, b, f[:, 1].copy(), 1, d)[0]
            cap = 2 * (sca - ((s[1] / ng) * nu0 + (s[2] / ng) * nu1)) / ((m0 / ng) * nu0 + (m1 / ng) * nu1)
            assert cap == res[0], (name, d, cap, res[0])
            assert 2 * (sca - s[0] / ng) == res[1], (name, d)
            assert 2 * (sca - s0a / ng) == res[2], (name, d)
            assert 2 * (sca - s1a / ng) == res[3], (name, d)
            seen += 1
    assert seen == 60


def test_flux_phase_sums_invariants():
    """s1 + s2 = s0 to rounding; one phase alone carries all of it; swapping
    the fields swaps the two phase sums exactly."""
    d = load_ref("tp_walls_12x14x10")
    p = d["params"]
    T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
    f0, rho = tp_start(d, T)
    f, *_ = oracle.run_twophase(1, T.neighbors, T.node_types, T.bulk, f0, rho, 1.0, 1.0, 0.01, 0.7, 0,
                                [1e-5, 2e-6, 0], 40)
    for direction in range(3):
        s = oracle.flux(1, T.bulk, f, 2, direction)
        assert abs(s[1] + s[2] - s[0]) <= 1e-13 * np.abs(f[T.bulk]).sum()
        sw = oracle.flux(1, T.bulk, f[:, ::-1].copy(), 2, direction)
        assert sw[0] == s[0] and sw[1] == s[2] and sw[2] == s[1]
        one = f.copy()
        one[:, 0] += one[:, 1]
        one[:, 1] = 0.0
        so = oracle.flux(1, T.bulk, one, 2, direction)
        assert so[1] == so[0] and so[2] == 0.0


def _init_cases():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_init_fixture", os.path.join(os.path.dirname(__file__), "golden", "make_init_fixture.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_init_equilibrium_matches_reference_initiate():
    """oracle.init_equilibrium against the reference's own initiateLbField
    (LBinitiatefield.h:33-57) run by the compiled harness (init_ref.npz),
    bit for bit, on every case and field."""
    mod = _init_cases()
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "init_ref.npz"))
    for name, lattice, geo, periodic, n_lb, n_rho, calls, seed in mod.cases():
        T, rho, vel = mod.inputs(name, lattice, geo, periodic, n_lb, n_rho, seed)
        assert np.array_equal(rho, ref[name + "/rho"]) and np.array_equal(vel, ref[name + "/vel"])
        Q = 9 if lattice == "D2Q9" else 19
        f = np.zeros((T.n_nodes, n_lb, Q))
        for lb, rf, vf in calls:
            oracle.init_equilibrium(T.lattice, T.bulk, f, lb, rho, rf, vel, vf)
        assert np.array_equal(f, ref[name + "/f"]), name


@pytest.mark.parametrize("lattice", [0, 1])
def test_init_equilibrium_moments(lattice):
    """The equilibrium carries rho, rho u and the momentum flux
    rho (c2 I + u u) (closed forms of the second-order equilibrium)."""
    from oracle.vtklb_oracle import BASIS
    c = BASIS["D2Q9" if lattice == 0 else "D3Q19"].astype(float)
    D = c.shape[1]
    n = 40
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.5, 1.5, (n, 1))
    vel = rng.normal(0.0, 0.05, (n, 1, D))
    f = np.zeros((n, 1, c.shape[0]))
    oracle.init_equilibrium(lattice, np.arange(n), f, 0, rho, 0, vel, 0)
    F = f[:, 0]
    u = vel[:, 0]
    assert np.allclose(F.sum(1), rho[:, 0], rtol=0, atol=1e-15)
    assert np.allclose(F @ c, rho * u, rtol=0, atol=1e-15)
    Pi = np.einsum("nq,qi,qj->nij", F, c, c)
    want = rho[:, :, None] * (np.eye(D)[None] / 3.0 + u[:, :, None] * u[:, None, :])
    assert np.allclose(Pi, want, rtol=0, atol=1e-15)
    # rest: f = w rho
    f0 = np.zeros_like(f)
    oracle.init_equilibrium(lattice, np.arange(n), f0, 0, rho, 0, np.zeros_like(vel), 0)
    assert np.array_equal(f0[:, 0], rho * oracle.WEIGHTS[lattice][None, :])
