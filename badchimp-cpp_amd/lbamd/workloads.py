"""Seeded synthetic workloads: the batched sphere-pack ensemble and the
serving request stream (shared by the benchmark and the GPU tests)."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .ffi import sphere_pack
from .serve import Request

# colour-gradient parameters of the relative-permeability requests
TWO_PHASE = {"tau0": 1.0, "tau1": 1.0, "sigma": 0.01, "beta": 0.7, "wett": 1.0}


def _workers():
    return min(16, os.cpu_count() or 1)


def ensemble(n, seed=2024, indices=None):
    """Geometries of the batched ensemble: cubic sizes log-uniform in
    32..128, porosity U(0.30, 0.50), sphere radius U(3, 8), all from one seed
    (periodic x, walls on the y and z faces).  indices: build only these
    samples (the parameters of every sample are still drawn in order).
    Returns (geometries, parameters)."""
    rng = np.random.default_rng(seed)
    L = np.rint(np.exp(rng.uniform(np.log(32), np.log(128), n))).astype(int)
    phi = rng.uniform(0.30, 0.50, n)
    rad = rng.uniform(3.0, 8.0, n)
    seeds = rng.integers(1, 2**62, n)
    idx = list(range(n)) if indices is None else [int(i) for i in indices]

    def one(i):
        geo, _ = sphere_pack((L[i],) * 3, rad[i], phi[i], int(seeds[i]), periodic="x", walls=True)
        return geo

    with ThreadPoolExecutor(max_workers=_workers()) as ex:
        geos = list(ex.map(one, idx))
    return geos, {"L": L[idx], "porosity": phi[idx], "radius": rad[idx]}


def request_mix(n, seed, t_rate=None, tol=1e-6, max_steps=400000, indices=None, L_range=(48, 160)):
    """Serving requests: cubic sizes U(48, 160), porosity U(0.30, 0.50),
    radius U(3, 8); 70 % permeability (single-phase, tau 1) and 30 % relative
    permeability (colour gradient, phases drawn per pore cell with wetting
    saturation U(0.2, 0.8)); body force 1e-6 along x; Poisson arrivals at
    rate t_rate (all at t = 0 without one).  One stop rule for both kinds:
    the relative change of k (both phase permeabilities, relative to their
    sum, for two-phase) at most tol per check interval, at two consecutive
    checks.  indices: build only these requests."""
    rng = np.random.default_rng(seed)
    L = rng.integers(L_range[0], L_range[1] + 1, n)
    two = rng.random(n) < 0.3
    phi = rng.uniform(0.30, 0.50, n)
    rad = rng.uniform(3.0, 8.0, n)
    sw = rng.uniform(0.2, 0.8, n)
    seeds = rng.integers(1, 2**62, n)
    arrivals = np.cumsum(rng.exponential(1.0 / t_rate, n)) if t_rate else np.zeros(n)
    idx = list(range(n)) if indices is None else [int(i) for i in indices]
    tp = TWO_PHASE

    def one(i):
        geo, _ = sphere_pack((L[i],) * 3, rad[i], phi[i], int(seeds[i]), periodic="x", walls=True)
        if not two[i]:
            return Request(i, geo, "x", tau=1.0, force=(1e-6, 0.0, 0.0), tol=tol, max_steps=max_steps,
                           t_arrival=float(arrivals[i]))
        r = np.random.default_rng(int(seeds[i]) + 1)
        rho0 = np.where(geo > 0, (r.random(geo.shape) < sw[i]).astype(float), 0.0)
        return Request(i, geo, "x", force=(1e-6, 0.0, 0.0), tol=tol, max_steps=max_steps,
                       t_arrival=float(arrivals[i]), kind="relperm", rho0=rho0, wett=tp["wett"],
                       tau0=tp["tau0"], tau1=tp["tau1"], sigma=tp["sigma"], beta=tp["beta"])

    with ThreadPoolExecutor(max_workers=_workers()) as ex:
        return list(ex.map(one, idx))
