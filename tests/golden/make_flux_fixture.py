"""Run the reference's own global-forcing functions on two-phase fields and
store their results as a fixture (forcing_ref.npz).

Needs oracle/_ref/flux_ref, built by oracle/build_ref.sh from
oracle/flux_harness.cpp against the reference's headers (in the build
container only).  The harness calls calcCapNumbForceCartDir and
calcFluxForceCartDir (src/lbsolver/LBglobalforcing.h:8-98).

Fields: for each two-phase reference fixture (ref_tp_*.npz), the populations
after its full run (the oracle's colour-gradient loop from the fixture's
start, in the momentum-controlled mode the fixture was made with; those
runs are bit-identical to the compiled reference on rho and u), plus a
random positive field on the same node list.  The fixture keeps, per field,
a sha256 of the populations (so the test can check it regenerated the same
input), the list of cases (direction, sigma*Ca, nu0, nu1, n_global) and the
harness output per case: the capillary-number force, and the flux force of
the total field and of each of the two fields.

Run:  python tests/golden/make_flux_fixture.py
"""
import hashlib
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from helpers import OracleTopo, load_ref, ref_names, tp_start  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "flux_ref")
MPILIB = os.path.join(ROOT, "oracle", "_ref", "mpilib")


def fields():
    """(name, topology, populations (n_nodes, 2, 19)) for every field the
    fixture covers (the D3Q19 two-phase fixtures; the harness is built for
    D3Q19); deterministic."""
    for name in ref_names("tp_"):
        d = load_ref(name)
        p = d["params"]
        if p["lattice"] != "D3Q19":
            continue
        T = OracleTopo(d["geo"].astype(np.int64), p["lattice"], p["periodic"])
        f0, rho = tp_start(d, T)
        f, *_ = oracle.run_twophase(1, T.neighbors, T.node_types, T.bulk, f0, rho, p["tau0"], p["tau1"],
                                    p["sigma"], p["beta"], 1, [p["momx"]] + p["force"][1:], p["steps"] + 1)
        yield name, T, f
        rng = np.random.default_rng(len(name))
        fr = np.zeros_like(f)
        fr[T.bulk] = rng.uniform(0.0, 0.1, (T.bulk.size, 2, 19))
        yield name + "_random", T, fr


def cases(n_bulk):
    """(dir, sigma*Ca, nu0, nu1, n_global): each phase-weighted sum alone
    (nu of the other phase 0), both together, and with a nonzero sigma*Ca."""
    out = []
    for d in range(3):
        out += [(d, 0.0, 1.0, 0.0, n_bulk), (d, 0.0, 0.0, 1.0, n_bulk), (d, 1e-3, 1.0 / 6.0, 1.0 / 3.0, n_bulk)]
    out.append((0, 2.5e-4, 0.1, 0.7, 3 * n_bulk))
    return out


def digest(f, bulk):
    return hashlib.sha256(np.ascontiguousarray(f[bulk], dtype="<f8").tobytes()).hexdigest()


def run_harness(T, f, cs):
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "in.bin")
        with open(path, "wb") as fh:
            fh.write(struct.pack("<iii", 1, T.n_nodes, T.bulk.size))
            fh.write(np.asarray(T.bulk, dtype="<i4").tobytes())
            fh.write(np.ascontiguousarray(f, dtype="<f8").tobytes())
            fh.write(struct.pack("<i", len(cs)))
            for d, s, n0, n1, ng in cs:
                fh.write(struct.pack("<idddi", d, s, n0, n1, ng))
        env = dict(os.environ, LD_LIBRARY_PATH=MPILIB)
        out = subprocess.run([HARNESS, path], capture_output=True, text=True, env=env, check=True).stdout
    return np.array([[float(v) for v in line.split()] for line in out.strip().split("\n")])


def main():
    if not os.path.exists(HARNESS):
        sys.exit("make_flux_fixture: %s absent; run oracle/build_ref.sh first" % HARNESS)
    store = {}
    names = []
    for name, T, f in fields():
        cs = cases(int(T.bulk.size))
        res = run_harness(T, f, cs)
        names.append(name)
        store[name + "/cases"] = np.array(cs, dtype=np.float64)
        store[name + "/result"] = res
        store[name + "/sha256"] = np.array(digest(f, T.bulk))
        print("%-28s %d bulk nodes, %d cases" % (name, T.bulk.size, len(cs)))
    store["names"] = np.array(names)
    np.savez(os.path.join(HERE, "forcing_ref.npz"), **store)
    print("wrote forcing_ref.npz")


if __name__ == "__main__":
    main()
