"""Run the reference's own initiateLbField on given density and velocity
fields and store the populations it sets as a fixture (init_ref.npz).

Needs oracle/_ref/init_ref, built by oracle/build_ref.sh from
oracle/init_harness.cpp against the reference's headers (in the build
container only).  The harness calls initiateLbField
(src/lbsolver/LBinitiatefield.h:33-57, the overload with separate
population, density and velocity field numbers).

Cases (node lists of committed .vtklb fixtures through the oracle's writer):
  d3q19_two_fields   D3Q19 6x5x4 pack, two population fields set from the
                     two density fields and one velocity field
  d3q19_walls        D3Q19 12x14x10 pack with walls, one field, |u| up to 0.1
  d2q9_random        D2Q9 7x5 random geometry, one field
Densities are drawn in [0.5, 1.5], velocities N(0, 0.04) per component.

Run:  python tests/golden/make_init_fixture.py
"""
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

from helpers import OracleTopo, load_ref  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "init_ref")
VT = os.path.join(HERE, "vtklb")


def cases():
    """(name, lattice, geo, periodic, n_lb, n_rho, calls, seed); calls are
    (lbFieldNo, rhoFieldNo, velFieldNo)."""
    g = load_ref("d3q19_pack_6x5x4_pxyz")["geo"]
    yield "d3q19_two_fields", "D3Q19", g, "xyz", 2, 2, [(0, 0, 0), (1, 1, 0)], 1
    g = load_ref("tp_walls_12x14x10")["geo"]
    yield "d3q19_walls", "D3Q19", g, "x", 1, 1, [(0, 0, 0)], 2
    g = load_ref("d2q9_random_7x5")["geo"]
    yield "d2q9_random", "D2Q9", g, "xy", 1, 1, [(0, 0, 0)], 3


def inputs(name, lattice, geo, periodic, n_lb, n_rho, seed):
    T = OracleTopo(np.asarray(geo).astype(np.int64), lattice, periodic)
    nd = 2 if lattice == "D2Q9" else 3
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.5, 1.5, (T.n_nodes, n_rho))
    vel = rng.normal(0.0, 0.04, (T.n_nodes, 1, nd))
    if name == "d3q19_walls":
        vel = np.clip(vel, -0.1, 0.1)
    return T, rho, vel


def run_harness(T, rho, vel, n_lb, calls):
    nd = vel.shape[2]
    lat = 0 if nd == 2 else 1
    Q = 9 if nd == 2 else 19
    with tempfile.TemporaryDirectory() as tmp:
        path, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(path, "wb") as fh:
            fh.write(struct.pack("<iii", lat, T.n_nodes, T.bulk.size))
            fh.write(np.asarray(T.bulk, dtype="<i4").tobytes())
            fh.write(struct.pack("<iii", n_lb, rho.shape[1], vel.shape[1]))
            fh.write(np.ascontiguousarray(rho, dtype="<f8").tobytes())
            fh.write(np.ascontiguousarray(vel, dtype="<f8").tobytes())
            fh.write(struct.pack("<i", len(calls)))
            for c in calls:
                fh.write(struct.pack("<iii", *c))
        subprocess.run([HARNESS, path, out], check=True)
        return np.fromfile(out, dtype="<f8").reshape(T.n_nodes, n_lb, Q)


def main():
    if not os.path.exists(HARNESS):
        sys.exit("make_init_fixture: %s absent; run oracle/build_ref.sh first" % HARNESS)
    store = {}
    names = []
    for name, lattice, geo, periodic, n_lb, n_rho, calls, seed in cases():
        T, rho, vel = inputs(name, lattice, geo, periodic, n_lb, n_rho, seed)
        f = run_harness(T, rho, vel, n_lb, calls)
        names.append(name)
        store[name + "/rho"] = rho
        store[name + "/vel"] = vel
        store[name + "/calls"] = np.array(calls, dtype=np.int32)
        store[name + "/f"] = f
        print("%-18s %s %d nodes, %d bulk" % (name, lattice, T.n_nodes, T.bulk.size))
    store["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "init_ref.npz"), **store)
    print("wrote init_ref.npz")


if __name__ == "__main__":
    main()
