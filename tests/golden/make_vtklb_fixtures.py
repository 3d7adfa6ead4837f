"""Write small .vtklb golden fixtures with the reference's own geometry writer.

The node numbering and the neighbour index of BADChIMP are defined only by the
reference's Python writer (PythonScripts/vtklb.py:6-44, 70-124, 182-218,
239-345).  This script imports that writer from the read-only reference tree
(in the build container only), runs it on a handful of tiny seeded geometries
and stores the files it writes under tests/golden/vtklb/.  The files are data:
inputs (the geometry arrays, stored beside them as .npy) and outputs (the
.vtklb text).  The tests compare the oracle's restatement of the writer, and
the library's native node-list builder, against these files.

Run:  python tests/golden/make_vtklb_fixtures.py   (needs /root/reference)
"""
import contextlib
import io
import os
import sys

import numpy as np

REF_SCRIPTS = "/root/reference/PythonScripts"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vtklb")


def cases():
    """Yield (name, geo, lattice, periodic, attributes) for every fixture."""
    rng = np.random.default_rng(20240611)

    # D2Q9 channel 8x6, solid rows y=0 and y=5, periodic in x.
    geo = np.ones((8, 6), dtype=int)
    geo[:, 0] = 0
    geo[:, -1] = 0
    yield "d2q9_channel_8x6_px", geo, "D2Q9", "x", {
        "init_rho": np.ones((8, 6), dtype=float)}

    # D2Q9 random solids, fully periodic.
    geo = (rng.random((7, 5)) > 0.3).astype(int)
    yield "d2q9_random_7x5_pxy", geo, "D2Q9", "xy", {
        "init_rho": 1.0 + 0.01 * rng.random((7, 5))}

    # D2Q9 two ranks split in x, solid walls, periodic in x.
    geo = np.ones((8, 6), dtype=int)
    geo[4:, :] = 2
    geo[:, 0] = 0
    geo[:, -1] = 0
    geo[2, 2] = 0
    yield "d2q9_2rank_8x6_px", geo, "D2Q9", "x", {}

    # D3Q19 random pack, fully periodic.
    geo = (rng.random((6, 5, 4)) > 0.35).astype(int)
    yield "d3q19_pack_6x5x4_pxyz", geo, "D3Q19", "xyz", {
        "init_rho": np.ones((6, 5, 4)),
        "rho0": (rng.random((6, 5, 4)) > 0.5).astype(float),
        "wettability": np.ones((6, 5, 4)),
        "source": np.zeros((6, 5, 4), dtype=int)}

    # D3Q19 random pack, periodic in x, solid walls on the y and z faces.
    geo = (rng.random((6, 5, 4)) > 0.35).astype(int)
    geo[:, 0, :] = 0
    geo[:, -1, :] = 0
    geo[:, :, 0] = 0
    geo[:, :, -1] = 0
    yield "d3q19_pack_6x5x4_px_walls", geo, "D3Q19", "x", {
        "init_rho": np.ones((6, 5, 4))}

    # D3Q19 open (non-periodic) box: fluid links to the ghost node 0 exist.
    geo = (rng.random((4, 4, 3)) > 0.25).astype(int)
    yield "d3q19_open_4x4x3", geo, "D3Q19", "", {}

    # D3Q19 two ranks split in y, fully periodic.
    geo = (rng.random((5, 6, 4)) > 0.3).astype(int)
    geo[:, 3:, :] *= 2
    yield "d3q19_2rank_5x6x4_pxyz", geo, "D3Q19", "xyz", {}


def main():
    sys.path.insert(0, REF_SCRIPTS)
    import vtklb  # noqa: E402  (reference writer, read-only tree)

    os.makedirs(OUT, exist_ok=True)
    for name, geo, lattice, periodic, attrs in cases():
        np.save(os.path.join(OUT, name + "_geo.npy"), geo)
        with contextlib.redirect_stdout(io.StringIO()):
            w = vtklb.vtklb(geo, lattice, periodic, name, OUT + "/")
            for key in sorted(attrs):
                w.append_data_set(key, attrs[key])
        for key in sorted(attrs):
            np.save(os.path.join(OUT, "%s_attr_%s.npy" % (name, key)), attrs[key])
        print("wrote", name, "ranks:", int(np.amax(geo)))


if __name__ == "__main__":
    main()
