"""Run the compiled reference std_case with its domain split over MPI ranks and
store every rank's fields, the fixtures of the halo exchange
(tests/test_gpu_halo.py).

Needs the reference binaries built by oracle/build_ref.sh and MPICH's
mpiexec (in the build container only; /root/reference is not available
elsewhere).  For every case:

  1. the fluid cells of the geometry are labelled 1..R by x slab, the
     reference's own multi-rank layout (oracle.ref_run.rank_slabs), and each
     rank's node list is written as input/mpi/tmp<rank>.vtklb with the
     library's writer (byte-identical to PythonScripts/vtklb.py,
     tests/test_topology.py), init_rho = 1;
  2. std_case (D2Q9 as shipped, or its D3Q19 build) runs on R ranks under
     mpiexec, so the ranks exchange their halo through BndMpi /
     MonLatMpi::communicateLbField (LBbndmpi.h:206-315, LBmonlatmpi.h:236-300)
     every step (std_case/main.cpp:147);
  3. each rank's VTU of the last iteration is parsed: rho and vel on its bulk
     nodes, in its bulk order;
  4. the rank-labelled geometry, the parameters and the fields of every rank
     are stored in ref_mpi_<case>.npz.

Run:  python tests/golden/make_mpi_fixtures.py
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "badchimp-cpp_amd"), HERE]

import lbamd  # noqa: E402
from make_ref_fixtures import last_vtu, parse_vtu  # noqa: E402
from oracle.ref_run import MPIEXEC, REF, rank_slabs  # noqa: E402

LATTICE = {"D2Q9": lbamd.D2Q9, "D3Q19": lbamd.D3Q19}


def cases():
    """(name, lattice, fluid geometry, periodic, ranks, tau, force, steps)"""
    geo, _ = lbamd.sphere_pack((24, 16, 12), 3.0, 0.45, 5, periodic="x", walls=True)
    yield "d3q19_pack_2ranks", "D3Q19", (geo > 0).astype(np.int32), "x", 2, 0.87, [2e-5, -1e-5, 5e-6], 40
    geo, _ = lbamd.sphere_pack((30, 14, 12), 3.0, 0.45, 7, periodic="x", walls=True)
    yield "d3q19_pack_3ranks", "D3Q19", (geo > 0).astype(np.int32), "x", 3, 1.0, [1e-5, 0.0, 0.0], 30
    rng = np.random.default_rng(5)
    geo = (rng.random((40, 18)) > 0.3).astype(np.int32)
    geo[:, 0] = 0
    geo[:, -1] = 0
    yield "d2q9_random_2ranks", "D2Q9", geo, "x", 2, 0.8, [1e-5, -3e-6], 40


def run(name, lat, fluid, periodic, ranks, tau, force, steps):
    g = rank_slabs(fluid, ranks)
    root = tempfile.mkdtemp(prefix="lbmpi_")
    try:
        for d in ("input/mpi", "output", "run"):
            os.makedirs(os.path.join(root, d))
        for r in range(1, ranks + 1):
            t = lbamd.Topology.from_geometry(g, LATTICE[lat], periodic, rank=r)
            t.write_vtklb(os.path.join(root, "input", "mpi", "tmp%d.vtklb" % (r - 1)),
                          [("init_rho", np.ones(t.n_nodes))])
            t.close()
        with open(os.path.join(root, "input", "input.dat"), "w") as fh:
            fh.write("<iterations>\n    max %d\n    write %d\n<end>\n<fluid>\n    tau %r\n    bodyforce %s\n<end>\n"
                     % (steps, steps, float(tau), " ".join(repr(float(x)) for x in force)))
        env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(REF, "mpilib"), OMP_NUM_THREADS="1")
        binary = os.path.join(REF, "std_case_" + lat.lower())
        res = subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(ranks), binary], cwd=os.path.join(root, "run"),
                             env=env, capture_output=True, text=True, timeout=600)
        if res.returncode != 0:
            raise RuntimeError("%s failed: %s %s" % (name, res.stdout[-2000:], res.stderr[-2000:]))
        out = {"geo": g, "params": json.dumps({"lattice": lat, "periodic": periodic, "ranks": ranks, "tau": tau,
                                               "force": force, "steps": steps})}
        od = os.path.join(root, "output")
        for r in range(ranks):
            data = parse_vtu(last_vtu(od, "lb_run").replace("/0000_lb_run_", "/%04d_lb_run_" % r))
            out["rho_%d" % r] = data["rho"]
            out["vel_%d" % r] = data["vel"]
        np.savez_compressed(os.path.join(HERE, "ref_mpi_%s.npz" % name), **out)
        print(name, ranks, "ranks:", [len(out["rho_%d" % r]) for r in range(ranks)], "bulk nodes")
    finally:
        shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    for c in cases():
        run(*c)
