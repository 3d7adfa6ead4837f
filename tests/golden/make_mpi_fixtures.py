"""Run the compiled reference std_case and twophase with the domain split over
MPI ranks and store every rank's fields, the fixtures of the halo exchange
(tests/test_gpu_halo.py).

Needs the reference binaries built by oracle/build_ref.sh and MPICH's
mpiexec (in the build container only; /root/reference is not available
elsewhere).  For every case:

  1. the fluid cells of the geometry are labelled 1..R by x slab, the
     reference's own multi-rank layout (oracle.ref_run.rank_slabs), and each
     rank's node list is written as input/mpi/tmp<rank>.vtklb: for std_case
     with the library's writer (byte-identical to PythonScripts/vtklb.py,
     tests/test_topology.py), init_rho = 1; for twophase with the oracle's
     restatement of that writer, as make_ref_fixtures.py writes the one-rank
     twophase cases, with the attributes rho0, rho1, wettability and source;
  2. std_case (D2Q9 as shipped, or its D3Q19 build) or twophase (D3Q19 as
     shipped, or its D2Q9 build) runs on R ranks under mpiexec, so the
     ranks exchange their halo through BndMpi /
     MonLatMpi::communicateLbField (LBbndmpi.h:206-315,
     LBmonlatmpi.h:236-300) every step (std_case/main.cpp:147,
     main_TWOPHASE.cpp:386-387), twophase also the colour field
     (communciateScalarField, main_TWOPHASE.cpp:287) and the flux sum
     (MPI_Allreduce, main_TWOPHASE.cpp:297);
  3. each rank's VTU of the last iteration is parsed: rho (rho0, rho1) and
     vel on its bulk nodes, in its bulk order; for twophase also the body
     force of force.dat (written by rank 0);
  4. the rank-labelled geometry, the parameters and the fields of every rank
     are stored in ref_mpi_<case>.npz (ref_mpi_tp_<case>.npz for twophase).

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
from make_ref_fixtures import fmt, input_dat, last_vtu, parse_vtu  # noqa: E402
from oracle.ref_run import MPIEXEC, REF, rank_blocks, rank_slabs  # noqa: E402
from oracle.vtklb_oracle import Geometry  # noqa: E402

LATTICE = {"D2Q9": lbamd.D2Q9, "D3Q19": lbamd.D3Q19}


def cases():
    """(name, lattice, fluid geometry, periodic, ranks, tau, force, steps,
    blocks per axis or None for x slabs)"""
    geo, _ = lbamd.sphere_pack((24, 16, 12), 3.0, 0.45, 5, periodic="x", walls=True)
    yield "d3q19_pack_2ranks", "D3Q19", (geo > 0).astype(np.int32), "x", 2, 0.87, [2e-5, -1e-5, 5e-6], 40, None
    geo, _ = lbamd.sphere_pack((30, 14, 12), 3.0, 0.45, 7, periodic="x", walls=True)
    yield "d3q19_pack_3ranks", "D3Q19", (geo > 0).astype(np.int32), "x", 3, 1.0, [1e-5, 0.0, 0.0], 30, None
    rng = np.random.default_rng(5)
    geo = (rng.random((40, 18)) > 0.3).astype(np.int32)
    geo[:, 0] = 0
    geo[:, -1] = 0
    yield "d2q9_random_2ranks", "D2Q9", geo, "x", 2, 0.8, [1e-5, -3e-6], 40, None
    # 2 x 2 blocks over x and y, periodic along both: ranks that share only an
    # edge exchange the diagonal directions
    geo, _ = lbamd.sphere_pack((20, 16, 10), 3.0, 0.5, 11, periodic="xy", walls=True)
    yield ("d3q19_pack_4ranks_2x2", "D3Q19", (geo > 0).astype(np.int32), "xy", 4, 0.9, [1e-5, 4e-6, -2e-6], 30,
           (2, 2))
    rng = np.random.default_rng(9)
    geo = (rng.random((24, 20)) > 0.3).astype(np.int32)
    yield "d2q9_random_4ranks_2x2", "D2Q9", geo, "xy", 4, 0.8, [1e-5, -3e-6], 40, (2, 2)


def run(name, lat, fluid, periodic, ranks, tau, force, steps, blocks=None):
    g = rank_slabs(fluid, ranks) if blocks is None else rank_blocks(fluid, blocks)
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
                                               "force": force, "steps": steps, "blocks": blocks})}
        od = os.path.join(root, "output")
        for r in range(ranks):
            data = parse_vtu(last_vtu(od, "lb_run").replace("/0000_lb_run_", "/%04d_lb_run_" % r))
            out["rho_%d" % r] = data["rho"]
            out["vel_%d" % r] = data["vel"]
        np.savez_compressed(os.path.join(HERE, "ref_mpi_%s.npz" % name), **out)
        print(name, ranks, "ranks:", [len(out["rho_%d" % r]) for r in range(ranks)], "bulk nodes")
    finally:
        shutil.rmtree(root, ignore_errors=True)


def tp_cases():
    """(name, lattice, fluid geometry, periodic, ranks, rho0, wettability,
    tau0, tau1, sigma, beta, momx, force, steps)"""
    rng = np.random.default_rng(31)
    geo, _ = lbamd.sphere_pack((20, 12, 10), 3.0, 0.5, 3, periodic="xyz", walls=False)
    fluid = (geo > 0).astype(np.int32)
    r0 = rng.choice([0.0, 0.3, 1.0], size=fluid.shape)
    yield ("pack_2ranks", "D3Q19", fluid, "xyz", 2, r0, np.full(fluid.shape, 0.8), 0.8, 1.2, 0.01, 0.7, 1e-4,
           [0.0, 1e-6, -2e-6], 60)
    geo, _ = lbamd.sphere_pack((24, 12, 10), 3.0, 0.55, 9, periodic="x", walls=True)
    fluid = (geo > 0).astype(np.int32)
    r0 = (rng.random(fluid.shape) > 0.5).astype(float)
    yield ("walls_3ranks", "D3Q19", fluid, "x", 3, r0, np.full(fluid.shape, 1.0), 1.0, 1.0, 0.01, 0.7, 5e-5,
           [0.0, 0.0, 0.0], 40)
    fluid = (rng.random((36, 16)) > 0.3).astype(np.int32)
    fluid[:, 0] = 0
    fluid[:, -1] = 0
    r0 = (rng.random(fluid.shape) > 0.5).astype(float)
    yield ("d2q9_walls_2ranks", "D2Q9", fluid, "x", 2, r0, np.full(fluid.shape, 0.8), 0.8, 1.2, 0.01, 0.7, 5e-5,
           [0.0, -1e-6], 60)


def run_tp(name, lat, fluid, periodic, ranks, r0, wett, tau0, tau1, sigma, beta, momx, force, steps):
    g = rank_slabs(fluid, ranks)
    geo = Geometry(g.astype(np.int64), lat, periodic)
    attrs = [("rho0", r0), ("rho1", 1.0 - r0), ("wettability", wett),
             ("source", np.zeros(g.shape, dtype=np.int64))]
    root = tempfile.mkdtemp(prefix="lbmpi_tp_")
    try:
        os.makedirs(os.path.join(root, "input", "mpi"))
        os.makedirs(os.path.join(root, "output", "tp"))
        for r in range(1, ranks + 1):
            with open(os.path.join(root, "input", "mpi", "tmp%d.vtklb" % (r - 1)), "w") as fh:
                fh.write(geo.vtklb_text(r, attrs))
        with open(os.path.join(root, "input", "input.dat"), "w") as fh:
            fh.write(input_dat(steps, ["tau %s" % fmt([tau0, tau1]), "bodyforce %s" % fmt(force),
                                       "sigma %s" % fmt(sigma), "beta %s" % fmt(beta), "momx %s" % fmt(momx)],
                               extra="outdir tp/\n"))
        env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(REF, "mpilib"), OMP_NUM_THREADS="1")
        binary = os.path.join(REF, "twophase" if lat == "D3Q19" else "twophase_d2q9")
        res = subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(ranks), binary],
                             cwd=root, env=env, capture_output=True, text=True, timeout=600)
        if res.returncode != 0:
            raise RuntimeError("%s failed: %s %s" % (name, res.stdout[-2000:], res.stderr[-2000:]))
        od = os.path.join(root, "output", "tp")
        params = dict(case="twophase", lattice=lat, periodic=periodic, ranks=ranks, tau0=tau0, tau1=tau1,
                      sigma=sigma, beta=beta, momx=momx, force=force, steps=steps)
        out = {"geo": g.astype(np.int8), "params": json.dumps(params), "rho0_attr": r0, "wett_attr": wett,
               "force_x": np.loadtxt(os.path.join(od, "force.dat"), ndmin=2)}
        for r in range(ranks):
            data = parse_vtu(last_vtu(od, "fluid").replace("/0000_fluid_", "/%04d_fluid_" % r))
            for k in ("rho0", "rho1", "vel"):
                out["%s_%d" % (k, r)] = data[k]
        np.savez_compressed(os.path.join(HERE, "ref_mpi_tp_%s.npz" % name), **out)
        print("tp_" + name, ranks, "ranks:", [len(out["rho0_%d" % r]) for r in range(ranks)], "bulk nodes,",
              "force", out["force_x"][-1])
    finally:
        shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    for c in cases():
        run(*c)
    for c in tp_cases():
        run_tp(*c)
