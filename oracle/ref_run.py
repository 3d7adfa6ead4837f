"""Timing of the compiled reference (oracle/_ref/std_case_d3q19, built from
src/std_case/main.cpp by oracle/build_ref.sh) on the host's cores, as the
CPU baseline of bench.py.  Test infrastructure: only that leg runs it.

The reference runs as it does in production: one MPI rank per core (MPICH's
mpiexec, local fork launcher), each rank reading its own node list
input/mpi/tmp<rank>.vtklb (std_case/main.cpp:34-45) and the input.dat
blocks <iterations> max/write and <fluid> tau/bodyforce
(std_case/main.cpp:53-59).  The domain is cut into R slabs along x, the
reference's own multi-rank layout (PythonScripts/vtklb.py:92-94 labels cells
by rank).  Two runs of n_a and n_b steps are timed from the outside and
differenced, which removes start-up, the node-list read and the VTK writes
(std_case/main.cpp:151 writes at step 0 and at the last step in both).
"""
import os
import shutil
import subprocess
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "_ref")
BINARY = os.path.join(REF, "std_case_d3q19")
MPIEXEC = "/opt/conda/bin/mpiexec"


def available():
    """The reference binary, its MPI runtime and a launcher are present."""
    return (os.path.exists(BINARY) and os.path.exists(MPIEXEC)
            and os.path.exists(os.path.join(REF, "mpilib", "libmpi.so.12")))


def rank_slabs(geo, ranks):
    """Fluid cells of geo (1 fluid / 0 solid) labelled 1..ranks by x slab."""
    g = np.ascontiguousarray(geo, dtype=np.int32).copy()
    for r, xs in enumerate(np.array_split(np.arange(g.shape[0]), ranks)):
        sl = g[xs[0]:xs[-1] + 1]
        sl[sl > 0] = r + 1
    return g


def rank_blocks(geo, counts):
    """Fluid cells of geo labelled 1..prod(counts) by block: counts[a] blocks
    along axis a (axes beyond len(counts) are not cut), numbered in C order of
    the block index."""
    g = np.ascontiguousarray(geo, dtype=np.int32).copy()
    fluid = g > 0
    label = np.zeros(g.shape, dtype=np.int32)
    stride = 1
    for a in reversed(range(len(counts))):
        idx = np.zeros(g.shape[a], dtype=np.int32)
        for b, xs in enumerate(np.array_split(np.arange(g.shape[a]), counts[a])):
            idx[xs] = b
        shape = [1] * g.ndim
        shape[a] = g.shape[a]
        label += stride * idx.reshape(shape)
        stride *= counts[a]
    g[fluid] = label[fluid] + 1
    return g


def _input_dat(steps, tau, force):
    f = " ".join(repr(float(x)) for x in force)
    return ("<iterations>\n    max %d\n    write %d\n<end>\n<fluid>\n    tau %r\n    bodyforce %s\n<end>\n"
            % (steps, steps, float(tau), f))


def time_std_case(geo, periodic, tau, force, ranks, n_a, n_b, timeout=600):
    """Seconds per (n_b - n_a) steps of the reference std_case (D3Q19) on
    geo split into `ranks` MPI ranks, and the bulk node count.  The node
    lists are written with the library's vtklb writer (same file as
    PythonScripts/vtklb.py writes; tests/test_topology.py)."""
    import lbamd
    root = tempfile.mkdtemp(prefix="lbref_")
    try:
        os.makedirs(os.path.join(root, "input", "mpi"))
        os.makedirs(os.path.join(root, "output"))
        os.makedirs(os.path.join(root, "run"))
        g = rank_slabs(geo, ranks)
        n_bulk = 0
        for r in range(1, ranks + 1):
            T = lbamd.Topology.from_geometry(g, lbamd.D3Q19, periodic, rank=r)
            n_bulk += T.n_bulk
            T.write_vtklb(os.path.join(root, "input", "mpi", "tmp%d.vtklb" % (r - 1)),
                          [("init_rho", np.ones(T.n_nodes))])
            del T
        env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(REF, "mpilib"), OMP_NUM_THREADS="1")
        wall = {}
        for n in (n_a, n_b):
            with open(os.path.join(root, "input", "input.dat"), "w") as fh:
                fh.write(_input_dat(n, tau, force))
            t0 = time.perf_counter()
            res = subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(ranks), BINARY],
                                 cwd=os.path.join(root, "run"), env=env, capture_output=True, text=True,
                                 timeout=timeout)
            wall[n] = time.perf_counter() - t0
            if res.returncode != 0:
                raise RuntimeError("reference std_case (%d ranks, %d steps) failed: %s %s"
                                   % (ranks, n, res.stdout[-500:], res.stderr[-500:]))
        return wall[n_b] - wall[n_a], n_bulk, wall
    finally:
        shutil.rmtree(root, ignore_errors=True)
