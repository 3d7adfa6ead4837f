"""Oracle: numpy restatement of the reference geometry writer (node list + neighbour index).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module; only
tests/, tests/golden/ scripts, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` may use it, and only as the checker.

Follows PythonScripts/vtklb.py of the reference:
  * padding with a rim of -1 and periodic rim copies x -> y -> z   (vtklb.py:84-114)
  * per-rank labels 1..N_r in C order over the padded array        (vtklb.py:92-94)
  * node list: own fluid (label order), then solid (geo == 0) cells
    adjacent to own fluid in C order, then other-rank cells adjacent to
    own fluid, one block per rank in set-iteration order          (vtklb.py:182-218)
  * neighbour index: local label of r + c_q, 0 off the array, the
    per-rank label when the cell is an own-rank rim copy           (vtklb.py:6-44)
  * points written as padded index - 1                             (vtklb.py:239-241)
  * PROCESSOR blocks [local label, label on the other rank]        (vtklb.py:271-281)
  * nodetype = geo > 0, attributes padded with 0 + periodic copies (vtklb.py:289-345)

Pinned against files written by the reference writer itself
(tests/golden/vtklb/*.vtklb, produced by tests/golden/make_vtklb_fixtures.py):
the text this module writes must equal those files byte for byte.
"""
import numpy as np

BASIS = {
    "D2Q9": np.array([[1, 0], [1, 1], [0, 1], [-1, 1], [-1, 0], [-1, -1],
                      [0, -1], [1, -1], [0, 0]], dtype=int),
    "D3Q19": np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, -1, 0],
                       [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1], [-1, 0, 0],
                       [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0],
                       [-1, 0, -1], [-1, 0, 1], [0, -1, -1], [0, -1, 1],
                       [0, 0, 0]], dtype=int),
}


def _periodic_copy(a, periodic):
    """Copy the far interior layer into the opposite rim, axis by axis (x, y, z)."""
    for axis, name in enumerate("xyz"[: a.ndim]):
        if name in periodic:
            lo = [slice(None)] * a.ndim
            hi = [slice(None)] * a.ndim
            lo[axis], hi[axis] = 0, -2
            a[tuple(lo)] = a[tuple(hi)]
            lo[axis], hi[axis] = -1, 1
            a[tuple(lo)] = a[tuple(hi)]
    return a


class RankTopology:
    """Node list and neighbour index of one rank, as the reference writer defines it."""

    def __init__(self, points, neighbors, nodetype, processors, ind_local, rank):
        self.points = points            # int[(N) , nD]  (node 1..N; node 0 not listed)
        self.neighbors = neighbors      # int[(N) , Q]   (file basis order)
        self.nodetype = nodetype        # int[(N)]       1 fluid, 0 solid
        self.processors = processors    # list of (rank0based, int[n,2])
        self.ind_local = ind_local      # tuple of padded-index arrays
        self.rank = rank                # 1-based

    @property
    def n_points(self):
        return self.points.shape[0]


class Geometry:
    """Restatement of the reference writer's set-up for a whole geometry."""

    def __init__(self, geo, lattice, periodic=""):
        geo = np.asarray(geo)
        self.nd = geo.ndim
        self.shape = geo.shape
        self.lattice = lattice
        self.basis = BASIS[lattice]
        self.periodic = periodic.lower()
        pad = tuple(n + 2 for n in geo.shape)
        inner = tuple([slice(1, -1)] * self.nd)
        self.geo = np.full(pad, -1, dtype=int)
        self.geo[inner] = geo
        self.interior = np.zeros(pad, dtype=bool)
        self.interior[inner] = True
        self.label = np.zeros(pad, dtype=int)
        self.n_ranks = int(np.amax(geo))
        for r in range(1, self.n_ranks + 1):
            sel = (self.geo == r) & self.interior
            self.label[sel] = np.arange(1, np.count_nonzero(sel) + 1)
        _periodic_copy(self.geo, self.periodic)
        _periodic_copy(self.label, self.periodic)

    def _shifted(self, cells, c):
        return tuple(ax + d for ax, d in zip(cells, c))

    def rank_topology(self, rank):
        """Build the node list of ``rank`` (1-based)."""
        fluid = np.nonzero((self.geo == rank) & self.interior)   # C order == label order
        blocks = [fluid]
        # mark solid cells reached from own fluid through any lattice link
        mark = np.zeros(self.geo.shape, dtype=bool)
        for c in self.basis:
            nb = self._shifted(fluid, c)
            hit = self.geo[nb] == 0
            mark[tuple(a[hit] for a in nb)] = True
        blocks.append(np.nonzero(mark))
        # neighbouring ranks, iterated in the same set order as the writer
        ranks = set()
        for c in self.basis:
            ranks = ranks.union(set(self.geo[self._shifted(fluid, c)]))
        for bad in (-1, 0, rank):
            ranks.discard(bad)
        order = list(ranks)
        for r in order:
            mark[:] = False
            for c in self.basis:
                nb = self._shifted(fluid, c)
                hit = self.geo[nb] == r
                mark[tuple(a[hit] for a in nb)] = True
            blocks.append(np.nonzero(mark))
        ind_local = tuple(np.concatenate([b[d] for b in blocks]) for d in range(self.nd))
        n = len(ind_local[0])
        local_label = np.zeros(self.geo.shape, dtype=int)
        local_label[ind_local] = np.arange(1, n + 1)

        pts = np.stack(ind_local, axis=1)
        neighbors = np.zeros((n, len(self.basis)), dtype=int)
        dims = np.array(self.geo.shape)
        for q, c in enumerate(self.basis):
            tgt = pts + c
            inside = np.all((tgt >= 0) & (tgt < dims), axis=1)
            idx = tuple(tgt[inside].T)
            val = local_label[idx]
            rim = (val == 0) & (self.geo[idx] == rank)
            val = np.where(rim, self.label[idx], val)
            neighbors[inside, q] = val
        nodetype = (self.geo[ind_local] > 0).astype(int)
        processors = []
        for r in order:
            sel = self.geo[ind_local] == r
            cells = tuple(a[sel] for a in ind_local)
            processors.append((int(r) - 1, np.stack([local_label[cells], self.label[cells]], axis=1)))
        return RankTopology(pts - 1, neighbors, nodetype, processors, ind_local, rank)

    def padded_attribute(self, val):
        """Pad an attribute with zeros and apply the periodic rim copies (vtklb.py:327-339)."""
        val = np.asarray(val)
        out = np.zeros(tuple(n + 2 for n in val.shape), dtype=val.dtype)
        out[tuple([slice(1, -1)] * self.nd)] = val
        return _periodic_copy(out, self.periodic)

    # ------------------------------------------------------------------ text
    def vtklb_text(self, rank, attributes=()):
        """The file the reference writer produces for ``rank`` (1-based), with
        ``attributes`` = [(name, array), ...] appended in the given order."""
        t = self.rank_topology(rank)
        out = []
        out.append("# BADChIMP vtklb Version na\n")
        out.append("Geometry file for process {}\n".format(rank - 1))
        out.append("ASCII\n")
        out.append("DATASET UNSTRUCTURED_LB_GRID\n")
        out.append("NUM_DIMENSIONS {}\n".format(self.nd))
        out.append("GLOBAL_DIMENSIONS" + "".join(" {}".format(s) for s in self.geo.shape) + "\n")
        out.append("USE_ZERO_GHOST_NODE\n")
        out.append("POINTS {} int\n".format(t.n_points))
        out.extend(" ".join(str(int(v)) for v in row) + "\n" for row in t.points)
        out.append("LATTICE {} int\n".format(len(self.basis)))
        out.extend(" ".join(str(int(v)) for v in row) + "\n" for row in self.basis)
        out.append("NEIGHBORS int\n")
        out.extend(" ".join(str(int(v)) for v in row) + "\n" for row in t.neighbors)
        out.append("PARALLEL_COMPUTING {}\n".format(rank - 1))
        for r0, pairs in t.processors:
            out.append("PROCESSOR {} {}\n".format(len(pairs), r0))
            out.extend("{} {}\n".format(int(a), int(b)) for a, b in pairs)
        out.append("POINT_DATA {}\n".format(t.n_points))
        out.append("SCALARS nodetype int\n")
        out.extend("{}\n".format(int(v)) for v in t.nodetype)
        for name, val in attributes:
            pv = self.padded_attribute(val)[t.ind_local]
            if np.issubdtype(pv.dtype, np.integer):
                out.append("SCALARS {} int\n".format(name))
                out.extend("{}\n".format(int(v)) for v in pv)
            else:
                out.append("SCALARS {} float\n".format(name))
                out.extend("{:.18e}\n".format(float(v)) for v in pv)
        return "".join(out)

    def write(self, path_prefix, attributes=()):
        """Write one file per rank: ``path_prefix + str(rank0) + '.vtklb'``."""
        for r in range(1, self.n_ranks + 1):
            with open("{}{}.vtklb".format(path_prefix, r - 1), "w") as fh:
                fh.write(self.vtklb_text(r, attributes))
