"""The timing harness of the compiled reference (oracle/ref_run.py), used as
the CPU baseline of bench.py: the rank slabs it cuts, and a two-rank run of
the reference std_case on node lists the library writes (skipped where the
reference was not built)."""
import numpy as np
import pytest

import lbamd
from oracle import ref_run


def test_rank_slabs_cover_the_fluid_in_x_slabs():
    geo, _ = lbamd.sphere_pack((23, 12, 10), 3.0, 0.5, 2, periodic="x", walls=True)
    g = ref_run.rank_slabs(geo, 4)
    assert np.array_equal(g > 0, geo > 0)          # the same fluid cells
    assert sorted(np.unique(g[g > 0])) == [1, 2, 3, 4]
    for r in range(1, 5):                          # each rank is one x slab
        xs = np.nonzero((g == r).any(axis=(1, 2)))[0]
        assert xs.max() - xs.min() + 1 == xs.size
    bounds = [np.nonzero((g == r).any(axis=(1, 2)))[0] for r in range(1, 5)]
    assert all(bounds[r][-1] < bounds[r + 1][0] for r in range(3))


@pytest.mark.skipif(not ref_run.available(), reason="compiled reference or MPICH runtime absent")
def test_reference_runs_on_two_ranks():
    geo, _ = lbamd.sphere_pack((24, 16, 16), 3.0, 0.45, 3, periodic="x", walls=True)
    dt, n_bulk, wall = ref_run.time_std_case(geo, "x", 1.0, [1e-6, 0.0, 0.0], 2, 2, 12, timeout=120)
    one = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    assert n_bulk == one.n_bulk                    # the slabs hold every bulk node once
    assert wall[2] > 0 and wall[12] > 0
