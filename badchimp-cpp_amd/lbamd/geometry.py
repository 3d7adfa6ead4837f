"""Synthetic geometries used by the cases: channels and sphere packs.

A geometry array follows the reference writer's convention
(PythonScripts/vtklb.py:70-80): ``geo[x, y(, z)]`` is 0 for solid and the
1-based rank for fluid (1 on one GPU).  Sphere packs come from the native
generator (lb_sphere_pack in include/badchimp_amd.h).
"""
import numpy as np

from .ffi import sphere_pack  # noqa: F401  (re-export)


def channel_2d(nx, ny):
    """D2Q9 channel: fluid everywhere, solid rows at y=0 and y=ny-1."""
    geo = np.ones((nx, ny), dtype=np.int32)
    geo[:, 0] = 0
    geo[:, -1] = 0
    return geo


def n_box(geo):
    """Number of lattice cells of the box (the permeability normalisation)."""
    return int(np.asarray(geo).size)
