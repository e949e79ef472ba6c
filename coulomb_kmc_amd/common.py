"""
Shared definitions of the drop-in interface (names and values as in
coulomb_kmc/common.py:6-81 so user code that imports them keeps working).
"""

from enum import Enum
import itertools

import numpy as np


class BCType(Enum):
    """Boundary condition (coulomb_kmc/common.py:36-49)."""

    PBC = "pbc"
    FREE_SPACE = "free_space"
    NEAREST = "27"
    FF_ONLY = "ff-only"


# 27-cell stencil, x fastest (coulomb_kmc/common.py:51-81)
cell_offsets = tuple((x, y, z) for z, y, x in itertools.product((-1, 0, 1), repeat=3))


def spherical(xyz):
    """(radius, polar angle, azimuth) of xyz (tuple or (n, 3) array); coulomb_kmc/common.py:6-33."""
    if isinstance(xyz, tuple):
        x, y, z = xyz
        xy = x * x + y * y
        return np.array([np.sqrt(xy + z * z), np.arctan2(np.sqrt(xy), z), np.arctan2(y, x)])
    xyz = np.asarray(xyz, dtype=np.float64)
    xy = xyz[:, 0] ** 2 + xyz[:, 1] ** 2
    out = np.zeros(xyz.shape)
    out[:, 0] = np.sqrt(xy + xyz[:, 2] ** 2)
    out[:, 1] = np.arctan2(np.sqrt(xy), xyz[:, 2])
    out[:, 2] = np.arctan2(xyz[:, 1], xyz[:, 0])
    return out


class Domain:
    """Minimal cubic domain: ``extent`` (3,) centred on the origin.

    Stands in for ppmd's BaseDomainHalo, of which the reference uses only
    ``extent`` on this path (cubic, asserted at coulomb_kmc/kmc_inject_extract.py:67-69).
    """

    def __init__(self, extent):
        e = np.atleast_1d(np.asarray(extent, dtype=np.float64))
        if e.size == 1:
            e = np.repeat(e, 3)
        self.extent = e


def leaf_of(pos, extent, ncps):
    """Leaf index per axis: trunc((p + E/2) / w) clamped to [0, ncps-1]
    (coulomb_kmc/kmc_fmm.py:229-242; the device kernels use the same expression)."""
    pos = np.asarray(pos, dtype=np.float64)
    w = extent / ncps
    c = ((0.5 * extent + pos) / w).astype(np.int64)
    return np.clip(c, 0, ncps - 1)
