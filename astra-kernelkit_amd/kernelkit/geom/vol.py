"""Volume geometry (mirrors ``kernelkit/geom/vol.py`` of the reference).

A volume is an axis-aligned box ``[extent_min, extent_max]`` in world (x, y, z)
order, split into ``shape`` voxels, optionally rotated by (roll, pitch, yaw).
Fields are stored as float32/int32 tuples like the reference (``vol.py:58-62``).
"""
from __future__ import annotations

import numpy as np


def _voxel_size(shape, extent_min, extent_max) -> tuple:
    """Per-axis voxel length from the extents (``vol.py:4-22``)."""
    return tuple((float(extent_max[i]) - float(extent_min[i]))
                 / float(shape[i]) for i in range(3))


class VolumeGeometry:
    """Geometry of one reconstruction volume (``vol.py:25-147``)."""

    def __init__(self, shape, voxel_size, extent_min, extent_max,
                 rotation=(0.0, 0.0, 0.0)):
        self.shape = tuple(int(s) for s in np.array(shape, dtype=np.int32))
        self.voxel_size = tuple(np.array(voxel_size, dtype=np.float32))
        self.extent_min = tuple(np.array(extent_min, dtype=np.float32))
        self.extent_max = tuple(np.array(extent_max, dtype=np.float32))
        self.rotation = tuple(np.array(rotation, dtype=np.float32))

    def has_isotropic_voxels(self, atol: float = 1e-8) -> bool:
        return bool(np.allclose(self.voxel_size, self.voxel_size[0],
                                atol=atol))

    def voxel_volume(self) -> float:
        """Volume of one voxel, from extents and shape (``vol.py:79-89``)."""
        vs = _voxel_size(self.shape, self.extent_min, self.extent_max)
        return float(vs[0] * vs[1] * vs[2])

    def check_consistency(self) -> bool:
        """Validate dimensions, ordering and ``n * voxel_size == extent``
        (``vol.py:91-147``)."""
        if len(self.shape) not in (2, 3):
            raise ValueError(
                f"Shape must have 2 or 3 dimensions, got {len(self.shape)}.")
        for name in ("voxel_size", "extent_min", "extent_max"):
            if len(getattr(self, name)) != len(self.shape):
                raise ValueError(f"{name} must match the shape dimensions, "
                                 f"got {len(getattr(self, name))} dimensions.")
        if len(self.rotation) != 3:
            raise ValueError("Rotation must have 3 dimensions, "
                             f"got {len(self.rotation)} dimensions.")
        for lo, hi in zip(self.extent_min, self.extent_max):
            if lo >= hi:
                raise ValueError(f"Extent min {lo} must be less than extent "
                                 f"max {hi}.")
        for n, vs in zip(self.shape, self.voxel_size):
            if n <= 0:
                raise ValueError(f"Shape dimension {n} must be positive.")
            if vs <= 0:
                raise ValueError(f"Voxel size {vs} must be positive.")
        for n, vs, lo, hi in zip(self.shape, self.voxel_size,
                                 self.extent_min, self.extent_max):
            if not np.isclose(n * vs, hi - lo):
                raise ValueError(f"Expected extent {n * vs} does not match "
                                 f"actual extent {hi - lo} for dimension.")
        return True

    def __repr__(self):
        return (f"VolumeGeometry(shape={self.shape}, "
                f"voxel_size={self.voxel_size}, extent_min={self.extent_min},"
                f" extent_max={self.extent_max}, rotation={self.rotation})")
