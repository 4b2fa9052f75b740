"""Geometry package: projection/volume geometry and the volume-frame
normalisation that turns world geometry into kernel (voxel-unit) geometry.

Mirrors ``kernelkit/geom/__init__.py`` of the reference:
``resolve_volume_geometry`` (:8-254), ``suggest_volume_extent`` (:257-302)
and ``normalize_`` (:305-330).
"""
from __future__ import annotations

import copy
import warnings

import numpy as np

from kernelkit.geom.proj import (ANGLES_CONVENTION, Beam, Detector,  # noqa: F401
                                 ProjectionGeometry,
                                 ProjectionGeometrySequence, angles2mat,
                                 process_beam_type, rotate, rotate_, scale,
                                 scale_, shift, shift_)
from kernelkit.geom.vol import VolumeGeometry


def _as3(arg):
    if arg is None or np.isscalar(arg):
        return [arg] * 3
    out = list(arg)
    if len(out) != 3:
        raise ValueError("Expected a scalar or a sequence of length 3.")
    return out


def resolve_volume_geometry(shape=None, extent_min=None, extent_max=None,
                            voxel_size=None, projection_geometry=None,
                            verbose=False):
    """Infer a full ``VolumeGeometry`` from partial information.

    Per dimension ``n * voxel_size = extent_max - extent_min``; any two of
    (shape, extents, voxel size) determine the third.  Unknown voxel sizes
    are copied from the smallest known one (isotropy), an unknown y shape is
    copied from x and vice versa, and a projection geometry, if given, fills
    in extents through :func:`suggest_volume_extent`.  Behaviour follows the
    reference (``geom/__init__.py:8-254``), including its error cases.
    """
    shape = _as3(shape)
    ext_min = _as3(extent_min)
    ext_max = _as3(extent_max)
    vox = _as3(voxel_size)
    done = [False] * 3

    def solve(n, lo, hi, sz, dim):
        if n is not None:
            if lo is not None and hi is not None:
                span = hi - lo
                if sz is not None:
                    if not np.allclose(n * sz, span):
                        raise ValueError(
                            f"The {dim}-th dimension contains conflicting "
                            f"parameter values: nr_voxels={n}, voxel_size="
                            f"{sz}, extent_min={lo}, extent_max={hi}. It must"
                            " hold that nr_voxels * voxel_size = extent_max -"
                            " extent_min.")
                else:
                    sz = span / n
                return n, lo, hi, sz
            if lo is None and hi is None:
                if sz is None:
                    return None
                return n, -n * sz / 2, n * sz / 2, sz
            raise ValueError(f"Volume extent in dim {dim} must be given with "
                             "both max and min value, or inferred "
                             "automatically.")
        if lo is not None and hi is not None:
            if sz is None:
                return None
            n_f = (hi - lo) / sz
            if not np.allclose(n_f, np.round(n_f)):
                raise ValueError(
                    f"Number of voxels in dim {dim} was inferred to be {n_f}, "
                    "which is not a rounded number. To have isotropic voxels,"
                    " please give in larger volume size, or leave the volume "
                    f"extent out in dim {dim} to have it inferred "
                    "automatically.")
            return int(np.round(n_f)), lo, hi, sz
        if lo is None and hi is None:
            return None
        raise ValueError(f"Volume extent in dim {dim} must be given with both"
                         " max and min value, or inferred automatically.")

    def resolve(dims, lo_vec, hi_vec):
        for d in dims:
            if done[d]:
                continue
            res = solve(shape[d], lo_vec[d], hi_vec[d], vox[d], d)
            if res is not None:
                shape[d], ext_min[d], ext_max[d], vox[d] = res
                done[d] = True
                if verbose:
                    print(f"Solved dim {d}: shape={res[0]}, extent_min="
                          f"{res[1]}, extent_max={res[2]}, voxel size={res[3]}")

    resolve(range(3), ext_min, ext_max)

    if not any(done):
        if projection_geometry is None:
            raise ValueError(
                "Not enough information is provided to infer a voxel size, "
                "number of voxels and volume extent of at least a single "
                "dimension. Consider passing a geometry for automatic "
                "inference.")
        s_min, s_max = suggest_volume_extent(projection_geometry[0],
                                             (0.0, 0.0, 0.0))
        for d in range(3):
            trial_lo, trial_hi = list(ext_min), list(ext_max)
            trial_lo[d], trial_hi[d] = s_min[d], s_max[d]
            resolve([d], trial_lo, trial_hi)
            if done[d]:
                break

    known = [v for v in vox if v is not None]
    if known:
        smallest = min(known)
        vox = [smallest if v is None else v for v in vox]
    resolve(range(3), ext_min, ext_max)

    if done[0] and not done[1]:
        shape[1] = shape[0]
    if done[1] and not done[0]:
        shape[0] = shape[1]
    resolve(range(2), ext_min, ext_max)

    if not all(done) and projection_geometry is not None:
        s_min, s_max = suggest_volume_extent(projection_geometry[0],
                                             (0.0, 0.0, 0.0))
        for d in range(3):
            if done[d]:
                continue
            trial_lo, trial_hi = list(ext_min), list(ext_max)
            need = (s_max[d] - s_min[d]) / vox[d]
            need = np.round(need) if np.allclose(np.round(need), need) \
                else np.ceil(need)
            trial_lo[d] = -need / 2 * vox[d]
            trial_hi[d] = need / 2 * vox[d]
            resolve([d], trial_lo, trial_hi)

    if verbose:
        print("The final deduced volume parameters are: "
              f"\n - shape: \t\t {shape}\n - extent min: \t {ext_min}"
              f"\n - extent max: \t {ext_max}\n - voxel size: \t {vox}")
    if not all(done):
        missing = [d for d in range(3) if not done[d]]
        raise RuntimeError(
            "Could not resolve the volume geometry. The following dimensions "
            f"cannot be inferred: {missing}. Provide two of the three "
            "unknowns and/or one of the projection geometries to deduce the "
            "dimensions.")
    return VolumeGeometry(shape=shape, voxel_size=vox, extent_min=ext_min,
                          extent_max=ext_max)


def suggest_volume_extent(geometry: ProjectionGeometry,
                          object_position=(0.0, 0.0, 0.0)):
    """Volume box that the detector just sees at the object position.

    The detector size is demagnified by SOD/SDD (``geom/__init__.py:257-302``):
    returns ``([-w/2, -w/2, -h/2], [w/2, w/2, h/2])``.
    """
    g = copy.deepcopy(geometry)
    shift_(g, -np.asarray(object_position, dtype=np.float64))
    src = g.source_position
    ray = g.detector_position - g.source_position
    if np.linalg.matrix_rank(np.stack([src, ray])) != 1:
        warnings.warn("Volume extents may not be suggested correctly when "
                      " the geometry does not project through the origin.")
    sod = np.linalg.norm(src)
    sdd = np.linalg.norm(ray)
    width = sod * g.detector.width / sdd
    height = sod * g.detector.height / sdd
    return (np.array([-width / 2, -width / 2, -height / 2]),
            np.array([width / 2, width / 2, height / 2]))


def normalize_(geometries, volume_geometry: VolumeGeometry) -> None:
    """Express geometries in the volume's voxel frame, in place.

    Shift by minus the volume centre, divide lengths by the voxel size, then
    apply the inverse volume rotation (``geom/__init__.py:305-330``).  After
    this the volume occupies ``[-n/2, n/2]`` in every axis with unit voxels.
    All arithmetic is float64.
    """
    centre = -(np.asarray(volume_geometry.extent_min, dtype=np.float64)
               + np.asarray(volume_geometry.extent_max, dtype=np.float64)) / 2
    shift_(geometries, centre)
    scale_(geometries, np.asarray(volume_geometry.voxel_size,
                                  dtype=np.float64))
    rotate_(geometries, *[float(r) for r in volume_geometry.rotation])
