"""Synthetic cone-beam workloads: seeded ellipsoid phantoms and scan
geometries of the benchmark configurations.

Not part of the reference's API (its examples build geometries by hand,
``example/algorithms/fdk.py:9-24``); shared by the tests and the benchmark
so both use identical inputs.  Everything is generated on the host with
numpy from an explicit seed.

Configurations (voxel size 1, volume centred at the origin, circular orbit
``theta_k = 2 pi k / N`` about z, endpoint excluded):

* ``small``: 128^3, 180 angles, SOD 512, ODD 512, detector 160^2 @ 1.5;
* ``full``:  512^3, 720 angles, SOD 1000, ODD 500, detector 768^2 @ 1.0;
* ``diff``:  256^3, 360 angles, SOD 1000, ODD 500, detector 384^2 @ 1.0;
* ``large``: 1024^3, 1440 angles, SOD 2000, ODD 1000, detector 1536^2 @ 1.0,
  with seeded +-1 degree wobble and +-3 pixel detector drift.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from kernelkit.geom import resolve_volume_geometry
from kernelkit.geom.proj import Detector, ProjectionGeometry, angles2mat


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniformly random rotation matrix (unit quaternion from 4 normals)."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@dataclass
class Ellipsoid:
    centre: np.ndarray      # voxel-index coordinates
    semi_axes: np.ndarray   # voxels
    rotation: np.ndarray    # 3x3, columns are the ellipsoid axes
    density: float


def random_ellipsoids(shape, n: int = 10, seed: int = 0,
                      axes_range=(5.0, 40.0), ref_size: int = 128) -> list:
    """``n`` ellipsoids: centres uniform in the inner 60 % of each axis,
    semi-axes uniform in ``axes_range * (size / ref_size)``, uniformly
    random orientation, density uniform in [0.1, 1.0]."""
    rng = np.random.default_rng(seed)
    shape = np.asarray(shape, dtype=np.float64)
    scale = shape.min() / ref_size
    out = []
    for _ in range(n):
        centre = rng.uniform(0.2 * shape, 0.8 * shape)
        semi = rng.uniform(axes_range[0], axes_range[1], size=3) * scale
        rot = _random_rotation(rng)
        dens = float(rng.uniform(0.1, 1.0))
        out.append(Ellipsoid(centre, semi, rot, dens))
    return out


def ellipsoid_phantom(shape, n: int = 10, seed: int = 0,
                      ellipsoids: list | None = None) -> np.ndarray:
    """float32 volume ``vol[x, y, z]`` with additive ellipsoid densities
    evaluated at voxel centres (index ``i`` has centre ``i``)."""
    shape = tuple(int(s) for s in shape)
    vol = np.zeros(shape, dtype=np.float32)
    ells = ellipsoids if ellipsoids is not None else random_ellipsoids(
        shape, n, seed)
    for e in ells:
        r = float(e.semi_axes.max())
        lo = np.maximum(np.floor(e.centre - r).astype(int), 0)
        hi = np.minimum(np.ceil(e.centre + r).astype(int) + 1, shape)
        if np.any(hi <= lo):
            continue
        ys = np.arange(lo[1], hi[1], dtype=np.float64) - e.centre[1]
        zs = np.arange(lo[2], hi[2], dtype=np.float64) - e.centre[2]
        Y, Z = np.meshgrid(ys, zs, indexing="ij")
        inv = e.rotation / e.semi_axes[None, :]  # project onto scaled axes
        for ix in range(lo[0], hi[0]):
            X = float(ix) - e.centre[0]
            q = np.zeros_like(Y)
            for k in range(3):
                t = X * inv[0, k] + Y * inv[1, k] + Z * inv[2, k]
                q += t * t
            inside = q <= 1.0
            if inside.any():
                vol[ix, lo[1]:hi[1], lo[2]:hi[2]] += np.where(
                    inside, np.float32(e.density), np.float32(0))
    return vol


def ellipsoid_phantom_device(shape, n: int = 10, seed: int = 0,
                             device=None):
    """:func:`ellipsoid_phantom` evaluated on a torch device (float64
    quadratic forms over each ellipsoid's bounding box), for volumes too
    large to build on the host.  Same ellipsoids; voxels within rounding of
    an ellipsoid surface may differ from the host version."""
    import torch
    shape = tuple(int(s) for s in shape)
    vol = torch.zeros(shape, dtype=torch.float32, device=device)
    for e in random_ellipsoids(shape, n, seed):
        r = float(e.semi_axes.max())
        lo = np.maximum(np.floor(e.centre - r).astype(int), 0)
        hi = np.minimum(np.ceil(e.centre + r).astype(int) + 1, shape)
        if np.any(hi <= lo):
            continue
        inv = torch.tensor(e.rotation / e.semi_axes[None, :],
                           dtype=torch.float64, device=device)
        ax = [torch.arange(int(lo[d]), int(hi[d]), dtype=torch.float64,
                           device=device) - float(e.centre[d])
              for d in range(3)]
        q = torch.zeros((len(ax[0]), len(ax[1]), len(ax[2])),
                        dtype=torch.float64, device=device)
        for k in range(3):
            t = (ax[0][:, None, None] * inv[0, k]
                 + ax[1][None, :, None] * inv[1, k]
                 + ax[2][None, None, :] * inv[2, k])
            q += t * t
        vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += torch.where(
            q <= 1.0, float(e.density), 0.0).to(torch.float32)
        del q
    return vol


def circular_geometry(n_angles: int, sod: float, odd: float, rows: int,
                      cols: int, pixel: float, wobble_deg: float = 0.0,
                      drift_px: float = 0.0, tilt_deg: float = 0.0,
                      seed: int = 0) -> list:
    """Circular cone orbit about z: ``s = Rz(t)(-sod, 0, 0)``,
    ``d = Rz(t)(odd, 0, 0)``, ``u = Rz(t)(0, 1, 0)``, ``v = (0, 0, 1)``.

    Optional seeded perturbations per angle: orbit wobble (roll/pitch of
    the whole source-detector pair, degrees), detector tilt (rotation of
    ``u``/``v`` about the detector normal, degrees) and detector drift
    (shift along ``u``/``v``, pixels).
    """
    rng = np.random.default_rng(seed)
    det = Detector(rows=rows, cols=cols, pixel_height=pixel,
                   pixel_width=pixel)
    out = []
    for k in range(n_angles):
        theta = 2 * np.pi * k / n_angles
        s = np.array([-sod, 0.0, 0.0])
        d = np.array([odd, 0.0, 0.0])
        u = np.array([0.0, 1.0, 0.0])
        v = np.array([0.0, 0.0, 1.0])
        if tilt_deg:
            a = np.deg2rad(rng.uniform(-tilt_deg, tilt_deg))
            R = angles2mat(a, 0.0, 0.0)  # about the detector normal (x)
            u, v = R @ u, R @ v
        if drift_px:
            du, dv = rng.uniform(-drift_px, drift_px, size=2) * pixel
            d = d + du * u + dv * v
        R = angles2mat(0.0, 0.0, theta)
        if wobble_deg:
            r, p = np.deg2rad(rng.uniform(-wobble_deg, wobble_deg, size=2))
            R = angles2mat(r, p, 0.0) @ R
        out.append(ProjectionGeometry(R @ s, R @ d, det, u=R @ u, v=R @ v))
    return out


CONFIGS = {
    "small": dict(shape=(128, 128, 128), n_angles=180, sod=512.0, odd=512.0,
                  rows=160, cols=160, pixel=1.5),
    "full": dict(shape=(512, 512, 512), n_angles=720, sod=1000.0, odd=500.0,
                 rows=768, cols=768, pixel=1.0),
    "diff": dict(shape=(256, 256, 256), n_angles=360, sod=1000.0, odd=500.0,
                 rows=384, cols=384, pixel=1.0),
    "large": dict(shape=(1024, 1024, 1024), n_angles=1440, sod=2000.0,
                  odd=1000.0, rows=1536, cols=1536, pixel=1.0,
                  wobble_deg=1.0, drift_px=3.0),
}


def config_geometry(name: str, n_angles: int | None = None, seed: int = 0):
    """``(projection geometries, volume geometry)`` of a named config."""
    c = dict(CONFIGS[name])
    shape = c.pop("shape")
    n = c.pop("n_angles") if n_angles is None else n_angles
    c.pop("n_angles", None)
    pgs = circular_geometry(n, seed=seed, **c)
    vg = resolve_volume_geometry(shape=shape, voxel_size=1.0)
    return pgs, vg


_BATCH_SIZES = (64, 96, 128, 160, 192, 224, 256)


def batch_problems(n: int = 64, seed: int = 0, scale: float = 1.0) -> list:
    """``n`` independent cone problems of uneven size, as
    ``(projection geometries, volume geometry)`` pairs.

    Volumes are cubic or anisotropic (``a x a x b``) with sides from 64 to
    256 voxels (times ``scale``); 90-360 angles; source-origin distance
    uniform in [300, 800] and magnification in [1.2, 2.5]; detectors of
    96-384 pixels per side sized to the magnified volume; per-angle detector
    tilt of up to 2 degrees and shifts of up to 5 pixels (cone_vec-style
    geometry).
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = int(rng.choice(_BATCH_SIZES))
        b = a if rng.uniform() < 0.6 else int(rng.choice(_BATCH_SIZES))
        shape = tuple(max(8, int(round(s * scale))) for s in (a, a, b))
        n_angles = int(rng.integers(90, 361))
        sod = float(rng.uniform(300.0, 800.0))
        mag = float(rng.uniform(1.2, 2.5))
        odd = sod * (mag - 1.0)
        # detector: the magnified volume (its xy diagonal across, z up)
        # on 96-384 pixels
        diag = float(np.hypot(shape[0], shape[1]))
        cols = int(np.clip(8 * round(1.5 * diag / 8), 96 * scale, 384 * scale))
        rows = int(np.clip(8 * round(1.5 * shape[2] / 8), 96 * scale,
                           384 * scale))
        cols, rows = max(cols, 8), max(rows, 8)
        # the near side of the volume is magnified more: leave 15 % margin
        pixel = 1.15 * mag * max(diag / cols, shape[2] / rows)
        pgs = circular_geometry(n_angles, sod, odd, rows, cols, pixel,
                                tilt_deg=2.0, drift_px=5.0,
                                seed=int(rng.integers(1 << 31)))
        vg = resolve_volume_geometry(shape=shape, voxel_size=1.0)
        out.append((pgs, vg))
    return out
