"""The largest workload (1024^3 volume, 1536^2 detector, seeded wobble and
drift) on the GPU, against the oracle on bounded samples: the forward
projection on detector row bands of eight angles (including the three
around the 45-degree ray-axis switch), the backprojection of those angles
on z-slabs of the volume (both exact restrictions: each ray and each voxel
is computed independently)."""
import numpy as np
import pytest

import oracle
from conftest import assert_parity

pytestmark = pytest.mark.gpu

kk = pytest.importorskip("kernelkit")
from kernelkit import workloads  # noqa: E402
from kernelkit.geom.proj import Detector, ProjectionGeometry  # noqa: E402


def _band(g, r0, band):
    """The rows [r0, r0 + band) of detector g as a detector of their own."""
    rows = int(g.detector.rows)
    shift = (r0 + band / 2.0 - rows / 2.0) * float(g.detector.pixel_height)
    return ProjectionGeometry(
        g.source_position,
        np.asarray(g.detector_position) + shift * np.asarray(g.v),
        Detector(band, int(g.detector.cols), g.detector.pixel_height,
                 g.detector.pixel_width), u=g.u, v=g.v)


@pytest.fixture(scope="module")
def large():
    pgs, vg = workloads.config_geometry("large")
    # 0.25 degrees apart: 179-181 straddle 45 degrees (the x/y ray-axis
    # switch under the wobble), the others spread over the orbit
    return [pgs[k] for k in (0, 179, 180, 181, 611, 719, 1080, 1439)], vg


def test_large_fp_bands(large, device):
    import torch
    pgs, vg = large
    n = vg.shape[0]
    # smooth synthetic volume built on the device (no 1024^3 phantom on
    # the host), copied back once for the oracle
    t = torch.linspace(-1.0, 1.0, n, device="cuda")
    x = (torch.cos(3.1 * t)[:, None, None] * torch.sin(2.3 * t + 0.4)
         [None, :, None] + 0.5 * torch.cos(4.7 * t)[None, None, :] + 1.0)
    x = x.contiguous()
    A = kk.XrayTransform(pgs, vg)
    y = A(x).cpu().numpy()
    vol = x.cpu().numpy().astype(np.float64)   # converted once
    del x
    for r0 in (764, 40, 1490):
        band = [_band(g, r0, 4) for g in pgs]
        ref = oracle.cone_fp(vol, band, vg,
                             axes=oracle.fp_major_axes(
                                 oracle.normalized_frames(pgs, vg)))
        assert_parity(y[:, r0:r0 + 4], ref, label=f"large fp rows {r0}")


def test_large_bp_slab(large, device):
    import torch
    pgs, vg = large
    rng = np.random.default_rng(0)
    proj = rng.uniform(size=(len(pgs), 1536, 1536)).astype(np.float32)
    A = kk.XrayTransform(pgs, vg)
    v = A.T(torch.from_numpy(proj).cuda())
    n = vg.shape[2]
    for z0 in (511, 0, 1022):
        got = v[:, :, z0:z0 + 2].cpu().numpy()
        lo = np.asarray(vg.extent_min, np.float64).copy()
        hi = np.asarray(vg.extent_max, np.float64).copy()
        vs = float(vg.voxel_size[2])
        lo[2] = vg.extent_min[2] + z0 * vs
        hi[2] = lo[2] + 2 * vs
        sub = kk.VolumeGeometry((vg.shape[0], vg.shape[1], 2),
                                vg.voxel_size, lo, hi)
        ref = oracle.cone_bp(proj.astype(np.float64), pgs, sub)
        assert_parity(got, ref, label=f"large bp z {z0}")
    assert n == 1024


def test_large_fp_noise_bands(large, device):
    """White-noise volume: every sub-voxel coordinate error shows, so this
    checks the float64 sample coordinates at 1024 slices."""
    import torch
    pgs, vg = large
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(vg.shape, device="cuda", generator=g)
    A = kk.XrayTransform(pgs, vg)
    y = A(x).cpu().numpy()
    vol = x.cpu().numpy().astype(np.float64)
    del x
    for r0 in (766, 3):
        band = [_band(g_, r0, 2) for g_ in pgs]
        axes = oracle.fp_major_axes(oracle.normalized_frames(pgs, vg))
        assert_parity(y[:, r0:r0 + 2], oracle.cone_fp(vol, band, vg,
                                                      axes=axes),
                      label=f"large fp noise rows {r0}")
