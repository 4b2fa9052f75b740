"""The differentiation workload on the GPU: 256^3 volume, 360 angles over
360 degrees, 384^2 detector, measurements
``y = -log(Poisson(I0 exp(-A x_ph)) / I0)`` with ``I0 = 1e4``.

* the dot-product adjoint test ``<A x, y> = <x, A^T y>`` on 10 random
  volume/sinogram pairs at full size;
* the gradient of ``||A x - y||^2`` through the exact transpose against
  central finite differences of the FP64 oracle's loss at 1000 seeded
  voxels, for the angles 0, 60, ..., 300 degrees (every angle contributes
  independently; the oracle over all 360 would take about an hour).  The
  loss is quadratic, so the central difference at voxel i is exactly
  ``2 <A x - y, A e_i>`` for any step.  ``A e_i`` vanishes outside the
  rays through the voxel's neighbourhood, so it is evaluated on a pixel box
  around the voxel's projection per angle, on a 5^3 crop of the volume
  (the same ray sums as the full volume, checked here for a few voxels);
  ``A x - y`` comes from the oracle on the full detector.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

kk = pytest.importorskip("kernelkit")
from kernelkit import workloads  # noqa: E402
from kernelkit.geom.proj import Detector, ProjectionGeometry  # noqa: E402

_W = 4              # box half-width (pixels) around a voxel's projection
_B = 2 * _W + 2     # box size


def _box(g, X):
    """The _B x _B pixel box of detector g centred on the projection of
    the world point X, as a detector geometry of its own (same pixel grid),
    and its first row / column on the full detector."""
    s = np.asarray(g.source_position, float)
    d = np.asarray(g.detector_position, float)
    u, v = np.asarray(g.u, float), np.asarray(g.v, float)
    rows, cols = int(g.detector.rows), int(g.detector.cols)
    pw, ph = float(g.detector.pixel_width), float(g.detector.pixel_height)
    n = np.cross(u, v)
    t = np.dot(d - s, n) / np.dot(X - s, n)
    P = s + t * (X - s) - d
    col = np.dot(P, u) / pw + cols / 2 - 0.5
    row = np.dot(P, v) / ph + rows / 2 - 0.5
    c0 = int(np.clip(np.floor(col) - _W, 0, cols - _B))
    r0 = int(np.clip(np.floor(row) - _W, 0, rows - _B))
    shift = (r0 + _B / 2 - rows / 2) * ph * v + (c0 + _B / 2 - cols / 2) * pw * u
    return (ProjectionGeometry(g.source_position, d + shift,
                               Detector(_B, _B, ph, pw), u=g.u, v=g.v),
            r0, c0)


def _crop(vg, idx, r=2):
    vs = np.asarray(vg.voxel_size, float)
    emin = np.asarray(vg.extent_min, float) + (np.asarray(idx) - r) * vs
    return kk.VolumeGeometry((2 * r + 1,) * 3, vs, emin,
                             emin + (2 * r + 1) * vs)


@pytest.fixture(scope="module")
def diff():
    return workloads.config_geometry("diff")


def test_diff_dot_product_adjoint(diff, device):
    import torch
    pgs, vg = diff
    A = kk.XrayTransform(pgs, vg)
    g = torch.Generator(device="cuda").manual_seed(11)
    worst = 0.0
    for _ in range(10):
        x = torch.rand(vg.shape, generator=g, device="cuda") * 2 - 1
        y = torch.rand(A.range_shape, generator=g, device="cuda") * 2 - 1
        Ax = A(x)
        ATy = A.matched_T(y)
        lhs = float((Ax.double() * y.double()).sum())
        rhs = float((x.double() * ATy.double()).sum())
        scale = float(Ax.double().norm() * y.double().norm())
        worst = max(worst, abs(lhs - rhs) / scale)
    assert worst <= 1e-5, worst


def test_diff_gradient_vs_oracle_central_differences(diff, device):
    import torch
    from kernelkit.torch_support import XrayForwardProjection
    pgs, vg = diff
    sub = [pgs[k] for k in range(0, 360, 60)]
    rng = np.random.default_rng(2024)
    # Poisson measurements of the phantom, line integrals scaled to <= 2
    A = kk.XrayTransform(sub, vg)
    x_ph = torch.from_numpy(workloads.ellipsoid_phantom(
        vg.shape, n=10, seed=0)).cuda()
    p = A(x_ph)
    x_ph *= 2.0 / float(p.max())
    p = A(x_ph).cpu().numpy().astype(np.float64)
    counts = rng.poisson(1e4 * np.exp(-p))
    y = -np.log(np.maximum(counts, 1) / 1e4)
    x0 = 0.8 * x_ph
    # GPU gradient through the exact transpose
    fp = XrayForwardProjection(sub, vg, matched=True)
    xg = x0[None, None].clone().requires_grad_(True)
    loss = ((fp(xg) - torch.from_numpy(y.astype(np.float32)).cuda()
             [None, None]) ** 2).sum()
    loss.backward()
    grad = xg.grad[0, 0].cpu().numpy().astype(np.float64)
    # oracle residual on the full detector
    x64 = x0.cpu().numpy().astype(np.float64)
    r = oracle.cone_fp(x64, sub, vg) - y
    n = np.asarray(vg.shape)
    voxels = rng.integers(2, n - 2, size=(1000, 3))
    vs = np.asarray(vg.voxel_size, float)
    emin = np.asarray(vg.extent_min, float)
    e5 = np.zeros((5, 5, 5))
    e5[2, 2, 2] = 1.0
    fd = np.empty(len(voxels))
    for m, idx in enumerate(voxels):
        X = emin + (idx + 0.5) * vs
        boxes = [_box(g, X) for g in sub]
        Ae = oracle.cone_fp(e5, [b[0] for b in boxes], _crop(vg, idx))
        fd[m] = 2.0 * sum(float(np.vdot(r[k, r0:r0 + _B, c0:c0 + _B],
                                        Ae[k]))
                          for k, (_, r0, c0) in enumerate(boxes))
        if m < 3:   # the crop and the box lose nothing
            e = np.zeros(vg.shape)
            e[tuple(idx)] = 1.0
            Af = oracle.cone_fp(e, [b[0] for b in boxes], vg)
            np.testing.assert_allclose(Ae, Af, rtol=0,
                                       atol=1e-12 * np.abs(Af).max())
            assert np.abs(Af[:, [0, -1], :]).max() == 0.0
            assert np.abs(Af[:, :, [0, -1]]).max() == 0.0
    got = grad[tuple(voxels.T)]
    tol = 1e-4 * np.abs(fd) + 1e-4 * np.abs(fd).max()
    worst = float(np.max(np.abs(got - fd) / (np.abs(fd) + np.abs(fd).max())))
    assert np.all(np.abs(got - fd) <= tol), worst
    assert np.count_nonzero(fd) > 900
