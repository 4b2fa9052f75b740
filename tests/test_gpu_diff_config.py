"""The differentiation workload on the GPU: 256^3 volume, 360 angles over
360 degrees, 384^2 detector, measurements
``y = -log(Poisson(I0 exp(-A x_ph)) / I0)`` with ``I0 = 1e4``.

* the dot-product adjoint test ``<A x, y> = <x, A^T y>`` on 10 random
  volume/sinogram pairs at full size;
* the gradient of ``||A x - y||^2`` through the exact transpose against
  central finite differences of the FP64 oracle's loss at 1000 seeded
  voxels, for 36 angles (0, 10, ..., 350 degrees; every angle contributes
  independently, and the oracle over all 360 would take hours).  The loss
  is quadratic, so the central difference at voxel i is exactly
  ``2 <A x - y, A e_i>`` for any step.  ``A e_i`` vanishes outside the
  rays through the voxel's neighbourhood, so it is evaluated on a pixel box
  around the voxel's projection per angle, on a 5^3 crop of the volume
  (the same ray sums as the full volume, checked here for a few voxels);
  ``A x - y`` comes from the oracle on the full detector (process pool).
  Each voxel's gradient must be within 1e-4 of its own finite difference,
  relative, plus a floor set by float32 rounding (see the test).
"""
import numpy as np
import pytest

import oracle
import oracle_pool

pytestmark = pytest.mark.gpu

kk = pytest.importorskip("kernelkit")
from kernelkit import workloads  # noqa: E402
from kernelkit.geom.proj import Detector, ProjectionGeometry  # noqa: E402

_W = 5              # box half-width (pixels) around a voxel's projection
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
    sub = [pgs[k] for k in range(0, 360, 10)]
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
    Ax32 = fp(xg)
    loss = ((Ax32 - torch.from_numpy(y.astype(np.float32)).cuda()
             [None, None]) ** 2).sum()
    loss.backward()
    # the GPU's own residual (float32), for the diagnosis of a failure
    r_gpu = (Ax32[0, 0].detach().cpu().numpy().astype(np.float64)
             - y.astype(np.float32).astype(np.float64))
    grad = xg.grad[0, 0].cpu().numpy().astype(np.float64)
    # oracle residual on the full detector
    x64 = x0.cpu().numpy().astype(np.float64)
    Ax = oracle_pool.cone_fp(x64, sub, vg)
    r = Ax - y
    n = np.asarray(vg.shape)
    voxels = rng.integers(2, n - 2, size=(1000, 3))
    vs = np.asarray(vg.voxel_size, float)
    emin = np.asarray(vg.extent_min, float)
    e5 = np.zeros((5, 5, 5))
    e5[2, 2, 2] = 1.0
    fd = np.empty(len(voxels))
    fd_gpu = np.empty(len(voxels))
    floor = np.empty(len(voxels))
    # float32 rounding per pixel of the GPU residual: the forward sum of
    # 256 slice samples (sqrt(256) = 16 roundings of |A x|, the usual
    # random-walk bound), y rounded to float32 and the subtraction
    mag = 16.0 * np.abs(Ax) + np.abs(y) + np.abs(r)
    # the ray axis of each angle comes from the full detector's centre: a
    # box far from it would otherwise pick its own (on angles near 45
    # degrees), which samples differently
    axes = oracle.fp_major_axes(oracle.normalized_frames(sub, vg))
    for m, idx in enumerate(voxels):
        X = emin + (idx + 0.5) * vs
        boxes = [_box(g, X) for g in sub]
        Ae = oracle.cone_fp(e5, [b[0] for b in boxes], _crop(vg, idx),
                            axes=axes)
        fd[m] = 2.0 * sum(float(np.vdot(r[k, r0:r0 + _B, c0:c0 + _B],
                                        Ae[k]))
                          for k, (_, r0, c0) in enumerate(boxes))
        fd_gpu[m] = 2.0 * sum(float(np.vdot(
            r_gpu[k, r0:r0 + _B, c0:c0 + _B], Ae[k]))
            for k, (_, r0, c0) in enumerate(boxes))
        floor[m] = 2.0 * sum(float(np.vdot(mag[k, r0:r0 + _B, c0:c0 + _B],
                                           np.abs(Ae[k])))
                             for k, (_, r0, c0) in enumerate(boxes))
        if m < 3:   # the crop and the box lose nothing
            e = np.zeros(vg.shape)
            e[tuple(idx)] = 1.0
            Af = oracle.cone_fp(e, [b[0] for b in boxes], vg, axes=axes)
            np.testing.assert_allclose(Ae, Af, rtol=0,
                                       atol=1e-12 * np.abs(Af).max())
            # the box's border rows and columns see nothing, except where
            # the box lies on the detector's own border (no pixels beyond)
            rows, cols = int(pgs[0].detector.rows), int(pgs[0].detector.cols)
            for k, (_, r0, c0) in enumerate(boxes):
                edges = [Af[k, 0] if r0 > 0 else 0.0,
                         Af[k, -1] if r0 + _B < rows else 0.0,
                         Af[k, :, 0] if c0 > 0 else 0.0,
                         Af[k, :, -1] if c0 + _B < cols else 0.0]
                assert max(float(np.abs(e).max()) for e in edges) == 0.0, k
    got = grad[tuple(voxels.T)]
    # Per voxel: 1e-4 of its own finite difference, plus what float32
    # rounding of the residual alone can cause, carried through the
    # gradient's weights: three times 2^-24 per unit of ``mag`` above.  It
    # matters only where the terms of a voxel's gradient cancel.
    floor *= 3.0 * 2.0 ** -24
    tol = 1e-4 * np.abs(fd) + floor
    err = np.abs(got - fd)
    worst = float(np.max(err / tol))
    bad = np.argsort(err / tol)[::-1][:5]
    detail = "; ".join(
        f"voxel {tuple(voxels[b])}: fd {fd[b]:.6e} got {got[b]:.6e} "
        f"(gpu residual through the oracle's A e_i: {fd_gpu[b]:.6e}), "
        f"tol {tol[b]:.3e}" for b in bad)
    assert np.all(err <= tol), f"worst err/tol {worst:.3g}: {detail}"
    assert np.count_nonzero(fd) > 900
    # the relative bound is the binding one for almost every voxel
    assert np.mean(floor < 1e-4 * np.abs(fd)) > 0.9, np.mean(
        floor < 1e-4 * np.abs(fd))
