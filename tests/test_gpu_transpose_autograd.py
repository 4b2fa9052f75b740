"""GPU tests of the exact transpose of the forward projection and of the
autograd paths built on it.

* parity of ``kk_cone_fp_transpose`` with the oracle's exact transpose;
* the dot-product adjoint test <A x, y> = <x, A^T y> on random pairs;
* autograd gradients of ||A x - y||^2 against central finite differences
  of the FP64 oracle's loss (exact for a quadratic) at seeded voxels and
  along random directions, within 1e-4 relative;
* the unrolled reconstruction: checkpointing changes nothing, and its
  backward pass is the transpose of its (affine) forward map.
"""
import itertools

import numpy as np
import pytest

import oracle
from conftest import assert_parity

pytestmark = pytest.mark.gpu

kk = pytest.importorskip("kernelkit")
from kernelkit import workloads  # noqa: E402


def _torch():
    import torch
    return torch


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _host(t):
    return t.detach().cpu().numpy()


def _case(n_angles=16, shape=(24, 28, 20), rows=30, cols=34, seed=6, **kw):
    pgs = workloads.circular_geometry(n_angles, sod=120.0, odd=60.0,
                                      rows=rows, cols=cols, pixel=1.2,
                                      seed=seed, **kw)
    vg = kk.resolve_volume_geometry(shape=shape, voxel_size=1.0)
    return pgs, vg


CASES = {
    "circular": dict(),
    "tilted": dict(tilt_deg=2.0, drift_px=5.0),
    "wobble": dict(wobble_deg=1.0, drift_px=3.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_transpose_matches_oracle(case, device):
    pgs, vg = _case(**CASES[case])
    y = np.random.default_rng(1).uniform(size=(len(pgs), 30, 34))
    A = kk.XrayTransform(pgs, vg)
    got = _host(A.matched_T(_dev(y)))
    assert_parity(got, oracle.cone_fp_transpose(y, pgs, vg),
                  label=f"fp_t/{case}")


def test_transpose_large_tiles_and_mixed_axes(device):
    """A volume spanning several tiles in every direction, with x-, y- and
    z-major angles."""
    base = workloads.circular_geometry(10, sod=150.0, odd=80.0, rows=72,
                                       cols=90, pixel=1.1, seed=2,
                                       tilt_deg=1.0)
    pgs = base + [kk.rotate(g, pitch=np.pi / 2) for g in base[:3]]
    vg = kk.resolve_volume_geometry(shape=(70, 45, 38), voxel_size=0.9)
    y = np.random.default_rng(2).uniform(size=(len(pgs), 72, 90))
    A = kk.XrayTransform(pgs, vg)
    assert A.projector.kernel.frame_mask == 0b111
    got = _host(A.matched_T(_dev(y)))
    assert_parity(got, oracle.cone_fp_transpose(y, pgs, vg), label="fp_t/big")


@pytest.mark.parametrize("va,pa", [((2, 0, 1), (0, 1, 2)),
                                   ((0, 1, 2), (1, 0, 2)),
                                   ((1, 2, 0), (2, 1, 0))])
def test_transpose_axes(va, pa, device):
    pgs, vg = _case(n_angles=8)
    y = np.random.default_rng(3).uniform(size=(8, 30, 34))
    ref = oracle.cone_fp_transpose(y, pgs, vg)
    A = kk.XrayTransform(pgs, vg, volume_axes=va, projection_axes=pa)
    got = _host(A.matched_T(_dev(np.transpose(y, pa))))
    assert_parity(np.transpose(got, np.argsort(va)), ref,
                  label=f"fp_t/axes {va} {pa}")


def test_transpose_additive(device):
    pgs, vg = _case(n_angles=6)
    A = kk.XrayTransform(pgs, vg)
    y = _dev(np.random.default_rng(4).uniform(size=(6, 30, 34)))
    v = A.matched_T(y)
    v2 = v.clone()
    A.matched_T(y, out=v2, additive=True)
    np.testing.assert_allclose(_host(v2), 2 * _host(v), rtol=1e-6,
                               atol=1e-6 * float(v.abs().max()))


def test_dot_product_adjoint(device):
    """<A x, y> = <x, A^T y> on 10 random volume/sinogram pairs."""
    pgs, vg = _case(n_angles=20, tilt_deg=1.5)
    A = kk.XrayTransform(pgs, vg)
    rng = np.random.default_rng(5)
    for _ in range(10):
        x = rng.uniform(-1, 1, size=vg.shape)
        y = rng.uniform(-1, 1, size=(20, 30, 34))
        Ax = _host(A(_dev(x))).astype(np.float64)
        ATy = _host(A.matched_T(_dev(y))).astype(np.float64)
        lhs = float(np.vdot(Ax, y))
        rhs = float(np.vdot(x, ATy))
        scale = np.linalg.norm(Ax) * np.linalg.norm(y)
        assert abs(lhs - rhs) <= 1e-5 * scale, (lhs, rhs)


def test_autograd_gradient_vs_oracle_finite_differences(device):
    """grad ||A x - y||^2 (matched backward) against central differences of
    the FP64 oracle loss at seeded voxels and along random directions."""
    torch = _torch()
    from kernelkit.torch_support import XrayForwardProjection
    pgs, vg = _case(n_angles=12, shape=(16, 18, 14), rows=22, cols=26)
    rng = np.random.default_rng(7)
    x0 = rng.uniform(0, 1, size=vg.shape)
    y = rng.uniform(0, 5, size=(12, 22, 26))
    fp = XrayForwardProjection(pgs, vg, matched=True)
    x = _dev(x0)[None, None].requires_grad_(True)
    loss = ((fp(x) - _dev(y)[None, None]) ** 2).sum()
    loss.backward()
    g = _host(x.grad[0, 0]).astype(np.float64)

    def L(v):
        return float(((oracle.cone_fp(v, pgs, vg) - y) ** 2).sum())

    h = 1e-2
    flat = rng.choice(x0.size, size=24, replace=False)
    for f in flat:
        idx = np.unravel_index(f, x0.shape)
        e = np.zeros_like(x0)
        e[idx] = h
        fd = (L(x0 + e) - L(x0 - e)) / (2 * h)
        assert g[idx] == pytest.approx(fd, rel=1e-4,
                                       abs=1e-4 * np.abs(g).max())
    for _ in range(3):
        d = rng.normal(size=x0.shape)
        fd = (L(x0 + h * d) - L(x0 - h * d)) / (2 * h)
        assert float(np.vdot(g, d)) == pytest.approx(fd, rel=1e-4)


def test_autograd_batch_and_unmatched_default(device):
    torch = _torch()
    from kernelkit.torch_support import (AutogradOperator,
                                         MatchedAutogradOperator)
    pgs, vg = _case(n_angles=6, shape=(12, 12, 10), rows=16, cols=18)
    A = kk.XrayTransform(pgs, vg)
    x = torch.rand(2, 3, *vg.shape, device="cuda", requires_grad=True)
    out = MatchedAutogradOperator.apply(x, A)
    assert tuple(out.shape) == (2, 3, 6, 16, 18)
    w = torch.rand_like(out)
    (out * w).sum().backward()
    for b, c in itertools.product(range(2), range(3)):
        ref = A.matched_T(w[b, c].contiguous())
        np.testing.assert_allclose(_host(x.grad[b, c]), _host(ref),
                                   rtol=1e-6, atol=1e-6)
    # the reference's operator keeps the voxel-driven BP in its backward
    x.grad = None
    out = AutogradOperator.apply(x, A)
    (out * w).sum().backward()
    ref = A.T(w[1, 2].contiguous())
    np.testing.assert_allclose(_host(x.grad[1, 2]), _host(ref), rtol=1e-6,
                               atol=1e-6)


@pytest.mark.parametrize("weighting", ["sirt", "none"])
def test_unrolled_checkpoint_and_adjoint(weighting, device):
    torch = _torch()
    from kernelkit.torch_support import unrolled_reconstruction
    pgs, vg = _case(n_angles=10, shape=(14, 14, 12), rows=20, cols=22)
    A = kk.XrayTransform(pgs, vg)
    rng = np.random.default_rng(8)
    step = None if weighting == "sirt" else 2e-3
    y1 = _dev(rng.uniform(size=(10, 20, 22)))
    w = _dev(rng.normal(size=vg.shape))

    def run(y, ckpt):
        return unrolled_reconstruction(y, A, iters=4, step=step,
                                       weighting=weighting,
                                       checkpoint=ckpt)

    ya = y1.clone().requires_grad_(True)
    (run(ya, True) * w).sum().backward()
    yb = y1.clone().requires_grad_(True)
    (run(yb, False) * w).sum().backward()
    np.testing.assert_allclose(_host(ya.grad), _host(yb.grad), rtol=1e-5,
                               atol=1e-6 * float(yb.grad.abs().max()))
    # x_K(y) = M y from x0 = 0 (linear, no clamp): <M y1, w> = <y1, M^T w>
    with torch.no_grad():
        My = run(y1, True)
    lhs = float((My.double() * w.double()).sum())
    rhs = float((y1.double() * ya.grad.double()).sum())
    assert lhs == pytest.approx(rhs, rel=1e-4)


def test_transpose_rejects_a_workspace_angle_of_4_gib(device):
    """The gather addresses one angle's padded workspace with 32-bit byte
    offsets: a 30000 x 40000 detector (fewer than 2^31 pixels, but
    cols * (rows + 32) * 4 bytes >= 2^32) is refused with
    KK_ERR_UNSUPPORTED before anything is launched."""
    import torch
    from kernelkit import _native
    from kernelkit.kernels.cone_fp import detector_shape
    pgs = workloads.circular_geometry(4, sod=150.0, odd=80.0, rows=20,
                                      cols=24, pixel=1.1, seed=3)
    vg = kk.resolve_volume_geometry(shape=(12, 14, 10), voxel_size=1.0)
    op = kk.XrayTransform(pgs, vg)
    K = op.projector.kernel
    rows, cols = detector_shape(op.projector._proj_geom_seq)
    K._upload(rows, cols, vg)
    y = torch.zeros(16, device="cuda")
    vol = torch.zeros(vg.shape, device="cuda")
    work = torch.zeros(16, device="cuda")
    lib = _native.lib()
    with pytest.raises(NotImplementedError, match="padded pixels"):
        _native.check(lib.kk_cone_fp_transpose(
            K._plan.ptr, _native._vp(y.data_ptr()), _native.i64((1, 1, 1)),
            30000, 40000, _native._vp(vol.data_ptr()),
            _native.i32(vg.shape), _native.i64(vol.stride()), 0,
            _native._vp(work.data_ptr()), _native.stream_ptr()))
    assert 40000 * (30000 + 32) * 4 >= 2 ** 32 > 30000 * 40000
