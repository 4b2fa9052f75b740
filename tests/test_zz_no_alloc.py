"""The allocation-free transpose entry point, through the C ABI.

``kk_cone_fp_transpose_static`` must give what ``kk_cone_fp_transpose``
gives with a caller workspace, call after call on the same device (the
static workspace keeps the previous call's contents), and reject problems
whose workspace exceeds the static array before launching anything.
"""
import numpy as np
import pytest

kk = pytest.importorskip("kernelkit")
from kernelkit import _native, workloads  # noqa: E402
from kernelkit.kernels.cone_fp import _logical_strides, detector_shape  # noqa

pytestmark = pytest.mark.gpu


def _operator(n, rows, cols, shape, wobble=0.0):
    pgs = workloads.circular_geometry(n, sod=150.0, odd=80.0, rows=rows,
                                      cols=cols, pixel=1.1, wobble_deg=wobble,
                                      drift_px=2.0 if wobble else 0.0, seed=3)
    vg = kk.resolve_volume_geometry(shape=shape, voxel_size=1.0)
    return kk.XrayTransform(pgs, vg), pgs, vg


def _both(device, op, seed):
    import torch
    K = op.projector.kernel
    seq = op.projector._proj_geom_seq
    vg = op.projector.volume_geometry
    rows, cols = detector_shape(seq)
    n = len(seq)
    K._upload(rows, cols, vg)
    g = torch.Generator(device="cpu").manual_seed(seed)
    y = torch.rand((n, rows, cols), generator=g).to(device)
    world = tuple(int(s) for s in vg.shape)
    out_ref = torch.full(world, float("nan"), device=device)
    out_new = torch.full(world, float("nan"), device=device)
    lib = _native.lib()
    st = _native.stream_ptr()
    pst = _native.i64(_logical_strides(y, (0, 1, 2)))
    vst = _native.i64(out_ref.stride())
    n_work = int(lib.kk_cone_fp_transpose_work_elems(K._plan.ptr, rows, cols))
    assert n_work >= n * rows * cols
    work = torch.empty(n_work, dtype=torch.float32, device=device)
    _native.check(lib.kk_cone_fp_transpose(
        K._plan.ptr, _native._vp(y.data_ptr()), pst, rows, cols,
        _native._vp(out_ref.data_ptr()), _native.i32(world), vst, 0,
        _native._vp(work.data_ptr()), st))
    _native.check(lib.kk_cone_fp_transpose_static(
        K._plan.ptr, _native._vp(y.data_ptr()), pst, rows, cols,
        _native._vp(out_new.data_ptr()), _native.i32(world), vst, 0, st))
    torch.cuda.synchronize()
    return out_ref.cpu().numpy(), out_new.cpu().numpy()


def test_static_transpose_equals_workspace_transpose(device):
    print("no-alloc transpose: 40 angles, 48x56 detector, 40x36x32", flush=True)
    op1, _, _ = _operator(40, 48, 56, (40, 36, 32))
    ref1, new1 = _both(device, op1, seed=1)
    assert np.isfinite(ref1).all() and np.abs(ref1).max() > 0
    # the same kernels on the same data: bit-identical
    np.testing.assert_array_equal(new1, ref1)
    # a second, smaller call: the static workspace still holds the first
    # call's larger contents, which must not leak into the result
    print("no-alloc transpose: 12 angles, 20x24 detector, 16x18x14, "
          "wobble", flush=True)
    op2, _, _ = _operator(12, 20, 24, (16, 18, 14), wobble=1.0)
    ref2, new2 = _both(device, op2, seed=2)
    assert np.abs(ref2).max() > 0
    np.testing.assert_array_equal(new2, ref2)


def test_static_transpose_through_the_kernel_api(device):
    import torch
    op, _, _ = _operator(24, 32, 40, (24, 24, 24))
    y = torch.rand(op.range_shape, device=device)
    K = op.projector.kernel
    a = torch.zeros(op.domain_shape, device=device)
    b = torch.zeros(op.domain_shape, device=device)
    K.transpose(y, op.projector._proj_geom_seq, op.projector.volume_geometry,
                a)
    K.transpose(y, op.projector._proj_geom_seq, op.projector.volume_geometry,
                b, allocate=False)
    np.testing.assert_array_equal(b.cpu().numpy(), a.cpu().numpy())
    # additive accumulation through the static entry
    K.transpose(y, op.projector._proj_geom_seq, op.projector.volume_geometry,
                b, additive=True, allocate=False)
    np.testing.assert_allclose(b.cpu().numpy(), 2 * a.cpu().numpy(),
                               rtol=1e-6, atol=0)


def test_static_transpose_rejects_oversized_workspace(device):
    import torch
    cap = int(_native.lib().kk_cone_fp_transpose_static_capacity())
    assert cap == 1 << 26
    # 1100 angles x 256 x 256 pixels = 72 Mi floats > 64 Mi: rejected by the
    # host check before any launch (the inputs are real, full-size buffers)
    n, rows, cols = 1100, 256, 256
    assert n * rows * cols > cap
    op, _, _ = _operator(n, rows, cols, (8, 8, 8))
    K = op.projector.kernel
    y = torch.zeros((n, rows, cols), device=device)
    v = torch.zeros((8, 8, 8), device=device)
    with pytest.raises(NotImplementedError, match="static scratch"):
        K.transpose(y, op.projector._proj_geom_seq,
                    op.projector.volume_geometry, v, allocate=False)
