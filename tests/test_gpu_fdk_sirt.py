"""GPU parity of FDK pre-processing (pre-weight, row filter), FDK and SIRT
against the FP64 oracle (relative L2 <= 1e-5, pointwise <= 1e-4 max|ref|).
"""
import numpy as np
import pytest

import oracle
from conftest import assert_parity

pytestmark = pytest.mark.gpu

kk = pytest.importorskip("kernelkit")
from kernelkit import workloads  # noqa: E402


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _host(t):
    return t.detach().cpu().numpy()


def _case(n_angles=24, shape=(32, 36, 28), rows=40, cols=48, seed=21, **kw):
    pgs = workloads.circular_geometry(n_angles, sod=120.0, odd=60.0,
                                      rows=rows, cols=cols, pixel=1.2,
                                      seed=seed, **kw)
    vg = kk.resolve_volume_geometry(shape=shape, voxel_size=1.0)
    vol = workloads.ellipsoid_phantom(shape, n=6, seed=seed)
    return pgs, vg, vol


@pytest.mark.parametrize("name", ["ramlak", "ramp", "cosine", "hann",
                                  "hamming", "shepp-logan"])
@pytest.mark.parametrize("cols", [31, 48, 160, 700, 1536, 2500, 5000])
def test_filter_rows(name, cols, device):
    """Row widths up to 5000 (filter length N = 16384): N <= 4096 keeps the
    twiddles in LDS, N = 8192 and 16384 read them from global memory, and
    N = 16384 holds 128 KiB of signal in LDS."""
    rng = np.random.default_rng(cols)
    proj = rng.uniform(size=(3, 5, cols))
    ref = oracle.filter_rows(proj, name)
    got = _dev(proj)
    kk.processing.filter(got, filter=name)
    assert_parity(_host(got), ref, label=f"filter/{name}/{cols}")


def test_filter_list_input(device):
    rng = np.random.default_rng(3)
    proj = rng.uniform(size=(4, 6, 50))
    ref = oracle.filter_rows(proj, "ramlak")
    lst = [_dev(p) for p in proj]
    kk.processing.filter(lst)
    assert_parity(np.stack([_host(p) for p in lst]), ref, label="filter/list")


def test_preweight(device):
    pgs, vg, vol = _case(n_angles=6, tilt_deg=2.0, drift_px=3.0)
    rng = np.random.default_rng(4)
    proj = rng.uniform(size=(6, 40, 48))
    ref = oracle.preweight(proj, pgs)
    got = _dev(proj)
    kk.processing.preweight(got, pgs)
    assert_parity(_host(got), ref, label="preweight")


@pytest.mark.parametrize("name", ["ramlak", "hann", "cosine"])
def test_fdk_matches_oracle(name, device):
    pgs, vg, vol = _case()
    proj = oracle.cone_fp(vol, pgs, vg)
    ref = oracle.fdk(proj, pgs, vg, filter=name)
    got = kk.fdk(proj.astype(np.float32), pgs, vg, filter=name)
    assert_parity(got, ref, label=f"fdk/{name}")


def test_fdk_device_input_not_modified(device):
    pgs, vg, vol = _case(n_angles=8)
    proj = oracle.cone_fp(vol, pgs, vg).astype(np.float32)
    dev = _dev(proj)
    kk.fdk(dev, pgs, vg)
    np.testing.assert_array_equal(_host(dev), proj)


def test_bp_with_preproc_and_device_out(device):
    import torch
    pgs, vg, vol = _case(n_angles=8)
    proj = oracle.cone_fp(vol, pgs, vg)
    ref = oracle.cone_bp(2.0 * proj, pgs, vg)

    def double(p, geoms):
        p.mul_(2.0)

    out = torch.zeros(vg.shape, dtype=torch.float32, device="cuda")
    kk.bp(proj.astype(np.float32), pgs, vg, preproc_fn=double, out=out)
    assert_parity(_host(out), ref, label="bp/preproc")


@pytest.mark.parametrize("opts", [
    {},
    {"min_constraint": 0.0},
    {"min_constraint": 0.0, "max_constraint": 0.8, "use_mask": True},
])
def test_sirt_matches_oracle(opts, device):
    pgs, vg, vol = _case(n_angles=20, shape=(24, 28, 20), rows=32, cols=36)
    proj = oracle.cone_fp(vol, pgs, vg)
    opts = dict(opts)
    mask = None
    if opts.pop("use_mask", False):
        mask = np.zeros(vg.shape)
        mask[2:-2, 2:-2, 1:-1] = 1.0
    ref = oracle.sirt(proj, pgs, vg, iters=6, mask=mask, **opts)
    got = kk.sirt(proj.astype(np.float32), pgs, vg, iters=6, verbose=False,
                  mask=mask, **opts)
    assert_parity(got, ref, label=f"sirt/{opts}")


def test_sirt_callback_and_return_gpu(device):
    pgs, vg, vol = _case(n_angles=12, shape=(16, 16, 16), rows=24, cols=24)
    proj = oracle.cone_fp(vol, pgs, vg)
    seen = []

    def cb(i, x, res):
        seen.append((i, float(res.abs().sum())))

    x = kk.sirt(proj.astype(np.float32), pgs, vg, iters=4, verbose=False,
                callback=cb, return_gpu=True)
    assert x.is_cuda and tuple(x.shape) == vg.shape
    assert [i for i, _ in seen] == [0, 1, 2, 3]
    # the weighted residual decreases on consistent data
    assert seen[-1][1] < seen[0][1]


def test_sirt_residual_fixed_point(device):
    """Data that is exactly A x* for the GPU operator: SIRT reduces the
    residual monotonically."""
    pgs, vg, vol = _case(n_angles=16, shape=(16, 16, 16), rows=24, cols=24)
    A = kk.XrayTransform(pgs, vg)
    y = A(_dev(vol))
    norms = []
    kk.sirt(_host(y), pgs, vg, iters=8, verbose=False,
            callback=lambda i, x, r: norms.append(float(r.norm())))
    assert all(b <= a * 1.0001 for a, b in zip(norms, norms[1:]))


@pytest.mark.parametrize("pert", [{}, {"wobble_deg": 1.0, "drift_px": 2.0}])
def test_fused_residual_equals_two_step(pert, device):
    """The fused forward-projection residual epilogue (kk_sirt_fp_residual,
    and the batch kernel's kk_batch_add_sirt_fp) writes exactly the BP
    texture of the two-step form algo.sirt uses (plain kk_cone_fp into a
    buffer, then kk_residual_to_bp_texture): the same float operations."""
    import torch
    from kernelkit import _native
    from kernelkit.kernel import ProjectionTexture
    pgs, vg, vol = _case(**pert)
    op = kk.XrayTransform(pgs, vg)
    x = _dev(vol)
    n, rows, cols = len(pgs), pgs[0].detector.rows, pgs[0].detector.cols
    rng = np.random.default_rng(5)
    y = _dev(rng.uniform(0.0, 5.0, (n, rows, cols)))
    R = _dev(rng.uniform(0.1, 2.0, (n, rows, cols)))
    op.projector.volume = x
    xtex = op.projector._texture
    fk = op.projector.kernel
    lib = _native.lib()
    st = _native.stream_ptr()
    texs = [ProjectionTexture(n, rows, cols, device=x.device)
            for _ in range(3)]
    for t in texs:
        t.zero()
    _native.check(lib.kk_sirt_fp_residual(
        fk._plan.ptr, _native._vp(xtex.ptr), _native._vp(y.data_ptr()),
        _native.i64(y.stride()), _native._vp(R.data_ptr()),
        _native.i64(R.stride()), _native._vp(texs[0].ptr), n, rows, cols,
        st))
    p = torch.empty((n, rows, cols), dtype=torch.float32, device=x.device)
    _native.check(lib.kk_cone_fp(
        fk._plan.ptr, _native._vp(xtex.ptr), _native._vp(p.data_ptr()),
        _native.i64(p.stride()), rows, cols, 0, st))
    _native.check(lib.kk_residual_to_bp_texture(
        _native._vp(p.data_ptr()), _native.i64(p.stride()),
        _native._vp(y.data_ptr()), _native.i64(y.stride()),
        _native._vp(R.data_ptr()), _native.i64(R.stride()), n, rows, cols,
        _native._vp(texs[1].ptr), st))
    b = _native.Batch()
    _native.check(lib.kk_batch_add_sirt_fp(
        b.ptr, fk._plan.ptr, _native._vp(xtex.ptr),
        _native._vp(y.data_ptr()), _native.i64(y.stride()),
        _native._vp(R.data_ptr()), _native.i64(R.stride()),
        _native._vp(texs[2].ptr), n, rows, cols))
    _native.check(lib.kk_batch_configure(b.ptr, 0, 8))
    _native.check(lib.kk_batch_fp(b.ptr, 0, st))
    torch.cuda.synchronize()
    ref = (_host(p) - _host(y)) * _host(R)
    assert float(np.abs(ref).max()) > 0
    assert torch.equal(texs[0].data, texs[1].data)
    assert torch.equal(texs[2].data, texs[1].data)


@pytest.mark.gpu
def test_fused_filter_tables_are_reused_and_keyed_by_preweight():
    """``filter_into`` keeps its device tables per (device, length, filter,
    pre-weight table): a repeated call reuses them and writes the same
    texture bit for bit; another geometry or filter gets tables of its own
    and matches a fresh call; a geometry count that differs from the
    projection count raises before any launch."""
    import torch
    from kernelkit import processing
    from kernelkit.kernel import ProjectionTexture
    pgs = workloads.circular_geometry(12, sod=200.0, odd=100.0, rows=20,
                                      cols=30, pixel=1.0, seed=1)
    # another source-to-detector distance (the cosine weights depend only
    # on it and the pixel positions)
    pgs2 = workloads.circular_geometry(12, sod=210.0, odd=110.0, rows=20,
                                       cols=30, pixel=1.0, seed=2)
    y = torch.from_numpy(np.random.default_rng(5).uniform(
        size=(12, 20, 30)).astype(np.float32)).cuda()

    def run(geoms, name):
        tex = ProjectionTexture(*y.shape, like=y)
        tex.zero()
        processing.filter_into(y, tex, geoms=geoms, filter=name)
        torch.cuda.synchronize()
        return tex

    processing._TABLES.clear()
    a = run(pgs, "ramlak")
    assert len(processing._TABLES) == 1
    b = run(pgs, "ramlak")
    assert len(processing._TABLES) == 1
    assert torch.equal(a.data, b.data)
    c = run(pgs2, "ramlak")
    d = run(pgs, "hann")
    assert len(processing._TABLES) == 3
    processing._TABLES.clear()
    assert torch.equal(run(pgs2, "ramlak").data, c.data)
    assert torch.equal(run(pgs, "hann").data, d.data)
    assert not torch.equal(a.data, c.data) and not torch.equal(a.data, d.data)
    with pytest.raises(ValueError, match="11 geometries for 12"):
        run(pgs[:11], "ramlak")
