"""Batched problems (one launch per operator) against the per-problem
path, the oracle and a HIP-graph replay of SIRT."""
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


def _problems(n=5, seed=1):
    return workloads.batch_problems(n, seed=seed, scale=0.25)


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _host(t):
    return t.detach().cpu().numpy()


def _volumes(problems, seed=0):
    return [workloads.ellipsoid_phantom(vg.shape, n=4, seed=seed + k)
            for k, (_, vg) in enumerate(problems)]


def test_batch_problem_sizes_are_uneven(device):
    P = workloads.batch_problems(64)
    shapes = {vg.shape for _, vg in P}
    angles = {len(pg) for pg, _ in P}
    dets = {(pg[0].detector.rows, pg[0].detector.cols) for pg, _ in P}
    assert len(shapes) > 10 and len(angles) > 30 and len(dets) > 5
    assert all(90 <= len(pg) <= 360 for pg, _ in P)


@pytest.mark.parametrize("tune", [False, True])
def test_batch_fp_bp_match_single(tune, device):
    from kernelkit.autotune import BatchTuner
    from kernelkit.batch import ConeBatch
    P = _problems()
    B = ConeBatch(P, tuner=BatchTuner(reps=1))
    vols = [_dev(v) for v in _volumes(P)]
    ys = B.fp(vols, tune=tune)
    for (pg, vg), v, y in zip(P, vols, ys):
        ref = kk.XrayTransform(pg, vg)(v)
        np.testing.assert_allclose(_host(y), _host(ref), rtol=1e-6,
                                   atol=1e-6 * float(ref.abs().max()))
    xs = B.bp(ys, tune=tune)
    for (pg, vg), y, x in zip(P, ys, xs):
        ref = kk.XrayTransform(pg, vg).T(y)
        np.testing.assert_allclose(_host(x), _host(ref), rtol=1e-6,
                                   atol=1e-6 * float(ref.abs().max()))


def test_tuner_cache_file_round_trip(tmp_path, device):
    """A tuned choice is written atomically to the cache file and reused by
    a second tuner without timing again; keys carry the library build."""
    import json
    from kernelkit import _native
    from kernelkit.autotune import BatchTuner
    from kernelkit.batch import ConeBatch
    path = tmp_path / "tune.json"
    P = _problems(3, seed=2)
    B = ConeBatch(P, tuner=BatchTuner(reps=1, cache_file=str(path)))
    vols = [_dev(v) for v in _volumes(P)]
    ys = B.fp(vols, tune=True)
    data = json.loads(path.read_text())
    assert len(data) == 1 and _native.build_id() in next(iter(data))
    assert not list(tmp_path.glob("*.tmp"))
    t2 = BatchTuner(reps=1, cache_file=str(path))
    t2._tune = None          # a cache hit must not time anything
    B2 = ConeBatch(P, tuner=t2)
    ys2 = B2.fp(vols, tune=True)
    for a, b in zip(ys, ys2):
        assert bool((a == b).all())


def test_batch_matches_oracle(device):
    from kernelkit.batch import ConeBatch
    # a few angles each keep the oracle fast
    P = [(pg[::max(1, len(pg) // 16)], vg) for pg, vg in _problems(3, seed=4)]
    vols = _volumes(P, seed=3)
    B = ConeBatch(P)
    ys = B.fp([_dev(v) for v in vols], tune=False)
    for (pg, vg), v, y in zip(P, vols, ys):
        assert_parity(_host(y), oracle.cone_fp(v, pg, vg), label="batch fp")
    xs = B.bp(ys, tune=False)
    for (pg, vg), y, x in zip(P, ys, xs):
        assert_parity(_host(x), oracle.cone_bp(_host(y), pg, vg),
                      label="batch bp")


def test_batch_sirt_matches_plain_and_graph_replays(device):
    torch = _torch()
    from kernelkit.batch import ConeBatch
    P = _problems(4, seed=7)
    vols = _volumes(P, seed=5)
    ys = [kk.XrayTransform(pg, vg)(_dev(v)) for (pg, vg), v in zip(P, vols)]
    refs = [kk.sirt(y, pg, vg, iters=6, verbose=False, min_constraint=0.0,
                    return_gpu=True) for (pg, vg), y in zip(P, ys)]
    B = ConeBatch(P)
    st = B.sirt_state(ys, min_constraint=0.0)
    st.configure(tune=True)
    st.run(3, graph=False)
    st.run(3, graph=False)
    for x, r in zip(st.x, refs):
        np.testing.assert_allclose(_host(x), _host(r), rtol=2e-5,
                                   atol=2e-5 * float(r.abs().max()))
    # graph: capture 3 iterations, replay twice with eager work between
    st.reset()
    st.capture(3)
    st._graph[1].replay()
    torch.cuda.synchronize()
    eager = [r.clone() for r in refs]  # keep references alive
    _ = torch.ones(10, device="cuda").sum().item()
    st._graph[1].replay()
    torch.cuda.synchronize()
    for x, r in zip(st.x, eager):
        np.testing.assert_allclose(_host(x), _host(r), rtol=2e-5,
                                   atol=2e-5 * float(r.abs().max()))
    # a third replay continues iterating: matches 9 plain iterations
    st._graph[1].replay()
    torch.cuda.synchronize()
    for (pg, vg), y, x in zip(P, ys, st.x):
        r = kk.sirt(y, pg, vg, iters=9, verbose=False, min_constraint=0.0,
                    return_gpu=True)
        np.testing.assert_allclose(_host(x), _host(r), rtol=2e-5,
                                   atol=2e-5 * float(r.abs().max()))


def test_batch_sirt_state_owns_its_textures(device):
    """Operator calls on the batch between SIRT iterations (which rebuild
    the operators' own projector textures) leave the SIRT state's forward
    textures alone: the iterates are bit-equal to an uninterrupted run."""
    torch = _torch()
    from kernelkit.batch import ConeBatch
    P = _problems(3, seed=11)
    vols = [_dev(v) for v in _volumes(P, seed=2)]
    B = ConeBatch(P)
    ys = B.fp(vols, tune=False)
    ys = [y.clone() for y in ys]
    st = B.sirt_state(ys)
    st.configure(tune=False)
    st.run(4, graph=False)
    ref = [x.clone() for x in st.x]
    st.reset()
    st.run(2, graph=False)
    other = [torch.rand_like(v) for v in vols]
    B.fp(other, tune=False)                    # rebuilds projector textures
    for op, v in zip(B.ops, other):
        op(v)
    st.run(2, graph=False)
    torch.cuda.synchronize()
    for x, r in zip(st.x, ref):
        assert torch.equal(x, r)


def test_batch_shape_errors(device):
    from kernelkit.batch import ConeBatch
    P = _problems(2)
    B = ConeBatch(P)
    with pytest.raises(ValueError):
        B.fp([_dev(np.zeros((3, 3, 3)))] * 2)
    with pytest.raises(ValueError):
        B.fp([_dev(np.zeros(P[0][1].shape))])


def test_batch_texture_builds_match_single(device):
    """kk_batch_residual / kk_batch_volume_textures write, bit for bit, the
    textures of the per-problem kk_residual_to_bp_texture /
    kk_volume_to_fp_texture calls, over jobs of different shapes, strided
    inputs and frame masks."""
    torch = _torch()
    from kernelkit import _native
    lib = _native.lib()
    st = _native.stream_ptr()
    g = torch.Generator(device="cuda").manual_seed(3)
    b = _native.Batch()
    res_cases, vol_cases = [], []
    keep = []      # the batch holds raw pointers: keep its inputs alive
    for n, rows, cols in ((3, 7, 11), (5, 17, 300), (2, 1, 1)):
        base = torch.rand((n, rows, cols + 2), device="cuda", generator=g)
        p = base[:, :, 1:cols + 1]          # row-strided
        y = torch.rand((n, rows, cols), device="cuda", generator=g)
        R = torch.rand((cols, rows, n), device="cuda",
                       generator=g).permute(2, 1, 0)   # transposed strides
        size = n * (rows + 3) * (cols + 3) * 2
        t_batch = torch.full((size,), 7.0, device="cuda")
        t_single = torch.full((size,), 7.0, device="cuda")
        _native.check(lib.kk_batch_add_residual(
            b.ptr, _native._vp(p.data_ptr()), _native.i64(p.stride()),
            _native._vp(y.data_ptr()), _native.i64(y.stride()),
            _native._vp(R.data_ptr()), _native.i64(R.stride()), n, rows,
            cols, _native._vp(t_batch.data_ptr())))
        _native.check(lib.kk_residual_to_bp_texture(
            _native._vp(p.data_ptr()), _native.i64(p.stride()),
            _native._vp(y.data_ptr()), _native.i64(y.stride()),
            _native._vp(R.data_ptr()), _native.i64(R.stride()), n, rows,
            cols, _native._vp(t_single.data_ptr()), st))
        res_cases.append((t_batch, t_single))
        keep += [base, y, R]
    for shape, mask in (((9, 5, 6), 7), ((4, 33, 2), 3), ((1, 1, 70), 4)):
        v = torch.rand(tuple(reversed(shape)), device="cuda",
                       generator=g).permute(2, 1, 0)    # world strides
        elems = lib.kk_fp_texture_elems(_native.i32(shape), mask)
        t_batch = torch.full((elems,), 7.0, device="cuda")   # floats
        t_single = torch.full((elems,), 7.0, device="cuda")
        _native.check(lib.kk_batch_add_volume_texture(
            b.ptr, _native._vp(v.data_ptr()), _native.i32(shape),
            _native.i64(v.stride()), mask, _native._vp(t_batch.data_ptr())))
        _native.check(lib.kk_volume_to_fp_texture(
            _native._vp(v.data_ptr()), _native.i32(shape),
            _native.i64(v.stride()), mask, _native._vp(t_single.data_ptr()),
            st))
        vol_cases.append((t_batch, t_single))
        keep.append(v)
    assert lib.kk_batch_size(b.ptr, 2) == 3
    assert lib.kk_batch_size(b.ptr, 3) == 3
    # not configured yet
    with pytest.raises(Exception):
        _native.check(lib.kk_batch_residual(b.ptr, st))
    _native.check(lib.kk_batch_configure(b.ptr, 0, 8))
    _native.check(lib.kk_batch_residual(b.ptr, st))
    _native.check(lib.kk_batch_volume_textures(b.ptr, st))
    torch.cuda.synchronize()
    for t_batch, t_single in res_cases + vol_cases:
        assert torch.equal(t_batch, t_single)
