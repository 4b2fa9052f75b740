"""GPU parity at the named workloads' own sizes.

* small: 128^3, 180 angles over 360 degrees, 160 x 160 detector of
  1.5-unit pixels, source-origin 512, origin-detector 512, seeded
  10-ellipsoid phantom.  Forward projection, backprojection and Ram-Lak
  FDK against the FP64 oracle in full (the oracle runs in a pool of host
  processes); 20 SIRT iterations against the oracle's fixture
  (``tests/golden/small_config_sirt.npz``).
* full: 512^3, 720 angles, 768 x 768 detector, source-origin 1000,
  origin-detector 500.  Forward projection on detector row bands of six
  angles (next to the 45 and 135 degree switches of the ray axis),
  backprojection and Hann FDK on voxel blocks at the bottom, middle and top
  (exact restrictions, ``tests/restrict.py``), 50 SIRT iterations by their
  properties.
* batched: the 64 uneven problems at full scale.  Batched forward and
  backprojection equal the per-problem launches bit for bit for every
  problem; four problems (the largest, an anisotropic one, the smallest and
  one more) against the oracle on row bands and voxel blocks; 30 SIRT
  iterations replayed from a HIP graph equal the eager ones and the
  per-problem ``kk.sirt``.
* large: 1024^3, 1536^2 detector, wobbling orbit with detector drift: the
  exact transpose on voxel blocks against the oracle's transpose, and the
  transpose over 45 angles (launches of 20 angles) against the sum over
  three 15-angle subsets (linearity, relative L2 <= 1e-6).

Bar: relative L2 <= 1e-5 and pointwise <= 1e-4 max|ref| (``assert_parity``);
bit equality where the same kernels run on the same data.
"""
import os

import numpy as np
import pytest

import oracle
import oracle_pool
import restrict
from conftest import assert_parity, rel_l2

pytestmark = pytest.mark.gpu

kk = pytest.importorskip("kernelkit")
from kernelkit import workloads  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _torch():
    import torch
    return torch


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _host(t):
    return t.detach().cpu().numpy()


def _say(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ small


@pytest.fixture(scope="module")
def small():
    pgs, vg = workloads.config_geometry("small")
    vol = workloads.ellipsoid_phantom(vg.shape, n=10, seed=0)
    _say("small: oracle forward projection (process pool)")
    y64 = oracle_pool.cone_fp(vol, pgs, vg)
    return pgs, vg, vol, y64


def test_small_fp_in_full(small, device):
    pgs, vg, vol, y64 = small
    got = kk.XrayTransform(pgs, vg)(_dev(vol))
    assert_parity(_host(got), y64, label="small fp")


def test_small_bp_in_full(small, device):
    pgs, vg, _, y64 = small
    proj = y64.astype(np.float32)
    got = kk.XrayTransform(pgs, vg).T(_dev(proj))
    _say("small: oracle backprojection (process pool)")
    ref = oracle_pool.cone_bp(proj.astype(np.float64), pgs, vg)
    assert_parity(_host(got), ref, label="small bp")


def test_small_fdk_ramlak_in_full(small, device):
    pgs, vg, _, y64 = small
    proj = y64.astype(np.float32)
    got = kk.fdk(proj, pgs, vg)          # Ram-Lak, the default filter
    _say("small: oracle FDK (process pool)")
    ref = oracle_pool.fdk(proj.astype(np.float64), pgs, vg, "ramlak")
    assert_parity(got, ref, label="small fdk")


def test_small_sirt20_against_oracle_fixture(small, device):
    pgs, vg, _, y64 = small
    fx = np.load(os.path.join(GOLDEN, "small_config_sirt.npz"))
    # the measurements are the ones the fixture was written from
    assert abs(float(y64.sum()) - float(fx["y_sum"])) <= \
        1e-12 * abs(float(fx["y_sum"]))
    norms = []
    x = kk.sirt(y64.astype(np.float32), pgs, vg, iters=int(fx["iters"]),
                verbose=False, return_gpu=True,
                callback=lambda i, x, r: norms.append(float(r.double()
                                                            .norm())))
    x = _host(x).astype(np.float64)
    planes = {"z20": x[:, :, 20], "z64": x[:, :, 64], "z100": x[:, :, 100],
              "x64": x[64], "y40": x[:, 40]}
    for k, got in planes.items():
        assert_parity(got, fx[f"plane_{k}"], label=f"small sirt plane {k}")
    for k, axes in (("sum_yz", (1, 2)), ("sum_xz", (0, 2)),
                    ("sum_xy", (0, 1))):
        assert rel_l2(x.sum(axis=axes), fx[k]) <= 1e-5, k
    assert rel_l2((x * x).sum(axis=(0, 1)), fx["sumsq_z"]) <= 1e-5
    assert rel_l2(np.array(norms), fx["res_norms"]) <= 1e-5


# ------------------------------------------------------------------- full


@pytest.fixture(scope="module")
def full(device):
    torch = _torch()
    pgs, vg = workloads.config_geometry("full")
    _say("full: phantom and forward projection on the device")
    x = workloads.ellipsoid_phantom_device(vg.shape, n=10, seed=0,
                                           device="cuda")
    A = kk.XrayTransform(pgs, vg)
    y = A(x)
    torch.cuda.synchronize()
    state = dict(pgs=pgs, vg=vg, A=A, x=x, y=y)
    yield state
    state.clear()
    torch.cuda.empty_cache()


_FULL_ANGLES = (0, 89, 90, 91, 270, 401)     # 0, 44.5, 45, 45.5, 135, 200.5
_FULL_BLOCKS = (((224, 224, 0), (64, 64, 2)),
                ((40, 300, 255), (64, 64, 2)),
                ((400, 100, 510), (64, 64, 2)))


def test_full_fp_row_bands(full):
    pgs, vg = full["pgs"], full["vg"]
    vol = _host(full["x"]).astype(np.float64)
    idx = list(_FULL_ANGLES)
    sub = [pgs[k] for k in idx]
    y = full["y"][idx]
    for r0 in (0, 380, 760):
        _say(f"full: fp rows {r0}..{r0 + 8}, angles {idx}")
        ref = oracle.cone_fp(vol, [restrict.row_band(g, r0, 8) for g in sub],
                             vg, axes=restrict.ray_axes(sub, vg))
        assert_parity(_host(y[:, r0:r0 + 8]), ref, label=f"full fp {r0}")


def test_full_bp_blocks(full):
    pgs, vg = full["pgs"], full["vg"]
    v = _host(full["A"].T(full["y"]))
    yh = _host(full["y"])
    for lo, n in _FULL_BLOCKS:
        _say(f"full: bp block {lo} + {n}")
        nr, nc, org = restrict.window(pgs, vg, lo, n)
        geoms, part = restrict.windowed(pgs, yh, nr, nc, org)
        ref = oracle.cone_bp(part.astype(np.float64), geoms,
                             restrict.sub_volume(vg, lo, n))
        blk = tuple(slice(a, a + k) for a, k in zip(lo, n))
        assert_parity(v[blk], ref, label=f"full bp {lo}")


def test_full_fdk_hann_blocks(full):
    pgs, vg = full["pgs"], full["vg"]
    got = kk.fdk(full["y"], pgs, vg, filter="hann")
    yh = _host(full["y"])
    cols = int(pgs[0].detector.cols)
    for lo, n in _FULL_BLOCKS:
        _say(f"full: fdk block {lo} + {n}")
        nr, _, org = restrict.window(pgs, vg, lo, n)
        geoms, part = restrict.windowed(pgs, yh, nr, cols,
                                        [(r0, 0) for r0, _ in org])
        ref = restrict.fdk(part.astype(np.float64), pgs, geoms,
                           restrict.sub_volume(vg, lo, n), "hann")
        blk = tuple(slice(a, a + k) for a, k in zip(lo, n))
        assert_parity(got[blk], ref, label=f"full fdk {lo}")


def test_full_sirt50_properties(full):
    """50 iterations on consistent data: the weighted residual decreases
    at every iteration and so does the distance to the phantom; with a
    mask (broadcast along z) and box constraints the iterate stays in the
    box and is zero outside the mask."""
    torch = _torch()
    pgs, vg, x_ph, y = full["pgs"], full["vg"], full["x"], full["y"]
    norms, errs = [], []

    def cb(i, x, r):
        norms.append(float(r.double().norm()))
        errs.append(float((x - x_ph).double().norm()))

    _say("full: 50 SIRT iterations")
    kk.sirt(y, pgs, vg, iters=50, verbose=False, callback=cb,
            return_gpu=True)
    _say(f"full sirt residual norms {norms[0]:.6e} -> {norms[-1]:.6e}, "
         f"error {errs[0]:.6e} -> {errs[-1]:.6e}")
    assert all(b <= a * (1 + 1e-6) for a, b in zip(norms, norms[1:])), norms
    assert all(b <= a * (1 + 1e-6) for a, b in zip(errs, errs[1:])), errs
    assert norms[-1] < 0.2 * norms[0]
    assert errs[-1] < float(x_ph.double().norm())
    n = vg.shape[0]
    t = torch.arange(n, device="cuda", dtype=torch.float32) - (n - 1) / 2
    mask = ((t[:, None] ** 2 + t[None, :] ** 2) <
            (0.42 * n) ** 2).float()[:, :, None]     # (n, n, 1): broadcast
    _say("full: 50 SIRT iterations, mask and box constraints")
    x = kk.sirt(y, pgs, vg, iters=50, verbose=False, return_gpu=True,
                mask=mask, min_constraint=0.0, max_constraint=0.75)
    assert float(x.min()) >= 0.0 and float(x.max()) <= 0.75
    assert float(x.max()) == 0.75          # the upper bound is active
    outside = (mask == 0).expand_as(x)
    assert int(torch.count_nonzero(x[outside])) == 0
    assert int(torch.count_nonzero(x[~outside])) > 0.5 * int(
        (~outside).sum())


# ---------------------------------------------------------------- batched


@pytest.fixture(scope="module")
def batch64(device):
    torch = _torch()
    from kernelkit.autotune import BatchTuner
    from kernelkit.batch import ConeBatch
    P = workloads.batch_problems(64)
    B = ConeBatch(P, tuner=BatchTuner(reps=1))
    vols = [workloads.ellipsoid_phantom_device(vg.shape, n=4, seed=k,
                                               device="cuda")
            for k, (_, vg) in enumerate(P)]
    state = dict(P=P, B=B, vols=vols)
    yield state
    state.clear()
    torch.cuda.empty_cache()


def _pick(P):
    """The largest problem (voxels x angles), an anisotropic one, the
    smallest and one in between."""
    work = [int(np.prod(vg.shape)) * len(pg) for pg, vg in P]
    big, small = int(np.argmax(work)), int(np.argmin(work))
    aniso = next(k for k, (_, vg) in enumerate(P)
                 if vg.shape[2] != vg.shape[0] and k not in (big, small))
    mid = next(k for k in range(len(P)) if k not in (big, small, aniso))
    return [big, aniso, small, mid]


def test_batch64_equals_per_problem_and_oracle(batch64):
    torch = _torch()
    P, B, vols = batch64["P"], batch64["B"], batch64["vols"]
    assert len(P) == 64 and max(max(vg.shape) for _, vg in P) == 256
    _say("batched: 64 forward projections in one launch")
    ys = B.fp(vols, tune=False)
    ys = [y.clone() for y in ys]
    for k, (op, v) in enumerate(zip(B.ops, vols)):
        ref = op(v)
        assert torch.equal(ys[k], ref), f"fp problem {k}"
    _say("batched: 64 backprojections in one launch")
    xs = B.bp(ys, tune=False)
    xs = [x.clone() for x in xs]
    for k, (op, y) in enumerate(zip(B.ops, ys)):
        ref = op.T(y)
        assert torch.equal(xs[k], ref), f"bp problem {k}"
    for k in _pick(P):
        pg, vg = P[k]
        _say(f"batched: oracle on problem {k}: {vg.shape}, {len(pg)} angles")
        vol = _host(vols[k]).astype(np.float64)
        na = len(pg)
        idx = [0, na // 3, 2 * na // 3]
        rows = int(pg[0].detector.rows)
        r0 = rows // 2 - 4
        ref = oracle.cone_fp(vol, [restrict.row_band(pg[i], r0, 8)
                                   for i in idx], vg,
                             axes=restrict.ray_axes([pg[i] for i in idx], vg))
        assert_parity(_host(ys[k][idx, r0:r0 + 8]), ref,
                      label=f"batch fp problem {k}")
        lo = tuple(int(s) // 2 - 12 for s in vg.shape[:2]) + (
            int(vg.shape[2]) // 2 - 1,)
        n = (24, 24, 2)
        nr, nc, org = restrict.window(pg, vg, lo, n)
        geoms, part = restrict.windowed(pg, _host(ys[k]), nr, nc, org)
        ref = oracle.cone_bp(part.astype(np.float64), geoms,
                             restrict.sub_volume(vg, lo, n))
        blk = tuple(slice(a, a + m) for a, m in zip(lo, n))
        assert_parity(_host(xs[k])[blk], ref, label=f"batch bp problem {k}")


def test_batch64_sirt30_graph_equals_eager_and_per_problem(batch64):
    torch = _torch()
    P, B, vols = batch64["P"], batch64["B"], batch64["vols"]
    ys = [y.clone() for y in B.fp(vols, tune=False)]
    st = B.sirt_state(ys, min_constraint=0.0)
    st.configure(tune=False)
    _say("batched: 30 SIRT iterations, eager")
    st.run(30, graph=False)
    eager = [x.clone() for x in st.x]
    st.reset()
    _say("batched: 30 SIRT iterations, HIP graph")
    st.run(30, graph=True)
    for k in range(len(P)):
        assert torch.equal(st.x[k], eager[k]), f"graph problem {k}"
    for k in _pick(P)[:3]:
        pg, vg = P[k]
        _say(f"batched: per-problem kk.sirt of problem {k}")
        ref = kk.sirt(ys[k], pg, vg, iters=30, verbose=False,
                      min_constraint=0.0, return_gpu=True)
        assert torch.equal(eager[k], ref), f"per-problem sirt {k}"


# ------------------------------------------------------------------ large


def test_large_transpose_blocks(device):
    torch = _torch()
    pgs, vg = workloads.config_geometry("large")
    sub = [pgs[k] for k in (0, 611, 1200)]
    A = kk.XrayTransform(sub, vg)
    g = torch.Generator(device="cuda").manual_seed(5)
    y = torch.rand(A.range_shape, generator=g, device="cuda")
    _say("large: exact transpose of 3 wobbling angles into 1024^3")
    t = A.matched_T(y)
    yh = _host(y).astype(np.float64)
    blocks = (((488, 488, 0), (48, 48, 2)),
              ((100, 700, 511), (48, 48, 2)),
              ((800, 200, 1022), (48, 48, 2)))
    for lo, n in blocks:
        _say(f"large: transpose block {lo} + {n}")
        nr, nc, org = restrict.window(sub, vg, lo, n, grow=1.5, margin=1)
        geoms, part = restrict.windowed(sub, yh, nr, nc, org)
        ref = oracle.cone_fp_transpose(part, geoms,
                                       restrict.sub_volume(vg, lo, n),
                                       axes=restrict.ray_axes(sub, vg))
        blk = tuple(slice(a, a + k) for a, k in zip(lo, n))
        got = _host(t[blk])
        assert float(np.abs(ref).max()) > 0
        assert_parity(got, ref, label=f"large transpose {lo}")
    del t
    torch.cuda.empty_cache()


def test_large_transpose_angle_chunks_are_linear(device):
    """At 1536^2 an angle's transpose workspace is 9.6 MB, so a launch
    takes 20 angles: 45 wobbling angles run in several launches per ray
    axis, the later ones adding to the volume.  Their transpose equals the
    sum of the transposes of three interleaved 15-angle subsets (one
    launch per axis each), up to float32 summation order."""
    torch = _torch()
    pgs, vg = workloads.config_geometry("large")
    sub = [pgs[k] for k in range(0, 1440, 32)]
    assert len(sub) == 45
    A = kk.XrayTransform(sub, vg)
    g = torch.Generator(device="cuda").manual_seed(7)
    y = torch.rand(A.range_shape, generator=g, device="cuda")
    _say("large: transpose of 45 wobbling angles (chunked launches)")
    t = A.matched_T(y)
    acc = torch.zeros_like(t)
    for part in range(3):
        ks = list(range(part, len(sub), 3))
        _say(f"large: transpose of the angle subset {part}")
        Ak = kk.XrayTransform([sub[k] for k in ks], vg)
        acc += Ak.matched_T(y[ks].contiguous())
        del Ak
    torch.cuda.synchronize()
    d = (t - acc).double()
    rel = float(d.norm() / acc.double().norm())
    point = float(d.abs().max() / acc.abs().max())
    assert float(acc.abs().max()) > 0
    assert rel < 1e-6 and point < 1e-5, (rel, point)
    del t, acc, d
    torch.cuda.empty_cache()


def teardown_module(module):
    oracle_pool.close()
