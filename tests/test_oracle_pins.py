"""Pins of the FP64 oracle (CPU).

Two kinds of check, chosen so that a plausible slip in the oracle (a dropped
term, a wrong sign, index or axis, a transposed operand) fails one of them:

* the reference's own known-answer tests, restated against the oracle with
  their tolerances (``tests/kernels/test_cone_fp.py``,
  ``tests/kernels/test_cone_bp.py`` of the reference);
* closed forms the mathematics fixes: exact line integrals of linear
  volumes (the Joseph midpoint rule is exact for them), the ray/detector
  plane intersection for the backprojector's detector coordinates, the
  magnification weight of one-voxel volumes, the dot-product identity of the
  transpose, direct spatial convolution with the Kak-Slaney kernel for the
  filter, the cosine law for the pre-weight, and textbook rotations.
"""
import numpy as np
import pytest

import oracle
from oracle.geometry import normalized_frames, fp_major_axes, rotation_sxyz

kk = pytest.importorskip("kernelkit")
from kernelkit import Detector, ProjectionGeometry, rotate  # noqa: E402
from kernelkit import resolve_volume_geometry, VolumeGeometry  # noqa: E402
from kernelkit import workloads  # noqa: E402


# --------------------------------------------------------------- rotations

def test_rotation_textbook():
    x, y, z = np.eye(3)
    np.testing.assert_allclose(rotation_sxyz(0, 0, np.pi / 2) @ x, y,
                               atol=1e-15)
    np.testing.assert_allclose(rotation_sxyz(0, np.pi / 2, 0) @ x, -z,
                               atol=1e-15)
    np.testing.assert_allclose(rotation_sxyz(np.pi / 2, 0, 0) @ y, z,
                               atol=1e-15)
    rng = np.random.default_rng(0)
    for _ in range(5):
        r, p, w = rng.uniform(-3, 3, 3)
        R = rotation_sxyz(r, p, w)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1) < 1e-14
        # static-frame x, then y, then z
        np.testing.assert_allclose(
            R, rotation_sxyz(0, 0, w) @ rotation_sxyz(0, p, 0)
            @ rotation_sxyz(r, 0, 0), atol=1e-14)


def test_product_rotation_matches_oracle():
    from kernelkit.geom.proj import angles2mat
    rng = np.random.default_rng(1)
    for _ in range(5):
        a = rng.uniform(-3, 3, 3)
        np.testing.assert_allclose(angles2mat(*a), rotation_sxyz(*a),
                                   atol=1e-15)


# ------------------------------------------ reference KATs: forward project

@pytest.mark.parametrize("axis", (0, 1, 2))
def test_kat_fp_projection_axes(axis):
    """test_cone_fp.py:74-169: the 20x10-voxel square lands on 128 pixels
    of the lower half of the detector, along the expected ray axis."""
    det = Detector(rows=128, cols=128, pixel_height=8. / 128,
                   pixel_width=8. / 128)
    SOD = 1e6
    if axis == 0:
        gt = np.zeros((1, 100, 100))
        gt[..., 80:100, 90:100] = 1.0
        g = ProjectionGeometry([-SOD, 0, 0], [0, 0, 0], det, u=[0, 1, 0],
                               v=[0, 0, 1])
    elif axis == 1:
        gt = np.zeros((100, 1, 100))
        gt[80:100, ..., 90:100] = 1.0
        g = ProjectionGeometry([0, -SOD, 0], [0, 0, 0], det, u=[1, 0, 0],
                               v=[0, 0, 1])
    else:
        gt = np.zeros((100, 100, 1))
        gt[80:100, 90:100] = 1.0
        g = ProjectionGeometry([0, 0, -SOD], [0, 0, 0], det, u=[1, 0, 0],
                               v=[0, 1, 0])
    vg = resolve_volume_geometry(shape=gt.shape, voxel_size=[0.05] * 3)
    assert fp_major_axes(normalized_frames([g], vg))[0] == axis
    p = oracle.cone_fp(gt, [g], vg)
    out = (p[0] > 1e-8).astype(np.int64)
    out[:64, :] = 0
    assert out.sum() == 128


@pytest.mark.parametrize("n", (1, 7, 100))
def test_kat_fp_unit_cube(n):
    """test_cone_fp.py:172-221: the central ray through a unit cube of ones
    integrates to 1 whatever the discretisation (decimal 7)."""
    det = Detector(1, 1, 1. / 1024, 1. / 1024)
    g = ProjectionGeometry([-4., 0, 0], [0.4, 0, 0], det)
    vg = resolve_volume_geometry(shape=[n] * 3, extent_min=[-.5] * 3,
                                 extent_max=[.5] * 3)
    p = oracle.cone_fp(np.ones([n] * 3), [g], vg)
    np.testing.assert_almost_equal(p[0, 0, 0], 1.0, decimal=7)


@pytest.mark.parametrize("direction", (0, 1, 2))
def test_kat_fp_line(direction):
    """test_cone_fp.py:224-302: a line of ones along the ray integrates to 1
    at the central pixel; neighbouring values are not seen."""
    det = Detector(9, 9, 1. / 128, 1. / 128)
    g = ProjectionGeometry([-4., 0, 0], [0.4, 0, 0], det)
    gt = np.zeros((9, 9, 9))
    if direction == 0:
        gt[:, 4, 4] = 1.
    elif direction == 1:
        g = rotate(g, yaw=np.pi / 2)
        gt[4, :, 4] = 1.
    else:
        g = rotate(g, pitch=np.pi / 2)
        gt[4, 4] = 1.
    vg = resolve_volume_geometry(shape=gt.shape, extent_min=[-.5] * 3,
                                 extent_max=[.5] * 3)
    assert fp_major_axes(normalized_frames([g], vg))[0] == direction
    p = oracle.cone_fp(gt, [g], vg)
    np.testing.assert_almost_equal(p[0, 4, 4], 1.0, decimal=7)


def test_kat_fp_anisotropic_mass():
    """test_cone_fp.py:438-483: near-parallel projection of ones conserves
    mass, sum(p) * pixel area = sum(vol) * voxel volume (decimal 6)."""
    fact = 8
    det = Detector(30 * fact, 10 * fact, 1. / fact, 1. / fact)
    g = ProjectionGeometry([-1e6, 0, 0], [0, 0, 0], det)
    vg = resolve_volume_geometry(shape=[16] * 3, voxel_size=[1.0, 0.5, 1.5])
    p = oracle.cone_fp(np.ones([16] * 3), [g], vg)
    np.testing.assert_almost_equal(p.sum() / fact / fact,
                                   16 ** 3 * np.prod(vg.voxel_size),
                                   decimal=6)


# --------------------------------------------- reference KATs: backproject

@pytest.mark.parametrize("axis", (0, 1, 2))
def test_kat_bp_axes(axis):
    """test_cone_bp.py:84-142: a one-pixel near-parallel projection of ones
    backprojects onto a line of ones through the 3^3 volume (decimal 4)."""
    det = Detector(1, 1, 1., 1.)
    g0 = ProjectionGeometry([-1e5, 0, 0], [0, 0, 0], det)
    exp = np.zeros((3, 3, 3))
    if axis == 0:
        g = rotate(g0, yaw=0.0)
        exp[:, 1, 1] = 1.
    elif axis == 1:
        g = rotate(g0, yaw=np.pi / 2)
        exp[1, :, 1] = 1.
    else:
        g = rotate(g0, pitch=np.pi / 2)
        exp[1, 1, :] = 1.
    vg = resolve_volume_geometry(shape=(3, 3, 3), voxel_size=(1., 1., 1.))
    v = oracle.cone_bp(np.ones((1, 1, 1)), [g], vg)
    np.testing.assert_array_almost_equal(v, exp, decimal=4)


# ------------------------------------------------- closed forms: Joseph FP

def _ray_cases():
    """(ray axis, geometry, volume geometry, f) with linear f on local
    voxel-centre coordinates and rays that stay inside the interpolation
    support of every slice."""
    c = np.array([1.0, -0.5, 0.3])
    cases = []
    for axis in (0, 1, 2):
        perm = {0: (0, 1, 2), 1: (1, 0, 2), 2: (2, 1, 0)}[axis]
        local_shape = (6, 22, 20)           # (a, b, c) extents in voxels
        shape = [0, 0, 0]
        for k, ax in enumerate(perm):
            shape[ax] = local_shape[k]
        vs = 0.7
        half = np.array(shape) * vs / 2

        def w(a, b, cc):
            out = np.zeros(3)
            out[perm[0]], out[perm[1]], out[perm[2]] = a, b, cc
            return out
        for rot in ((0.0, 0.0, 0.0), (0.05, -0.08, 0.1)):
            R = rotation_sxyz(*rot)
            s = c + R @ w(-50.0, 1.5, -1.0)
            d = c + R @ w(30.0, 0.0, 0.0)
            u = R @ w(0, 1, 0)
            v = R @ w(0, 0, 1)
            g = ProjectionGeometry(s, d, Detector(3, 4, 0.6, 0.5), u=u, v=v)
            vg = VolumeGeometry(shape, [vs] * 3, c - half, c + half,
                                rotation=rot)
            coef = np.array([0.3, -1.1, 0.8]) * (axis + 1)
            cases.append((axis, g, vg, R, c, coef))
    return cases


@pytest.mark.parametrize("case", range(6))
def test_fp_exact_on_linear_volume(case):
    """For f linear in the volume's local coordinates the Joseph sum equals
    the exact integral over the major-axis slab: length x f(midpoint)."""
    axis, g, vg, R, c, coef = _ray_cases()[case]
    shape = np.array(vg.shape)
    # the volume geometry stores float32 extents, sizes and angles; the
    # voxel grid is centred on the extents with the stored voxel size
    lo = np.asarray(vg.extent_min, np.float64)
    hi = np.asarray(vg.extent_max, np.float64)
    c = (lo + hi) / 2
    vs = float(vg.voxel_size[0])
    R = rotation_sxyz(*[float(r) for r in vg.rotation])
    idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    local = [(idx[k] + 0.5) * vs - shape[k] * vs / 2 for k in range(3)]
    vol = 2.0 + sum(coef[k] * local[k] for k in range(3))
    assert fp_major_axes(normalized_frames([g], vg))[0] == axis
    p = oracle.cone_fp(vol, [g], vg)
    det = g.detector
    s_l = R.T @ (np.asarray(g.source_position) - c)
    dmin = (np.asarray(g.detector_position) - np.asarray(g.v) * det.rows / 2
            * det.pixel_height - np.asarray(g.u) * det.cols / 2
            * det.pixel_width)
    half_a = shape[axis] * vs / 2
    for r in range(det.rows):
        for col in range(det.cols):
            D = (dmin + (col + 0.5) * det.pixel_width * np.asarray(g.u)
                 + (r + 0.5) * det.pixel_height * np.asarray(g.v))
            D_l = R.T @ (D - c)
            t0 = (-half_a - s_l[axis]) / (D_l[axis] - s_l[axis])
            t1 = (half_a - s_l[axis]) / (D_l[axis] - s_l[axis])
            P0 = s_l + t0 * (D_l - s_l)
            P1 = s_l + t1 * (D_l - s_l)
            mid = (P0 + P1) / 2
            # the ray stays inside the voxel-centre hull in b and c
            for k in range(3):
                if k != axis:
                    lim = shape[k] * vs / 2 - vs / 2
                    assert max(abs(P0[k]), abs(P1[k])) < lim
            expect = np.linalg.norm(P1 - P0) * (2.0 + coef @ mid)
            assert p[0, r, col] == pytest.approx(expect, rel=1e-12)


def test_fp_linearity_and_zero():
    pgs = workloads.circular_geometry(3, sod=40.0, odd=20.0, rows=7,
                                         cols=8, pixel=1.3, seed=2,
                                         tilt_deg=2.0)
    vg = resolve_volume_geometry(shape=(9, 8, 7), voxel_size=1.1)
    rng = np.random.default_rng(5)
    a, b = rng.uniform(size=(2,) + vg.shape)
    pa, pb = oracle.cone_fp(a, pgs, vg), oracle.cone_fp(b, pgs, vg)
    np.testing.assert_allclose(oracle.cone_fp(2 * a - 3 * b, pgs, vg),
                               2 * pa - 3 * pb, rtol=1e-12, atol=1e-12)
    assert not oracle.cone_fp(np.zeros(vg.shape), pgs, vg).any()


# ---------------------------------------------------- closed forms: BP

def _tilted_geometries(n=5, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        g = ProjectionGeometry([-30.0, rng.uniform(-2, 2), rng.uniform(-2, 2)],
                               [18.0, rng.uniform(-1, 1), rng.uniform(-1, 1)],
                               Detector(11, 13, 1.1, 0.9))
        g = rotate(g, roll=rng.uniform(-0.1, 0.1),
                   pitch=rng.uniform(-0.1, 0.1), yaw=2 * np.pi * k / n)
        # tilt the detector about its own centre
        Rt = rotation_sxyz(*rng.uniform(-0.05, 0.05, 3))
        g.u = Rt @ np.asarray(g.u)
        g.v = Rt @ np.asarray(g.v)
        out.append(g)
    return out


def test_bp_detector_coordinates_are_plane_intersections():
    """u = numU/den and v = numV/den equal the intersection of the line
    source -> voxel centre with the detector plane, in pixel steps from the
    corner of pixel (0, 0)."""
    pgs = _tilted_geometries()
    vg = VolumeGeometry((7, 6, 5), [0.8] * 3, [-2.0, -3.0, -1.5],
                        [3.6, 1.8, 2.5], rotation=(0.02, 0.05, -0.1))
    fr = normalized_frames(pgs, vg)
    P = oracle.bp_params(fr, vg.voxel_size, with_adjoint_scaling=True)
    rng = np.random.default_rng(7)
    for k in range(len(pgs)):
        for _ in range(10):
            X = rng.uniform(-3, 3, 3)
            num_u = P[k, 0, 3] + P[k, 0, :3] @ X
            num_v = P[k, 1, 3] + P[k, 1, :3] @ X
            den = P[k, 2, 3] + P[k, 2, :3] @ X
            # s + t (X - s) = d0 + u U + v V
            M = np.stack([X - fr["s"][k], -fr["U"][k], -fr["V"][k]], axis=1)
            t, u, v = np.linalg.solve(M, fr["d0"][k] - fr["s"][k])
            assert num_u / den == pytest.approx(u, rel=1e-11, abs=1e-11)
            assert num_v / den == pytest.approx(v, rel=1e-11, abs=1e-11)


@pytest.mark.parametrize("offset", (-3.0, 0.0, 4.5))
def test_bp_one_voxel_magnification_weight(offset):
    """Perpendicular detector, one voxel on the central ray at distance z
    from the source: BP of ones = vs^3 * (SDD / z)^2 / (pw * ph)."""
    vs, pw, ph, SOD, ODD = 0.5, 0.8, 1.1, 20.0, 12.0
    g = ProjectionGeometry([-SOD, 0, 0], [ODD, 0, 0], Detector(5, 5, ph, pw))
    lo = np.array([offset - vs / 2, -vs / 2, -vs / 2])
    vg = VolumeGeometry((1, 1, 1), [vs] * 3, lo, lo + vs)
    v = oracle.cone_bp(np.ones((1, 5, 5)), [g], vg)
    z = SOD + offset
    expect = vs ** 3 * ((SOD + ODD) / z) ** 2 / (pw * ph)
    assert v[0, 0, 0] == pytest.approx(expect, rel=1e-12)


def test_bp_linearity():
    pgs = _tilted_geometries(3, seed=9)
    vg = resolve_volume_geometry(shape=(6, 7, 5), voxel_size=0.9)
    rng = np.random.default_rng(11)
    a, b = rng.uniform(size=(2, 3, 11, 13))
    np.testing.assert_allclose(
        oracle.cone_bp(a - 0.5 * b, pgs, vg),
        oracle.cone_bp(a, pgs, vg) - 0.5 * oracle.cone_bp(b, pgs, vg),
        rtol=1e-12, atol=1e-12)


# ------------------------------------------------ closed forms: transpose

@pytest.mark.parametrize("seed", range(3))
def test_fp_transpose_dot_product(seed):
    """<A x, y> = <x, A^T y> for the exact transpose of the Joseph FP."""
    pgs = _tilted_geometries(4, seed=seed)
    pgs += [rotate(pgs[0], pitch=np.pi / 2)]      # a z-major angle
    vg = VolumeGeometry((8, 7, 6), [0.9] * 3, [-3.6, -3.1, -2.7],
                        [3.6, 3.2, 2.7])
    rng = np.random.default_rng(seed)
    x = rng.uniform(size=vg.shape)
    y = rng.uniform(size=(5, 11, 13))
    lhs = float(np.vdot(oracle.cone_fp(x, pgs, vg), y))
    rhs = float(np.vdot(x, oracle.cone_fp_transpose(y, pgs, vg)))
    assert lhs == pytest.approx(rhs, rel=1e-12)


# ---------------------------------------------------- closed forms: filter

def _kak_slaney(num):
    """Spatial Ram-Lak kernel on the circular distance n: 1/4 at 0,
    -1/(pi n)^2 at odd n, 0 at even n."""
    n = np.minimum(np.arange(num), num - np.arange(num))
    g = np.where(n % 2 == 1, -1.0 / (np.pi * np.maximum(n, 1)) ** 2, 0.0)
    g[0] = 0.25
    return g


@pytest.mark.parametrize("cols", (5, 31, 32, 33, 100))
def test_filter_is_spatial_convolution(cols):
    """The Ram-Lak row filter equals zero-padded circular convolution with
    4 g (g real and even, so H = 4 Re DFT(g) is the DFT of 4 g)."""
    N = oracle.padded_length(cols)
    assert N == max(64, 1 << int(np.ceil(np.log2(2 * cols))))
    rng = np.random.default_rng(cols)
    p = rng.uniform(size=(2, 3, cols))
    g4 = 4 * _kak_slaney(N)
    direct = np.zeros_like(p)
    for j in range(cols):
        for m in range(cols):
            direct[..., j] += p[..., m] * g4[(j - m) % N]
    np.testing.assert_allclose(oracle.filter_rows(p, "ramlak"), direct,
                               rtol=1e-12, atol=1e-13)


def test_filter_response_shape():
    N = 2048
    H = oracle.filter_response(N, "ramlak")
    k = np.arange(N)
    kk_ = np.minimum(k, N - k)
    # a ramp 4|k|/N up to the Nyquist value 2, with a small DC lift
    np.testing.assert_allclose(H[64:N // 2], 4 * kk_[64:N // 2] / N,
                               atol=2e-3)
    # H[0] = 4 (1/4 - S) and H[N/2] = 4 (1/4 + S) with S the truncated odd
    # series sum 2 sum 1/(pi n)^2 -> 1/4; its tail is ~ 2 / (pi^2 N)
    assert H[0] + H[N // 2] == pytest.approx(2.0, rel=1e-12)
    assert H[0] == pytest.approx(8 / (np.pi ** 2 * N), rel=1e-3)
    hann = oracle.filter_response(N, "hann")
    assert abs(hann[N // 2]) < 1e-12 and hann[0] == pytest.approx(H[0])
    cos = oracle.filter_response(N, "cosine")
    assert abs(cos[N // 2]) < 1e-12
    assert cos[N // 4] == pytest.approx(H[N // 4] * np.sin(np.pi / 4),
                                        rel=1e-12)
    for name in ("ramp", "hamming", "shepp-logan"):
        assert np.all(np.isfinite(oracle.filter_response(N, name)))
    with pytest.raises(ValueError):
        oracle.filter_response(N, "nope")


@pytest.mark.parametrize("name", ("ramlak", "ramp", "cosine", "hann",
                                  "hamming", "shepp-logan"))
@pytest.mark.parametrize("num", (64, 512, 2048))
def test_product_filter_response_matches_oracle(name, num):
    from kernelkit.processing import _filter
    np.testing.assert_allclose(_filter(num, name), oracle.filter_response(
        num, name), rtol=1e-13, atol=1e-13)


# ------------------------------------------------- closed forms: preweight

def test_preweight_cosine_law():
    """Perpendicular detector at distance D from the source: the weight of
    pixel corner (r, c) is D / sqrt(D^2 + x^2 + y^2)."""
    rows, cols, ph, pw, SOD, ODD = 6, 9, 1.3, 0.7, 30.0, 20.0
    g = rotate(ProjectionGeometry([-SOD, 0, 0], [ODD, 0, 0],
                                  Detector(rows, cols, ph, pw)), yaw=0.7)
    w = oracle.preweight(np.ones((1, rows, cols)), [g])[0]
    D = SOD + ODD
    y = (np.arange(rows)[:, None] - rows / 2) * ph
    x = (np.arange(cols)[None, :] - cols / 2) * pw
    np.testing.assert_allclose(w, D / np.sqrt(D * D + x * x + y * y),
                               rtol=1e-13)


# -------------------------------------------------------------- FDK / SIRT

def test_fdk_is_bp_of_filtered_preweighted():
    pgs = workloads.circular_geometry(6, sod=40.0, odd=20.0, rows=7,
                                         cols=9, pixel=1.2)
    vg = resolve_volume_geometry(shape=(6, 6, 5), voxel_size=1.0)
    y = np.random.default_rng(0).uniform(size=(6, 7, 9))
    expect = oracle.cone_bp(oracle.filter_rows(oracle.preweight(y, pgs),
                                               "hann"), pgs, vg)
    np.testing.assert_allclose(oracle.fdk(y, pgs, vg, filter="hann"), expect,
                               rtol=1e-13, atol=1e-13)


def test_sirt_residual_decreases_and_constraints():
    pgs = workloads.circular_geometry(12, sod=40.0, odd=20.0, rows=10,
                                         cols=12, pixel=1.2)
    vg = resolve_volume_geometry(shape=(8, 8, 6), voxel_size=1.0)
    x_true = workloads.ellipsoid_phantom(vg.shape, n=3, seed=1)
    y = oracle.cone_fp(x_true.astype(np.float64), pgs, vg)
    norms = []
    oracle.sirt(y, pgs, vg, iters=8,
                callback=lambda i, x, r: norms.append(np.linalg.norm(r)))
    assert all(b < a for a, b in zip(norms, norms[1:]))
    # zero data is a fixed point; constraints clip; the mask zeroes
    assert not oracle.sirt(np.zeros_like(y), pgs, vg, iters=3).any()
    x = oracle.sirt(y, pgs, vg, iters=3, min_constraint=0.0,
                    max_constraint=0.05)
    assert x.min() >= 0.0 and x.max() <= 0.05
    mask = np.zeros(vg.shape)
    mask[2:6, 2:6, 1:5] = 1.0
    x = oracle.sirt(y, pgs, vg, iters=3, mask=mask)
    assert not x[mask == 0].any()



@pytest.mark.parametrize("name", ("hann", "hamming", "cosine",
                                  "shepp-logan"))
def test_filter_even_part_suffices(name):
    """Re(IDFT(DFT(p) H)) = IDFT(DFT(p) He) with He[k] = (H[k] + H[-k]) / 2
    for real rows; the device filter keeps only He (fftshift(hanning(N)) is
    not even), which lets it transform two real rows at once."""
    cols = 37
    N = oracle.padded_length(cols)
    H = oracle.filter_response(N, name)
    He = 0.5 * (H + np.roll(H[::-1], 1))
    p = np.random.default_rng(1).uniform(size=(4, cols))
    q = np.fft.ifft(np.fft.fft(p, n=N, axis=-1) * He, axis=-1)
    assert np.abs(q.imag).max() < 1e-14
    np.testing.assert_allclose(q.real[:, :cols], oracle.filter_rows(p, name),
                               rtol=1e-12, atol=1e-14)
    # two rows packed as one complex row give both results at once
    z = np.fft.ifft(np.fft.fft(p[0] + 1j * p[1], n=N) * He)[:cols]
    np.testing.assert_allclose(z.real, q.real[0, :cols], atol=1e-14)
    np.testing.assert_allclose(z.imag, q.real[1, :cols], atol=1e-14)
