"""GPU parity of the HIP forward projector and backprojector against the
FP64 oracle (relative L2 <= 1e-5, pointwise <= 1e-4 * max|ref|).

Covers circular and perturbed (cone_vec-like) orbits, every volume_axes and
projection_axes permutation, shifted/rotated/anisotropic volumes, the
``additive`` flag, and edge cases (rays missing the volume, one-pixel
detectors, one-voxel-thick volumes).
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


def _small_case(n_angles=24, shape=(40, 44, 36), rows=48, cols=56, seed=3,
                **pert):
    pgs = workloads.circular_geometry(n_angles, sod=120.0, odd=60.0,
                                      rows=rows, cols=cols, pixel=1.2,
                                      seed=seed, **pert)
    vg = kk.resolve_volume_geometry(shape=shape, voxel_size=1.0)
    vol = workloads.ellipsoid_phantom(shape, n=6, seed=seed)
    return pgs, vg, vol


CASES = {
    "circular": dict(),
    "tilted": dict(tilt_deg=2.0, drift_px=5.0),
    "wobble": dict(wobble_deg=1.0, drift_px=3.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp_matches_oracle(case, device):
    pgs, vg, vol = _small_case(**CASES[case])
    ref = oracle.cone_fp(vol, pgs, vg)
    A = kk.XrayTransform(pgs, vg)
    got = _host(A(_dev(vol)))
    assert_parity(got, ref, label=f"fp/{case}")


@pytest.mark.parametrize("case", sorted(CASES))
def test_bp_matches_oracle(case, device):
    pgs, vg, vol = _small_case(**CASES[case])
    proj = oracle.cone_fp(vol, pgs, vg)
    ref = oracle.cone_bp(proj, pgs, vg)
    A = kk.XrayTransform(pgs, vg)
    got = _host(A.T(_dev(proj)))
    assert_parity(got, ref, label=f"bp/{case}")


@pytest.mark.parametrize("shape", [(64, 64, 32), (64, 60, 40)])
def test_bp_workgroup_groups_match_oracle(shape, device):
    """On a circular orbit whose tile counts (16 x 16 columns, z-runs of 8)
    divide by 4 the backprojector deals its workgroups to the XCDs in
    groups of 4 x 4 x 4 tiles (64^3-class volumes here: 4 x 4 x 4 tiles);
    otherwise (64 x 60 x 40: 5 z tiles) in runs along x.  Both orders give
    the oracle's backprojection; the FP of the same case, 20 angles in
    groups of 16 (padding ids), does too."""
    pgs, vg, vol = _small_case(n_angles=20, shape=shape, rows=72, cols=80)
    proj = oracle.cone_fp(vol, pgs, vg)
    A = kk.XrayTransform(pgs, vg)
    assert_parity(_host(A(_dev(vol))), proj, label=f"fp/groups {shape}")
    ref = oracle.cone_bp(proj, pgs, vg)
    got = _host(A.T(_dev(proj)))
    assert_parity(got, ref, label=f"bp/groups {shape}")


def test_bp_noise_input(device):
    pgs, vg, _ = _small_case(n_angles=16, seed=5)
    rng = np.random.default_rng(7)
    proj = rng.uniform(0.0, 1.0, size=(16, 48, 56))
    ref = oracle.cone_bp(proj, pgs, vg)
    got = _host(kk.XrayTransform(pgs, vg).T(_dev(proj)))
    assert_parity(got, ref, label="bp/noise")


@pytest.mark.parametrize("axes", list(itertools.permutations((0, 1, 2))))
def test_volume_axes(axes, device):
    pgs, vg, vol = _small_case(n_angles=8, shape=(20, 24, 28), rows=32,
                               cols=36)
    ref_fp = oracle.cone_fp(vol, pgs, vg)
    A = kk.XrayTransform(pgs, vg, volume_axes=axes)
    got_fp = _host(A(_dev(np.transpose(vol, axes))))
    assert_parity(got_fp, ref_fp, label=f"fp/volume_axes={axes}")
    ref_bp = oracle.cone_bp(ref_fp, pgs, vg)
    got_bp = _host(A.T(_dev(ref_fp)))
    assert_parity(np.transpose(got_bp, np.argsort(axes)), ref_bp,
                  label=f"bp/volume_axes={axes}")


@pytest.mark.parametrize("axes", list(itertools.permutations((0, 1, 2))))
def test_projection_axes(axes, device):
    pgs, vg, vol = _small_case(n_angles=8, shape=(20, 24, 28), rows=32,
                               cols=36)
    ref_fp = oracle.cone_fp(vol, pgs, vg)
    A = kk.XrayTransform(pgs, vg, projection_axes=axes)
    got_fp = _host(A(_dev(vol)))
    assert_parity(np.transpose(got_fp, np.argsort(axes)), ref_fp,
                  label=f"fp/projection_axes={axes}")
    ref_bp = oracle.cone_bp(ref_fp, pgs, vg)
    got_bp = _host(A.T(_dev(np.transpose(ref_fp, axes))))
    assert_parity(got_bp, ref_bp, label=f"bp/projection_axes={axes}")


@pytest.fixture(scope="module")
def axis_matrix_case():
    """A perturbed orbit (per-angle detector tilt and shift) with its oracle
    FP, BP and exact transpose in world order, computed once."""
    pgs, vg, vol = _small_case(n_angles=8, shape=(20, 24, 28), rows=32,
                               cols=36, tilt_deg=2.0, drift_px=3.0)
    fp = oracle.cone_fp(vol, pgs, vg)
    bp = oracle.cone_bp(fp, pgs, vg)
    y = np.random.default_rng(9).uniform(size=fp.shape)
    fpt = oracle.cone_fp_transpose(y, pgs, vg)
    return pgs, vg, vol, fp, bp, y, fpt


@pytest.mark.parametrize("vaxes", list(itertools.permutations((0, 1, 2))))
@pytest.mark.parametrize("paxes", list(itertools.permutations((0, 1, 2))))
def test_axis_order_matrix(axis_matrix_case, vaxes, paxes, device):
    """Every (volume_axes, projection_axes) pair, ASTRA's (2, 1, 0) /
    (1, 0, 2) among them (the reference specialises its kernels per pair,
    kernels/cone_fp.py:43-52, cuda/cone_fp.cu:94-135, cuda/cone_bp.cu:89-91;
    here they are strides): FP, BP and the exact transpose."""
    pgs, vg, vol, fp, bp, y, fpt = axis_matrix_case
    A = kk.XrayTransform(pgs, vg, volume_axes=vaxes, projection_axes=paxes)
    got = _host(A(_dev(np.transpose(vol, vaxes))))
    assert_parity(np.transpose(got, np.argsort(paxes)), fp,
                  label=f"fp/{vaxes}/{paxes}")
    got = _host(A.T(_dev(np.transpose(fp, paxes))))
    assert_parity(np.transpose(got, np.argsort(vaxes)), bp,
                  label=f"bp/{vaxes}/{paxes}")
    got = _host(A.matched_T(_dev(np.transpose(y, paxes))))
    assert_parity(np.transpose(got, np.argsort(vaxes)), fpt,
                  label=f"fp_t/{vaxes}/{paxes}")


def test_shifted_rotated_volume(device):
    pgs, _, _ = _small_case(n_angles=12)
    shape = (30, 34, 26)
    vg = kk.VolumeGeometry(shape=shape, voxel_size=[0.9] * 3,
                           extent_min=[-12.5, -16.3, -11.0],
                           extent_max=[-12.5 + 27.0, -16.3 + 30.6,
                                       -11.0 + 23.4],
                           rotation=(0.05, -0.1, 0.3))
    vol = workloads.ellipsoid_phantom(shape, n=5, seed=11)
    ref = oracle.cone_fp(vol, pgs, vg)
    A = kk.XrayTransform(pgs, vg)
    assert_parity(_host(A(_dev(vol))), ref, label="fp/shifted-rotated")
    ref_bp = oracle.cone_bp(ref, pgs, vg)
    assert_parity(_host(A.T(_dev(ref))), ref_bp, label="bp/shifted-rotated")


def test_anisotropic_fp(device):
    pgs, _, _ = _small_case(n_angles=10)
    shape = (24, 30, 20)
    vg = kk.resolve_volume_geometry(shape=shape, voxel_size=[1.0, 0.7, 1.3])
    vol = workloads.ellipsoid_phantom(shape, n=4, seed=2)
    ref = oracle.cone_fp(vol, pgs, vg)
    got = _host(kk.XrayTransform(pgs, vg)(_dev(vol)))
    assert_parity(got, ref, label="fp/anisotropic")


def test_anisotropic_bp_raises(device):
    pgs, _, _ = _small_case(n_angles=4)
    vg = kk.resolve_volume_geometry(shape=(8, 8, 8),
                                    voxel_size=[1.0, 0.5, 1.5])
    A = kk.XrayTransform(pgs, vg)
    with pytest.raises(NotImplementedError):
        A.T(_dev(np.ones(A.range_shape)))


def test_additive_flag(device):
    torch = _torch()
    pgs, vg, vol = _small_case(n_angles=6)
    A = kk.XrayTransform(pgs, vg)
    x = _dev(vol)
    y = A(x)
    y2 = y.clone()
    A(x, out=y2, additive=True)
    np.testing.assert_allclose(_host(y2), 2 * _host(y), rtol=1e-6,
                               atol=1e-6 * float(y.abs().max()))
    v = A.T(y)
    v2 = v.clone()
    A.T(y, out=v2, additive=True)
    np.testing.assert_allclose(_host(v2), 2 * _host(v), rtol=1e-6,
                               atol=1e-6 * float(v.abs().max()))
    # non-additive overwrites garbage
    y3 = torch.full_like(y, 1e6)
    A(x, out=y3)
    np.testing.assert_array_equal(_host(y3), _host(y))


def test_rays_missing_volume(device):
    # detector much larger than the volume's shadow: outer rays miss it
    pgs = workloads.circular_geometry(5, sod=80.0, odd=40.0, rows=64,
                                      cols=72, pixel=2.0)
    vg = kk.resolve_volume_geometry(shape=(10, 12, 8), voxel_size=1.0)
    vol = np.random.default_rng(1).uniform(size=(10, 12, 8))
    ref = oracle.cone_fp(vol, pgs, vg)
    got = _host(kk.XrayTransform(pgs, vg)(_dev(vol)))
    assert_parity(got, ref, label="fp/miss")
    assert np.count_nonzero(ref == 0) > 0


def test_thin_volume_and_single_pixel(device):
    pgs = workloads.circular_geometry(7, sod=50.0, odd=50.0, rows=1, cols=1,
                                      pixel=0.5)
    vg = kk.resolve_volume_geometry(shape=(9, 9, 1), voxel_size=1.0)
    vol = np.random.default_rng(2).uniform(size=(9, 9, 1))
    A = kk.XrayTransform(pgs, vg)
    ref = oracle.cone_fp(vol, pgs, vg)
    assert_parity(_host(A(_dev(vol))), ref, label="fp/single-pixel")
    ref_bp = oracle.cone_bp(ref, pgs, vg)
    assert_parity(_host(A.T(_dev(ref))), ref_bp, label="bp/single-pixel")


def test_z_major_rays(device):
    """Rays along z (a top-down orbit) use the z frame texture."""
    g0 = kk.rotate(workloads.circular_geometry(
        1, sod=120.0, odd=60.0, rows=40, cols=36, pixel=1.2)[0],
        pitch=np.pi / 2)
    pgs = [kk.rotate(g0, yaw=t) for t in np.linspace(0, 2 * np.pi, 6, False)]
    vg = kk.resolve_volume_geometry(shape=(24, 28, 20), voxel_size=1.0)
    vol = workloads.ellipsoid_phantom(vg.shape, n=4, seed=9)
    A = kk.XrayTransform(pgs, vg)
    ref = oracle.cone_fp(vol, pgs, vg)
    assert_parity(_host(A(_dev(vol))), ref, label="fp/z-major")
    ref_bp = oracle.cone_bp(ref, pgs, vg)
    assert_parity(_host(A.T(_dev(ref))), ref_bp, label="bp/z-major")


def test_mixed_ray_axes(device):
    """Angles of all three ray axes in one operator (three frames)."""
    base = workloads.circular_geometry(8, sod=120.0, odd=60.0, rows=40,
                                       cols=36, pixel=1.2)
    pgs = base + [kk.rotate(g, pitch=np.pi / 2) for g in base[:3]]
    vg = kk.resolve_volume_geometry(shape=(24, 28, 20), voxel_size=1.0)
    vol = workloads.ellipsoid_phantom(vg.shape, n=4, seed=10)
    A = kk.XrayTransform(pgs, vg)
    assert A.projector.kernel.frame_mask == 0b111
    ref = oracle.cone_fp(vol, pgs, vg)
    assert_parity(_host(A(_dev(vol))), ref, label="fp/mixed")


def test_volume_set_before_geometry(device):
    """The texture is rebuilt when the geometry needs frames it lacks."""
    g0 = kk.rotate(workloads.circular_geometry(
        1, sod=120.0, odd=60.0, rows=40, cols=36, pixel=1.2)[0],
        pitch=np.pi / 2)
    vg = kk.resolve_volume_geometry(shape=(24, 28, 20), voxel_size=1.0)
    vol = workloads.ellipsoid_phantom(vg.shape, n=4, seed=11)
    fp = kk.ForwardProjector()
    fp.volume = _dev(vol)            # frames default to x and y
    fp.volume_geometry = vg
    fp.projection_geometry = [g0]    # needs the z frame
    out = _dev(np.zeros((1, 40, 36)))
    fp.projections = out
    fp()
    assert_parity(_host(out), oracle.cone_fp(vol, [g0], vg),
                  label="fp/late-geometry")
