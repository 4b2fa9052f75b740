"""The exact restrictions used by the large-workload GPU tests, checked on
the CPU with the oracle on small problems: a row band of the forward
projection, a voxel block of the backprojection, of FDK and of the exact
transpose each equal the same part of the full result."""
import numpy as np
import pytest

import oracle
import restrict

kk = pytest.importorskip("kernelkit")
from kernelkit import workloads  # noqa: E402


@pytest.fixture(scope="module")
def problem():
    pgs = workloads.circular_geometry(7, sod=60.0, odd=40.0, rows=22,
                                      cols=26, pixel=1.3, wobble_deg=1.0,
                                      drift_px=2.0, tilt_deg=1.5, seed=4)
    vg = kk.resolve_volume_geometry(shape=(18, 16, 14), voxel_size=1.0)
    rng = np.random.default_rng(7)
    return pgs, vg, rng.uniform(size=vg.shape), rng.uniform(size=(7, 22, 26))


def test_row_band_forward_projection(problem):
    pgs, vg, vol, _ = problem
    full = oracle.cone_fp(vol, pgs, vg)
    for r0, nr in ((0, 3), (9, 4), (19, 3)):
        band = oracle.cone_fp(vol, [restrict.row_band(g, r0, nr)
                                    for g in pgs], vg)
        np.testing.assert_allclose(band, full[:, r0:r0 + nr], rtol=1e-12,
                                   atol=1e-12 * np.abs(full).max())


@pytest.mark.parametrize("lo,n", [((0, 0, 0), (5, 4, 3)),
                                  ((7, 6, 5), (4, 4, 2)),
                                  ((13, 11, 11), (5, 5, 3))])
def test_block_backprojection_and_fdk(problem, lo, n):
    pgs, vg, _, proj = problem
    sub = restrict.sub_volume(vg, lo, n)
    blk = tuple(slice(a, a + k) for a, k in zip(lo, n))
    full = oracle.cone_bp(proj, pgs, vg)
    nr, nc, org = restrict.window(pgs, vg, lo, n)
    assert nr < 22 or nc < 26       # the window is a real restriction
    geoms, part = restrict.windowed(pgs, proj, nr, nc, org)
    got = oracle.cone_bp(part, geoms, sub)
    np.testing.assert_allclose(got, full[blk], rtol=1e-12,
                               atol=1e-12 * np.abs(full).max())
    # FDK filters whole rows: restrict rows only
    full_fdk = oracle.fdk(proj, pgs, vg, filter="hann")
    geoms, part = restrict.windowed(pgs, proj, nr, 26,
                                    [(r0, 0) for r0, _ in org])
    got = restrict.fdk(part, pgs, geoms, sub, filter="hann")
    np.testing.assert_allclose(got, full_fdk[blk], rtol=1e-12,
                               atol=1e-12 * np.abs(full_fdk).max())


@pytest.mark.parametrize("lo,n", [((0, 0, 0), (5, 4, 3)),
                                  ((6, 5, 6), (5, 5, 2)),
                                  ((14, 12, 11), (4, 4, 3))])
def test_block_transpose(problem, lo, n):
    pgs, vg, _, proj = problem
    sub = restrict.sub_volume(vg, lo, n)
    blk = tuple(slice(a, a + k) for a, k in zip(lo, n))
    full = oracle.cone_fp_transpose(proj, pgs, vg)
    nr, nc, org = restrict.window(pgs, vg, lo, n, grow=1.5, margin=1)
    geoms, part = restrict.windowed(pgs, proj, nr, nc, org)
    got = oracle.cone_fp_transpose(part, geoms, sub,
                                   axes=restrict.ray_axes(pgs, vg))
    np.testing.assert_allclose(got, full[blk], rtol=1e-12,
                               atol=1e-12 * np.abs(full).max())


def test_window_keeps_the_full_detectors_ray_axes():
    """Near 45 degrees a window at a detector corner has its own centre ray
    on the other side of the diagonal: left to itself the oracle would pick
    the other ray axis there (a different sampling).  With the full
    detector's axes the windowed forward projection and transpose equal the
    full ones; without them they differ."""
    from kernelkit import workloads
    pgs = workloads.circular_geometry(8, sod=60.0, odd=60.0, rows=20,
                                      cols=60, pixel=1.0)
    pgs = [pgs[1], pgs[3]]                        # 45 and 135 degrees
    pgs = [kk.rotate(g, yaw=np.deg2rad(-3.0)) for g in pgs]
    vg = kk.resolve_volume_geometry(shape=(24, 24, 8), voxel_size=1.0)
    rng = np.random.default_rng(1)
    vol = rng.uniform(size=vg.shape)
    proj = rng.uniform(size=(2, 20, 60))
    full_fp = oracle.cone_fp(vol, pgs, vg)
    full_t = oracle.cone_fp_transpose(proj, pgs, vg)
    lo, n = (0, 0, 3), (4, 4, 2)
    nr, nc, org = restrict.window(pgs, vg, lo, n, grow=1.5, margin=1)
    geoms, part = restrict.windowed(pgs, proj, nr, nc, org)
    axes = restrict.ray_axes(pgs, vg)
    assert (restrict.ray_axes(geoms, vg) != axes).any()   # the trap is set
    blk = tuple(slice(a, a + k) for a, k in zip(lo, n))
    sub = restrict.sub_volume(vg, lo, n)
    got = oracle.cone_fp_transpose(part, geoms, sub, axes=axes)
    np.testing.assert_allclose(got, full_t[blk], rtol=1e-12,
                               atol=1e-12 * np.abs(full_t).max())
    bad = oracle.cone_fp_transpose(part, geoms, sub)
    assert np.abs(bad - full_t[blk]).max() > 1e-3 * np.abs(full_t).max()
    fp = oracle.cone_fp(vol, geoms, vg, axes=axes)
    ref = np.stack([full_fp[k, r0:r0 + nr, c0:c0 + nc]
                    for k, (r0, c0) in enumerate(org)])
    np.testing.assert_allclose(fp, ref, rtol=1e-12,
                               atol=1e-12 * np.abs(full_fp).max())


def test_window_too_small_is_detected(problem):
    """A window a few pixels short of the block's projection loses
    contributions: the check above would catch a wrong window."""
    pgs, vg, _, proj = problem
    lo, n = (7, 6, 5), (4, 4, 2)
    sub = restrict.sub_volume(vg, lo, n)
    full = oracle.cone_bp(proj, pgs, vg)[7:11, 6:10, 5:7]
    nr, nc, org = restrict.window(pgs, vg, lo, n, margin=-2)
    geoms, part = restrict.windowed(pgs, proj, nr, nc, org)
    got = oracle.cone_bp(part, geoms, sub)
    assert np.abs(got - full).max() > 1e-3 * np.abs(full).max()
