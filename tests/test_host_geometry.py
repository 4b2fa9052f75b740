"""Host-side geometry of the product against the oracle (CPU): the kernel
parameter tables built by ``geoms2params`` equal the oracle's normalised
frames and affine forms in float64, the caller's geometry is not mutated,
and the projector/operator plumbing validates like the reference.
"""
import copy

import numpy as np
import pytest

import oracle

kk = pytest.importorskip("kernelkit")
from kernelkit import workloads  # noqa: E402
from kernelkit.kernels.cone_bp import VoxelDrivenConeBP  # noqa: E402
from kernelkit.kernels.cone_fp import RayDrivenConeFP  # noqa: E402


def _case():
    pgs = workloads.circular_geometry(7, sod=90.0, odd=45.0, rows=12,
                                      cols=14, pixel=1.3, seed=4,
                                      tilt_deg=2.0, drift_px=3.0,
                                      wobble_deg=1.0)
    vg = kk.VolumeGeometry((10, 12, 9), [0.9] * 3, [-4.0, -6.1, -3.3],
                           [5.0, 4.7, 4.8], rotation=(0.03, -0.02, 0.2))
    return pgs, vg


def test_fp_params_match_oracle_frames():
    pgs, vg = _case()
    seq = kk.ProjectionGeometrySequence.fromList(pgs)
    before = copy.deepcopy(seq)
    P = RayDrivenConeFP.geoms2params(seq, vg)
    fr = oracle.normalized_frames(pgs, vg)
    got = np.stack(P[:12], axis=1)
    want = np.concatenate([fr["s"], fr["d0"], fr["U"], fr["V"]], axis=1)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(P.ray_axes, oracle.fp_major_axes(fr))
    # the caller's sequence is left as it was
    np.testing.assert_array_equal(seq.source_position, before.source_position)
    np.testing.assert_array_equal(seq.u, before.u)


def test_bp_params_match_oracle_forms():
    pgs, vg = _case()
    P = np.asarray(VoxelDrivenConeBP.geoms2params(pgs, vg)).reshape(-1, 3, 4)
    fr = oracle.normalized_frames(pgs, vg)
    want = oracle.bp_params(fr, vg.voxel_size, with_adjoint_scaling=True)
    np.testing.assert_allclose(P, want, rtol=1e-12,
                               atol=1e-12 * np.abs(want).max())
    P1 = np.asarray(VoxelDrivenConeBP.geoms2params(
        pgs, vg, with_adjoint_scaling=False)).reshape(-1, 3, 4)
    want1 = oracle.bp_params(fr, vg.voxel_size, with_adjoint_scaling=False)
    np.testing.assert_allclose(P1, want1, rtol=1e-12,
                               atol=1e-12 * np.abs(want1).max())


def test_geometry_sequence_roundtrip():
    pgs, _ = _case()
    seq = kk.ProjectionGeometrySequence.fromList(pgs)
    assert len(seq) == len(pgs)
    for k in (0, 3, 6):
        g = seq.geometry(k)
        np.testing.assert_allclose(g.source_position, pgs[k].source_position)
        np.testing.assert_allclose(g.u, pgs[k].u)
        assert g.detector.rows == pgs[k].detector.rows
        sub = seq[k]                       # take(): a one-angle sequence
        assert len(sub) == 1
        np.testing.assert_allclose(sub.source_position[0],
                                   pgs[k].source_position)
    sub = seq.take([1, 4])
    assert len(sub) == 2
    np.testing.assert_allclose(sub.v[1], pgs[4].v)


def test_unit_vector_validation():
    g = workloads.circular_geometry(1, sod=10.0, odd=5.0, rows=2, cols=2,
                                    pixel=1.0)[0]
    with pytest.raises(ValueError):
        g.u = (0.0, 2.0, 0.0)


def test_fdk_rejects_non_cone():
    pgs, vg = _case()
    g = copy.deepcopy(pgs[0])
    g.beam = kk.Beam.PARALLEL
    with pytest.raises(NotImplementedError):
        kk.fdk(np.zeros((1, 12, 14), np.float32), [g], vg)


def test_fbp_not_implemented():
    pgs, vg = _case()
    with pytest.raises(NotImplementedError):
        kk.fbp(np.zeros((7, 12, 14), np.float32), pgs, vg)


def test_sirt_shape_mismatch():
    pgs, vg = _case()
    with pytest.raises(ValueError):
        kk.sirt(np.zeros((3, 12, 14), np.float32), pgs, vg, iters=1,
                verbose=False)


def test_filter_names_and_padding():
    from kernelkit.processing import _filter, padded_length
    assert padded_length(768) == 2048
    assert padded_length(10) == 64
    assert padded_length(1536) == 4096
    with pytest.raises(ValueError):
        _filter(64, "nope")
    with pytest.raises(ValueError):
        _filter(63)


def test_device_phantom_matches_host_phantom():
    """The device phantom builder (used for the 1024^3 workload) evaluates
    the same ellipsoids as the host builder."""
    import torch
    from kernelkit import workloads
    for shape, seed in (((40, 48, 36), 3), ((33, 20, 25), 7)):
        host = workloads.ellipsoid_phantom(shape, n=6, seed=seed)
        dev = workloads.ellipsoid_phantom_device(shape, n=6, seed=seed,
                                                 device="cpu").numpy()
        assert host.sum() > 0
        np.testing.assert_array_equal(dev, host)


def test_preweight_table_reproduces_the_oracle_weights():
    """``processing.preweight_params`` (the per-angle table the pre-weight
    kernel reads: corner ``d_min - s``, pixel steps ``U``, ``V`` and the
    central ray length) gives, for every pixel, the oracle's cosine weight
    ``|d - s| / |corner(r, c) - s|`` on a wobbling, tilted, drifting orbit;
    with piercing points the central length is ``|p - s|`` per angle."""
    from kernelkit.processing import preweight_params
    rows, cols = 9, 13
    pgs = workloads.circular_geometry(11, sod=120.0, odd=60.0, rows=rows,
                                      cols=cols, pixel=1.3, wobble_deg=2.0,
                                      drift_px=1.5, tilt_deg=3.0, seed=8)
    P = preweight_params(pgs)
    assert P.shape == (11, 10) and P.dtype == np.float64
    rr = np.arange(rows, dtype=np.float64)[:, None, None]
    cc = np.arange(cols, dtype=np.float64)[None, :, None]
    want = oracle.preweight(np.ones((11, rows, cols)), pgs)
    for k in range(11):
        ray = P[k, 0:3] + cc * P[k, 3:6] + rr * P[k, 6:9]
        got = P[k, 9] / np.linalg.norm(ray, axis=2)
        np.testing.assert_allclose(got, want[k], rtol=1e-14, atol=0)
    pier = [np.asarray(g.detector_position) + (0.5 * k, -0.25 * k, 0.1)
            for k, g in enumerate(pgs)]
    Pp = preweight_params(pgs, pier)
    np.testing.assert_array_equal(Pp[:, :9], P[:, :9])
    lens = [np.linalg.norm(p - np.asarray(g.source_position))
            for p, g in zip(pier, pgs)]
    np.testing.assert_allclose(Pp[:, 9], lens, rtol=1e-15, atol=0)
