"""The CPU baseline's process-pool restatement of FP, BP and SIRT adds up
to the oracle's own results (FP and BP are sums over angles), so its
timings time the oracle's algorithm."""
import numpy as np
import pytest

import oracle

bench = pytest.importorskip("bench")
kk = pytest.importorskip("kernelkit")
from kernelkit import workloads  # noqa: E402


def test_pool_matches_oracle():
    pgs = workloads.circular_geometry(7, sod=60.0, odd=30.0, rows=9,
                                      cols=11, pixel=1.5, tilt_deg=2.0,
                                      seed=2)
    vg = kk.resolve_volume_geometry(shape=(8, 9, 10), voxel_size=1.0)
    vol = workloads.ellipsoid_phantom(vg.shape, n=3, seed=1).astype(
        np.float64)
    y = oracle.cone_fp(vol, pgs, vg)
    pool = bench._SmallPool(pgs, vg, 3)
    try:
        assert len(pool.chunks) == 3
        np.testing.assert_allclose(pool.fp(vol), y, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(pool.bp(y), oracle.cone_bp(y, pgs, vg),
                                   rtol=1e-12, atol=1e-9)
        t, x = bench._small_run(pool, vol, y, 2)
    finally:
        pool.close()
    assert set(t) == {"fp", "bp", "fdk", "sirt"}
    np.testing.assert_allclose(x, oracle.sirt(y, pgs, vg, iters=2),
                               rtol=1e-10, atol=1e-12)
