"""Write ``small_config_sirt.npz``: 20 SIRT iterations of the small
workload, from the FP64 oracle.

Workload: 128^3 volume (voxel size 1) holding the seeded 10-ellipsoid
phantom (seed 0), circular cone orbit of 180 angles over 360 degrees,
source-origin 512, origin-detector 512, 160 x 160 detector of 1.5-unit
pixels; measurements ``y = FP(phantom)`` from the oracle.  The iterations
follow ``oracle.sirt`` (``kernelkit/algo.py:283-365``) with every forward
projection and backprojection split over angle subsets in a process pool
(both are exactly linear in the angles).

At about 20 CPU-minutes this is too slow for a test, so the fixture keeps
what a GPU run of the same workload is compared against:

* five full planes of ``x_20`` (z = 20, 64, 100; x = 64; y = 40), float32;
* the plane sums of ``x_20`` along each axis and the sum of squares of
  every z-plane (float64: they cover every voxel);
* the norm of the weighted residual ``R (FP(x_i) - y)`` of every iteration.

The test regenerates ``y`` with the oracle, which is deterministic.  Run
from the repository root (minutes, all cores):

    python tests/golden/make_small_config.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "astra-kernelkit_amd"),
          os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

ITERS = 20
PLANES = {"z20": (slice(None), slice(None), 20),
          "z64": (slice(None), slice(None), 64),
          "z100": (slice(None), slice(None), 100),
          "x64": (64, slice(None), slice(None)),
          "y40": (slice(None), 40, slice(None))}


def workload():
    from kernelkit import workloads
    pgs, vg = workloads.config_geometry("small")
    vol = workloads.ellipsoid_phantom(vg.shape, n=10, seed=0)
    return pgs, vg, vol


def _recip(a):
    with np.errstate(divide="ignore"):
        r = 1.0 / a
    r[r == np.inf] = 0.0
    return r


def main(path=os.path.join(HERE, "small_config_sirt.npz")):
    import oracle_pool as op
    pgs, vg, vol = workload()
    t0 = time.time()
    y = op.cone_fp(vol, pgs, vg)
    shape = tuple(int(n) for n in vg.shape)
    C = _recip(op.cone_bp(np.ones_like(y), pgs, vg))
    R = _recip(op.cone_fp(np.ones(shape), pgs, vg))
    x = np.zeros(shape)
    res_norms = []
    for i in range(ITERS):
        res = (op.cone_fp(x, pgs, vg) - y) * R
        res_norms.append(float(np.linalg.norm(res)))
        x = x - C * op.cone_bp(res, pgs, vg)
        print(f"iteration {i}: |res| {res_norms[-1]:.6e} "
              f"({time.time() - t0:.0f} s)", flush=True)
    op.close()
    data = {f"plane_{k}": x[s].astype(np.float32) for k, s in PLANES.items()}
    data.update(
        iters=ITERS,
        sum_yz=x.sum(axis=(1, 2)), sum_xz=x.sum(axis=(0, 2)),
        sum_xy=x.sum(axis=(0, 1)), sumsq_z=(x * x).sum(axis=(0, 1)),
        res_norms=np.array(res_norms),
        y_sum=float(y.sum()), y_sumsq=float((y * y).sum()))
    np.savez_compressed(path, **data)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
