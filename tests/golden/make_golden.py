"""Write the golden fixtures ``cone_small.npz`` from the FP64 oracle.

The reference ships no data files for this path (its tests assert analytic
values), so these vectors come from the oracle, which the CPU tests pin
against the reference's known-answer tests and closed forms.  The fixture
holds the inputs (volume, projections, geometry as plain arrays) and the
FP64 outputs of the forward projection, backprojection, exact transpose,
row filters, pre-weight and FDK.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "astra-kernelkit_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle  # noqa: E402
from kernelkit import workloads, rotate, VolumeGeometry  # noqa: E402


def geometry():
    pgs = workloads.circular_geometry(4, sod=40.0, odd=25.0, rows=9, cols=10,
                                      pixel=1.1, seed=5, tilt_deg=2.0,
                                      drift_px=2.0, wobble_deg=1.0)
    pgs.append(rotate(pgs[1], pitch=np.pi / 2))     # a z-major angle
    vg = VolumeGeometry((8, 9, 7), [0.8] * 3, [-3.0, -3.9, -2.6],
                        [3.4, 3.3, 3.0], rotation=(0.02, -0.03, 0.1))
    return pgs, vg


def main(path=os.path.join(HERE, "cone_small.npz")):
    pgs, vg = geometry()
    rng = np.random.default_rng(2024)
    vol = rng.uniform(size=vg.shape)
    y = rng.uniform(size=(len(pgs), 9, 10))
    fp = oracle.cone_fp(vol, pgs, vg)
    data = dict(
        vol=vol, y=y,
        src=np.array([g.source_position for g in pgs]),
        det=np.array([g.detector_position for g in pgs]),
        u=np.array([g.u for g in pgs]), v=np.array([g.v for g in pgs]),
        rows=9, cols=10,
        pixel=np.array([[g.detector.pixel_height, g.detector.pixel_width]
                        for g in pgs]),
        shape=np.array(vg.shape), voxel_size=np.array(vg.voxel_size),
        extent_min=np.array(vg.extent_min), extent_max=np.array(vg.extent_max),
        rotation=np.array(vg.rotation),
        fp=fp,
        bp=oracle.cone_bp(y, pgs, vg),
        fp_transpose=oracle.cone_fp_transpose(y, pgs, vg),
        filter_ramlak=oracle.filter_rows(y, "ramlak"),
        filter_hann=oracle.filter_rows(y, "hann"),
        preweight=oracle.preweight(y, pgs),
        fdk=oracle.fdk(fp, pgs, vg, filter="ramlak"),
    )
    np.savez_compressed(path, **data)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
