"""Golden fixtures (``tests/golden/cone_small.npz``, written by
``tests/golden/make_golden.py``): the oracle reproduces them (CPU), and the
HIP path matches them under the parity bar (GPU)."""
import os

import numpy as np
import pytest

import oracle
from conftest import assert_parity

kk = pytest.importorskip("kernelkit")

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "cone_small.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(PATH, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    rows, cols = int(g["rows"]), int(g["cols"])
    pgs = [kk.ProjectionGeometry(
        g["src"][k], g["det"][k],
        kk.Detector(rows, cols, float(g["pixel"][k, 0]),
                    float(g["pixel"][k, 1])),
        u=g["u"][k], v=g["v"][k]) for k in range(len(g["src"]))]
    vg = kk.VolumeGeometry(tuple(int(s) for s in g["shape"]),
                           g["voxel_size"], g["extent_min"], g["extent_max"],
                           rotation=g["rotation"])
    return g, pgs, vg


def test_oracle_reproduces_golden(golden):
    g, pgs, vg = golden
    close = dict(rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oracle.cone_fp(g["vol"], pgs, vg), g["fp"],
                               **close)
    np.testing.assert_allclose(oracle.cone_bp(g["y"], pgs, vg), g["bp"],
                               **close)
    np.testing.assert_allclose(oracle.cone_fp_transpose(g["y"], pgs, vg),
                               g["fp_transpose"], **close)
    np.testing.assert_allclose(oracle.filter_rows(g["y"], "ramlak"),
                               g["filter_ramlak"], **close)
    np.testing.assert_allclose(oracle.filter_rows(g["y"], "hann"),
                               g["filter_hann"], **close)
    np.testing.assert_allclose(oracle.preweight(g["y"], pgs), g["preweight"],
                               **close)
    np.testing.assert_allclose(oracle.fdk(g["fp"], pgs, vg), g["fdk"],
                               **close)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.mark.gpu
def test_hip_matches_golden(golden, device):
    g, pgs, vg = golden
    A = kk.XrayTransform(pgs, vg)
    assert_parity(A(_dev(g["vol"])).cpu().numpy(), g["fp"], label="fp")
    assert_parity(A.T(_dev(g["y"])).cpu().numpy(), g["bp"], label="bp")
    for name in ("ramlak", "hann"):
        t = _dev(g["y"])
        kk.processing.filter(t, filter=name)
        assert_parity(t.cpu().numpy(), g[f"filter_{name}"],
                      label=f"filter/{name}")
    t = _dev(g["y"])
    kk.processing.preweight(t, pgs)
    assert_parity(t.cpu().numpy(), g["preweight"], label="preweight")
    assert_parity(kk.fdk(g["fp"].astype(np.float32), pgs, vg), g["fdk"],
                  label="fdk")
