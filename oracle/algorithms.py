"""FP64 FDK and SIRT for the oracle (test infrastructure only).

* ``fdk``  restates ``kernelkit/algo.py:94-126`` and ``:210-223``: pre-weight,
  row filter, adjoint-scaled voxel-driven BP into a zero volume.  No FDK
  distance weighting and no pi/N normalisation, as in the reference.
* ``sirt`` restates ``kernelkit/algo.py:283-365``.
"""
from __future__ import annotations

import numpy as np

from oracle.processing import filter_rows, preweight
from oracle.projectors import cone_bp, cone_fp


def fdk(proj: np.ndarray, pgs, vg, filter: str = "ramlak") -> np.ndarray:
    pgs = list(pgs)
    y = filter_rows(preweight(proj, pgs), filter)
    return cone_bp(y, pgs, vg)


def _safe_reciprocal(a: np.ndarray) -> np.ndarray:
    """``1/a`` with ``+inf -> 0`` (``algo.py:324-325``, ``:331-332``)."""
    with np.errstate(divide="ignore"):
        r = 1.0 / a
    r[r == np.inf] = 0.0
    return r


def sirt(proj: np.ndarray, pgs, vg, iters: int = 100, mask=None,
         min_constraint=None, max_constraint=None, callback=None):
    """``x <- clip(mask * (x - C * BP(R * (FP(x) - y))))`` from ``x0 = 0``,
    with ``C = 1/BP(1)`` and ``R = 1/FP(1)``."""
    pgs = list(pgs)
    y = np.asarray(proj, dtype=np.float64)
    shape = tuple(int(n) for n in vg.shape)
    x = np.zeros(shape)
    C = _safe_reciprocal(cone_bp(np.ones_like(y), pgs, vg))
    R = _safe_reciprocal(cone_fp(np.ones(shape), pgs, vg))
    m = None if mask is None else np.asarray(mask, dtype=np.float64)
    for i in range(iters):
        res = (cone_fp(x, pgs, vg) - y) * R
        x = x - C * cone_bp(res, pgs, vg)
        if m is not None:
            x = x * m
        if min_constraint is not None or max_constraint is not None:
            x = np.clip(x, min_constraint, max_constraint)
        if callback is not None:
            callback(i, x, res)
    return x
