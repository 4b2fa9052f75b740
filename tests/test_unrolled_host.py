"""``torch_support.unrolled_reconstruction`` on the CPU with a dense
matrix standing in for the X-ray transform (its exact transpose is the
matrix transpose): step-size gradients against central finite differences,
checkpointing changes nothing, and the callback sees every iterate.  No
GPU: the host logic of the unrolled iteration only."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
from kernelkit.torch_support import unrolled_reconstruction  # noqa: E402


class _DenseOp:
    """Operator protocol used by unrolled_reconstruction (call with out=,
    matched_T, domain/range shapes) over a dense float64 matrix."""

    def __init__(self, A, T=None):
        self.A = A
        self._T = T

    @property
    def domain_shape(self):
        return (self.A.shape[1],)

    @property
    def range_shape(self):
        return (self.A.shape[0],)

    def __call__(self, x, out=None, additive=False):
        r = (self.A @ x.double()).to(x.dtype)
        if out is None:
            return r
        out.copy_(r)
        return out

    @property
    def matched_T(self):
        if self._T is None:
            self._T = _DenseOp(self.A.T.contiguous(), self)
        return self._T


@pytest.fixture
def problem():
    g = torch.Generator().manual_seed(1)
    A = torch.rand(40, 25, generator=g, dtype=torch.float64)
    xt = torch.rand(25, generator=g, dtype=torch.float64)
    return _DenseOp(A), (A @ xt).float(), xt.float()


@pytest.mark.parametrize("weighting", ["sirt", "none"])
def test_step_gradients_match_finite_differences(problem, weighting):
    op, y, xt = problem
    steps0 = torch.tensor([1.0, 0.9, 0.7, 1.1, 0.8], dtype=torch.float64)
    if weighting == "none":
        steps0 = steps0 * 1e-3

    def loss(steps, ckpt=True):
        x = unrolled_reconstruction(y, op, iters=5, step=steps,
                                    weighting=weighting, checkpoint=ckpt)
        return ((x.double() - xt.double()) ** 2).sum()

    s = steps0.clone().requires_grad_(True)
    loss(s).backward()
    s2 = steps0.clone().requires_grad_(True)
    loss(s2, ckpt=False).backward()
    np.testing.assert_allclose(s.grad.numpy(), s2.grad.numpy(), rtol=1e-6)
    h = 1e-4 * float(steps0.abs().max())
    fd = []
    for k in range(5):
        e = torch.zeros(5, dtype=torch.float64)
        e[k] = h
        with torch.no_grad():
            fd.append((float(loss(steps0 + e)) - float(loss(steps0 - e)))
                      / (2 * h))
    np.testing.assert_allclose(s.grad.numpy(), fd, rtol=2e-2,
                               atol=1e-3 * max(abs(v) for v in fd))
    # distinct steps give distinct gradients (equal steps would not: the
    # iteration maps commute)
    assert len({round(float(v), 6) for v in s.grad}) > 1


def test_callback_sees_every_iterate(problem):
    op, y, _ = problem
    seen = []
    x = unrolled_reconstruction(y, op, iters=3, callback=lambda k, xk:
                                seen.append((k, xk.detach().clone())))
    assert [k for k, _ in seen] == [0, 1, 2]
    assert torch.equal(seen[-1][1], x.detach())
