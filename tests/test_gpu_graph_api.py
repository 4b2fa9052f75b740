"""The single-problem path under HIP graph capture, and the entry points of
the reference's algorithm and torch API not covered elsewhere.

* ``ptor.projections = y; ptor()`` of a ``BackProjector`` captured in a
  graph and replayed with new projections (``example/cuda/graph.py:31-58``),
  the same for a ``ForwardProjector`` (``ptor.volume = x; ptor()``), and a
  backprojection module captured in ``torch.cuda.graph``
  (``example/pytorch/pytorch_forward_graph.py:56-94``).  Every replay is
  compared with the eager result, with eager work and synchronisation
  between replays;
* ``kk.sirt(graph=True)`` equals the eager iterations bit for bit;
* ``kk.fp`` with host input and host, numpy ``out`` and device ``out``
  (``algo.py:18-49``); ``XrayBackprojection`` and its gradient
  (``torch_support.py:48-56``) against the oracle.
"""
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


def _case(n=18, shape=(30, 28, 26), rows=34, cols=40, seed=8, **kw):
    pgs = workloads.circular_geometry(n, sod=110.0, odd=70.0, rows=rows,
                                      cols=cols, pixel=1.2, seed=seed, **kw)
    vg = kk.resolve_volume_geometry(shape=shape, voxel_size=1.0)
    return pgs, vg


@pytest.mark.parametrize("pert", [{}, {"wobble_deg": 1.0, "drift_px": 2.0}])
def test_backprojector_graph_replays(pert, device):
    torch = _torch()
    pgs, vg = _case(**pert)
    rng = np.random.default_rng(1)
    projs = _dev(rng.uniform(size=(len(pgs), 34, 40)))
    out = torch.zeros(vg.shape, device="cuda")
    ptor = kk.BackProjector(texture_type="array")
    ptor.projection_geometry = pgs
    ptor.volume_geometry = vg
    ptor.projections = projs          # warm-up: texture and tables exist
    ptor.volume = out
    ptor()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        ptor.projections = projs      # the texture transfer is in the graph
        ptor()
    torch.cuda.current_stream().wait_stream(s)
    eager_op = kk.XrayTransform(pgs, vg)
    for k in range(3):
        new = rng.uniform(size=projs.shape)
        projs.copy_(_dev(new))
        out.fill_(-1.0)
        g.replay()
        ref = eager_op.T(projs)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), f"replay {k}"
        if k == 0:
            assert_parity(_host(out), oracle.cone_bp(new, pgs, vg),
                          label="graph bp vs oracle")


def test_forward_projector_graph_replays(device):
    torch = _torch()
    pgs, vg = _case(wobble_deg=1.0)
    x = torch.zeros(vg.shape, device="cuda")
    y = torch.zeros((len(pgs), 34, 40), device="cuda")
    ptor = kk.ForwardProjector()
    ptor.projection_geometry = pgs
    ptor.volume_geometry = vg
    ptor.volume = x
    ptor.projections = y
    ptor()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        ptor.volume = x
        ptor()
    torch.cuda.current_stream().wait_stream(s)
    op = kk.XrayTransform(pgs, vg)
    for k in range(3):
        vol = workloads.ellipsoid_phantom(vg.shape, n=4, seed=10 + k)
        x.copy_(_dev(vol))
        y.fill_(float("nan"))
        g.replay()
        ref = op(x)
        torch.cuda.synchronize()
        assert torch.equal(y, ref), f"replay {k}"
        if k == 0:
            assert_parity(_host(y), oracle.cone_fp(vol, pgs, vg),
                          label="graph fp vs oracle")


def test_backprojection_module_graph(device):
    torch = _torch()
    from kernelkit.torch_support import XrayBackprojection
    pgs, vg = _case(n=8, shape=(24, 24, 6), rows=12, cols=36)
    B = 3
    model = XrayBackprojection(pgs, vg)
    static_y = torch.ones(B, 1, len(pgs), 12, 36, device="cuda")
    static_x = torch.zeros(B, 1, *vg.shape, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):                       # warm-up
            static_x[...] = model(static_y)
    torch.cuda.current_stream().wait_stream(s)
    static_x.fill_(-1.0)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        static_x[...] = model(static_y)
    assert float(static_x.min()) == -1.0        # capture ran nothing
    rng = np.random.default_rng(4)
    for k in range(3):
        new = rng.normal(size=static_y.shape)
        static_y.copy_(_dev(new))
        g.replay()
        ref = model(static_y)
        torch.cuda.synchronize()
        assert torch.equal(static_x, ref), f"replay {k}"
    for b in range(B):
        assert_parity(_host(static_x[b, 0]),
                      oracle.cone_bp(new[b, 0], pgs, vg),
                      label=f"module bp {b}")


@pytest.mark.parametrize("opts", [{}, {"min_constraint": 0.0,
                                       "max_constraint": 0.7}])
def test_sirt_graph_equals_eager(opts, device):
    torch = _torch()
    pgs, vg = _case(n=24, shape=(32, 32, 24), rows=32, cols=44)
    vol = workloads.ellipsoid_phantom(vg.shape, n=5, seed=2)
    y = kk.XrayTransform(pgs, vg)(_dev(vol))
    seen = []
    a = kk.sirt(y, pgs, vg, iters=12, verbose=False, return_gpu=True,
                **opts)
    b = kk.sirt(y, pgs, vg, iters=12, verbose=False, return_gpu=True,
                graph=True, callback=lambda i, x, r: seen.append(
                    float(r.norm())), **opts)
    assert torch.equal(a, b)
    assert len(seen) == 12 and seen[-1] < seen[0]


def test_fp_returns(device):
    torch = _torch()
    pgs, vg = _case(n=10)
    vol = workloads.ellipsoid_phantom(vg.shape, n=5, seed=3)
    ref = oracle.cone_fp(vol, pgs, vg)
    # host in, host out (the reference returns a pinned host array)
    got = kk.fp(vol, pgs, vg)
    assert isinstance(got, np.ndarray)
    assert_parity(got, ref, label="kk.fp host")
    # numpy out: filled and returned
    out = np.full(ref.shape, np.nan, dtype=np.float32)
    r = kk.fp(vol, pgs, vg, out=out)
    assert r is out
    assert_parity(out, ref, label="kk.fp numpy out")
    # device out: filled and returned, from a device volume
    dout = torch.full(ref.shape, float("nan"), device="cuda")
    r = kk.fp(_dev(vol), pgs, vg, out=dout)
    assert r is dout
    assert_parity(_host(dout), ref, label="kk.fp device out")


def test_xray_backprojection_module_and_gradient(device):
    torch = _torch()
    from kernelkit.torch_support import XrayBackprojection
    pgs, vg = _case(n=12)
    rng = np.random.default_rng(9)
    y = rng.uniform(size=(2, 1, len(pgs), 34, 40))
    mod = XrayBackprojection(pgs, vg)
    yt = _dev(y).requires_grad_(True)
    x = mod(yt)
    assert tuple(x.shape) == (2, 1, *vg.shape)
    for b in range(2):
        assert_parity(_host(x[b, 0]), oracle.cone_bp(y[b, 0], pgs, vg),
                      label=f"XrayBackprojection {b}")
    # the gradient of <BP y, w> with respect to y is the operator's .T.T,
    # the forward projection (torch_support.py:27-34)
    w = rng.uniform(size=x.shape)
    (x * _dev(w)).sum().backward()
    for b in range(2):
        assert_parity(_host(yt.grad[b, 0]), oracle.cone_fp(w[b, 0], pgs, vg),
                      label=f"XrayBackprojection grad {b}")


def test_library_graph_capture_of_a_projector_call(device):
    """``_native.Graph.capture`` (the library's own stream capture, no
    torch.cuda.graph) records ``ptor.projections = y; ptor()`` of a
    backprojector; three replays with new projections, eager work and a
    synchronisation between them, each bit-equal to the eager call
    (``example/cuda/graph.py:46-51`` records the same call through CuPy)."""
    torch = _torch()
    from kernelkit import _native
    pgs, vg = _case(wobble_deg=1.0, drift_px=2.0)
    rng = np.random.default_rng(4)
    projs = _dev(rng.uniform(size=(len(pgs), 34, 40)))
    out = torch.zeros(vg.shape, device="cuda")
    ptor = kk.BackProjector(texture_type="array")
    ptor.projection_geometry = pgs
    ptor.volume_geometry = vg
    ptor.projections = projs
    ptor.volume = out
    ptor()                            # warm-up: tables and texture exist
    torch.cuda.synchronize()

    def call():
        ptor.projections = projs
        ptor()
    g = _native.Graph.capture(call)
    eager_op = kk.XrayTransform(pgs, vg)
    for k in range(3):
        projs.copy_(_dev(rng.uniform(size=projs.shape)))
        out.fill_(-1.0)
        g.replay()
        torch.cuda.synchronize()
        ref = eager_op.T(projs)
        _ = (ref * 2.0).sum().item()  # eager work between replays
        assert torch.equal(out, ref), f"replay {k}"
