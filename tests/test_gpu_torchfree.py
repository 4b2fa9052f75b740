"""The operators without PyTorch.

The reference computes on CuPy arrays (``kernelkit/kernel.py:47-119``); the
drop-in takes any device array.  One test runs FP, BP, FDK, SIRT (eager and
replayed from a HIP graph), the exact transpose and a batch of two problems
(FP, BP, tuned tile shapes, graph-replayed SIRT) in a child process in which
``import torch`` fails, with host (numpy) inputs and
:class:`kernelkit.DeviceArray` inputs, and checks them against the oracle;
another feeds arrays that only expose
``__cuda_array_interface__`` (as CuPy arrays do) and checks the results
against the torch path.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import assert_parity

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_CHILD = r"""
import json, sys
sys.modules["torch"] = None          # any `import torch` raises ImportError
sys.path[:0] = [sys.argv[1], sys.argv[1] + "/astra-kernelkit_amd"]
import numpy as np
import kernelkit as kk
from kernelkit import workloads
import oracle

def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))

def point(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())

pgs = workloads.circular_geometry(24, sod=110.0, odd=50.0, rows=36,
                                  cols=44, pixel=1.25, tilt_deg=1.0,
                                  drift_px=1.0, seed=3)
vg = kk.resolve_volume_geometry(shape=(28, 30, 32), voxel_size=1.0)
vol = workloads.ellipsoid_phantom(vg.shape, n=6, seed=2).astype(np.float32)
res = {}

# algorithms on host inputs
y = kk.fp(vol, pgs, vg)
y_ref = oracle.cone_fp(vol.astype(np.float64), pgs, vg)
res["fp"] = (rel(y, y_ref), point(y, y_ref))
x = kk.bp(y, pgs, vg)
x_ref = oracle.cone_bp(y.astype(np.float64), pgs, vg)
res["bp"] = (rel(x, x_ref), point(x, x_ref))
f = kk.fdk(y, pgs, vg)
f_ref = oracle.fdk(y.astype(np.float64), pgs, vg)
res["fdk"] = (rel(f, f_ref), point(f, f_ref))
s = kk.sirt(y, pgs, vg, iters=4, verbose=False)
s_ref = oracle.sirt(y.astype(np.float64), pgs, vg, iters=4)
res["sirt"] = (rel(s, s_ref), point(s, s_ref))

# the operator on the package's own device arrays
A = kk.XrayTransform(pgs, vg)
dv = kk.device.from_host(vol)
dy = A(dv)
assert isinstance(dy, kk.DeviceArray), type(dy)
res["op_fp"] = (rel(dy.get(), y_ref), point(dy.get(), y_ref))
w = np.random.default_rng(5).random(y.shape).astype(np.float32)
xt = A.matched_T(kk.device.from_host(w))
xt_ref = oracle.cone_fp_transpose(w.astype(np.float64), pgs, vg)
res["transpose"] = (rel(xt.get(), xt_ref), point(xt.get(), xt_ref))
xb = A.T(dy)
xb_ref = oracle.cone_bp(dy.get().astype(np.float64), pgs, vg)
res["op_bp"] = (rel(xb.get(), xb_ref), point(xb.get(), xb_ref))

# views: a non-contiguous slab of the volume projects like its copy
sub = dv[:, 1:29]
assert not sub.is_contiguous()
vg_sub = kk.resolve_volume_geometry(shape=(28, 28, 32), voxel_size=1.0)
ys = kk.XrayTransform(pgs, vg_sub)(sub).get()
ys_ref = oracle.cone_fp(vol[:, 1:29].astype(np.float64), pgs, vg_sub)
res["view_fp"] = (rel(ys, ys_ref), point(ys, ys_ref))

# a batch of two problems of different shapes: one launch per operator,
# tile shapes timed with the library's events, SIRT replayed from the
# library's own graph capture
from kernelkit.autotune import BatchTuner
from kernelkit.batch import ConeBatch
cus, name = kk._native.device_info()
assert cus > 0 and name, (cus, name)
pgs1 = pgs[::2]
vol1 = np.ascontiguousarray(vol[:, 1:29])
B = ConeBatch([(pgs, vg), (pgs1, vg_sub)], tuner=BatchTuner(reps=1))
yb = B.fp([dv, kk.device.from_host(vol1)], tune=True)
assert all(isinstance(t, kk.DeviceArray) for t in yb)
y1_ref = oracle.cone_fp(vol1.astype(np.float64), pgs1, vg_sub)
for k, ref in enumerate((y_ref, y1_ref)):
    res[f"batch_fp{k}"] = (rel(yb[k].get(), ref), point(yb[k].get(), ref))
xb2 = B.bp(yb, tune=True)
for k, (g, vgk) in enumerate(((pgs, vg), (pgs1, vg_sub))):
    ref = oracle.cone_bp(yb[k].get().astype(np.float64), g, vgk)
    res[f"batch_bp{k}"] = (rel(xb2[k].get(), ref), point(xb2[k].get(), ref))
y1 = yb[1].get()
st = B.sirt_state([kk.device.from_host(y), kk.device.from_host(y1)])
st.configure(tune=True)
st.run(2, graph=True)
st.run(2, graph=True)          # a second replay continues iterating
s1_ref = oracle.sirt(y1.astype(np.float64), pgs1, vg_sub, iters=4)
for k, ref in enumerate((s_ref, s1_ref)):
    res[f"batch_sirt{k}"] = (rel(st.x[k].get(), ref),
                             point(st.x[k].get(), ref))
sg = kk.sirt(y, pgs, vg, iters=4, verbose=False, graph=True)
res["sirt_graph"] = (rel(sg, s_ref), point(sg, s_ref))
a, b = kk._native.Event(), kk._native.Event()
a.record()
A(dv, out=dy)
b.record()
b.synchronize()
assert a.elapsed_ms(b) > 0.0

assert sys.modules.get("torch") is None, "torch was imported"
print("RESULT " + json.dumps(res), flush=True)
"""


def test_without_torch_against_oracle(tmp_path):
    script = tmp_path / "child.py"
    script.write_text(_CHILD)
    env = dict(os.environ)
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    out = subprocess.run([sys.executable, "-u", str(script), ROOT],
                         capture_output=True, text=True, timeout=600,
                         env=env, cwd=ROOT)
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-3000:]
    line = [s for s in out.stdout.splitlines() if s.startswith("RESULT ")]
    assert line, out.stdout[-3000:]
    res = json.loads(line[-1][len("RESULT "):])
    print(res)
    assert set(res) == {"fp", "bp", "fdk", "sirt", "op_fp", "transpose",
                        "op_bp", "view_fp", "batch_fp0", "batch_fp1",
                        "batch_bp0", "batch_bp1", "batch_sirt0",
                        "batch_sirt1", "sirt_graph"}
    for name, (e, p) in res.items():
        assert e <= 1e-5, (name, e)
        assert p <= 1e-4, (name, p)


class _CaiOnly:
    """An array that shows nothing but ``__cuda_array_interface__``."""

    def __init__(self, t):
        self._t = t
        self.__cuda_array_interface__ = t.__cuda_array_interface__


def test_cuda_array_interface_inputs():
    torch = pytest.importorskip("torch")
    import kernelkit as kk
    from kernelkit import workloads
    import oracle

    pgs = workloads.circular_geometry(20, sod=100.0, odd=60.0, rows=30,
                                      cols=34, pixel=1.3, seed=0)
    vg = kk.resolve_volume_geometry(shape=(24, 26, 28), voxel_size=1.0)
    vol = workloads.ellipsoid_phantom(vg.shape, n=5, seed=4)
    A = kk.XrayTransform(pgs, vg)
    tv = torch.from_numpy(vol.astype(np.float32)).cuda()
    y_t = A(tv)
    y_c = A(_CaiOnly(tv))
    assert isinstance(y_c, kk.DeviceArray)
    y_host = y_c.get()
    np.testing.assert_array_equal(y_host, y_t.cpu().numpy())
    assert_parity(y_host, oracle.cone_fp(vol.astype(np.float64), pgs, vg),
                  label="fp via __cuda_array_interface__")
    # output into a foreign array, in place
    out_t = torch.full_like(y_t, 7.0)
    A(_CaiOnly(tv), out=_CaiOnly(out_t))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out_t.cpu().numpy(), y_t.cpu().numpy())
    x_c = A.T(_CaiOnly(y_t))
    np.testing.assert_array_equal(x_c.get(), A.T(y_t).cpu().numpy())
    # a non-contiguous foreign view (every other angle)
    sub = y_t[::2]
    assert not sub.is_contiguous()
    a = kk.device.asarray(_CaiOnly(sub))
    assert a.stride() == sub.stride()
    np.testing.assert_array_equal(a.get(), sub.cpu().numpy())
