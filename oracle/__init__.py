"""FP64 NumPy oracle for the cone-beam hot path -- TEST INFRASTRUCTURE ONLY.

This package is a CPU restatement of the reference's arithmetic (KernelKit's
cone-beam forward projector, backprojector, FDK filter/pre-weight and SIRT),
written from the reference files read as text; every function cites the
reference file:line it follows.  It is the parity checker for the HIP path
and the timed CPU baseline ("port").  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import
it, and only as the checker -- never as the product path.  The product
(``kernelkit``) fails loudly when its HIP library is missing; it never falls
back to this code.

Inputs are duck-typed: any projection geometry object with
``source_position``, ``detector_position``, ``u``, ``v`` and
``detector.{rows, cols, pixel_width, pixel_height}``, and any volume geometry
with ``shape``, ``voxel_size``, ``extent_min``, ``extent_max``, ``rotation``.
Arrays are in world order: volumes ``vol[x, y, z]`` and projections
``proj[angle, row, col]``; axis permutations are applied by the caller.

How each part is pinned (tests/test_oracle_*.py):

* geometry/params -- rotation closed forms, detector coordinates against a
  brute-force ray/plane intersection, BP weight against ``vs^3 M^2 / A_pix``;
* forward projection -- the reference's known-answer tests
  (tests/kernels/test_cone_fp.py:74-374, :438-483), exactness on linear
  volumes, and a pure-Python ray-marching brute force on tiny inputs;
* backprojection -- the reference's known-answer tests
  (tests/kernels/test_cone_bp.py:84-224) and exactness on linear detector
  images against independently intersected coordinates;
* matched transpose -- dot-product test against the forward projection;
* filter -- direct circular convolution with the closed-form Kak-Slaney
  kernel; Hann/cosine windows against their textbook definitions;
* pre-weight -- closed form per pixel;
* FDK/SIRT -- compositions of the above; SIRT fixed point and residual
  decrease on consistent data.
"""
from oracle.geometry import (bp_params, fp_major_axes, fp_output_scale,  # noqa: F401
                             normalized_frames, rotation_sxyz)
from oracle.projectors import (cone_bp, cone_fp, cone_fp_transpose,  # noqa: F401
                               bilinear_border_zero)
from oracle.processing import (filter_response, filter_rows,  # noqa: F401
                               padded_length, preweight)
from oracle.algorithms import fdk, sirt  # noqa: F401
