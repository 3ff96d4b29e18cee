"""CPU oracle for the HI-SLAM2 dense-tracking hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the ``droid_backends``
extension, the HIP library, the ``hislam2_amd`` package) may import, call or
link anything under ``oracle/``.  The only legitimate users are ``tests/``,
``__graft_entry__.smoke()`` (as the checker) and the ``cpu_baseline`` leg of
``bench.py``.

The oracle is a float64 numpy restatement of the reference algorithm.  Each
function cites the reference file:line it restates (paths relative to the
HI-SLAM2 source tree).  Modules:

* ``se3``  -- SE3 algebra in the reference's conventions (xyzw quaternions,
  tangent order [tau, phi], left retraction).
* ``ba``   -- the per-edge linearization and the Gauss-Newton / Schur step in
  the semantics of ``ba_cuda`` (src/droid_kernels.cu:1820-1937), plus the
  dense full-system check used to pin the Schur algebra.
* ``corr`` -- all-pairs correlation, avg-pool pyramid, radius-r lookup and its
  adjoint (src/correlation_kernels.cu), and the on-the-fly variant
  (src/altcorr_kernel.cu).

How each function is pinned (closed forms, finite differences, brute force,
library routines) is stated in the test files under ``tests/``.
"""
