"""Mirror of the reference's ``hislam2/geom`` package for the hot path: one differentiable
Gauss-Newton step (``ba.BA``, ``ba.MoBA``) and the projective transform
(``projective_ops``), both running on the gfx950 kernels of ``droid_backends``.

Poses are ``[..., 7]`` tensors ``[tx, ty, tz, qx, qy, qz, qw]`` where the reference passes
lietorch ``SE3`` objects (lietorch is a CUDA extension); ``se3`` holds the few group
operations the autograd glue needs.
"""
