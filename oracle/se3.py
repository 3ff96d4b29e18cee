"""SE3 algebra in the conventions of the reference's BA kernels (float64).

Pose layout: ``[tx, ty, tz, qx, qy, qz, qw]`` (translation, then an xyzw unit
quaternion), as stored in ``DepthVideo.poses`` (hislam2/depth_video.py:29) and
read by ``projective_transform_kernel`` (src/droid_kernels.cu:406-414).
Tangent vectors are ``[tau, phi]``: translation part first
(src/droid_kernels.cu:219-220).  The retraction is a left update
``T <- Exp(xi) * T`` (src/droid_kernels.cu:1342-1360).

All functions broadcast over leading dimensions.  Test infrastructure only.
"""
import numpy as np


def quat_mul(a, b):
    """Hamilton product a*b of xyzw quaternions (src/droid_kernels.cu:1350-1353)."""
    ax, ay, az, aw = np.moveaxis(np.asarray(a, np.float64), -1, 0)
    bx, by, bz, bw = np.moveaxis(np.asarray(b, np.float64), -1, 0)
    return np.stack([
        aw * bx + ax * bw + ay * bz - az * by,
        aw * by + ay * bw + az * bx - ax * bz,
        aw * bz + az * bw + ax * by - ay * bx,
        aw * bw - ax * bx - ay * by - az * bz,
    ], axis=-1)


def quat_conj(q):
    q = np.asarray(q, np.float64)
    return np.concatenate([-q[..., :3], q[..., 3:]], axis=-1)


def quat_to_mat(q):
    """Rotation matrix of a unit xyzw quaternion."""
    x, y, z, w = np.moveaxis(np.asarray(q, np.float64), -1, 0)
    R = np.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y),
    ], axis=-1)
    return R.reshape(R.shape[:-1] + (3, 3))


def act_so3(q, X):
    """Rotate 3-vectors X by q (src/droid_kernels.cu:74-84).

    The reference uses the two-cross-product form; for a unit quaternion it is
    exactly R(q) X, which is what is computed here."""
    return np.einsum('...ij,...j->...i', quat_to_mat(q), np.asarray(X, np.float64))


def relative(pi, pj):
    """G_ij = G_j * G_i^-1 for poses [..., 7] (src/droid_kernels.cu:148-159).

    Returns (t_ij, q_ij)."""
    pi = np.asarray(pi, np.float64)
    pj = np.asarray(pj, np.float64)
    qij = quat_mul(pj[..., 3:], quat_conj(pi[..., 3:]))
    tij = pj[..., :3] - act_so3(qij, pi[..., :3])
    return tij, qij


def exp_so3(phi):
    """SO3 exponential as an xyzw quaternion (src/droid_kernels.cu:179-201).

    Keeps the reference's Taylor branch for theta^2 < 1e-8."""
    phi = np.asarray(phi, np.float64)
    th2 = np.sum(phi * phi, axis=-1)
    th = np.sqrt(th2)
    small = th2 < 1e-8
    th_safe = np.where(small, 1.0, th)
    imag = np.where(small, 0.5 - th2 / 48.0 + th2 * th2 / 3840.0,
                    np.sin(0.5 * th_safe) / th_safe)
    real = np.where(small, 1.0 - th2 / 8.0 + th2 * th2 / 384.0, np.cos(0.5 * th_safe))
    return np.concatenate([imag[..., None] * phi, real[..., None]], axis=-1)


def exp_se3(xi):
    """SE3 exponential of [tau, phi] (src/droid_kernels.cu:216-244).

    t = V(phi) tau with V = I + a [phi]x + b [phi]x^2, a = (1-cos)/th^2,
    b = (th - sin)/th^3; the reference drops the V correction when
    theta <= 1e-4 and that branch is kept here.  Returns (t, q)."""
    xi = np.asarray(xi, np.float64)
    tau, phi = xi[..., :3], xi[..., 3:]
    q = exp_so3(phi)
    th2 = np.sum(phi * phi, axis=-1)
    th = np.sqrt(th2)
    big = th > 1e-4
    th_s = np.where(big, th, 1.0)
    a = np.where(big, (1 - np.cos(th_s)) / (th_s * th_s), 0.0)
    b = np.where(big, (th_s - np.sin(th_s)) / (th_s ** 3), 0.0)
    c1 = np.cross(phi, tau)
    c2 = np.cross(phi, c1)
    t = tau + a[..., None] * c1 + b[..., None] * c2
    return t, q


def retr(xi, pose):
    """Left retraction Exp(xi) * T (src/droid_kernels.cu:1342-1360)."""
    pose = np.asarray(pose, np.float64)
    dt, dq = exp_se3(xi)
    q1 = quat_mul(dq, pose[..., 3:])
    t1 = act_so3(dq, pose[..., :3]) + dt
    return np.concatenate([t1, q1], axis=-1)


def adj_T(tij, qij, J):
    """Adj(G)^T applied to row Jacobians J [..., 6] (src/droid_kernels.cu:104-119).

    With Adj = [[R, [t]x R], [0, R]], Adj^T J = [R^T J_tau ; R^T (J_phi - t x J_tau)]."""
    qinv = quat_conj(qij)
    Jt, Jr = J[..., :3], J[..., 3:]
    tb = np.broadcast_to(tij, Jt.shape)
    top = act_so3(qinv, Jt)
    bot = act_so3(qinv, Jr - np.cross(tb, Jt))
    return np.concatenate([top, bot], axis=-1)


def pose_to_mat(pose):
    """4x4 homogeneous matrix of a pose [..., 7]."""
    pose = np.asarray(pose, np.float64)
    T = np.zeros(pose.shape[:-1] + (4, 4))
    T[..., :3, :3] = quat_to_mat(pose[..., 3:])
    T[..., :3, 3] = pose[..., :3]
    T[..., 3, 3] = 1.0
    return T


def hat(xi):
    """4x4 twist matrix of [tau, phi]."""
    xi = np.asarray(xi, np.float64)
    tau, phi = xi[..., :3], xi[..., 3:]
    M = np.zeros(xi.shape[:-1] + (4, 4))
    M[..., 0, 1] = -phi[..., 2]
    M[..., 0, 2] = phi[..., 1]
    M[..., 1, 0] = phi[..., 2]
    M[..., 1, 2] = -phi[..., 0]
    M[..., 2, 0] = -phi[..., 1]
    M[..., 2, 1] = phi[..., 0]
    M[..., :3, 3] = tau
    return M
