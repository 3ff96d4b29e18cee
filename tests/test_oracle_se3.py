"""Pins the SE3 oracle against closed forms and library routines (scipy)."""
import numpy as np
import pytest
from scipy.linalg import expm
from scipy.spatial.transform import Rotation

from oracle import se3


def rand_pose(rng, n, scale=0.5):
    q = Rotation.random(n, random_state=rng.integers(1 << 31)).as_quat()    # xyzw
    t = rng.normal(scale=scale, size=(n, 3))
    return np.concatenate([t, q], -1)


def test_quat_rotation_matches_scipy():
    rng = np.random.default_rng(0)
    p = rand_pose(rng, 20)
    X = rng.normal(size=(20, 3))
    ref = Rotation.from_quat(p[:, 3:]).apply(X)
    assert np.allclose(se3.act_so3(p[:, 3:], X), ref, atol=1e-12)


@pytest.mark.parametrize('scale', [1e-3, 1e-1, 1.0])
def test_exp_matches_matrix_exponential(scale):
    rng = np.random.default_rng(1)
    xi = rng.normal(scale=scale, size=(50, 6))
    t, q = se3.exp_se3(xi)
    T = se3.pose_to_mat(np.concatenate([t, q], -1))
    for k in range(50):
        assert np.allclose(T[k], expm(se3.hat(xi[k])), atol=1e-12)


def test_exp_small_angle_branch():
    """theta <= 1e-4: the reference keeps t = tau (src/droid_kernels.cu:231);
    the deviation from expm is bounded by |phi x tau| / 2 + O(theta^2)."""
    rng = np.random.default_rng(2)
    xi = rng.normal(size=(20, 6))
    xi[:, 3:] *= 1e-5 / np.linalg.norm(xi[:, 3:], axis=1, keepdims=True)
    t, q = se3.exp_se3(xi)
    assert np.array_equal(t, xi[:, :3])
    for k in range(20):
        Tm = expm(se3.hat(xi[k]))
        bound = 0.5 * np.linalg.norm(np.cross(xi[k, 3:], xi[k, :3])) + 1e-12
        assert np.linalg.norm(Tm[:3, 3] - t[k]) <= 1.01 * bound
        assert np.allclose(se3.quat_to_mat(q[k]), Tm[:3, :3], atol=1e-14)
    # theta^2 < 1e-8 Taylor branch of exp_so3 agrees with sin/cos to 1e-16
    phi = rng.normal(size=(20, 3)) * 5e-5
    assert np.allclose(se3.exp_so3(phi), Rotation.from_rotvec(phi).as_quat(), atol=1e-15)


def test_exp_zero_is_identity_and_retr_zero_is_noop():
    rng = np.random.default_rng(3)
    p = rand_pose(rng, 5)
    t, q = se3.exp_se3(np.zeros(6))
    assert np.array_equal(t, np.zeros(3)) and np.array_equal(q, [0, 0, 0, 1])
    assert np.allclose(se3.retr(np.zeros((5, 6)), p), p, atol=0, rtol=0)


def test_relative_pose_matches_matrices():
    rng = np.random.default_rng(4)
    pi, pj = rand_pose(rng, 10), rand_pose(rng, 10)
    tij, qij = se3.relative(pi, pj)
    Tij = se3.pose_to_mat(np.concatenate([tij, qij], -1))
    ref = se3.pose_to_mat(pj) @ np.linalg.inv(se3.pose_to_mat(pi))
    assert np.allclose(Tij, ref, atol=1e-12)
    # G_ij * G_i = G_j
    assert np.allclose(Tij @ se3.pose_to_mat(pi), se3.pose_to_mat(pj), atol=1e-12)


def test_retraction_is_left_multiplication():
    rng = np.random.default_rng(5)
    p = rand_pose(rng, 10)
    xi = rng.normal(scale=0.3, size=(10, 6))
    out = se3.pose_to_mat(se3.retr(xi, p))
    for k in range(10):
        assert np.allclose(out[k], expm(se3.hat(xi[k])) @ se3.pose_to_mat(p[k]), atol=1e-12)


def test_adjoint_transpose_identity():
    """adj_T(J) = Adj(G)^T J where Adj is defined by G hat(xi) G^-1 = hat(Adj xi)."""
    rng = np.random.default_rng(6)
    p = rand_pose(rng, 1)[0]
    G = se3.pose_to_mat(p)
    Adj = np.zeros((6, 6))
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1.0
        M = G @ se3.hat(e) @ np.linalg.inv(G)
        Adj[:, k] = [M[0, 3], M[1, 3], M[2, 3], M[2, 1], M[0, 2], M[1, 0]]
    J = rng.normal(size=(7, 6))
    got = se3.adj_T(p[:3], p[3:], J)
    assert np.allclose(got, J @ Adj, atol=1e-12)
