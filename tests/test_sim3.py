"""Sim3 group operations of hislam2_amd.geom.sim3 against 4x4 matrices (scipy expm / products)
and the pose-graph constraint blocks against their closed form at a consistent graph:
r = log(Gij Gi Gj^-1) has Jacobians Adj(Gij) (left perturbation of Gi) and -I (of Gj) there."""
import numpy as np
import torch
from scipy.linalg import expm

from hislam2_amd.geom import sim3


def mat(T):
    T = T.numpy()
    x, y, z, w = T[3:7]
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = T[7] * R
    M[:3, 3] = T[:3]
    return M


def gen(xi):
    tau, phi, sigma = xi[:3], xi[3:6], xi[6]
    G = np.zeros((4, 4))
    G[:3, :3] = np.array([[0, -phi[2], phi[1]], [phi[2], 0, -phi[0]], [-phi[1], phi[0], 0]]) + sigma * np.eye(3)
    G[:3, 3] = tau
    return G


def adj(T):
    """7x7 adjoint of [[s R, t], [0, 1]] on [tau, phi, sigma]."""
    M = mat(T)
    s = T[7].item()
    R = M[:3, :3] / s
    t = M[:3, 3]
    tx = np.array([[0, -t[2], t[1]], [t[2], 0, -t[0]], [-t[1], t[0], 0]])
    A = np.zeros((7, 7))
    A[:3, :3] = s * R
    A[:3, 3:6] = tx @ R
    A[:3, 6] = -t
    A[3:6, 3:6] = R
    A[6, 6] = 1.0
    return A


def random_xi(rng, scale=0.5):
    return torch.tensor(rng.normal(scale=scale, size=7), dtype=torch.float64)


def test_exp_matches_matrix_exponential():
    rng = np.random.default_rng(0)
    cases = [random_xi(rng) for _ in range(5)] + [random_xi(rng, 1e-9), torch.zeros(7, dtype=torch.float64)]
    cases.append(torch.tensor([0.1, -0.2, 0.3, 0, 0, 0, 0.4], dtype=torch.float64))     # no rotation
    cases.append(torch.tensor([0.1, -0.2, 0.3, 0.5, -0.1, 0.2, 0], dtype=torch.float64))  # no scale
    for xi in cases:
        assert np.allclose(mat(sim3.exp(xi)), expm(gen(xi.numpy())), atol=1e-12)


def test_log_inverts_exp():
    rng = np.random.default_rng(1)
    for _ in range(6):
        xi = random_xi(rng)
        assert torch.allclose(sim3.log(sim3.exp(xi)), xi, atol=1e-12)
    xi = random_xi(rng, 1e-10)
    assert torch.allclose(sim3.log(sim3.exp(xi)), xi, atol=1e-15, rtol=0)


def test_mul_inv_retr_are_matrix_products():
    rng = np.random.default_rng(2)
    a, b = sim3.exp(random_xi(rng)), sim3.exp(random_xi(rng))
    assert np.allclose(mat(sim3.mul(a, b)), mat(a) @ mat(b), atol=1e-12)
    assert np.allclose(mat(sim3.inv(a)), np.linalg.inv(mat(a)), atol=1e-12)
    xi = random_xi(rng, 0.1)
    assert np.allclose(mat(sim3.retr(a, xi)), expm(gen(xi.numpy())) @ mat(a), atol=1e-12)


def test_constraints_at_a_consistent_graph():
    rng = np.random.default_rng(3)
    N, Np = 6, 3
    poses = torch.stack([sim3.exp(random_xi(rng)) for _ in range(N)])[None]       # [1, N, 8]
    ii = torch.tensor([0, 1, 3])
    jj = torch.tensor([2, 4, 5])
    # consistent relative poses with unit scale: choose Gj = Gij Gi with an SE3 Gij
    rel = []
    for k in range(Np):
        g = sim3.exp(torch.cat([random_xi(rng)[:6], torch.zeros(1, dtype=torch.float64)]))
        poses[0, jj[k]] = sim3.mul(g, poses[0, ii[k]])
        rel.append(g[:7])
    rel = torch.stack(rel)[None]
    infos = torch.diag_embed(torch.tensor(rng.uniform(0.5, 2.0, size=(Np, 7)), dtype=torch.float32))
    pw = 1e-3
    Hsp, vsp, chi2, _ = sim3.global_relative_posesim3_constraints(ii, jj, poses, rel, infos, pw=pw)
    assert Hsp.shape == (4, 1, Np, 7, 7) and vsp.shape == (2, 1, Np, 7)
    assert float(chi2) < 1e-20 and vsp.abs().max() < 1e-6
    for k in range(Np):
        Ji = adj(rel[0, k].double().new_tensor(np.r_[rel[0, k].numpy(), 1.0]))
        Jj = -np.eye(7)
        L = infos[k].numpy()
        assert np.allclose(Hsp[0, 0, k].numpy(), pw * Ji.T @ L @ Ji, rtol=1e-5, atol=1e-7)
        assert np.allclose(Hsp[1, 0, k].numpy(), pw * Ji.T @ L @ Jj, rtol=1e-5, atol=1e-7)
        assert np.allclose(Hsp[2, 0, k].numpy(), pw * Jj.T @ L @ Ji, rtol=1e-5, atol=1e-7)
        assert np.allclose(Hsp[3, 0, k].numpy(), pw * Jj.T @ L @ Jj, rtol=1e-5, atol=1e-7)
