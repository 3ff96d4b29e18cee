"""Pins the torch-formulation oracle (oracle/torch_ba.py) on the CPU in float64.

Each check is chosen so that a plausible slip fails it: the analytic Jacobians against
autograd of the projection itself, one GN step against the independent numpy oracle of the
CUDA semantics in the regime where the two formulations coincide, dzcov against the diagonal
of the inverse of the full damped system, the solver adjoint against gradcheck, and the
gradient of the two-step differentiation loss against central finite differences.
"""
import numpy as np
import pytest
import scipy.linalg
import torch

from hislam2_amd import synth
from oracle import ba as oba
from oracle import torch_ba as tb


def _problem(h=12, w=16):
    p = synth.ba_diff16(h=h, w=w)
    T = lambda a: torch.from_numpy(np.asarray(a, np.float64))
    N = p.poses.shape[0]
    pz = T(p.poses)
    pz[:, 3:] /= pz[:, 3:].norm(dim=-1, keepdim=True)
    d = dict(p=p, N=N, h=h, w=w, poses=pz[None].clone(), disps=T(p.disps)[None].clone(),
             intr=T(p.intrinsics)[None, None].expand(1, N, 4).contiguous(),
             ii=torch.from_numpy(p.ii), jj=torch.from_numpy(p.jj),
             tg=T(p.targets).permute(0, 2, 3, 1)[None].contiguous(),
             wt=T(p.weights).permute(0, 2, 3, 1)[None].contiguous())
    d['M'] = len(torch.unique(d['ii']))
    d['eta'] = torch.full((1, d['M'], h * w), 1e-4, dtype=torch.float64)
    return d


def test_se3_exp_log_closed_forms():
    rng = np.random.default_rng(0)
    xi = torch.from_numpy(rng.normal(scale=0.7, size=(16, 6)))
    T = tb.se3_exp(xi)
    for k in range(16):
        M = scipy.linalg.expm(_twist(xi[k].numpy()))
        assert np.allclose(_mat(T[k].numpy()), M, atol=1e-12)
    assert torch.allclose(tb.se3_log(T), xi, atol=1e-10)
    assert torch.allclose(tb.se3_exp(torch.zeros(6, dtype=torch.float64)),
                          torch.tensor([0, 0, 0, 0, 0, 0, 1.0], dtype=torch.float64))


def _twist(xi):
    M = np.zeros((4, 4))
    t, p = xi[:3], xi[3:]
    M[:3, :3] = [[0, -p[2], p[1]], [p[2], 0, -p[0]], [-p[1], p[0], 0]]
    M[:3, 3] = t
    return M


def _mat(T):
    x, y, z, w = T[3:]
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = T[:3]
    return M


def test_projective_jacobians_match_autograd():
    d = _problem()
    _, _, (Ji, Jj, Jz) = tb.projective_transform(d['poses'], d['disps'], d['intr'], d['ii'], d['jj'],
                                                 jacobian=True)
    e = 5
    i, j = int(d['ii'][e]), int(d['jj'][e])

    def f(xi_i, xi_j, dd):
        P2 = d['poses'].clone()
        P2[0, i] = tb.se3_retr(d['poses'][0, i], xi_i)
        P2[0, j] = tb.se3_retr(d['poses'][0, j], xi_j)
        D2 = d['disps'].clone()
        D2[0, i] = d['disps'][0, i] + dd
        c, _ = tb.projective_transform(P2, D2, d['intr'], d['ii'][e:e + 1], d['jj'][e:e + 1])
        return c[0, 0]

    z6 = torch.zeros(6, dtype=torch.float64)
    Ja = torch.autograd.functional.jacobian(f, (z6, z6, torch.zeros(d['h'], d['w'], dtype=torch.float64)))
    assert (Ja[0] - Ji[0, e]).abs().max() < 1e-10
    assert (Ja[1] - Jj[0, e]).abs().max() < 1e-10
    yy, xx = torch.meshgrid(torch.arange(d['h']), torch.arange(d['w']), indexing='ij')
    assert (Ja[2][yy, xx, :, yy, xx] - Jz[0, e][..., 0]).abs().max() < 1e-12


def test_step_matches_cuda_semantics_without_damping():
    """With lm = ep = 0, eta + 1e-7 and no slot-0 skip the two formulations solve the same
    system: the torch formulation against the independent numpy oracle of ba_cuda."""
    d = _problem()
    p, fixedp = d['p'], 1
    h, w, N = d['h'], d['w'], d['N']
    _, _, _, dx, dz = tb.BA(d['tg'], d['wt'], d['eta'], d['poses'], d['disps'], d['intr'], d['ii'],
                            d['jj'], fixedp=fixedp, ep=0.0, lm=0.0, return_update=True)
    eta = np.full((d['M'], h * w), 1e-4 + 1e-7)
    _, _, dxo, dzo = oba.ba(d['poses'][0].numpy(), p.disps, p.intrinsics, p.targets, p.weights, eta,
                            p.ii, p.jj, fixedp, N, iterations=1, lm=0.0, ep=0.0, t0_skip=False)
    assert np.abs(dx[0].numpy() - dxo).max() <= 1e-8 * np.abs(dxo).max()
    assert np.abs(dz[0].numpy() - dzo).max() <= 1e-8 * np.abs(dzo).max()


def test_damping_before_schur_and_dzcov_is_the_marginal_variance():
    d = _problem(h=6, w=8)
    fixedp, ep, lm = 2, 0.1, 1e-4
    H, E, C, v, wv, kx = tb.linear_system(d['tg'], d['wt'], d['eta'], d['poses'], d['disps'], d['intr'],
                                          d['ii'], d['jj'], fixedp)
    dx, dz, dzcov = tb.schur_solve(H, E, C, v, wv, ep, lm)
    B, P, M, D, HW = E.shape
    Hm = H.permute(0, 1, 3, 2, 4).reshape(P * D, P * D)
    Hm = Hm + torch.diag(ep + lm * torch.diagonal(Hm))
    Em = E.permute(0, 1, 3, 2, 4).reshape(P * D, M * HW)
    K = torch.zeros(P * D + M * HW, P * D + M * HW, dtype=torch.float64)
    K[:P * D, :P * D] = Hm
    K[:P * D, P * D:] = Em
    K[P * D:, :P * D] = Em.T
    K[P * D:, P * D:] = torch.diag(C.reshape(-1))
    y = torch.linalg.solve(K, torch.cat([v.reshape(-1), wv.reshape(-1)]))
    assert torch.allclose(dx.reshape(-1), y[:P * D], rtol=1e-9, atol=1e-12)
    assert torch.allclose(dz.reshape(-1), y[P * D:], rtol=1e-9, atol=1e-12)
    cov = torch.diagonal(torch.linalg.inv(K))[P * D:]
    assert torch.allclose(dzcov.reshape(-1), cov, rtol=1e-8)


def test_cholesky_solver_adjoint_gradcheck():
    g = torch.Generator().manual_seed(1)
    A = torch.randn(7, 7, generator=g, dtype=torch.float64)
    H = (A @ A.T + 7 * torch.eye(7, dtype=torch.float64))[None].requires_grad_()
    b = torch.randn(1, 7, 1, generator=g, dtype=torch.float64, requires_grad=True)
    sym = lambda H, b: tb.CholeskySolver.apply(0.5 * (H + H.transpose(-1, -2)), b)[0]
    assert torch.autograd.gradcheck(sym, (H, b))


def test_failed_factorisation_returns_zero():
    H = -torch.eye(6, dtype=torch.float64)[None]
    b = torch.ones(1, 6, 1, dtype=torch.float64)
    x, _ = tb.CholeskySolver.apply(H, b)
    assert torch.count_nonzero(x) == 0


def two_step_loss(d, tg, wt, eta, fixedp=2):
    p = d['p']
    P, D = d['poses'], d['disps']
    for _ in range(2):
        P, D, _ = tb.BA(tg, wt, eta, P, D, d['intr'], d['ii'], d['jj'], fixedp=fixedp)
    pt = torch.from_numpy(np.asarray(p.poses_true, np.float64))[None]
    pt[..., 3:] /= pt[..., 3:].norm(dim=-1, keepdim=True)
    dt = torch.from_numpy(np.asarray(p.disps_true, np.float64))[None]
    return tb.pose_error_loss(P, pt, D, dt)


def test_two_step_gradients_match_central_differences():
    """The differentiation workload's loss after 2 unrolled steps: autograd through the
    torch formulation against float64 central differences on 64 random coordinates of
    targets, weights and eta.  Step h = 1e-4 for targets and weights; eta itself is 1e-4,
    so its step is 1e-6: the loss carries about 1e-12 of float64 rounding noise, which
    smaller steps amplify past the tolerance."""
    d = _problem(h=6, w=8)
    tg = d['tg'].clone().requires_grad_()
    wt = d['wt'].clone().requires_grad_()
    eta = d['eta'].clone().requires_grad_()
    L = two_step_loss(d, tg, wt, eta)
    grads = torch.autograd.grad(L, [tg, wt, eta])
    rng = np.random.default_rng(3)
    for k, (base, gr) in enumerate(zip([d['tg'], d['wt'], d['eta']], grads)):
        hstep = 1e-6 if k == 2 else 1e-4
        flat = base.reshape(-1)
        for idx in rng.choice(flat.numel(), size=64 // 3 + (k == 0), replace=False):
            args = [d['tg'].clone(), d['wt'].clone(), d['eta'].clone()]
            args[k].reshape(-1)[idx] += hstep
            lp = two_step_loss(d, *args)
            args[k].reshape(-1)[idx] -= 2 * hstep
            lm_ = two_step_loss(d, *args)
            fd = (lp - lm_) / (2 * hstep)
            an = gr.reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(float(an))), (k, int(idx), float(fd), float(an))
