"""GPU parity of the torch-formulation GN step (hislam2_amd.geom.ba over droid_backends.GNStep)
against the float64 torch oracle (oracle/torch_ba.py) on the differentiation workload
(16 keyframes, 80 edges, 48x64, fixedp = 2).

Tolerances (errors are measured against the float64 oracle run from the same float32 inputs):
* forward: dx, dz and dzcov within 1e-4 of the largest oracle value (relative, max norm);
* gradients of the two-step loss with respect to targets, weights and eta, and of a one-step
  loss with respect to the input poses and disparities: within 2e-3 of the largest oracle
  gradient (the forward runs FP32 Jacobians and the loss after two steps amplifies their
  rounding), and at 8 random coordinates against float64 central differences of the oracle.
"""
import numpy as np
import pytest
import torch

from hislam2_amd import synth
from oracle import torch_ba as tb

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


def _inputs(h=48, w=64, fixedp=2):
    p = synth.ba_diff16(h=h, w=w)
    N = p.poses.shape[0]
    f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32))
    d = dict(p=p, N=N, h=h, w=w, fixedp=fixedp,
             poses=f32(p.poses)[None], disps=f32(p.disps)[None],
             intr=f32(p.intrinsics)[None, None].expand(1, N, 4).contiguous(),
             ii=torch.from_numpy(p.ii), jj=torch.from_numpy(p.jj),
             tg=f32(p.targets).permute(0, 2, 3, 1)[None].contiguous(),
             wt=f32(p.weights).permute(0, 2, 3, 1)[None].contiguous())
    d['M'] = len(torch.unique(d['ii']))
    d['eta'] = torch.full((1, d['M'], h, w), 1e-4)
    return d


def _gpu(d, k):
    return d[k].to(DEV)


def _rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


def test_gn_step_forward_matches_oracle():
    from hislam2_amd.geom import ba as gba
    d = _inputs()
    P1, D1, cov = gba.BA(_gpu(d, 'tg'), _gpu(d, 'wt'), _gpu(d, 'eta'), _gpu(d, 'poses'), _gpu(d, 'disps'),
                         _gpu(d, 'intr'), d['ii'].to(DEV), d['jj'].to(DEV), fixedp=d['fixedp'])
    dd = lambda k: d[k].double()
    Pr, Dr, covr, dxr, dzr = tb.BA(dd('tg'), dd('wt'), dd('eta').reshape(1, d['M'], -1), dd('poses'),
                                   dd('disps'), dd('intr'), d['ii'], d['jj'], fixedp=d['fixedp'],
                                   return_update=True)
    # pose updates as tangent differences, disparity updates as differences
    e = _rel(P1[0].cpu() - d['poses'][0], Pr[0] - dd('poses')[0]); assert e <= 1e-4, e
    e = _rel(D1[0].cpu() - d['disps'][0], Dr[0] - dd('disps')[0]); assert e <= 1e-4, e
    e = _rel(cov, covr[0]); assert e <= 1e-4, e
    assert cov.shape == (d['M'], d['h'] * d['w'])


def test_moba_forward_matches_oracle():
    from hislam2_amd.geom import ba as gba
    d = _inputs()
    P1 = gba.MoBA(_gpu(d, 'tg'), _gpu(d, 'wt'), None, _gpu(d, 'poses'), _gpu(d, 'disps'), _gpu(d, 'intr'),
                  d['ii'].to(DEV), d['jj'].to(DEV), fixedp=d['fixedp'])
    dd = lambda k: d[k].double()
    Pr = tb.MoBA(dd('tg'), dd('wt'), None, dd('poses'), dd('disps'), dd('intr'), d['ii'], d['jj'],
                 fixedp=d['fixedp'])
    e = _rel(P1[0].cpu() - d['poses'][0], Pr[0] - dd('poses')[0]); assert e <= 1e-4, e
    assert torch.equal(P1[0, :d['fixedp']].cpu(), d['poses'][0, :d['fixedp']])


def test_moba_gradients_match_oracle():
    """MoBA's adjoint, including the block_solve damping of every 6x6 block (chol.py:36-37)."""
    from hislam2_amd.geom import ba as gba
    from hislam2_amd.geom import se3
    d = _inputs()
    ii, jj = d['ii'], d['jj']
    wP = torch.randn(1, d['N'], 7, generator=torch.Generator().manual_seed(7))

    def run(step, tg, wt, P, D, intr, ii_, jj_):
        return (step(tg, wt, None, P, D, intr, ii_, jj_, fixedp=d['fixedp']) * wP.to(P)).sum()

    args = [_gpu(d, 'tg').requires_grad_(), _gpu(d, 'wt').requires_grad_(), _gpu(d, 'poses').requires_grad_(),
            _gpu(d, 'disps').requires_grad_()]
    grads = torch.autograd.grad(run(gba.MoBA, *args, _gpu(d, 'intr'), ii.to(DEV), jj.to(DEV)), args)
    rargs = [d['tg'].double().requires_grad_(), d['wt'].double().requires_grad_(),
             d['poses'].double().requires_grad_(), d['disps'].double().requires_grad_()]
    gref = torch.autograd.grad(run(tb.MoBA, *rargs, d['intr'].double(), ii, jj), rargs)
    M = se3.tangent_basis(d['poses'][0].double())
    tan = lambda G_: (M.transpose(-1, -2) @ G_[0].double().cpu()[..., None])[..., 0]
    for name, g, gr in zip(['targets', 'weights', 'poses', 'disps'], grads, gref):
        e = _rel(tan(g), tan(gr)) if name == 'poses' else _rel(g, gr)
        assert e <= 2e-3, (name, e)


def _loss_two_steps(step, d, tg, wt, eta, poses, disps, intr, ii, jj):
    P, D = poses, disps
    for _ in range(2):
        P, D, _ = step(tg, wt, eta, P, D, intr, ii, jj, fixedp=d['fixedp'])
    p = d['p']
    pt = torch.from_numpy(np.asarray(p.poses_true, np.float64))[None].to(P.device, P.dtype)
    dt = torch.from_numpy(np.asarray(p.disps_true, np.float64))[None].to(P.device, P.dtype)
    return tb.pose_error_loss(P, pt, D, dt)


def test_two_step_gradients_match_oracle_autograd_and_differences():
    """The differentiation workload: loss = sum ||log(T T*^-1)||^2 + sum |d - d*| after two
    unrolled steps, backpropagated through the hand-written adjoint."""
    from hislam2_amd.geom import ba as gba
    d = _inputs()
    ii, jj = d['ii'], d['jj']
    g_in = [_gpu(d, 'tg').requires_grad_(), _gpu(d, 'wt').requires_grad_(), _gpu(d, 'eta').requires_grad_()]
    L = _loss_two_steps(gba.BA, d, *g_in, _gpu(d, 'poses'), _gpu(d, 'disps'), _gpu(d, 'intr'),
                        ii.to(DEV), jj.to(DEV))
    grads = torch.autograd.grad(L, g_in)
    r_in = [d['tg'].double().requires_grad_(), d['wt'].double().requires_grad_(),
            d['eta'].double().requires_grad_()]
    Lr = _loss_two_steps(tb.BA, d, *r_in, d['poses'].double(), d['disps'].double(), d['intr'].double(), ii, jj)
    gref = torch.autograd.grad(Lr, r_in)
    assert abs(float(L.detach()) - float(Lr.detach())) <= 1e-4 * abs(float(Lr.detach()))
    for name, g, gr in zip(['targets', 'weights', 'eta'], grads, gref):
        e = _rel(g, gr)
        assert e <= 2e-3, (name, e)
    # 8 coordinates against float64 central differences of the oracle (h = 1e-4; 1e-6 for eta)
    rng = np.random.default_rng(11)
    base = [d['tg'].double(), d['wt'].double(), d['eta'].double()]
    for k in range(3):
        h = 1e-6 if k == 2 else 1e-4
        scale = float(gref[k].abs().max())
        for idx in rng.choice(base[k].numel(), size=3 if k < 2 else 2, replace=False):
            args = [b.clone() for b in base]
            args[k].reshape(-1)[idx] += h
            lp = _loss_two_steps(tb.BA, d, *args, d['poses'].double(), d['disps'].double(),
                                 d['intr'].double(), ii, jj)
            args[k].reshape(-1)[idx] -= 2 * h
            lm_ = _loss_two_steps(tb.BA, d, *args, d['poses'].double(), d['disps'].double(),
                                  d['intr'].double(), ii, jj)
            fd = float((lp - lm_) / (2 * h))
            got = float(grads[k].reshape(-1)[idx])
            assert abs(got - fd) <= 2e-3 * scale, (k, int(idx), got, fd)


def test_one_step_state_gradients_match_oracle():
    """Gradients with respect to the input poses (7 stored numbers, along the manifold) and
    disparities of a loss after one step: what chains two unrolled steps together."""
    from hislam2_amd.geom import ba as gba
    d = _inputs()
    ii, jj = d['ii'], d['jj']
    g = torch.Generator().manual_seed(5)
    wP = torch.randn(1, d['N'], 7, generator=g)
    wD = torch.randn(1, d['N'], d['h'], d['w'], generator=g)

    def run(step, P, D, tg, wt, eta, intr, ii_, jj_):
        P1, D1, _ = step(tg, wt, eta, P, D, intr, ii_, jj_, fixedp=d['fixedp'])
        return (P1 * wP.to(P1)).sum() + (D1 * wD.to(D1)).sum()

    P = _gpu(d, 'poses').requires_grad_()
    D = _gpu(d, 'disps').requires_grad_()
    L = run(gba.BA, P, D, _gpu(d, 'tg'), _gpu(d, 'wt'), _gpu(d, 'eta'), _gpu(d, 'intr'), ii.to(DEV), jj.to(DEV))
    gP, gD = torch.autograd.grad(L, [P, D])
    Pr = d['poses'].double().requires_grad_()
    Dr = d['disps'].double().requires_grad_()
    Lr = run(tb.BA, Pr, Dr, d['tg'].double(), d['wt'].double(), d['eta'].double().reshape(1, d['M'], -1),
             d['intr'].double(), ii, jj)
    gPr, gDr = torch.autograd.grad(Lr, [Pr, Dr])
    # compare pose gradients along the manifold: M^T g (tangent components)
    from hislam2_amd.geom import se3
    M = se3.tangent_basis(d['poses'][0].double())
    tan = lambda G: (M.transpose(-1, -2) @ G[0].double().cpu()[..., None])[..., 0]
    e = _rel(tan(gP), tan(gPr)); assert e <= 2e-3, e
    e = _rel(gD, gDr); assert e <= 2e-3, e


def test_failed_factorisation_gives_zero_pose_update():
    """A damping that makes H indefinite: CholeskySolver's failure path (chol.py:14-18) gives
    dx = 0 while the depth update Q w is still applied; dzcov falls back to Q."""
    from hislam2_amd.geom import ba as gba
    d = _inputs(h=12, w=16)
    P1, D1, cov = gba.BA(_gpu(d, 'tg'), _gpu(d, 'wt'), _gpu(d, 'eta'), _gpu(d, 'poses'), _gpu(d, 'disps'),
                         _gpu(d, 'intr'), d['ii'].to(DEV), d['jj'].to(DEV), fixedp=d['fixedp'], ep=-1e6)
    assert torch.equal(P1.cpu(), d['poses'])
    H, E, C, v, w, kx = tb.linear_system(d['tg'].double(), d['wt'].double(),
                                         d['eta'].double().reshape(1, d['M'], -1), d['poses'].double(),
                                         d['disps'].double(), d['intr'].double(), d['ii'], d['jj'], d['fixedp'])
    dz = (w / C)[0].reshape(-1, d['h'], d['w'])
    ref = d['disps'][0].double().index_add(0, kx, dz)
    ref = torch.where(ref > 10, torch.zeros_like(ref), ref).clamp(min=0.001)
    e = _rel(D1[0].cpu(), ref); assert e <= 1e-5, e
    e = _rel(cov, 1.0 / C[0]); assert e <= 1e-5, e


def test_projective_transform_matches_oracle():
    from hislam2_amd.geom import projective_ops as pops
    d = _inputs()
    c, v, (Ji, Jj, Jz) = pops.projective_transform(_gpu(d, 'poses'), _gpu(d, 'disps'), _gpu(d, 'intr'),
                                                   d['ii'].to(DEV), d['jj'].to(DEV), jacobian=True)
    cr, vr, (Jir, Jjr, Jzr) = tb.projective_transform(d['poses'].double(), d['disps'].double(),
                                                      d['intr'].double(), d['ii'], d['jj'], jacobian=True)
    assert c.shape == cr.shape and v.shape == vr.shape and Ji.shape == Jir.shape and Jz.shape == Jzr.shape
    e = float((c.cpu().double() - cr).abs().max()); assert e <= 1e-3, e        # pixels
    assert float((v.cpu().double() != vr).double().mean()) <= 1e-4
    for a, b in ((Ji, Jir), (Jj, Jjr), (Jz, Jzr)):
        e = _rel(a, b); assert e <= 1e-5, e
    c3, _ = pops.projective_transform(_gpu(d, 'poses'), _gpu(d, 'disps'), _gpu(d, 'intr'), d['ii'].to(DEV),
                                      d['jj'].to(DEV), return_depth=True)
    assert c3.shape[-1] == 3 and torch.allclose(c3[..., :2], c)


def test_projective_transform_gradients_match_oracle():
    """Gradient of a weighted sum of coords (2 and 3 channels) with respect to the poses
    (along the manifold) and the disparities, against autograd through the oracle."""
    from hislam2_amd.geom import projective_ops as pops
    from hislam2_amd.geom import se3
    d = _inputs()
    for ch in (2, 3):
        g = torch.Generator().manual_seed(ch)
        W = torch.randn(1, len(d['ii']), d['h'], d['w'], ch, generator=g)
        P = _gpu(d, 'poses').requires_grad_()
        D = _gpu(d, 'disps').requires_grad_()
        c, _ = pops.projective_transform(P, D, _gpu(d, 'intr'), d['ii'].to(DEV), d['jj'].to(DEV),
                                         return_depth=ch == 3)
        gP, gD = torch.autograd.grad((c * W.to(DEV)).sum(), [P, D])
        Pr = d['poses'].double().requires_grad_()
        Dr = d['disps'].double().requires_grad_()
        X0 = tb.iproj(Dr[:, d['ii']], d['intr'].double()[:, d['ii']])
        G = tb.se3_mul(Pr[:, d['jj']], tb.se3_inv(Pr[:, d['ii']]))
        X1 = tb.se3_act4(G[:, :, None, None], X0)
        cr, _ = tb.proj(X1, d['intr'].double()[:, d['jj']])
        if ch == 3:
            Zs = torch.where(X1[..., 2] < 0.1, torch.ones_like(X1[..., 2]), X1[..., 2])
            cr = torch.cat([cr, (X1[..., 3] / Zs)[..., None]], -1)
        gPr, gDr = torch.autograd.grad((cr * W.double()).sum(), [Pr, Dr])
        M = se3.tangent_basis(d['poses'][0].double())
        tan = lambda G_: (M.transpose(-1, -2) @ G_[0].double().cpu()[..., None])[..., 0]
        e = _rel(tan(gP), tan(gPr)); assert e <= 1e-4, (ch, e)
        e = _rel(gD, gDr); assert e <= 1e-4, (ch, e)
