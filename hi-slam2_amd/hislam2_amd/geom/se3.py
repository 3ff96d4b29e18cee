"""SE3 operations on ``[..., 7]`` pose tensors for the autograd glue of the GN step.

Conventions of lietorch's SE3 (the reference's pose type): translation then an xyzw unit
quaternion, tangent vectors ``[tau, phi]``, left retraction ``retr(T, xi) = Exp(xi) * T``,
``Exp`` with the SO3 left Jacobian ``V``.  Only what the backward passes need: the retraction
(differentiated by torch autograd, it acts on [N, 7] tensors) and the map between
tangent-space gradients and gradients of the 7 stored numbers.
"""
import torch


def quat_mul(a, b):
    ax, ay, az, aw = a.unbind(-1)
    bx, by, bz, bw = b.unbind(-1)
    return torch.stack([aw * bx + ax * bw + ay * bz - az * by,
                        aw * by + ay * bw + az * bx - ax * bz,
                        aw * bz + az * bw + ax * by - ay * bx,
                        aw * bw - ax * bx - ay * by - az * bz], -1)


def quat_rotate(q, X):
    v, w = q[..., :3], q[..., 3:]
    uv = 2.0 * torch.cross(v, X, dim=-1)
    return X + w * uv + torch.cross(v, uv, dim=-1)


def exp(xi):
    """Exp of [tau, phi] as a [..., 7] pose; Taylor series of the coefficients near 0."""
    tau, phi = xi[..., :3], xi[..., 3:]
    th2 = (phi * phi).sum(-1, keepdim=True)
    small = th2 < 1e-4
    th = torch.sqrt(torch.where(small, torch.ones_like(th2), th2))
    imag = torch.where(small, 0.5 - th2 / 48.0 + th2 * th2 / 3840.0, torch.sin(0.5 * th) / th)
    real = torch.where(small, 1.0 - th2 / 8.0 + th2 * th2 / 384.0, torch.cos(0.5 * th))
    a = torch.where(small, 0.5 - th2 / 24.0 + th2 * th2 / 720.0, (1.0 - torch.cos(th)) / (th * th))
    b = torch.where(small, 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0,
                    (th - torch.sin(th)) / (th * th * th))
    c1 = torch.cross(phi, tau, dim=-1)
    t = tau + a * c1 + b * torch.cross(phi, c1, dim=-1)
    return torch.cat([t, imag * phi, real], -1)


def retr(T, xi):
    """Exp(xi) * T."""
    E = exp(xi)
    return torch.cat([E[..., :3] + quat_rotate(E[..., 3:], T[..., :3]), quat_mul(E[..., 3:], T[..., 3:])], -1)


def tangent_basis(T):
    """M [..., 7, 6] = d(Exp(xi) T) / d xi at xi = 0: translation rows [I, -[t]x], quaternion
    rows [0, (w I - [v]x; -v^T) / 2]."""
    t, v, w = T[..., :3], T[..., 3:6], T[..., 6:]
    M = T.new_zeros(T.shape[:-1] + (7, 6))
    eye = torch.eye(3, dtype=T.dtype, device=T.device)
    M[..., :3, :3] = eye

    def skew(u):
        o = torch.zeros_like(u[..., 0])
        return torch.stack([o, -u[..., 2], u[..., 1], u[..., 2], o, -u[..., 0],
                            -u[..., 1], u[..., 0], o], -1).view(u.shape[:-1] + (3, 3))
    M[..., :3, 3:] = -skew(t)
    M[..., 3:6, 3:] = 0.5 * (w[..., None] * eye - skew(v))
    M[..., 6, 3:] = -0.5 * v
    return M


def tangent_to_ambient(T, g):
    """Gradient g [..., 6] with respect to the left tangent of T -> gradient with respect to
    the 7 numbers of T, G = M (M^T M)^-1 g, so that M^T G = g (the part along the manifold is
    exact, the normal part is zero)."""
    M = tangent_basis(T.double())
    MtM = M.transpose(-1, -2) @ M
    return (M @ torch.linalg.solve(MtM, g.double()[..., None]))[..., 0].to(g.dtype)
