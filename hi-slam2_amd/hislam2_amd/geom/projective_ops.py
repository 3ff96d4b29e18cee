"""Mirror of hislam2/geom/projective_ops.py (with pinhole.py) on the gfx950 kernels.

``projective_transform(poses, depths, intrinsics, ii, jj, jacobian=False, return_depth=False)``
keeps the reference's signature and return convention (``(coords, valid)`` or
``(coords, valid, (Ji, Jj, Jz))``), semantics (validity X1.z > 0.2 and X0.z > 0.2,
``Z < 0.1 -> 1`` before the divide, per-frame intrinsics) and layouts (coords
[B, E, h, w, 2 or 3], valid [B, E, h, w, 1], Ji / Jj [B, E, h, w, 2, 6], Jz [B, E, h, w, 2, 1]).
Poses are [B, N, 7] tensors in place of lietorch SE3.  coords is differentiable with respect
to poses and depths (hand-written adjoint, ``hsl2_projective_transform_backward``); the
Jacobians are returned without gradient.
"""
import torch

from .. import backend
from . import se3


def coords_grid(ht, wd, **kwargs):
    """[ht, wd, 2] pixel grid (x, y) (projective_ops.py:8-13)."""
    y, x = torch.meshgrid(torch.arange(ht).to(**kwargs).float(), torch.arange(wd).to(**kwargs).float(),
                          indexing='ij')
    return torch.stack([x, y], dim=-1)


class _ProjFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, poses, disps, intrinsics, ii, jj, return_depth):
        db = backend()
        coords, valid, _, _, _ = db.projective_transform(poses, disps, intrinsics, ii, jj, False, return_depth)
        ctx.save_for_backward(poses, disps, intrinsics, ii, jj)
        ctx.return_depth = return_depth
        ctx.mark_non_differentiable(valid)
        return coords, valid

    @staticmethod
    def backward(ctx, gcoords, gvalid):
        poses, disps, intrinsics, ii, jj = ctx.saved_tensors
        gp, gd = backend().projective_transform_backward(poses, disps, intrinsics, ii, jj,
                                                         gcoords.float().contiguous(), ctx.return_depth)
        return se3.tangent_to_ambient(poses, gp), gd, None, None, None, None


def projective_transform(poses, depths, intrinsics, ii, jj, jacobian=False, return_depth=False):
    """Map the pixels of frames ii into frames jj (projective_ops.py:44-74)."""
    db = backend()
    B = poses.shape[0]
    f32 = lambda x: (x if x.dtype == torch.float32 else x.float()).contiguous()
    ii = ii.to(device=poses.device, dtype=torch.long).contiguous()
    jj = jj.to(device=poses.device, dtype=torch.long).contiguous()
    coords, valid, Jis, Jjs, Jzs = [], [], [], [], []
    for b in range(B):
        c, v = _ProjFn.apply(f32(poses[b]), f32(depths[b]), f32(intrinsics[b]), ii, jj, bool(return_depth))
        coords.append(c)
        valid.append(v)
        if jacobian:
            _, _, Ji, Jj, Jz = db.projective_transform(f32(poses[b]), f32(depths[b]), f32(intrinsics[b]), ii, jj,
                                                       True, False)
            Jis.append(Ji)
            Jjs.append(Jj)
            Jzs.append(Jz)
    coords = torch.stack(coords)
    valid = torch.stack(valid)
    if jacobian:
        return coords, valid, (torch.stack(Jis), torch.stack(Jjs), torch.stack(Jzs))
    return coords, valid
