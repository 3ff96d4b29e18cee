"""Accuracy metric of the BA output, evaluated on the GPU in float64 with torch ops.

``reprojection_rms`` = sqrt(sum wbar |target - pi(G_j G_i^-1 X_i)|^2 / sum wbar) with
wbar = weight * valid and valid = (transformed depth >= 0.25), the validity rule of
projective_transform_kernel (src/droid_kernels.cu:474-481).  This is a measurement
helper, not part of the hot path.
"""
import torch


def _quat_mul(a, b):
    ax, ay, az, aw = a.unbind(-1)
    bx, by, bz, bw = b.unbind(-1)
    return torch.stack([aw * bx + ax * bw + ay * bz - az * by,
                        aw * by + ay * bw + az * bx - ax * bz,
                        aw * bz + az * bw + ax * by - ay * bx,
                        aw * bw - ax * bx - ay * by - az * bz], -1)


def _rot(q):
    x, y, z, w = q.unbind(-1)
    return torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
                       -1).reshape(q.shape[:-1] + (3, 3))


@torch.no_grad()
def reprojection_rms(poses, disps, intrinsics, targets, weights, ii, jj, chunk=512):
    poses = poses.double()
    fx, fy, cx, cy = [float(v) for v in intrinsics[:4].tolist()]
    N, h, w = disps.shape
    dev = disps.device
    v, u = torch.meshgrid(torch.arange(h, device=dev, dtype=torch.float64),
                          torch.arange(w, device=dev, dtype=torch.float64), indexing='ij')
    X0 = torch.stack([(u - cx) / fx, (v - cy) / fy, torch.ones_like(u)]).reshape(3, -1)
    num = torch.zeros((), dtype=torch.float64, device=dev)
    den = torch.zeros((), dtype=torch.float64, device=dev)
    for s in range(0, ii.numel(), chunk):
        a, b = ii[s:s + chunk], jj[s:s + chunk]
        pi, pj = poses[a], poses[b]
        qinv = torch.cat([-pi[:, 3:6], pi[:, 6:]], -1)
        qij = _quat_mul(pj[:, 3:], qinv)
        R = _rot(qij)
        tij = pj[:, :3] - torch.einsum('eij,ej->ei', R, pi[:, :3])
        d = disps[a].reshape(a.numel(), -1).double()
        X = torch.einsum('eij,jp->eip', R, X0) + tij[:, :, None] * d[:, None]
        z = X[:, 2]
        valid = z >= 0.25
        zs = torch.where(valid, z, torch.ones_like(z))
        px = fx * X[:, 0] / zs + cx
        py = fy * X[:, 1] / zs + cy
        tg = targets[s:s + chunk].reshape(a.numel(), 2, -1).double()
        wt = weights[s:s + chunk].reshape(a.numel(), 2, -1).double() * valid[:, None]
        r2 = (tg[:, 0] - px) ** 2 * wt[:, 0] + (tg[:, 1] - py) ** 2 * wt[:, 1]
        num += torch.where(valid, r2, torch.zeros_like(r2)).sum()
        den += wt.sum()
    return float(torch.sqrt(num / den))
