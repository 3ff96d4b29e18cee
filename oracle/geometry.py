"""Edge-selection and point-cloud geometry oracle (float64): ``frame_distance``,
``covis_distance``, ``depth_filter``, ``iproj``.

Restates src/droid_kernels.cu:860-992 (frame_distance_kernel), 994-1092 (covis_distance_kernel),
1094-1200 (depth_filter_kernel) and 1204-1273 (iproj_kernel) with their host wrappers
2047-2143, in float64 numpy: one pinhole camera (fx, fy, cx, cy), pixel (u, v) = (column, row)
lifted to (u - cx) / fx, (v - cy) / fy, 1 with its disparity d, moved by the relative pose
G_ij = G_j G_i^-1 (or the frame's own pose for iproj) as R X + d t.  Kept from the reference on
purpose: frame_distance's translation-only term, its 0.25 validity on each term's own depth and
the 1000 for less than 75 % valid weight; covis_distance's neighbours ii - 1 and ii + 1 and no
depth check; depth_filter's neighbour offsets -1, -2, -3, +3, +4, +5 and the corner order.

Test infrastructure only.
"""
import numpy as np

from . import se3

MIN_DEPTH = 0.25  # src/droid_kernels.cu:26


def _lift(ht, wd, intr):
    fx, fy, cx, cy = [float(a) for a in np.asarray(intr).reshape(-1)[:4]]
    v, u = np.meshgrid(np.arange(ht, dtype=np.float64), np.arange(wd, dtype=np.float64), indexing='ij')
    X = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)]).reshape(3, -1)
    return X, u.reshape(-1), v.reshape(-1), (fx, fy, cx, cy)


def _move(t, q, X, d):
    """R(q) X + d t for X [3, HW], d [HW]."""
    R = se3.quat_to_mat(np.asarray(q, np.float64)[None])[0]
    return R @ X + np.asarray(t, np.float64)[:, None] * d[None]


def _proj(Y, cam):
    fx, fy, cx, cy = cam
    with np.errstate(divide='ignore', invalid='ignore'):
        return fx * (Y[0] / Y[2]) + cx, fy * (Y[1] / Y[2]) + cy


def frame_distance(poses, disps, intr, ii, jj, beta):
    poses = np.asarray(poses, np.float64)
    disps = np.asarray(disps, np.float64)
    N, ht, wd = disps.shape
    X, u, v, cam = _lift(ht, wd, intr)
    out = np.zeros(len(ii))
    for e, (i, j) in enumerate(zip(ii, jj)):
        t, q = se3.relative(poses[i][None], poses[j][None])
        d = disps[i].reshape(-1)
        acc = valid = total = 0.0
        for Y, wgt in ((_move(t[0], q[0], X, d), beta), (X + t[0][:, None] * d[None], 1.0 - beta)):
            x, y = _proj(Y, cam)
            flow = np.sqrt((x - u) ** 2 + (y - v) ** 2)
            ok = Y[2] > MIN_DEPTH
            total += wgt * len(d)
            acc += wgt * flow[ok].sum()
            valid += wgt * ok.sum()
        out[e] = 1000.0 if valid / (total + 1e-8) < 0.75 else acc / valid
    return out


def covis_distance(poses, disps, intr, ii):
    poses = np.asarray(poses, np.float64)
    disps = np.asarray(disps, np.float64)
    N, ht, wd = disps.shape
    X, u, v, cam = _lift(ht, wd, intr)
    out = np.zeros(len(ii))
    for e, i in enumerate(ii):
        d = disps[i].reshape(-1)
        outside = np.ones(len(d), bool)
        for j in (i - 1, i + 1):
            t, q = se3.relative(poses[i][None], poses[j][None])
            x, y = _proj(_move(t[0], q[0], X, d), cam)
            with np.errstate(invalid='ignore'):
                outside &= (x < 0) | (x >= wd) | (y < 0) | (y >= ht)
        out[e] = outside.sum() / (ht * wd)
    return out


def depth_filter(poses, disps, intr, inds, thresh):
    poses = np.asarray(poses, np.float64)
    disps = np.asarray(disps, np.float64)
    N, ht, wd = disps.shape
    X, u, v, cam = _lift(ht, wd, intr)
    counter = np.zeros((len(inds), ht * wd))
    for b, i in enumerate(inds):
        d = disps[i].reshape(-1)
        for n in range(6):
            j = i - n - 1 if n < 3 else i + n
            if j < 0 or j >= N:
                continue
            t, q = se3.relative(poses[i][None], poses[j][None])
            Y = _move(t[0], q[0], X, d)
            x, y = _proj(Y, cam)
            dj = d / Y[2]
            u0, v0 = np.floor(x), np.floor(y)
            ok = (u0 >= 0) & (v0 >= 0) & (u0 < wd - 1) & (v0 < ht - 1)
            u0 = np.where(ok, u0, 0).astype(np.int64)
            v0 = np.where(ok, v0, 0).astype(np.int64)
            dn = disps[j]
            hit = np.zeros(len(d), bool)
            for dv, du in ((0, 0), (0, 1), (1, 0), (1, 1)):
                with np.errstate(divide='ignore', invalid='ignore'):
                    hit |= np.abs(1.0 / dj - 1.0 / dn[v0 + dv, u0 + du]) < thresh[b]
            counter[b] += ok & hit
    return counter.reshape(len(inds), ht, wd)


def iproj(poses, disps, intr):
    poses = np.asarray(poses, np.float64)
    disps = np.asarray(disps, np.float64)
    N, ht, wd = disps.shape
    X, u, v, cam = _lift(ht, wd, intr)
    pts = np.zeros((N, ht * wd, 3))
    for n in range(N):
        d = disps[n].reshape(-1)
        pts[n] = (_move(poses[n, :3], poses[n, 3:], X, d) / d[None]).T
    return pts.reshape(N, ht, wd, 3)
