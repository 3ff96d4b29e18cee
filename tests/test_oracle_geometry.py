"""The edge-selection / point-cloud geometry oracle (oracle/geometry.py) pinned by closed forms
and by an independent 4x4-matrix round trip."""
import numpy as np

from oracle import geometry as og
from oracle import se3
from hislam2_amd import synth


def _poses(n, t=None):
    p = np.zeros((n, 7))
    p[:, 6] = 1.0
    if t is not None:
        p[:, :3] = t
    return p


def test_iproj_round_trips_through_the_inverse_pose():
    q = synth.ba_small()
    pts = og.iproj(q.poses, q.disps, q.intrinsics)
    fx, fy, cx, cy = q.intrinsics
    N, h, w = q.disps.shape
    v, u = np.meshgrid(np.arange(h), np.arange(w), indexing='ij')
    for n in range(N):
        T = np.linalg.inv(se3.pose_to_mat(q.poses[n].astype(np.float64)))
        P = pts[n].reshape(-1, 3) @ T[:3, :3].T + T[:3, 3]
        assert np.allclose(fx * P[:, 0] / P[:, 2] + cx, u.reshape(-1), atol=1e-9)
        assert np.allclose(fy * P[:, 1] / P[:, 2] + cy, v.reshape(-1), atol=1e-9)
        assert np.allclose(1.0 / P[:, 2], q.disps[n].reshape(-1), rtol=1e-9)


def test_frame_distance_of_a_sideways_translation():
    d, tx, fx = 0.5, 0.3, 100.0
    intr = np.array([fx, fx, 16.0, 12.0])
    poses = _poses(2)
    poses[1, 0] = tx
    disps = np.full((2, 24, 32), d)
    for beta in (0.0, 0.7, 1.0):
        assert np.allclose(og.frame_distance(poses, disps, intr, [0], [1], beta), fx * d * tx, rtol=1e-12)
    poses[1, 2] = -3.5                      # every transformed depth 1 - 1.75 < 0.25 -> invalid
    assert og.frame_distance(poses, disps, intr, [0], [1], 0.5)[0] == 1000.0


def test_covis_distance_extremes():
    intr = np.array([50.0, 50.0, 16.0, 12.0])
    disps = np.full((3, 24, 32), 0.5)
    assert og.covis_distance(_poses(3), disps, intr, [1])[0] == 0.0
    far = _poses(3, np.array([[100.0, 0, 0], [0, 0, 0], [-100.0, 0, 0]]))
    assert og.covis_distance(far, disps, intr, [1])[0] == 1.0
    half = _poses(3, np.array([[100.0, 0, 0], [0, 0, 0], [0, 0, 0]]))   # only one neighbour loses them
    assert og.covis_distance(half, disps, intr, [1])[0] == 0.0


def test_depth_filter_counts_the_neighbours_in_range():
    """Identity poses and equal constant disparities: a pixel maps onto itself (intrinsics with
    exact arithmetic), so every neighbour in range agrees except on the last row and column."""
    N, h, w = 8, 10, 12
    intr = np.array([2.0, 2.0, 0.0, 0.0])
    disps = np.full((N, h, w), 0.4)
    c = og.depth_filter(_poses(N), disps, intr, [0, 4, 7], np.full(3, 1e-3))
    expect = {0: 3, 4: 4, 7: 3}                # offsets -1, -2, -3, +3, +4, +5 within [0, 8)
    for b, i in enumerate([0, 4, 7]):
        assert np.all(c[b, :-1, :-1] == expect[i])
        assert np.all(c[b, -1, :] == 0) and np.all(c[b, :, -1] == 0)
