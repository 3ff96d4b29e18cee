"""GPU parity of droid_backends.frame_distance, covis_distance, depth_filter and iproj against
the float64 oracle (oracle/geometry.py) on the window problem (48x64).

Tolerances: frame_distance and iproj 1e-5 relative to the largest oracle value (FP32 pixel
arithmetic).  covis_distance and depth_filter count pixels through comparisons of FP32
projections; a pixel whose projection or depth lies within rounding of a bound may count
differently from the FP64 oracle, so at most 0.2 % of the pixels may differ, by at most one.
Index errors raise ValueError (out-of-range frames are undefined in the reference)."""
import numpy as np
import pytest
import torch

from hislam2_amd import synth
from oracle import geometry as og

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


@pytest.fixture(scope='module')
def db():
    from hislam2_amd import backend
    return backend()


@pytest.fixture(scope='module')
def prob():
    return synth.PROBLEMS['window']()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def rel_err(got, ref):
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref)) / max(np.max(np.abs(ref)), 1e-30))


def test_frame_distance(db, prob):
    for beta in (0.0, 0.7):
        d = db.frame_distance(dev(prob.poses), dev(prob.disps), dev(prob.intrinsics), dev(prob.ii), dev(prob.jj), beta)
        ref = og.frame_distance(prob.poses, prob.disps, prob.intrinsics, prob.ii, prob.jj, beta)
        assert rel_err(d.cpu().numpy(), ref) <= 1e-5


def test_covis_distance(db, prob):
    N = prob.poses.shape[0]
    ii = np.arange(1, N - 1)
    d = db.covis_distance(dev(prob.poses), dev(prob.disps), dev(prob.intrinsics), dev(ii))
    ref = og.covis_distance(prob.poses, prob.disps, prob.intrinsics, ii)
    assert np.max(np.abs(d.cpu().numpy() - ref)) <= 0.002
    with pytest.raises(ValueError):
        db.covis_distance(dev(prob.poses), dev(prob.disps), dev(prob.intrinsics), dev(np.array([0])))


def test_depth_filter(db, prob):
    N = prob.poses.shape[0]
    inds = np.arange(N)
    thresh = np.full(N, 0.05, np.float32)
    c = db.depth_filter(dev(prob.poses), dev(prob.disps), dev(prob.intrinsics), dev(inds), dev(thresh)).cpu().numpy()
    ref = og.depth_filter(prob.poses, prob.disps, prob.intrinsics, inds, thresh.astype(np.float64))
    diff = np.abs(c - ref)
    assert diff.max() <= 1 and (diff > 0).mean() <= 0.002
    assert ref.max() >= 1                     # the comparison is not vacuous


def test_iproj(db, prob):
    pts = db.iproj(dev(prob.poses), dev(prob.disps), dev(prob.intrinsics))
    ref = og.iproj(prob.poses, prob.disps, prob.intrinsics)
    assert pts.shape == ref.shape
    assert rel_err(pts.cpu().numpy(), ref) <= 1e-5
