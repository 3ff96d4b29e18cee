"""Writes the golden fixtures of the hot path from the float64 oracle.

    python tests/golden/make_fixtures.py

Each fixture stores its inputs and the oracle's outputs (float64), so the GPU
tests can compare against committed numbers without re-running the oracle, and
a CPU test re-derives them to catch any drift of the oracle itself.  The oracle
is pinned independently (closed forms, finite differences, brute force, torch
grid_sample / avg_pool2d) by tests/test_oracle_*.py.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, 'hi-slam2_amd')]

from hislam2_amd import synth          # noqa: E402
from oracle import ba as oba          # noqa: E402
from oracle import corr as oc         # noqa: E402


def ba_fixture():
    p = synth.ba_small(seed=11, h=24, w=32)
    poses, disps, dx, dz = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta,
                                  p.ii, p.jj, p.t0, p.t1, p.iterations, p.lm, p.ep)
    return dict(poses=p.poses, disps=p.disps, intrinsics=p.intrinsics, targets=p.targets,
                weights=p.weights, eta=p.eta, ii=p.ii, jj=p.jj,
                scalars=np.array([p.t0, p.t1, p.iterations, p.lm, p.ep]),
                out_poses=poses, out_disps=disps, out_dx=dx, out_dz=dz)


def ba_hires64_fixture():
    """First GN iteration of the 96x128 workload (64 keyframes, 800 edges): the inputs are
    regenerated from synth.ba_hires64() (seeded), so only input checksums, dx and dz at 4096
    seeded pixel positions are stored."""
    p = synth.ba_hires64()
    _, _, dx, dz = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta,
                          p.ii, p.jj, p.t0, p.t1, 1, p.lm, p.ep)
    idx = np.random.default_rng(0).choice(dz.size, size=4096, replace=False)
    return dict(checksums=np.array([p.poses.astype(np.float64).sum(), p.disps.astype(np.float64).sum(),
                                    p.targets.astype(np.float64).sum(), p.weights.astype(np.float64).sum()]),
                out_dx=dx, dz_index=idx, out_dz=dz.reshape(-1)[idx], dz_absmax=np.abs(dz).max())


def corr_fixture():
    fm = synth.fmaps(12, 2, c=32, h=8, w=16)
    coords = synth.lookup_coords(13, 1, h=8, w=16)
    pyr = oc.pyramid(oc.corr_allpairs(fm[:, :1], fm[:, 1:]), 4)
    out = oc.corrblock_lookup(pyr, coords.astype(np.float64), 3)
    return dict(fmap1=fm[:, :1], fmap2=fm[:, 1:], coords=coords, out=out)


def main():
    np.savez_compressed(os.path.join(HERE, 'ba_small_24x32.npz'), **ba_fixture())
    np.savez_compressed(os.path.join(HERE, 'corrblock_8x16.npz'), **corr_fixture())
    np.savez_compressed(os.path.join(HERE, 'ba_hires64_first_iteration.npz'), **ba_hires64_fixture())


if __name__ == '__main__':
    main()
