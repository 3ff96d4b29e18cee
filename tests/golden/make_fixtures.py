"""Writes the golden fixtures of the hot path from the float64 oracle.

    python tests/golden/make_fixtures.py [name ...]      (all fixtures when no name is given)

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


def _checksums(p):
    return np.array([p.poses.astype(np.float64).sum(), p.disps.astype(np.float64).sum(),
                     p.targets.astype(np.float64).sum(), p.weights.astype(np.float64).sum()])


def ba_global300_fixture():
    """The benchmarked call of the 300-keyframe / 3000-edge workload: all 8 GN iterations
    (t0 = 0, t1 = 300, lm = 1e-4, ep = 0.1).  Inputs are regenerated from synth.ba_global300()
    (seeded; checksums stored).  Stored: the final poses in full, the final disparities at 8192
    seeded positions, the last iteration's dx in full and dz at the same kind of positions, and
    the first iteration's dx (a separate one-iteration oracle run)."""
    p = synth.ba_global300()
    poses, disps, dx, dz = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta,
                                  p.ii, p.jj, p.t0, p.t1, p.iterations, p.lm, p.ep)
    _, _, dx1, dz1 = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta,
                            p.ii, p.jj, p.t0, p.t1, 1, p.lm, p.ep)
    rng = np.random.default_rng(1)
    di = rng.choice(disps.size, size=8192, replace=False)
    zi = rng.choice(dz.size, size=8192, replace=False)
    return dict(checksums=_checksums(p), iterations=np.array(p.iterations),
                out_poses=poses, disps_index=di, out_disps=disps.reshape(-1)[di],
                out_dx=dx, dz_index=zi, out_dz=dz.reshape(-1)[zi], dz_absmax=np.abs(dz).max(),
                out_dx_it1=dx1, out_dz_it1=dz1.reshape(-1)[zi], dz1_absmax=np.abs(dz1).max())


GLOBAL300_LOOKUP_EDGES = np.array([0, 1, 777, 1499, 2000, 2998, 2999])


def corr_global300_fixture():
    """The bf16 lookup of the 300-keyframe workload reads a pyramid of all 3000 edges
    (fmaps synth.fmaps(7, 300) indexed by the edges of synth.ba_global300(), coords
    synth.lookup_coords(8, 3000)); the oracle gives, for the edges GLOBAL300_LOOKUP_EDGES
    (first, middle and last, whose volume offsets pass 2^31 elements), the 196 channels at 128
    seeded pixels each."""
    p = synth.ba_global300()
    fm = synth.fmaps(7, p.poses.shape[0])
    coords = synth.lookup_coords(8, len(p.ii))
    rng = np.random.default_rng(2)
    h, w = fm.shape[-2:]
    pix = np.stack([np.sort(rng.choice(h * w, size=128, replace=False)) for _ in GLOBAL300_LOOKUP_EDGES])
    out = []
    for e, px in zip(GLOBAL300_LOOKUP_EDGES, pix):
        f1 = fm[:, p.ii[e]:p.ii[e] + 1].astype(np.float64)
        f2 = fm[:, p.jj[e]:p.jj[e] + 1].astype(np.float64)
        pyr = oc.pyramid(oc.corr_allpairs(f1, f2), 4)
        o = oc.corrblock_lookup(pyr, coords[:, e:e + 1].astype(np.float64), 3)   # [1, 1, 196, h, w]
        out.append(o[0, 0].reshape(196, h * w)[:, px])
    return dict(edges=GLOBAL300_LOOKUP_EDGES, pixels=pix, out=np.stack(out).astype(np.float32),
                fm_checksum=np.array(fm.astype(np.float64).sum()),
                coords_checksum=np.array(coords.astype(np.float64).sum()))


def corr_fixture():
    fm = synth.fmaps(12, 2, c=32, h=8, w=16)
    coords = synth.lookup_coords(13, 1, h=8, w=16)
    pyr = oc.pyramid(oc.corr_allpairs(fm[:, :1], fm[:, 1:]), 4)
    out = oc.corrblock_lookup(pyr, coords.astype(np.float64), 3)
    return dict(fmap1=fm[:, :1], fmap2=fm[:, 1:], coords=coords, out=out)


FIXTURES = {
    'ba_small_24x32': ba_fixture,
    'corrblock_8x16': corr_fixture,
    'ba_hires64_first_iteration': ba_hires64_fixture,
    'ba_global300_8_iterations': ba_global300_fixture,
    'corrblock_global300_bf16': corr_global300_fixture,
}


def main(names):
    for name in names or FIXTURES:
        np.savez_compressed(os.path.join(HERE, name + '.npz'), **FIXTURES[name]())


if __name__ == '__main__':
    main(sys.argv[1:])
