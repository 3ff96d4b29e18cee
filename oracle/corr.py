"""Correlation-volume oracle (float64): all-pairs build, pyramid, lookup, alt lookup.

Restates
* ``CorrBlock.corr`` (hislam2/modules/corr.py:63-71): corr = (f1/4)^T (f2/4);
* the pyramid of ``CorrBlock.__init__`` (hislam2/modules/corr.py:24-38):
  repeated 2x2 average pooling with stride 2 (floor for odd sizes);
* ``corr_index_forward_kernel`` / ``corr_index_backward_kernel``
  (src/correlation_kernels.cu:19-124): radius-r bilinear lookup into each
  pixel's own correlation plane, zero outside, output order
  [n][a = x-offset][b = y-offset][y][x];
* ``CorrBlock.__call__`` (hislam2/modules/corr.py:40-50): levels concatenated
  into L*(2r+1)^2 channels, coordinates divided by 2^level;
* ``altcorr_cuda_forward`` / ``altcorr_cuda_backward``
  (src/altcorr_kernel.cu:24-225): the same lookup computed directly from the
  feature maps, without a stored volume;
* ``AltCorrBlock`` (hislam2/modules/corr.py:89-117).

Test infrastructure only.
"""
import numpy as np


def corr_allpairs(fmap1, fmap2):
    """[B,N,C,h,w] x [B,N,C,h,w] -> [B,N,h,w,h,w] (hislam2/modules/corr.py:63-71)."""
    f1 = np.asarray(fmap1, np.float64)
    f2 = np.asarray(fmap2, np.float64)
    B, N, C, h, w = f1.shape
    a = f1.reshape(B, N, C, h * w) / 4.0
    b = f2.reshape(B, N, C, h * w) / 4.0
    out = np.einsum('bncp,bncq->bnpq', a, b)
    return out.reshape(B, N, h, w, h, w)


def avg_pool2(x):
    """2x2 average pooling, stride 2, floor, over the last two axes."""
    h, w = x.shape[-2:]
    h2, w2 = h // 2, w // 2
    x = x[..., :2 * h2, :2 * w2]
    x = x.reshape(x.shape[:-2] + (h2, 2, w2, 2))
    return x.mean(axis=(-3, -1))


def pyramid(corr, num_levels=4):
    """List of [B*N, h1, w1, h2/2^i, w2/2^i] levels (hislam2/modules/corr.py:32-38)."""
    B, N, h1, w1, h2, w2 = corr.shape
    c = np.asarray(corr, np.float64).reshape(B * N, h1, w1, h2, w2)
    out = []
    for _ in range(num_levels):
        out.append(c)
        c = avg_pool2(c)
    return out


def _taps(coords, r):
    """Integer tap origins and fractional parts, as the kernels compute them: the
    coordinates are float32 and ``dx = x0 - floor(x0)`` is a float32 subtraction
    (src/correlation_kernels.cu:38-39), which rounds for x0 in (-1, 0)."""
    x0 = np.asarray(coords[:, 0], np.float32)
    y0 = np.asarray(coords[:, 1], np.float32)
    fx = np.floor(x0)
    fy = np.floor(y0)
    return fx.astype(np.int64) - r, fy.astype(np.int64) - r, x0 - fx, y0 - fy


def _weights(dx, dy):
    """Bilinear corner weights formed in float32 like the reference's
    ``scalar_t(dx * dy)`` etc. (src/correlation_kernels.cu:50-63), then widened."""
    one = np.float32(1.0)
    w = [(one - dx) * (one - dy), dx * (one - dy), (one - dx) * dy, dx * dy]
    return [np.asarray(v, np.float32).astype(np.float64) for v in w]


def corr_index_forward(volume, coords, r):
    """Radius-r lookup (src/correlation_kernels.cu:19-70).

    volume [BN,h1,w1,h2,w2], coords [BN,2,h1,w1] (x, y) ->
    out [BN, 2r+1 (x-offset a), 2r+1 (y-offset b), h1, w1]."""
    vol = np.asarray(volume, np.float64)
    BN, h1, w1, h2, w2 = vol.shape
    xb, yb, dx, dy = _taps(np.asarray(coords), r)
    D = 2 * r + 2
    n_idx, y_idx, x_idx = np.meshgrid(np.arange(BN), np.arange(h1), np.arange(w1), indexing='ij')
    tap = np.zeros((BN, D, D, h1, w1))
    for i in range(D):          # x offset
        for j in range(D):      # y offset
            x1 = xb + i
            y1 = yb + j
            ok = (x1 >= 0) & (x1 < w2) & (y1 >= 0) & (y1 < h2)
            vals = vol[n_idx, y_idx, x_idx, np.clip(y1, 0, h2 - 1), np.clip(x1, 0, w2 - 1)]
            tap[:, i, j] = np.where(ok, vals, 0.0)
    w00, w10, w01, w11 = _weights(dx, dy)
    return (tap[:, :-1, :-1] * w00[:, None, None] + tap[:, 1:, :-1] * w10[:, None, None]
            + tap[:, :-1, 1:] * w01[:, None, None] + tap[:, 1:, 1:] * w11[:, None, None])


def corr_index_backward(volume_shape, coords, grad, r):
    """Adjoint of corr_index_forward (src/correlation_kernels.cu:73-124)."""
    BN, h1, w1, h2, w2 = volume_shape
    g = np.asarray(grad, np.float64)
    xb, yb, dx, dy = _taps(np.asarray(coords), r)
    D = 2 * r + 2
    w00, w10, w01, w11 = _weights(dx, dy)
    gt = np.zeros((BN, D, D, h1, w1))
    gt[:, :-1, :-1] += g * w00[:, None, None]
    gt[:, 1:, :-1] += g * w10[:, None, None]
    gt[:, :-1, 1:] += g * w01[:, None, None]
    gt[:, 1:, 1:] += g * w11[:, None, None]
    out = np.zeros((BN, h1, w1, h2, w2))
    n_idx, y_idx, x_idx = np.meshgrid(np.arange(BN), np.arange(h1), np.arange(w1), indexing='ij')
    for i in range(D):
        for j in range(D):
            x1 = xb + i
            y1 = yb + j
            ok = (x1 >= 0) & (x1 < w2) & (y1 >= 0) & (y1 < h2)
            np.add.at(out, (n_idx[ok], y_idx[ok], x_idx[ok], y1[ok], x1[ok]), gt[:, i, j][ok])
    return out


def corrblock_lookup(pyr, coords, r):
    """CorrBlock.__call__ (hislam2/modules/corr.py:40-50).

    pyr: list of [B*N, h, w, h2_i, w2_i]; coords [B, N, h, w, 2] ->
    [B, N, L*(2r+1)^2, h, w]."""
    B, N, h, w, _ = coords.shape
    c = np.moveaxis(np.asarray(coords, np.float64), -1, 2).reshape(B * N, 2, h, w)
    outs = []
    for i, vol in enumerate(pyr):
        o = corr_index_forward(vol, c / 2 ** i, r)
        outs.append(o.reshape(B, N, -1, h, w))
    return np.concatenate(outs, axis=2)


def altcorr_forward(fmap1, fmap2, coords, ii, jj, r):
    """On-the-fly lookup (src/altcorr_kernel.cu:24-75, 132-172).

    fmap1 [B,N1,C,H,W], fmap2 [B,N2,C,H2,W2], coords [B,M,2,H,W],
    ii/jj [M] index fmap1/fmap2 -> [B, M, 2r+1 (x-offset), 2r+1 (y-offset), H, W]."""
    f1 = np.asarray(fmap1, np.float64) / 4.0
    f2 = np.asarray(fmap2, np.float64) / 4.0
    B, M, _, H, W = coords.shape
    H2, W2 = f2.shape[-2:]
    D = 2 * r + 2
    out = np.zeros((B, M, 2 * r + 1, 2 * r + 1, H, W))
    for b in range(B):
        for m in range(M):
            a = f1[b, ii[m]]                          # [C,H,W]
            g = f2[b, jj[m]]                          # [C,H2,W2]
            cx = np.asarray(coords[b, m, 0], np.float32)
            cy = np.asarray(coords[b, m, 1], np.float32)
            fx0, fy0 = np.floor(cx), np.floor(cy)
            w00, w10, w01, w11 = _weights(cx - fx0, cy - fy0)   # float32 dx, dy (altcorr_kernel.cu:160-163)
            tap = np.zeros((D, D, H, W))               # [y-offset][x-offset]
            for ty in range(D):
                for tx in range(D):
                    y1 = fy0.astype(np.int64) + ty - r
                    x1 = fx0.astype(np.int64) + tx - r
                    ok = (y1 >= 0) & (y1 < H2) & (x1 >= 0) & (x1 < W2)
                    gv = g[:, np.clip(y1, 0, H2 - 1), np.clip(x1, 0, W2 - 1)]
                    tap[ty, tx] = np.where(ok, np.sum(a * gv, axis=0), 0.0)
            o = w00 * tap[:-1, :-1] + w10 * tap[:-1, 1:] + w01 * tap[1:, :-1] + w11 * tap[1:, 1:]
            out[b, m] = np.swapaxes(o, 0, 1)          # permute(0,1,3,2,4,5)
    return out


def altcorr_backward(fmap1, fmap2, coords, grad, ii, jj, r):
    """Adjoint of altcorr_forward w.r.t. fmap1 and fmap2 (src/altcorr_kernel.cu:78-129, 175-225).

    grad [B, M, 2r+1 (x-offset), 2r+1 (y-offset), H, W] -> (g1, g2) shaped like
    fmap1 / fmap2.  The reference binding passes its arguments permuted
    (src/droid.cpp:107-114 vs src/altcorr_kernel.cu:175-182); this is the
    mathematically intended adjoint."""
    f1 = np.asarray(fmap1, np.float64)
    f2 = np.asarray(fmap2, np.float64)
    g1 = np.zeros_like(f1)
    g2 = np.zeros_like(f2)
    B, M, _, H, W = coords.shape
    H2, W2 = f2.shape[-2:]
    D = 2 * r + 2
    for b in range(B):
        for m in range(M):
            go = np.swapaxes(np.asarray(grad[b, m], np.float64), 0, 1)   # [y-off][x-off]
            cx = np.asarray(coords[b, m, 0], np.float32)
            cy = np.asarray(coords[b, m, 1], np.float32)
            fx0, fy0 = np.floor(cx), np.floor(cy)
            w00, w10, w01, w11 = _weights(cx - fx0, cy - fy0)
            gt = np.zeros((D, D, H, W))
            gt[:-1, :-1] += w00 * go
            gt[:-1, 1:] += w10 * go
            gt[1:, :-1] += w01 * go
            gt[1:, 1:] += w11 * go
            yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing='ij')
            for ty in range(D):
                for tx in range(D):
                    y1 = fy0.astype(np.int64) + ty - r
                    x1 = fx0.astype(np.int64) + tx - r
                    ok = (y1 >= 0) & (y1 < H2) & (x1 >= 0) & (x1 < W2)
                    s = gt[ty, tx][ok] / 16.0
                    g2v = f2[b, jj[m]][:, y1[ok], x1[ok]]
                    g1[b, ii[m]][:, yy[ok], xx[ok]] += s * g2v
                    np.add.at(g2[b, jj[m]], (slice(None), y1[ok], x1[ok]),
                              s * f1[b, ii[m]][:, yy[ok], xx[ok]])
    return g1, g2


def altcorr_block(fmaps, coords, ii, jj, r=3, num_levels=4):
    """AltCorrBlock (hislam2/modules/corr.py:89-117).

    fmaps [B,N,C,H,W]; coords [B,M,H,W,2] -> [B, M, L*(2r+1)^2, H, W]."""
    f = np.asarray(fmaps, np.float64)
    c = np.moveaxis(np.asarray(coords, np.float64), -1, 2)
    levels = [f]
    for _ in range(num_levels - 1):
        levels.append(avg_pool2(levels[-1]))
    outs = []
    for i in range(num_levels):
        o = altcorr_forward(f, levels[i], c / 2 ** i, ii, jj, r)
        B, M = o.shape[:2]
        outs.append(o.reshape(B, M, -1, *o.shape[-2:]))
    return np.concatenate(outs, axis=2)
