"""Correlation blocks over the gfx950 kernels (mirror of hislam2/modules/corr.py).

Same classes, constructor arguments and call conventions as the reference:
``CorrSampler`` (autograd wrapper of corr_index_forward/backward, corr.py:6-20),
``CorrBlock`` (all-pairs volume + 4-level pyramid + radius-3 lookup, corr.py:23-71),
``CorrLayer`` / ``AltCorrBlock`` (on-the-fly variant, corr.py:74-117).

Differences, all deliberate:
* the volume and its pyramid are built by one MFMA kernel (``corr_build``) that
  writes every level once, instead of matmul + three avg_pool2d passes; the
  volume dtype is selectable (``dtype=torch.bfloat16`` gives the split-precision
  path, float32 the plain path);
* ``CorrLayer.backward`` returns the gradient of the forward the binding computes
  (the reference unpacks 5 saved tensors into 3 names and reads 5 outputs from a
  2-output binding, corr.py:78-87, so it cannot run);
* the volume build is differentiable (``CorrBuild``) so that gradients reach the
  feature maps as they do through the reference's matmul + avg_pool2d.
"""
import torch
import torch.nn.functional as F

from . import backend

_db = backend()


class CorrSampler(torch.autograd.Function):
    """corr.py:6-20."""

    @staticmethod
    def forward(ctx, volume, coords, radius):
        ctx.save_for_backward(volume, coords)
        ctx.radius = radius
        corr, = _db.corr_index_forward(volume, coords, radius)
        return corr

    @staticmethod
    def backward(ctx, grad_output):
        volume, coords = ctx.saved_tensors
        grad_output = grad_output.contiguous()
        grad_volume, = _db.corr_index_backward(volume, coords, grad_output, ctx.radius)
        return grad_volume, None, None


class CorrBuild(torch.autograd.Function):
    """All-pairs correlation + pyramid (corr.py:24-38, 63-71) with its adjoint.

    Forward: the fused MFMA kernel.  Backward: the adjoint of the pooling
    (each level's gradient spread evenly over its 2^l x 2^l block) summed over
    levels, then the two GEMM adjoints of corr = f1^T f2 / 16 (library GEMMs).
    """

    @staticmethod
    def forward(ctx, fmap1, fmap2, num_levels, out_dtype):
        ctx.save_for_backward(fmap1, fmap2)
        levels = _db.corr_build(fmap1.contiguous(), fmap2.contiguous(), num_levels, out_dtype)
        return tuple(levels)

    @staticmethod
    def backward(ctx, *grads):
        fmap1, fmap2 = ctx.saved_tensors
        B, N, C, h, w = fmap1.shape
        g = None
        for i, gl in enumerate(grads):
            if gl is None:
                continue
            gl = gl.float().reshape(B * N * h * w, 1, h >> i, w >> i)
            if i > 0:
                gl = F.interpolate(gl, scale_factor=2 ** i, mode='nearest') / float(4 ** i)
            g = gl if g is None else g + gl
        g = g.reshape(B * N, h * w, h * w)
        f1 = fmap1.float().reshape(B * N, C, h * w)
        f2 = fmap2.float().reshape(B * N, C, h * w)
        g1 = torch.bmm(f2, g.transpose(1, 2).contiguous()) / 16.0     # [BN, C, p]
        g2 = torch.bmm(f1, g) / 16.0                                  # [BN, C, q]
        return (g1.reshape(fmap1.shape).to(fmap1.dtype), g2.reshape(fmap2.shape).to(fmap2.dtype),
                None, None)


def pyramid_planes(level, i):
    """Level ``i`` of a tiled CorrBlock pyramid ([B*N, h, w, h >> i, w >> i], the layout
    ``corr_build(..., tiled=True)`` writes) as the reference's row-major planes (a copy)."""
    n, h, w, h2, w2 = level.shape
    if i > 0 and (h * w) % 64 == 0:
        if i <= 2 and h2 % 2 == 0 and w2 % 4 == 0:
            # [group][y / 2][x / 4][pixel][y % 2][x % 4]
            g = level.reshape(n * h * w // 64, h2 // 2, w2 // 4, 64, 2, 4)
            return g.permute(0, 3, 1, 4, 2, 5).reshape(level.shape).contiguous()
        if w2 % 8 == 0:
            g = level.reshape(n * h * w // 64, h2, w2 // 8, 64, 8)   # [group][y][x / 8][pixel][x % 8]
            return g.permute(0, 3, 1, 2, 4).reshape(level.shape).contiguous()
    if h2 % 8 == 0 and w2 % 8 == 0:
        t = level.reshape(n, h, w, h2 // 8, w2 // 8, 8, 8)       # [..][y / 8][x / 8][y % 8][x % 8]
        return t.permute(0, 1, 2, 3, 5, 4, 6).reshape(level.shape).contiguous()
    return level.clone()


class CorrBlock:
    """corr.py:23-71.  ``dtype`` selects the volume element type (default: fmap dtype).

    Without autograd the levels are stored in the tiled pyramid layout (``self.tiled``):
    level 0 as 8x8 tiles per plane, the pooled levels with the 2x4 (levels 1, 2) or 1x8
    (further levels) segments of each 64 consecutive pixels interleaved (``pyramid_planes``
    converts back), which the
    fused lookup reads with far fewer memory transactions than row-major planes; with
    autograd (feature maps that require grad) the reference's row-major planes are
    kept so that ``CorrSampler`` and its adjoint apply.  ``cat``/``__getitem__`` act
    on the first dimension and work for both layouts."""

    def __init__(self, fmap1, fmap2, num_levels=4, radius=3, dtype=None, tiled=None):
        self.num_levels = num_levels
        self.radius = radius
        grad = torch.is_grad_enabled() and (fmap1.requires_grad or fmap2.requires_grad)
        self.tiled = (not grad) if tiled is None else bool(tiled)
        if grad:
            if self.tiled:
                raise ValueError('the tiled layout has no autograd path; use tiled=False')
            self.corr_pyramid = list(CorrBuild.apply(fmap1, fmap2, num_levels, dtype))
        else:
            self.corr_pyramid = _db.corr_build(fmap1.contiguous(), fmap2.contiguous(),
                                               num_levels, dtype, self.tiled)

    def __call__(self, coords):
        if not (torch.is_grad_enabled() and any(v.requires_grad for v in self.corr_pyramid)):
            # all levels in one launch, written straight into the concatenated output
            return _db.corr_pyramid_lookup(self.corr_pyramid, coords.float().contiguous(),
                                           self.radius, self.tiled)
        out_pyramid = []
        batch, num, ht, wd, _ = coords.shape
        coords = coords.permute(0, 1, 4, 2, 3)
        coords = coords.contiguous().view(batch * num, 2, ht, wd)
        for i in range(self.num_levels):
            corr = CorrSampler.apply(self.corr_pyramid[i], coords / 2 ** i, self.radius)
            out_pyramid.append(corr.view(batch, num, -1, ht, wd))
        return torch.cat(out_pyramid, dim=2)

    def lookup_conv1x1(self, coords, weight, bias, relu=True):
        """The lookup followed by a 1x1 convolution (+ ReLU), fused: equals
        ``relu(conv2d(self(coords).view(B*N, 196, h, w), weight, bias))`` viewed as
        [B, N, C_out, h, w] -- corr_encoder's first layer (hislam2/modules/droid_net.py:86-90)
        -- without writing the 196 lookup channels.  bf16 pyramids of 4 levels, radius 3, no
        autograd; weight [C_out, 196(, 1, 1)] with C_out 64 or 128, bias [C_out]."""
        if self.num_levels != 4 or self.radius != 3:
            raise ValueError('lookup_conv1x1: 4 levels of radius 3 (196 channels)')
        w = weight.detach().reshape(weight.shape[0], -1).to(torch.bfloat16).contiguous()
        b = bias.detach().float().contiguous()
        return _db.corr_lookup_conv(self.corr_pyramid, coords.float().contiguous(), self.radius, self.tiled,
                                    w, b, bool(relu))

    def cat(self, other):
        for i in range(self.num_levels):
            self.corr_pyramid[i] = torch.cat([self.corr_pyramid[i], other.corr_pyramid[i]], 0)
        return self

    def __getitem__(self, index):
        for i in range(self.num_levels):
            self.corr_pyramid[i] = self.corr_pyramid[i][index]
        return self

    @staticmethod
    def corr(fmap1, fmap2):
        """All-pairs correlation [B, N, h, w, h, w] (corr.py:63-71), level 0 of corr_build."""
        batch, num, dim, ht, wd = fmap1.shape
        lvl0, = _db.corr_build(fmap1.contiguous(), fmap2.contiguous(), 1, None, False)
        return lvl0.view(batch, num, ht, wd, ht, wd)


class CorrLayer(torch.autograd.Function):
    """corr.py:74-87 (backward repaired, see module docstring)."""

    @staticmethod
    def forward(ctx, fmap1, fmap2, coords, ii, jj, r):
        ctx.r = r
        ctx.save_for_backward(fmap1, fmap2, coords, ii, jj)
        corr, = _db.altcorr_forward(fmap1, fmap2, coords, ii, jj, ctx.r)
        return corr

    @staticmethod
    def backward(ctx, grad_corr):
        fmap1, fmap2, coords, ii, jj = ctx.saved_tensors
        fmap1_grad, fmap2_grad = _db.altcorr_backward(fmap1, fmap2, coords,
                                                      grad_corr.contiguous(), ii, jj, ctx.r)
        return fmap1_grad, fmap2_grad, None, None, None, None


_MFMA_DTYPES = (torch.float32, torch.bfloat16, torch.float16)


class AltCorrBlock:
    """corr.py:89-117.

    Without autograd, for 128-channel maps, radius 3 or 4 and up to 4 levels (the reference's
    configuration), the feature pyramid is built once, channel-last (``altcorr_pyramid``), and
    every call runs all levels in one launch on the matrix cores (``altcorr_pyramid_forward``),
    writing the stacked [B, M, L (2r+1)^2, H, W] output directly.  ``dtype`` selects the element
    type of that pyramid and of the output (default: the maps' dtype): bfloat16 / float16 /
    float32.  Otherwise the reference's per-level ``CorrLayer`` path runs (with autograd).

    ``dtype=torch.float8_e4m3fn`` (an addition of this build for high-resolution frames) keeps
    every pyramid level as e4m3fn codes with per-channel FP32 scales (``fp8_quantize``, the
    levels pooled in FP32 first) and runs the lookup on them (``altcorr_forward_fp8``); the
    output is bf16.  That path has no autograd."""

    def __init__(self, fmaps, num_levels=4, radius=3, dtype=None):
        self.num_levels = num_levels
        self.radius = radius
        self.fp8 = dtype == torch.float8_e4m3fn
        B, N, C, H, W = fmaps.shape
        grad = torch.is_grad_enabled() and fmaps.requires_grad
        dt = fmaps.dtype if dtype is None else dtype
        self.mfma = (not self.fp8 and not grad and C == 128 and radius in (3, 4) and 1 <= num_levels <= 4
                     and dt in _MFMA_DTYPES and fmaps.dtype in _MFMA_DTYPES)
        if self.mfma:
            self.levels = _db.altcorr_pyramid(fmaps.detach(), num_levels, dt)
            return
        if dtype is not None and not self.fp8:
            fmaps = fmaps.to(dtype)
        if self.fp8:
            if torch.is_grad_enabled() and fmaps.requires_grad:
                raise ValueError('AltCorrBlock: the fp8 path has no autograd')
            f = fmaps.float().reshape(B * N, C, H, W)
            self.levels = []
            for i in range(self.num_levels):
                lvl = f.reshape(B, N, C, H // 2 ** i, W // 2 ** i).contiguous()
                self.levels.append(_db.fp8_quantize(lvl))
                f = F.avg_pool2d(f, 2, stride=2)
            return
        fmaps = fmaps.view(B * N, C, H, W)
        self.pyramid = []
        for i in range(self.num_levels):
            sz = (B, N, C, H // 2 ** i, W // 2 ** i)
            self.pyramid.append(fmaps.view(*sz))
            fmaps = F.avg_pool2d(fmaps, 2, stride=2)

    def __call__(self, coords, ii, jj):
        if self.mfma:
            return _db.altcorr_pyramid_forward(self.levels, coords.detach().float().contiguous(), ii, jj,
                                               self.radius)
        coords = coords.permute(0, 1, 4, 2, 3).contiguous()
        corr_list = []
        for i in range(self.num_levels):
            if self.fp8:
                (q1, s1), (q2, s2) = self.levels[0], self.levels[i]
                corr, = _db.altcorr_forward_fp8(q1, s1, q2, s2, (coords / 2 ** i).contiguous(),
                                                ii.contiguous(), jj.contiguous(), self.radius)
            else:
                corr = CorrLayer.apply(self.pyramid[0], self.pyramid[i].contiguous(),
                                       coords / 2 ** i, ii, jj, self.radius)
            corr_list.append(corr.flatten(2, 3))
        corr = torch.stack(corr_list, dim=2).flatten(2, 3)
        return corr
