// All-pairs correlation volume with a fused average-pool pyramid, on MFMA.
//
// Semantics of CorrBlock.__init__ / CorrBlock.corr (hislam2/modules/corr.py:24-38,
// 63-71): corr[n, p, q] = sum_c (f1[n, c, p] / 4) (f2[n, c, q] / 4) for pixels p of
// image 1 and q of image 2, then levels i = 1..L-1 by repeated 2x2 average pooling
// over q's (h, w) grid.
//
// One 256-thread block computes a 32 (p) x 8 image rows x WT columns (q) tile with
// v_mfma_f32_32x32x{8,8,2} (bf16 / f16 / f32 operands, f32 accumulation).  The feature maps
// arrive channel-last ([pixel][channel]), so every lane fetches its MFMA operand fragment
// (4 consecutive channels of one pixel: 8 bytes; 1 channel for f32) straight from memory:
// no LDS staging and no barrier in the channel loop.  Each wave owns four 32x32 output
// blocks.  Blocks are ordered so that each XCD works through whole edges (the operands of
// one edge stay in that XCD's L2).  The finished f32 tile is staged in LDS and written as
// level 0 plus every pooled level it fully contains (8 rows cover 2x2, 4x4 and 8x8 pooling),
// so the volume is written once and never read back.
#include "common.hpp"

namespace hsl2 {
namespace {

constexpr int kThreads = 256;
constexpr int kTP = 32;       // p rows per block
constexpr int kRows = 8;      // image-2 rows per block

typedef float acc16 __attribute__((ext_vector_type(16)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<__hip_bfloat16> {
  static constexpr int K = 8;                      // channels per instruction
  typedef bf16x4 Frag;
  static __device__ __forceinline__ Frag load(const __hip_bfloat16* p) { return *reinterpret_cast<const bf16x4*>(p); }
  static __device__ __forceinline__ acc16 run(Frag a, Frag b, acc16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<__half> {
  static constexpr int K = 8;
  typedef f16x4 Frag;
  static __device__ __forceinline__ Frag load(const __half* p) { return *reinterpret_cast<const f16x4*>(p); }
  static __device__ __forceinline__ acc16 run(Frag a, Frag b, acc16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  static constexpr int K = 2;
  typedef float Frag;
  static __device__ __forceinline__ Frag load(const float* p) { return *p; }
  static __device__ __forceinline__ acc16 run(Frag a, Frag b, acc16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
};

// Offset of (y, x) in a plane of width w stored as row-major 8x8 tiles (the layout the
// pyramid lookup reads when asked for tiled volumes).
__device__ __forceinline__ int plane_offset(int y, int x, int w) {
  return (((y >> 3) * (w >> 3) + (x >> 3)) << 6) + ((y & 7) << 3) + (x & 7);
}

// 1-D grid over (p tile, q tile, edge), p tile fastest; see the block order note above.
template <typename Ti, typename To>
__global__ void __launch_bounds__(kThreads)
corr_build_kernel(const Ti* __restrict__ f1, const Ti* __restrict__ f2, To* __restrict__ l0,
                  To* __restrict__ l1, To* __restrict__ l2, To* __restrict__ l3, int C, int h,
                  int w, int WT, int levels, int tiled, int ptiles, int qtiles, int64_t nblocks) {
  extern __shared__ __align__(16) unsigned char smem[];
  float* sout = reinterpret_cast<float*>(smem);    // [kTP][kRows*WT]

  // XCD-aware order: hardware block b runs on XCD b % 8; give each XCD a contiguous range of
  // logical blocks (whole edges), so an edge's feature maps are fetched into one L2 only.
  const int64_t b = blockIdx.x;
  const int64_t per = nblocks / 8;
  const int64_t lb = b < per * 8 ? (b % 8) * per + b / 8 : b;
  const int pt = static_cast<int>(lb % ptiles);
  const int qt = static_cast<int>((lb / ptiles) % qtiles);
  const int n = static_cast<int>(lb / ((int64_t)ptiles * qtiles));

  const int HW = h * w;
  const int p0 = pt * kTP;
  const int qtiles_x = w / WT;
  const int y0 = (qt / qtiles_x) * kRows;
  const int x0 = (qt % qtiles_x) * WT;
  const int NQ = kRows * WT;                       // columns of this tile
  const int t = threadIdx.x;
  const int wave = t / kWave, lane = t % kWave;
  const int per_wave = NQ / 32 / 4;                // NQ is a multiple of 128
  constexpr int K = Mfma<Ti>::K;
  // lane i of the 32-row MFMA operand; lanes 32-63 carry the second half of the K block
  const int i = lane % 32, kg = lane / 32;
  const Ti* ap = f1 + ((int64_t)n * HW + p0 + i) * C + kg * (K / 2);
  const Ti* bp[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int q = (wave * per_wave + (k < per_wave ? k : 0)) * 32 + i;
    const int qpix = (y0 + q / WT) * w + x0 + q % WT;
    bp[k] = f2 + ((int64_t)n * HW + qpix) * C + kg * (K / 2);
  }

  acc16 acc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[k][v] = 0.f;
#pragma unroll 4
  for (int kk = 0; kk < C; kk += K) {
    const typename Mfma<Ti>::Frag a = Mfma<Ti>::load(ap + kk);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < per_wave) acc[k] = Mfma<Ti>::run(a, Mfma<Ti>::load(bp[k] + kk), acc[k]);
  }

  // scatter the accumulators into the LDS tile [p][q] (scaled by 1/16)
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k < per_wave) {
      const int q = (wave * per_wave + k) * 32 + lane % 32;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int p = 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
        sout[p * NQ + q] = acc[k][v] * 0.0625f;
      }
    }
  }
  __syncthreads();

  // level 0: groups of 8 elements that are consecutive in memory (a row segment of the
  // tile, or one row of an 8x8 block in the tiled layout), one vector store each
  const bool tile0 = tiled && (h % 8 == 0) && (w % 8 == 0);
  struct alignas(8 * sizeof(To)) Out8 { To v[8]; };
  for (int g = t; g < kTP * NQ / 8; g += kThreads) {
    const int p = g / (NQ / 8), q8 = g % (NQ / 8);
    int qy, qx;
    if (tile0) {            // consecutive groups walk the 8x8 blocks of the 8-row band in memory order
      qy = y0 + (q8 & 7);
      qx = x0 + (q8 >> 3) * 8;
    } else {
      qy = y0 + (q8 * 8) / WT;
      qx = x0 + (q8 * 8) % WT;
    }
    const float* src = sout + p * NQ + (qy - y0) * WT + (qx - x0);
    Out8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o.v[e] = from_f<To>(src[e]);
    const int off = tile0 ? plane_offset(qy, qx, w) : qy * w + qx;
    *reinterpret_cast<Out8*>(l0 + ((int64_t)n * HW + p0 + p) * HW + off) = o;
  }
  // pooled levels: level L averages 2^L x 2^L blocks of level 0
  To* outs[3] = {l1, l2, l3};
  for (int L = 1; L < levels; ++L) {
    const int s = 1 << L;
    const int hl = h >> L, wl = w >> L;
    const int ry = kRows >> L, rx = WT >> L;       // pooled rows / cols inside the tile
    const int cnt = ry * rx;
    const float inv = 1.f / static_cast<float>(s * s);
    for (int idx = t; idx < kTP * cnt; idx += kThreads) {
      const int p = idx / cnt, rem = idx % cnt;
      const int yy = rem / rx, xx = rem % rx;
      float sum = 0.f;
      for (int dy = 0; dy < s; ++dy)
        for (int dx = 0; dx < s; ++dx) sum += sout[p * NQ + (yy * s + dy) * WT + xx * s + dx];
      const int gy = (y0 >> L) + yy, gx = (x0 >> L) + xx;
      const bool tl = tiled && (hl % 8 == 0) && (wl % 8 == 0);
      outs[L - 1][((int64_t)n * HW + p0 + p) * hl * wl + (tl ? plane_offset(gy, gx, wl) : gy * wl + gx)] =
          from_f<To>(sum * inv);
    }
  }
}

template <typename Ti, typename To>
int launch(const void* f1, const void* f2, void* const* levels, int64_t batch, int64_t C,
           int64_t h, int64_t w, int L, int tiled, hipStream_t s) {
  const int WT = w % 64 == 0 ? 64 : static_cast<int>(w);
  HSL2_REQUIRE((kRows * WT) % 128 == 0, HSL2_ERR_UNSUPPORTED,
               "corr build: width must be a multiple of 64 or 16..63 with 8*w % 128 == 0");
  const size_t lds = sizeof(float) * kTP * kRows * WT;
  HSL2_REQUIRE(lds <= 160 * 1024, HSL2_ERR_UNSUPPORTED, "corr build: LDS budget exceeded");
  const int ptiles = static_cast<int>(h * w / kTP);
  const int qtiles = static_cast<int>((h / kRows) * (w / WT));
  const int64_t nblocks = (int64_t)ptiles * qtiles * batch;
  if (nblocks == 0) return HSL2_OK;
  HSL2_REQUIRE(nblocks < (1ll << 31), HSL2_ERR_SHAPE, "corr build: grid too large");
  To* l[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int i = 0; i < L; ++i) l[i] = static_cast<To*>(levels[i]);
  hipLaunchKernelGGL((corr_build_kernel<Ti, To>), dim3(static_cast<unsigned>(nblocks)), dim3(kThreads), lds, s,
                     static_cast<const Ti*>(f1), static_cast<const Ti*>(f2), l[0], l[1], l[2], l[3],
                     static_cast<int>(C), static_cast<int>(h), static_cast<int>(w), WT, L, tiled, ptiles,
                     qtiles, nblocks);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

}  // namespace
}  // namespace hsl2

using namespace hsl2;

extern "C" int hsl2_corr_build(const void* fmap1, const void* fmap2, int in_dtype,
                               void* const* levels, int out_dtype, int64_t batch, int64_t channels,
                               int64_t h, int64_t w, int num_levels, int tiled, void* stream) {
  HSL2_REQUIRE(fmap1 && fmap2 && levels, HSL2_ERR_ARGUMENT, "corr build: null pointer");
  HSL2_REQUIRE(num_levels >= 1 && num_levels <= 4, HSL2_ERR_ARGUMENT, "corr build: 1..4 levels");
  for (int i = 0; i < num_levels; ++i)
    HSL2_REQUIRE(levels[i], HSL2_ERR_ARGUMENT, "corr build: null level pointer");
  HSL2_REQUIRE(batch >= 0 && channels > 0 && h > 0 && w > 0, HSL2_ERR_SHAPE, "corr build: bad sizes");
  HSL2_REQUIRE(channels % 8 == 0, HSL2_ERR_UNSUPPORTED, "corr build: channels must be a multiple of 8");
  HSL2_REQUIRE(h % kRows == 0 && w % 8 == 0, HSL2_ERR_UNSUPPORTED,
               "corr build: h and w must be multiples of 8");
  HSL2_REQUIRE(h * w < (1ll << 31) && (h * w) % kTP == 0, HSL2_ERR_SHAPE, "corr build: bad plane size");
  hipStream_t s = as_stream(stream);
  if (out_dtype == HSL2_BF16) {
    if (in_dtype == HSL2_BF16) return launch<__hip_bfloat16, __hip_bfloat16>(fmap1, fmap2, levels, batch, channels, h, w, num_levels, tiled, s);
    if (in_dtype == HSL2_F16) return launch<__half, __hip_bfloat16>(fmap1, fmap2, levels, batch, channels, h, w, num_levels, tiled, s);
    if (in_dtype == HSL2_F32) return launch<float, __hip_bfloat16>(fmap1, fmap2, levels, batch, channels, h, w, num_levels, tiled, s);
  } else if (out_dtype == HSL2_F32) {
    if (in_dtype == HSL2_BF16) return launch<__hip_bfloat16, float>(fmap1, fmap2, levels, batch, channels, h, w, num_levels, tiled, s);
    if (in_dtype == HSL2_F16) return launch<__half, float>(fmap1, fmap2, levels, batch, channels, h, w, num_levels, tiled, s);
    if (in_dtype == HSL2_F32) return launch<float, float>(fmap1, fmap2, levels, batch, channels, h, w, num_levels, tiled, s);
  } else if (out_dtype == HSL2_F16) {
    if (in_dtype == HSL2_F16) return launch<__half, __half>(fmap1, fmap2, levels, batch, channels, h, w, num_levels, tiled, s);
  }
  set_error("corr build: unsupported dtype combination");
  return HSL2_ERR_UNSUPPORTED;
}
