// All-pairs correlation volume with a fused average-pool pyramid, on MFMA.
//
// Semantics of CorrBlock.__init__ / CorrBlock.corr (hislam2/modules/corr.py:24-38,
// 63-71): corr[n, p, q] = sum_c (f1[n, c, p] / 4) (f2[n, c, q] / 4) for pixels p of
// image 1 and q of image 2, then levels i = 1..L-1 by repeated 2x2 average pooling
// over q's (h, w) grid.
//
// One 256-thread block computes a 32 (p) x 8 image rows x WT columns (q) tile with
// v_mfma_f32_32x32x{8,8,2} (bf16 / f16 / f32 operands, f32 accumulation): the
// operands are staged through LDS in 32-channel chunks in [pixel][channel] order
// (padded rows), each wave owns four 32x32 output blocks.  The finished f32 tile is
// staged in LDS and written as level 0 plus every pooled level it fully contains
// (8 rows cover 2x2, 4x4 and 8x8 pooling), so the volume is written once and never
// read back.
#include "common.hpp"

namespace hsl2 {
namespace {

constexpr int kThreads = 256;
constexpr int kTP = 32;       // p rows per block
constexpr int kRows = 8;      // image-2 rows per block
constexpr int kKC = 32;       // channels per LDS chunk

typedef float acc16 __attribute__((ext_vector_type(16)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<__hip_bfloat16> {
  static constexpr int K = 8;
  static __device__ __forceinline__ acc16 run(const __hip_bfloat16* a, const __hip_bfloat16* b, acc16 c) {
    const bf16x4 av = *reinterpret_cast<const bf16x4*>(a);
    const bf16x4 bv = *reinterpret_cast<const bf16x4*>(b);
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, c, 0, 0, 0);
  }
};
template <> struct Mfma<__half> {
  static constexpr int K = 8;
  static __device__ __forceinline__ acc16 run(const __half* a, const __half* b, acc16 c) {
    const f16x4 av = *reinterpret_cast<const f16x4*>(a);
    const f16x4 bv = *reinterpret_cast<const f16x4*>(b);
    return __builtin_amdgcn_mfma_f32_32x32x8f16(av, bv, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  static constexpr int K = 2;
  static __device__ __forceinline__ acc16 run(const float* a, const float* b, acc16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(*a, *b, c, 0, 0, 0);
  }
};

// Offset of (y, x) in a plane of width w stored as row-major 8x8 tiles (the layout the
// pyramid lookup reads when asked for tiled volumes).
__device__ __forceinline__ int plane_offset(int y, int x, int w) {
  return (((y >> 3) * (w >> 3) + (x >> 3)) << 6) + ((y & 7) << 3) + (x & 7);
}

// grid: (p tiles, q tiles = (h/8) * (w/WT), batch)
template <typename Ti, typename To>
__global__ void __launch_bounds__(kThreads)
corr_build_kernel(const Ti* __restrict__ f1, const Ti* __restrict__ f2, To* __restrict__ l0,
                  To* __restrict__ l1, To* __restrict__ l2, To* __restrict__ l3, int C, int h,
                  int w, int WT, int levels, int tiled) {
  constexpr int KP = kKC + 4;                      // padded LDS row (elements)
  extern __shared__ __align__(16) unsigned char smem[];
  Ti* s1 = reinterpret_cast<Ti*>(smem);            // [kTP][KP]
  Ti* s2 = s1 + kTP * KP;                          // [kRows*WT][KP]
  float* sout = reinterpret_cast<float*>(smem);    // [kTP][kRows*WT] after the K loop

  const int HW = h * w;
  const int n = blockIdx.z;
  const int p0 = blockIdx.x * kTP;
  const int qtiles_x = w / WT;
  const int y0 = (blockIdx.y / qtiles_x) * kRows;
  const int x0 = (blockIdx.y % qtiles_x) * WT;
  const int NQ = kRows * WT;                       // columns of this tile
  const Ti* a = f1 + (int64_t)n * C * HW;
  const Ti* b = f2 + (int64_t)n * C * HW;
  const int t = threadIdx.x;
  const int wave = t / kWave, lane = t % kWave;
  const int nblk = NQ / 32;                        // 32-wide column blocks
  const int per_wave = nblk / 4;                   // NQ is a multiple of 128

  acc16 acc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[k][v] = 0.f;

  constexpr int EPL = 16 / sizeof(Ti);           // elements per 16-byte load
  struct alignas(16) Vec { Ti v[EPL]; };
  const int qpr = WT / EPL;                        // 16-byte loads per image-2 row of the tile
  const int nb_loads = kKC * NQ / EPL;             // B-tile loads per chunk (a multiple of 256)
  for (int c0 = 0; c0 < C; c0 += kKC) {
    // stage [pixel][channel] chunks with 16-byte loads along pixels; all of a thread's
    // loads are issued before the transposing LDS stores
    if (t < kKC * kTP / EPL) {
      const int c = t / (kTP / EPL), pv = t % (kTP / EPL);
      const Vec va = *reinterpret_cast<const Vec*>(a + (int64_t)(c0 + c) * HW + p0 + pv * EPL);
#pragma unroll
      for (int e = 0; e < EPL; ++e) s1[(pv * EPL + e) * KP + c] = va.v[e];
    }
    for (int base = 0; base < nb_loads; base += 8 * kThreads) {
      Vec vb[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = base + u * kThreads + t;
        if (idx < nb_loads) {
          const int c = idx / (NQ / EPL), qv = idx % (NQ / EPL);
          const int qy = y0 + qv / qpr, qx = x0 + (qv % qpr) * EPL;
          vb[u] = *reinterpret_cast<const Vec*>(b + (int64_t)(c0 + c) * HW + qy * w + qx);
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = base + u * kThreads + t;
        if (idx < nb_loads) {
          const int c = idx / (NQ / EPL), qv = idx % (NQ / EPL);
#pragma unroll
          for (int e = 0; e < EPL; ++e) s2[(qv * EPL + e) * KP + c] = vb[u].v[e];
        }
      }
    }
    __syncthreads();
    const int i = lane % 32;
    const int kg = lane / 32;
#pragma unroll
    for (int kk = 0; kk < kKC; kk += Mfma<Ti>::K) {
      // lanes 0-31 carry the first K/2 channels of the MFMA's K block, lanes 32-63 the rest
      const int ko = kk + kg * (Mfma<Ti>::K / 2);
      const Ti* ap = s1 + i * KP + ko;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k < per_wave) {
          const int q = (wave * per_wave + k) * 32 + i;
          acc[k] = Mfma<Ti>::run(ap, s2 + q * KP + ko, acc[k]);
        }
      }
    }
    __syncthreads();
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
  constexpr int KP = kKC + 4;
  const size_t ops = sizeof(Ti) * (kTP + kRows * WT) * KP;
  const size_t outb = sizeof(float) * kTP * kRows * WT;
  const size_t lds = ops > outb ? ops : outb;
  HSL2_REQUIRE(lds <= 160 * 1024, HSL2_ERR_UNSUPPORTED, "corr build: LDS budget exceeded");
  dim3 grid(static_cast<unsigned>(h * w / kTP), static_cast<unsigned>((h / kRows) * (w / WT)),
            static_cast<unsigned>(batch));
  if (batch == 0) return HSL2_OK;
  To* l[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int i = 0; i < L; ++i) l[i] = static_cast<To*>(levels[i]);
  hipLaunchKernelGGL((corr_build_kernel<Ti, To>), grid, dim3(kThreads), lds, s,
                     static_cast<const Ti*>(f1), static_cast<const Ti*>(f2), l[0], l[1], l[2], l[3],
                     static_cast<int>(C), static_cast<int>(h), static_cast<int>(w), WT, L, tiled);
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
  HSL2_REQUIRE(channels % kKC == 0, HSL2_ERR_UNSUPPORTED, "corr build: channels must be a multiple of 32");
  HSL2_REQUIRE(h % kRows == 0 && w % 8 == 0, HSL2_ERR_UNSUPPORTED,
               "corr build: h and w must be multiples of 8");
  HSL2_REQUIRE(h * w < (1ll << 31) && (h * w) % kTP == 0, HSL2_ERR_SHAPE, "corr build: bad plane size");
  HSL2_REQUIRE(batch < 65536, HSL2_ERR_SHAPE, "corr build: batch too large for one launch");
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
