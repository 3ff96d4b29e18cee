// On-the-fly correlation lookup (no stored volume) and its adjoint.
//
// Semantics of altcorr_cuda_forward / corr_forward_kernel
// (src/altcorr_kernel.cu:24-75, 132-172): for edge m and pixel (y, x) of fmap1
// frame ii[m], the (2r+2)^2 taps s(ty, tx) = sum_c (f1/4)(f2/4) at
// (floor(y) - r + ty, floor(x) - r + tx) of fmap2 frame jj[m] (zero outside) are
// combined bilinearly; the output is [B, M, 2r+1 (x-offset), 2r+1 (y-offset), H, W].
//
// The kernels read channel-last feature maps [B, N, H, W, C] so that one tap's
// channel vector is contiguous; hsl2_channels_last converts the reference layout.
// A 256-thread block owns 64 consecutive pixels of one edge: the 64 fmap1 vectors
// are staged in LDS, each wave runs lanes over taps (one dot product per lane,
// 16-byte channel loads), and the bilinear combine writes [a][b][pixel] rows
// coalesced.  Sums are in float.
#include "common.hpp"

namespace hsl2 {
namespace {

constexpr int kPix = 64;
constexpr int kThreads = 256;

// ---- layout conversion [rows][C][P] -> [rows][P][C] (kTiles: fragment tiles) ----------------
// Offset of (pixel p, channel c) in one row of the converted layout.
template <typename T, bool kTiles>
__device__ __forceinline__ int64_t cl_offset(int p, int c, int C) {
  if constexpr (kTiles) return frag_tile_offset<T>(p, c, C);
  return (int64_t)p * C + c;
}

template <typename T, bool kTiles>
__global__ void __launch_bounds__(256)
channels_last_kernel(const T* __restrict__ in, T* __restrict__ out, int C, int P) {
  __shared__ T tile[32][33];
  const int64_t row = blockIdx.z;
  const int c0 = blockIdx.y * 32, p0 = blockIdx.x * 32;
  const T* src = in + row * (int64_t)C * P;
  T* dst = out + row * (int64_t)C * P;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;   // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + k, p = p0 + tx;
    if (c < C && p < P) tile[k][tx] = src[(int64_t)c * P + p];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int p = p0 + k, c = c0 + tx;
    if (c < C && p < P) dst[cl_offset<T, kTiles>(p, c, C)] = tile[tx][k];
  }
}

// The same conversion for P and C multiples of 64 and 16-byte aligned buffers: one block per
// 64 pixels x 64 channels, 16-byte loads along the pixels of a channel (one 128-byte line per
// channel row in bf16) and 16-byte stores along the channels of a pixel, transposed in LDS.
template <typename T, bool kTiles>
__global__ void __launch_bounds__(256)
channels_last_vec_kernel(const T* __restrict__ in, T* __restrict__ out, int C, int P) {
  constexpr int VE = 16 / sizeof(T);               // elements per 16-byte vector
  constexpr int LD = 64 + VE;                      // padded LDS row (pixel-major)
  __shared__ T tile[64][LD];                       // [pixel][channel]
  const int64_t row = blockIdx.z;
  const int c0 = blockIdx.y * 64, p0 = blockIdx.x * 64;
  const T* src = in + row * (int64_t)C * P;
  T* dst = out + row * (int64_t)C * P;
  const int t = threadIdx.x;
  constexpr int NV = 64 * 64 / VE;                 // vectors per tile
#pragma unroll
  for (int i = t; i < NV; i += 256) {
    const int c = i / (64 / VE), pv = (i % (64 / VE)) * VE;
    const uint4 v = *reinterpret_cast<const uint4*>(src + (int64_t)(c0 + c) * P + p0 + pv);
    const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
    for (int k = 0; k < VE; ++k) tile[pv + k][c] = e[k];
  }
  __syncthreads();
#pragma unroll
  for (int i = t; i < NV; i += 256) {
    const int p = i / (64 / VE), cv = (i % (64 / VE)) * VE;
    uint4 v;
    T* e = reinterpret_cast<T*>(&v);
#pragma unroll
    for (int k = 0; k < VE; ++k) e[k] = tile[p][cv + k];
    *reinterpret_cast<uint4*>(dst + cl_offset<T, kTiles>(p0 + p, c0 + cv, C)) = v;
  }
}

// [rows][P][C] f32 -> [rows][C][P] f32 (gradients back to the reference layout)
__global__ void __launch_bounds__(256)
channels_first_kernel(const float* __restrict__ in, float* __restrict__ out, int C, int P) {
  __shared__ float tile[32][33];
  const int64_t row = blockIdx.z;
  const int p0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const float* src = in + row * (int64_t)C * P;
  float* dst = out + row * (int64_t)C * P;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int k = ty; k < 32; k += 8) {
    const int p = p0 + k, c = c0 + tx;
    if (c < C && p < P) tile[k][tx] = src[(int64_t)p * C + c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + k, p = p0 + tx;
    if (c < C && p < P) dst[(int64_t)c * P + p] = tile[tx][k];
  }
}

template <typename T> struct Vec8 { T v[8]; };

template <typename T>
__device__ __forceinline__ void load8(const T* p, float out[8]) {
  const Vec8<T> x = *reinterpret_cast<const Vec8<T>*>(p);
#pragma unroll
  for (int k = 0; k < 8; ++k) out[k] = to_f<T>(x.v[k]);
}

struct PixCoord {
  int x0, y0;
  float dx, dy;
};

template <typename T>
__global__ void __launch_bounds__(kThreads)
altcorr_fwd_kernel(const T* __restrict__ f1, const T* __restrict__ f2,
                   const float* __restrict__ coords, const int64_t* __restrict__ ii,
                   const int64_t* __restrict__ jj, T* __restrict__ out, int N1, int N2, int C,
                   int HW, int W, int H2, int W2, int M, int r, int tiles) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int D = 2 * r + 2, D2 = D * D, R = 2 * r + 1;
  float* s_f1 = reinterpret_cast<float*>(smem_raw);          // [kPix][C]
  float* s_tap = s_f1 + kPix * C;                            // [kPix][D2]
  __shared__ PixCoord s_c[kPix];

  const int tile = blockIdx.x % tiles;
  const int bm = blockIdx.x / tiles;                         // b * M + m
  const int m = bm % M;
  const int b = bm / M;
  const int pix0 = tile * kPix;
  const int npix = min(kPix, HW - pix0);
  const int t = threadIdx.x;
  const int64_t i1 = ii[m], j2 = jj[m];
  // an edge naming a frame outside the buffers reads nothing and yields zeros
  const bool edge_ok = i1 >= 0 && i1 < N1 && j2 >= 0 && j2 < N2;
  const T* a = f1 + (((int64_t)b * N1 + (edge_ok ? i1 : 0)) * HW + pix0) * C;
  const T* g = f2 + ((int64_t)b * N2 + (edge_ok ? j2 : 0)) * (int64_t)H2 * W2 * C;

  for (int k = t; k < npix * C; k += kThreads) s_f1[k] = edge_ok ? to_f<T>(a[k]) : 0.f;
  if (t < kPix) {
    PixCoord pc{0, 0, 0.f, 0.f};
    if (t < npix) {
      const float* cb = coords + (int64_t)bm * 2 * HW;
      const float cx = cb[pix0 + t], cy = cb[HW + pix0 + t];
      const float fx = floorf(cx), fy = floorf(cy);
      pc.x0 = static_cast<int>(fx) - r;
      pc.y0 = static_cast<int>(fy) - r;
      pc.dx = cx - fx;
      pc.dy = cy - fy;
    }
    s_c[t] = pc;
  }
  __syncthreads();

  const int wave = t / kWave, lane = t % kWave;
  for (int p = wave; p < npix; p += kThreads / kWave) {
    const PixCoord pc = s_c[p];
    const float* av = s_f1 + p * C;
    for (int tap = lane; tap < D2; tap += kWave) {
      const int ty = tap / D, tx = tap - ty * D;
      const int y1 = pc.y0 + ty, x1 = pc.x0 + tx;
      float acc = 0.f;
      if (edge_ok && y1 >= 0 && y1 < H2 && x1 >= 0 && x1 < W2) {
        const T* gv = g + ((int64_t)y1 * W2 + x1) * C;
        for (int c = 0; c < C; c += 8) {
          float bv[8];
          load8<T>(gv + c, bv);
#pragma unroll
          for (int k = 0; k < 8; ++k) acc += av[c + k] * bv[k];
        }
      }
      s_tap[p * D2 + tap] = acc * 0.0625f;
    }
  }
  __syncthreads();

  T* ob = out + (int64_t)bm * R * R * HW + pix0;
  for (int idx = t; idx < npix * R * R; idx += kThreads) {
    const int o = idx / npix;
    const int p = idx - o * npix;
    const int ax = o / R, by = o - ax * R;
    const float fx = s_c[p].dx, fy = s_c[p].dy;
    const float* tp = s_tap + p * D2;
    const float v = (1 - fx) * (1 - fy) * tp[by * D + ax] + fx * (1 - fy) * tp[by * D + ax + 1] +
                    (1 - fx) * fy * tp[(by + 1) * D + ax] + fx * fy * tp[(by + 1) * D + ax + 1];
    ob[(int64_t)o * HW + p] = from_f<T>(v);
  }
}

// Adjoint.  g1, g2: channel-last f32 accumulators (must be zeroed by the caller).
template <typename T>
__global__ void __launch_bounds__(kThreads)
altcorr_bwd_kernel(const T* __restrict__ f1, const T* __restrict__ f2,
                   const float* __restrict__ coords, const float* __restrict__ grad,
                   const int64_t* __restrict__ ii, const int64_t* __restrict__ jj,
                   float* __restrict__ g1, float* __restrict__ g2, int N1, int N2, int C, int HW,
                   int W, int H2, int W2, int M, int r, int tiles) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int D = 2 * r + 2, D2 = D * D, R = 2 * r + 1;
  float* s_gt = reinterpret_cast<float*>(smem_raw);          // [kPix][D2]
  __shared__ PixCoord s_c[kPix];
  const int tile = blockIdx.x % tiles;
  const int bm = blockIdx.x / tiles;
  const int m = bm % M;
  const int b = bm / M;
  const int pix0 = tile * kPix;
  const int npix = min(kPix, HW - pix0);
  const int t = threadIdx.x;
  const int64_t i1 = ii[m], j2 = jj[m];
  if (!(i1 >= 0 && i1 < N1 && j2 >= 0 && j2 < N2)) return;   // whole block exits together
  if (t < kPix) {
    PixCoord pc{0, 0, 0.f, 0.f};
    if (t < npix) {
      const float* cb = coords + (int64_t)bm * 2 * HW;
      const float cx = cb[pix0 + t], cy = cb[HW + pix0 + t];
      const float fx = floorf(cx), fy = floorf(cy);
      pc.x0 = static_cast<int>(fx) - r;
      pc.y0 = static_cast<int>(fy) - r;
      pc.dx = cx - fx;
      pc.dy = cy - fy;
    }
    s_c[t] = pc;
  }
  __syncthreads();
  const float* gb = grad + (int64_t)bm * R * R * HW + pix0;
  for (int idx = t; idx < npix * D2; idx += kThreads) {
    const int p = idx % npix;
    const int tap = idx / npix;
    const int ty = tap / D, tx = tap - ty * D;
    const float fx = s_c[p].dx, fy = s_c[p].dy;
    float v = 0.f;
    // output (a = x-offset, b = y-offset) at o = a * R + b reads tap (b + j, a + i)
    if (tx < R && ty < R) v += (1 - fx) * (1 - fy) * gb[(int64_t)(tx * R + ty) * HW + p];
    if (tx > 0 && ty < R) v += fx * (1 - fy) * gb[(int64_t)((tx - 1) * R + ty) * HW + p];
    if (tx < R && ty > 0) v += (1 - fx) * fy * gb[(int64_t)(tx * R + ty - 1) * HW + p];
    if (tx > 0 && ty > 0) v += fx * fy * gb[(int64_t)((tx - 1) * R + ty - 1) * HW + p];
    s_gt[p * D2 + tap] = v * 0.0625f;
  }
  __syncthreads();

  const T* a = f1 + ((int64_t)b * N1 + i1) * HW * C;
  const T* g = f2 + ((int64_t)b * N2 + j2) * (int64_t)H2 * W2 * C;
  float* ga = g1 + ((int64_t)b * N1 + i1) * HW * C;
  float* gg = g2 + ((int64_t)b * N2 + j2) * (int64_t)H2 * W2 * C;
  const int wave = t / kWave, lane = t % kWave;
  for (int p = wave; p < npix; p += kThreads / kWave) {
    const PixCoord pc = s_c[p];
    const int pix = pix0 + p;
    for (int c = lane; c < C; c += kWave) {
      const float av = to_f<T>(a[(int64_t)pix * C + c]);
      float acc = 0.f;
      for (int tap = 0; tap < D2; ++tap) {
        const int ty = tap / D, tx = tap - ty * D;
        const int y1 = pc.y0 + ty, x1 = pc.x0 + tx;
        if (y1 < 0 || y1 >= H2 || x1 < 0 || x1 >= W2) continue;
        const float gt = s_gt[p * D2 + tap];
        const int64_t q = ((int64_t)y1 * W2 + x1) * C + c;
        acc += gt * to_f<T>(g[q]);
        atomicAdd(gg + q, gt * av);
      }
      atomicAdd(ga + (int64_t)pix * C + c, acc);
    }
  }
}

// ---- fp8 (OCP e4m3fn, gfx950's FP8 format) with per-channel FP32 scales -------------------
// Quantisation: scale_c = max |x_c| / 448 over every vector of the tensor (1 where the channel
// is all zero), q = e4m3(x / scale_c) rounded to nearest even by the hardware convert.
typedef float f32x2 __attribute__((ext_vector_type(2)));
constexpr float kFp8Max = 448.f;

__device__ __forceinline__ void fp8x4_to_f32(uint32_t w, float o[4]) {
  const f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(w), false);
  const f32x2 hi = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(w), true);
  o[0] = lo[0];
  o[1] = lo[1];
  o[2] = hi[0];
  o[3] = hi[1];
}

__device__ __forceinline__ float fp8_scale(uint32_t amax_bits) {
  const float a = __uint_as_float(amax_bits);
  return a > 0.f ? a / kFp8Max : 1.f;
}

__global__ void fill_u32(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// per-channel max |x| over `vectors` channel-last vectors of C channels; one block per chunk
template <typename T>
__global__ void __launch_bounds__(256)
fp8_amax_kernel(const T* __restrict__ x, int64_t vectors, int C, int64_t chunk, uint32_t* __restrict__ amax) {
  const int64_t v0 = (int64_t)blockIdx.x * chunk;
  const int64_t v1 = min(vectors, v0 + chunk);
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    float m = 0.f;
    for (int64_t v = v0; v < v1; ++v) m = fmaxf(m, fabsf(to_f<T>(x[v * C + c])));
    // non-negative floats order like their bit patterns
    atomicMax(amax + c, __float_as_uint(m));
  }
}

// four consecutive channels per thread -> one 32-bit store
template <typename T>
__global__ void __launch_bounds__(256)
fp8_quantize_kernel(const T* __restrict__ x, int64_t n4, int C, const uint32_t* __restrict__ amax,
                    uint32_t* __restrict__ q) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  const int c = static_cast<int>((i * 4) % C);
  float v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = to_f<T>(x[i * 4 + k]) / fp8_scale(amax[c + k]);
  int w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], w, true);
  q[i] = static_cast<uint32_t>(w);
}

__global__ void fp8_scales_kernel(uint32_t* __restrict__ amax_then_scale, int C) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) amax_then_scale[c] = __float_as_uint(fp8_scale(amax_then_scale[c]));
}

// altcorr forward on fp8 feature maps: fmap1 vectors are dequantised into LDS with both
// channel scales folded in (s1_c s2_c), the taps of fmap2 are read 16 channels (16 bytes) at a
// time and converted by the hardware, the bilinear combine is that of altcorr_fwd_kernel.
template <typename To>
__global__ void __launch_bounds__(kThreads)
altcorr_fwd_fp8_kernel(const uint8_t* __restrict__ f1, const float* __restrict__ s1,
                       const uint8_t* __restrict__ f2, const float* __restrict__ s2,
                       const float* __restrict__ coords, const int64_t* __restrict__ ii,
                       const int64_t* __restrict__ jj, To* __restrict__ out, int N1, int N2, int C,
                       int HW, int W, int H2, int W2, int M, int r, int tiles) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int D = 2 * r + 2, D2 = D * D, R = 2 * r + 1;
  float* s_f1 = reinterpret_cast<float*>(smem_raw);          // [kPix][C]
  float* s_tap = s_f1 + kPix * C;                            // [kPix][D2]
  __shared__ PixCoord s_c[kPix];

  const int tile = blockIdx.x % tiles;
  const int bm = blockIdx.x / tiles;
  const int m = bm % M;
  const int b = bm / M;
  const int pix0 = tile * kPix;
  const int npix = min(kPix, HW - pix0);
  const int t = threadIdx.x;
  const int64_t i1 = ii[m], j2 = jj[m];
  const bool edge_ok = i1 >= 0 && i1 < N1 && j2 >= 0 && j2 < N2;
  const uint8_t* a = f1 + (((int64_t)b * N1 + (edge_ok ? i1 : 0)) * HW + pix0) * C;
  const uint8_t* g = f2 + ((int64_t)b * N2 + (edge_ok ? j2 : 0)) * (int64_t)H2 * W2 * C;

  for (int k = 4 * t; k < npix * C; k += 4 * kThreads) {
    float v[4] = {0.f, 0.f, 0.f, 0.f};
    if (edge_ok) fp8x4_to_f32(*reinterpret_cast<const uint32_t*>(a + k), v);
    const int c = k % C;
#pragma unroll
    for (int u = 0; u < 4; ++u) s_f1[k + u] = v[u] * (s1[c + u] * s2[c + u]);
  }
  if (t < kPix) {
    PixCoord pc{0, 0, 0.f, 0.f};
    if (t < npix) {
      const float* cb = coords + (int64_t)bm * 2 * HW;
      const float cx = cb[pix0 + t], cy = cb[HW + pix0 + t];
      const float fx = floorf(cx), fy = floorf(cy);
      pc.x0 = static_cast<int>(fx) - r;
      pc.y0 = static_cast<int>(fy) - r;
      pc.dx = cx - fx;
      pc.dy = cy - fy;
    }
    s_c[t] = pc;
  }
  __syncthreads();

  const int wave = t / kWave, lane = t % kWave;
  for (int p = wave; p < npix; p += kThreads / kWave) {
    const PixCoord pc = s_c[p];
    const float* av = s_f1 + p * C;
    for (int tap = lane; tap < D2; tap += kWave) {
      const int ty = tap / D, tx = tap - ty * D;
      const int y1 = pc.y0 + ty, x1 = pc.x0 + tx;
      float acc = 0.f;
      if (edge_ok && y1 >= 0 && y1 < H2 && x1 >= 0 && x1 < W2) {
        const uint8_t* gv = g + ((int64_t)y1 * W2 + x1) * C;
        for (int c = 0; c < C; c += 16) {
          const uint4 w4 = *reinterpret_cast<const uint4*>(gv + c);
          float bv[16];
          fp8x4_to_f32(w4.x, bv);
          fp8x4_to_f32(w4.y, bv + 4);
          fp8x4_to_f32(w4.z, bv + 8);
          fp8x4_to_f32(w4.w, bv + 12);
#pragma unroll
          for (int k = 0; k < 16; ++k) acc += av[c + k] * bv[k];
        }
      }
      s_tap[p * D2 + tap] = acc * 0.0625f;
    }
  }
  __syncthreads();

  To* ob = out + (int64_t)bm * R * R * HW + pix0;
  for (int idx = t; idx < npix * R * R; idx += kThreads) {
    const int o = idx / npix;
    const int p = idx - o * npix;
    const int ax = o / R, by = o - ax * R;
    const float fx = s_c[p].dx, fy = s_c[p].dy;
    const float* tp = s_tap + p * D2;
    const float v = (1 - fx) * (1 - fy) * tp[by * D + ax] + fx * (1 - fy) * tp[by * D + ax + 1] +
                    (1 - fx) * fy * tp[(by + 1) * D + ax] + fx * fy * tp[(by + 1) * D + ax + 1];
    ob[(int64_t)o * HW + p] = from_f<To>(v);
  }
}

int check_alt(const void* f1, const void* f2, const float* coords, const int64_t* ii,
              const int64_t* jj, int64_t B, int64_t N1, int64_t N2, int64_t C, int64_t H,
              int64_t W, int64_t H2, int64_t W2, int64_t M, int radius) {
  HSL2_REQUIRE(f1 && f2 && coords && ii && jj, HSL2_ERR_ARGUMENT, "altcorr: null pointer");
  HSL2_REQUIRE(B >= 0 && N1 > 0 && N2 > 0 && C > 0 && H > 0 && W > 0 && H2 > 0 && W2 > 0 && M >= 0,
               HSL2_ERR_SHAPE, "altcorr: bad sizes");
  HSL2_REQUIRE(C % 8 == 0 && C <= 512, HSL2_ERR_UNSUPPORTED,
               "altcorr: channel count must be a multiple of 8 and at most 512");
  HSL2_REQUIRE(radius >= 0 && radius <= 6, HSL2_ERR_ARGUMENT, "altcorr: radius must be in [0, 6]");
  HSL2_REQUIRE(B * M * ((H * W + kPix - 1) / kPix) < (1ll << 31), HSL2_ERR_SHAPE, "altcorr: grid too large");
  return HSL2_OK;
}

}  // namespace
}  // namespace hsl2

using namespace hsl2;

namespace hsl2 {
namespace {

template <bool kTiles>
int convert_channel_major(const void* in, void* out, int dtype, int64_t rows, int64_t channels, int64_t pixels,
                          void* stream, const char* what) {
  HSL2_REQUIRE(in && out, HSL2_ERR_ARGUMENT, std::string(what) + ": null pointer");
  HSL2_REQUIRE(rows >= 0 && channels > 0 && pixels > 0 && rows < 65536 && channels * pixels < (1ll << 31),
               HSL2_ERR_SHAPE, std::string(what) + ": bad sizes");
  HSL2_REQUIRE(dtype == HSL2_F32 || dtype == HSL2_F16 || dtype == HSL2_BF16, HSL2_ERR_ARGUMENT,
               std::string(what) + ": unsupported dtype");
  if (kTiles) {
    const int G = dtype == HSL2_F32 ? frag_group<float>() : frag_group<__half>();
    HSL2_REQUIRE(pixels % kFragPixels == 0 && channels % G == 0, HSL2_ERR_UNSUPPORTED,
                 std::string(what) + ": pixels must be a multiple of 32 and channels of 8 (f32) or 16");
  }
  if (rows == 0) return HSL2_OK;
  hipStream_t s = as_stream(stream);
  const int C = static_cast<int>(channels), P = static_cast<int>(pixels);
  const bool vec = channels % 64 == 0 && pixels % 64 == 0 && reinterpret_cast<uintptr_t>(in) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(out) % 16 == 0;
  if (vec) {
    const dim3 g(static_cast<unsigned>(pixels / 64), static_cast<unsigned>(channels / 64), static_cast<unsigned>(rows));
    switch (dtype) {
      case HSL2_F32: hipLaunchKernelGGL((channels_last_vec_kernel<float, kTiles>), g, dim3(256), 0, s, static_cast<const float*>(in), static_cast<float*>(out), C, P); break;
      case HSL2_F16: hipLaunchKernelGGL((channels_last_vec_kernel<__half, kTiles>), g, dim3(256), 0, s, static_cast<const __half*>(in), static_cast<__half*>(out), C, P); break;
      default: hipLaunchKernelGGL((channels_last_vec_kernel<__hip_bfloat16, kTiles>), g, dim3(256), 0, s, static_cast<const __hip_bfloat16*>(in), static_cast<__hip_bfloat16*>(out), C, P); break;
    }
  } else {
    const dim3 g(static_cast<unsigned>((pixels + 31) / 32), static_cast<unsigned>((channels + 31) / 32),
                 static_cast<unsigned>(rows));
    switch (dtype) {
      case HSL2_F32: hipLaunchKernelGGL((channels_last_kernel<float, kTiles>), g, dim3(256), 0, s, static_cast<const float*>(in), static_cast<float*>(out), C, P); break;
      case HSL2_F16: hipLaunchKernelGGL((channels_last_kernel<__half, kTiles>), g, dim3(256), 0, s, static_cast<const __half*>(in), static_cast<__half*>(out), C, P); break;
      default: hipLaunchKernelGGL((channels_last_kernel<__hip_bfloat16, kTiles>), g, dim3(256), 0, s, static_cast<const __hip_bfloat16*>(in), static_cast<__hip_bfloat16*>(out), C, P); break;
    }
  }
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

}  // namespace
}  // namespace hsl2

extern "C" int hsl2_channels_last(const void* in, void* out, int dtype, int64_t rows,
                                  int64_t channels, int64_t pixels, void* stream) {
  return convert_channel_major<false>(in, out, dtype, rows, channels, pixels, stream, "channels_last");
}

extern "C" int hsl2_fmap_tiles(const void* in, void* out, int dtype, int64_t rows, int64_t channels,
                               int64_t pixels, void* stream) {
  return convert_channel_major<true>(in, out, dtype, rows, channels, pixels, stream, "fmap_tiles");
}

extern "C" int hsl2_channels_first_f32(const float* in, float* out, int64_t rows,
                                       int64_t channels, int64_t pixels, void* stream) {
  HSL2_REQUIRE(in && out, HSL2_ERR_ARGUMENT, "channels_first: null pointer");
  HSL2_REQUIRE(rows >= 0 && channels > 0 && pixels > 0 && rows < 65536, HSL2_ERR_SHAPE,
               "channels_first: bad sizes");
  if (rows == 0) return HSL2_OK;
  dim3 grid(static_cast<unsigned>((channels + 31) / 32), static_cast<unsigned>((pixels + 31) / 32),
            static_cast<unsigned>(rows));
  hipLaunchKernelGGL(channels_first_kernel, grid, dim3(256), 0, as_stream(stream), in, out,
                     static_cast<int>(channels), static_cast<int>(pixels));
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" int hsl2_altcorr_forward(const void* fmap1, const void* fmap2, const float* coords,
                                    const int64_t* ii, const int64_t* jj, void* corr, int dtype,
                                    int64_t B, int64_t N1, int64_t N2, int64_t C, int64_t H,
                                    int64_t W, int64_t H2, int64_t W2, int64_t M, int radius,
                                    void* stream) {
  int st = check_alt(fmap1, fmap2, coords, ii, jj, B, N1, N2, C, H, W, H2, W2, M, radius);
  if (st) return st;
  HSL2_REQUIRE(corr, HSL2_ERR_ARGUMENT, "altcorr: null output");
  const int HW = static_cast<int>(H * W);
  const int tiles = (HW + kPix - 1) / kPix;
  const int64_t blocks = B * M * tiles;
  if (blocks == 0) return HSL2_OK;
  const int D = 2 * radius + 2;
  const size_t lds = sizeof(float) * kPix * (C + D * D);
  HSL2_REQUIRE(lds <= 160 * 1024, HSL2_ERR_UNSUPPORTED, "altcorr: LDS budget exceeded");
  hipStream_t s = as_stream(stream);
#define HSL2_ALT_FWD(T)                                                                          \
  hipLaunchKernelGGL(altcorr_fwd_kernel<T>, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), \
                     lds, s, static_cast<const T*>(fmap1), static_cast<const T*>(fmap2), coords,  \
                     ii, jj, static_cast<T*>(corr), static_cast<int>(N1), static_cast<int>(N2),   \
                     static_cast<int>(C), HW, static_cast<int>(W), static_cast<int>(H2),          \
                     static_cast<int>(W2), static_cast<int>(M), radius, tiles)
  switch (dtype) {
    case HSL2_F32: HSL2_ALT_FWD(float); break;
    case HSL2_F16: HSL2_ALT_FWD(__half); break;
    case HSL2_BF16: HSL2_ALT_FWD(__hip_bfloat16); break;
    default: set_error("altcorr: unsupported dtype"); return HSL2_ERR_ARGUMENT;
  }
#undef HSL2_ALT_FWD
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" int hsl2_altcorr_backward(const void* fmap1, const void* fmap2, const float* coords,
                                     const float* corr_grad, const int64_t* ii, const int64_t* jj,
                                     float* fmap1_grad, float* fmap2_grad, int dtype, int64_t B,
                                     int64_t N1, int64_t N2, int64_t C, int64_t H, int64_t W,
                                     int64_t H2, int64_t W2, int64_t M, int radius, void* stream) {
  int st = check_alt(fmap1, fmap2, coords, ii, jj, B, N1, N2, C, H, W, H2, W2, M, radius);
  if (st) return st;
  HSL2_REQUIRE(corr_grad && fmap1_grad && fmap2_grad, HSL2_ERR_ARGUMENT, "altcorr: null pointer");
  hipStream_t s = as_stream(stream);
  HSL2_HIP(hipMemsetAsync(fmap1_grad, 0, sizeof(float) * B * N1 * C * H * W, s));
  HSL2_HIP(hipMemsetAsync(fmap2_grad, 0, sizeof(float) * B * N2 * C * H2 * W2, s));
  const int HW = static_cast<int>(H * W);
  const int tiles = (HW + kPix - 1) / kPix;
  const int64_t blocks = B * M * tiles;
  if (blocks == 0) return HSL2_OK;
  const int D = 2 * radius + 2;
  const size_t lds = sizeof(float) * kPix * D * D;
#define HSL2_ALT_BWD(T)                                                                          \
  hipLaunchKernelGGL(altcorr_bwd_kernel<T>, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), \
                     lds, s, static_cast<const T*>(fmap1), static_cast<const T*>(fmap2), coords,  \
                     corr_grad, ii, jj, fmap1_grad, fmap2_grad, static_cast<int>(N1),             \
                     static_cast<int>(N2), static_cast<int>(C), HW, static_cast<int>(W),          \
                     static_cast<int>(H2), static_cast<int>(W2), static_cast<int>(M), radius, tiles)
  switch (dtype) {
    case HSL2_F32: HSL2_ALT_BWD(float); break;
    case HSL2_F16: HSL2_ALT_BWD(__half); break;
    case HSL2_BF16: HSL2_ALT_BWD(__hip_bfloat16); break;
    default: set_error("altcorr: unsupported dtype"); return HSL2_ERR_ARGUMENT;
  }
#undef HSL2_ALT_BWD
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" int hsl2_fp8_quantize(const void* x, int dtype, int64_t vectors, int64_t channels, uint8_t* q,
                                 float* scales, void* stream) {
  HSL2_REQUIRE(x && q && scales, HSL2_ERR_ARGUMENT, "fp8_quantize: null pointer");
  HSL2_REQUIRE(vectors >= 0 && channels > 0 && channels % 4 == 0 && channels <= 4096, HSL2_ERR_SHAPE,
               "fp8_quantize: channels must be a positive multiple of 4 (at most 4096)");
  hipStream_t s = as_stream(stream);
  const int C = static_cast<int>(channels);
  uint32_t* amax = reinterpret_cast<uint32_t*>(scales);
  hipLaunchKernelGGL(fill_u32, dim3((C + 255) / 256), dim3(256), 0, s, amax, C);
  HSL2_LAUNCH_CHECK();
  if (vectors > 0) {
    const int64_t chunk = 64;
    const int64_t blocks = (vectors + chunk - 1) / chunk;
    HSL2_REQUIRE(blocks < (1ll << 31), HSL2_ERR_SHAPE, "fp8_quantize: too many vectors");
    const int64_t n4 = vectors * C / 4;
    const unsigned qb = static_cast<unsigned>((n4 + 255) / 256);
    switch (dtype) {
      case HSL2_F32:
        hipLaunchKernelGGL(fp8_amax_kernel<float>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                           static_cast<const float*>(x), vectors, C, chunk, amax);
        hipLaunchKernelGGL(fp8_quantize_kernel<float>, dim3(qb), dim3(256), 0, s, static_cast<const float*>(x),
                           n4, C, amax, reinterpret_cast<uint32_t*>(q));
        break;
      case HSL2_BF16:
        hipLaunchKernelGGL(fp8_amax_kernel<__hip_bfloat16>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                           static_cast<const __hip_bfloat16*>(x), vectors, C, chunk, amax);
        hipLaunchKernelGGL(fp8_quantize_kernel<__hip_bfloat16>, dim3(qb), dim3(256), 0, s,
                           static_cast<const __hip_bfloat16*>(x), n4, C, amax, reinterpret_cast<uint32_t*>(q));
        break;
      default: set_error("fp8_quantize: unsupported dtype (f32 or bf16)"); return HSL2_ERR_ARGUMENT;
    }
    HSL2_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(fp8_scales_kernel, dim3((C + 255) / 256), dim3(256), 0, s, amax, C);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" int hsl2_altcorr_forward_fp8(const uint8_t* fmap1, const float* scale1, const uint8_t* fmap2,
                                        const float* scale2, const float* coords, const int64_t* ii,
                                        const int64_t* jj, void* corr, int out_dtype, int64_t B, int64_t N1,
                                        int64_t N2, int64_t C, int64_t H, int64_t W, int64_t H2, int64_t W2,
                                        int64_t M, int radius, void* stream) {
  int st = check_alt(fmap1, fmap2, coords, ii, jj, B, N1, N2, C, H, W, H2, W2, M, radius);
  if (st) return st;
  HSL2_REQUIRE(corr && scale1 && scale2, HSL2_ERR_ARGUMENT, "altcorr fp8: null pointer");
  HSL2_REQUIRE(C % 16 == 0, HSL2_ERR_UNSUPPORTED, "altcorr fp8: channel count must be a multiple of 16");
  const int HW = static_cast<int>(H * W);
  const int tiles = (HW + kPix - 1) / kPix;
  const int64_t blocks = B * M * tiles;
  if (blocks == 0) return HSL2_OK;
  const int D = 2 * radius + 2;
  const size_t lds = sizeof(float) * kPix * (C + D * D);
  HSL2_REQUIRE(lds <= 160 * 1024, HSL2_ERR_UNSUPPORTED, "altcorr fp8: LDS budget exceeded");
  hipStream_t s = as_stream(stream);
#define HSL2_ALT_FP8(To)                                                                              \
  hipLaunchKernelGGL(altcorr_fwd_fp8_kernel<To>, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), lds, \
                     s, fmap1, scale1, fmap2, scale2, coords, ii, jj, static_cast<To*>(corr),              \
                     static_cast<int>(N1), static_cast<int>(N2), static_cast<int>(C), HW,                  \
                     static_cast<int>(W), static_cast<int>(H2), static_cast<int>(W2), static_cast<int>(M),  \
                     radius, tiles)
  switch (out_dtype) {
    case HSL2_F32: HSL2_ALT_FP8(float); break;
    case HSL2_BF16: HSL2_ALT_FP8(__hip_bfloat16); break;
    default: set_error("altcorr fp8: output dtype must be f32 or bf16"); return HSL2_ERR_ARGUMENT;
  }
#undef HSL2_ALT_FP8
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}
