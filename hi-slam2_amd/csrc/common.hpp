// Shared helpers for the gfx950 kernels: status plumbing, element conversions,
// wave64 reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <string>

#include "hsl2.h"

namespace hsl2 {

void set_error(const std::string& msg);

#define HSL2_REQUIRE(cond, code, msg)          \
  do {                                         \
    if (!(cond)) {                             \
      ::hsl2::set_error(msg);                  \
      return (code);                           \
    }                                          \
  } while (0)

#define HSL2_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::hsl2::set_error(std::string(#call " failed: ") + hipGetErrorString(e_));    \
      return HSL2_ERR_HIP;                                                          \
    }                                                                               \
  } while (0)

#define HSL2_LAUNCH_CHECK() HSL2_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kWave = 64;

// ---- element <-> float -------------------------------------------------------
template <typename T> __device__ __forceinline__ float to_f(T v);
template <> __device__ __forceinline__ float to_f<float>(float v) { return v; }
template <> __device__ __forceinline__ float to_f<double>(double v) { return static_cast<float>(v); }
template <> __device__ __forceinline__ float to_f<__half>(__half v) { return __half2float(v); }
template <> __device__ __forceinline__ float to_f<__hip_bfloat16>(__hip_bfloat16 v) {
  return __bfloat162float(v);
}

template <typename T> __device__ __forceinline__ T from_f(float v);
template <> __device__ __forceinline__ float from_f<float>(float v) { return v; }
template <> __device__ __forceinline__ double from_f<double>(float v) { return v; }
template <> __device__ __forceinline__ __half from_f<__half>(float v) { return __float2half(v); }
template <> __device__ __forceinline__ __hip_bfloat16 from_f<__hip_bfloat16>(float v) {
  return __float2bfloat16(v);
}

// ---- wave64 reductions ---------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Reduce V per-thread values over a block of NT threads (NT a multiple of 64, <= 1024)
// through LDS in chunks of 16 values, in the type T of `lds` and `out` (the per-thread values
// are converted on the way in); the block writes the V sums to out[0..V) (plain stores).  `lds` must hold 16 * NT elements.  Every thread of the block must call this.
template <int V, int NT, typename T, typename S>
__device__ void block_reduce_store(const S (&vals)[V], T* lds, T* out) {
  constexpr int C = 16;
  const int t = threadIdx.x;
#pragma unroll
  for (int c0 = 0; c0 < V; c0 += C) {
#pragma unroll
    for (int c = 0; c < C; ++c)
      if (c0 + c < V) lds[c * NT + t] = static_cast<T>(vals[c0 + c]);
    __syncthreads();
    // NT threads: value c = t / (NT/16), segment = t % (NT/16)
    constexpr int SEG = NT / C;                 // threads per value
    const int c = t / SEG;
    const int s = t % SEG;
    T acc = 0;
    if (c0 + c < V) {
      for (int k = s; k < NT; k += SEG) acc += lds[c * NT + k];
    }
    // reduce SEG consecutive lanes (SEG is a power of two <= 64)
#pragma unroll
    for (int o = SEG / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
    if (s == 0 && c0 + c < V) out[c0 + c] = acc;
    __syncthreads();
  }
}

// ---- stored correlation-pyramid levels -------------------------------------------
// Element (pixel pix = n * hw1 + p, y, x) of a level with h2 x w2 planes, by layout:
//   kPlaneRows     row-major planes: pix * h2 w2 + y w2 + x (the reference's layout)
//   kPlaneTiles    each plane as row-major 8x8 tiles, each tile row-major (h2, w2 multiples of 8)
//   kPlaneGroups   the 64 pixels of a group g = pix / 64 interleaved in 16-byte segments of one
//                  row: group g holds [y][x / 8][pix % 64][x % 8] (w2 a multiple of 8)
//   kPlaneGroups24 the same with segments of 2 rows x 4 columns: group g holds
//                  [y / 2][x / 4][pix % 64][y % 2][x % 4] (h2 even, w2 a multiple of 4)
// In the group layouts the neighbouring pixels of a lookup block read neighbouring segments,
// so the 128-byte lines they fetch are shared (hw1 must be a multiple of 64).
enum { kPlaneRows = 0, kPlaneTiles = 1, kPlaneGroups = 2, kPlaneGroups24 = 3 };
constexpr int kPlaneGroup = 64;

// Layout of level `level` (planes h2 x w2, hw1 pixels per item) in a tiled pyramid
// (tiled != 0): level 0 in 8x8 tiles, levels 1 and 2 in 2x4 segments, the further levels in
// row segments of 8 (whichever the line counts of radius-3 windows favour); a level whose
// sizes do not allow it falls back to the next layout that they allow.
__host__ __device__ inline int pyramid_layout(int level, int64_t hw1, int64_t h2, int64_t w2, int tiled) {
  if (!tiled) return kPlaneRows;
  if (level > 0 && hw1 % kPlaneGroup == 0) {
    if (level <= 2 && h2 % 2 == 0 && w2 % 4 == 0) return kPlaneGroups24;
    if (w2 % 8 == 0) return kPlaneGroups;
  }
  return (h2 % 8 == 0 && w2 % 8 == 0) ? kPlaneTiles : kPlaneRows;
}

// ---- feature maps of the correlation build: MFMA fragment tiles -----------------------
// Per frame, [pixels / 32][channels / G][32 pixels][G channels] with G = 32 / sizeof(T) (two
// 16-byte pieces of each pixel's channels): the operand fragment of one 32-row MFMA step, 32
// pixels x G channels, is 1 KB of contiguous memory, so that one wave's 16-byte-per-lane load
// covers 8 whole cache lines instead of 32 quarter lines.
template <typename T> __host__ __device__ constexpr int frag_group() { return 32 / static_cast<int>(sizeof(T)); }
constexpr int kFragPixels = 32;

template <typename T>
__host__ __device__ inline int64_t frag_tile_offset(int64_t p, int c, int C) {
  constexpr int G = frag_group<T>();
  return ((p / kFragPixels) * (C / G) + c / G) * (kFragPixels * G) + (p % kFragPixels) * G + c % G;
}

__host__ __device__ inline int64_t pyramid_offset(int layout, int64_t pix, int y, int x, int w2, int64_t plane) {
  const int64_t g0 = (pix - pix % kPlaneGroup) * plane, pg = pix % kPlaneGroup;
  if (layout == kPlaneGroups24)
    return g0 + ((int64_t)((y >> 1) * (w2 >> 2) + (x >> 2)) * kPlaneGroup + pg) * 8 + ((y & 1) << 2) + (x & 3);
  if (layout == kPlaneGroups)
    return g0 + ((int64_t)(y * (w2 >> 3) + (x >> 3)) * kPlaneGroup + pg) * 8 + (x & 7);
  if (layout == kPlaneTiles)
    return pix * plane + ((((y >> 3) * (w2 >> 3) + (x >> 3)) << 6) + ((y & 7) << 3) + (x & 7));
  return pix * plane + (int64_t)y * w2 + x;
}

}  // namespace hsl2
