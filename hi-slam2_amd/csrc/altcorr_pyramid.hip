// On-the-fly correlation pyramid on the matrix cores: AltCorrBlock.__call__
// (hislam2/modules/corr.py:107-117 -- one altcorr_forward per level on coords / 2^level, then
// stack) in one launch, with the per-level semantics of corr_forward_kernel +
// altcorr_cuda_forward (src/altcorr_kernel.cu:24-75, 132-172): for edge m, pixel (y, x) of frame
// ii[m] and level l, the (2r+2)^2 taps s(ty, tx) = sum_c (f1_c / 4)(f2_c / 4) of fmap1's pixel
// vector with level l of frame jj[m] at (floor(y_l) - r + ty, floor(x_l) - r + tx) (zero outside
// the map), combined bilinearly with the fractions of (x_l, y_l) = coords / 2^l; output channel
// l (2r+1)^2 + a (2r+1) + b holds x-offset a, y-offset b (the reference's stack + flatten order).
//
// Mapping onto gfx950: a 256-thread block owns 64 pixels of one edge (a 4 x 16 patch), each wave
// a 4 x 4 sub-patch, i.e. the 16 rows of a v_mfma_f32_16x16x32_bf16 (16x16x32_f16, 16x16x4f32 for
// the other types).  Per level the block walks the bounding box of its pixels' windows in tiles
// of 2 x 8 map positions (the 16 columns of one MFMA), 64 positions per chunk staged channel-last
// in LDS (double-buffered, one barrier per chunk); a wave runs only the tiles that meet its own
// sub-patch's windows, K = 128 channels per tile, and scatters each of the 16 x 16 results into
// the window taps of its pixels (LDS, one row of (2r+2)^2 + 1 floats per pixel).  The taps are
// then combined bilinearly (the 1/16 folded into the weights, exact in FP32) and each wave writes
// one output channel of the 64 pixels.  Dot products, taps and the combine are FP32; the output
// is rounded once to the feature-map type.
//
// Blocks are ordered so that each XCD runs a contiguous range of (edge, tile) pairs: the tiles of
// an edge share fmap2's frame and the output lines of neighbouring tiles in one L2.
#include <climits>

#include "common.hpp"

namespace hsl2 {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int kApRows = 4;       // block patch: 4 rows x 16 columns of pixels
constexpr int kApCols = 16;
constexpr int kApPix = 64;
constexpr int kApThreads = 256;
// map positions staged per chunk: 64 (4 tiles of 2 x 8) for 2-byte types, 32 for FP32, so that
// the staging buffers and the taps stay within 64 KB of LDS
template <typename T> constexpr int ap_chunk() { return sizeof(T) == 4 ? 32 : 64; }
constexpr int kApC = 128;        // feature channels (the reference's fmaps are 128-d)
constexpr int kApMaxLevels = 4;

template <typename T>
struct ApLevels {
  const T* p[kApMaxLevels];
  int h[kApMaxLevels], w[kApMaxLevels];
  int L;
};

// One MFMA step on a lane's channel slice: `s` selects the instruction within the slice.
template <typename T> struct ApMma;
template <> struct ApMma<__hip_bfloat16> {
  static constexpr int kSteps = kApC / 32;               // 8 channels per lane per instruction
  __device__ static f32x4 step(const u32x4* a, const u32x4* b, int s, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[s]),
                                                     __builtin_bit_cast(bf16x8, b[s]), c, 0, 0, 0);
  }
};
template <> struct ApMma<__half> {
  static constexpr int kSteps = kApC / 32;
  __device__ static f32x4 step(const u32x4* a, const u32x4* b, int s, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a[s]),
                                                    __builtin_bit_cast(f16x8, b[s]), c, 0, 0, 0);
  }
};
template <> struct ApMma<float> {
  static constexpr int kSteps = kApC / 4;                // 1 channel per lane per instruction
  __device__ static f32x4 step(const u32x4* a, const u32x4* b, int s, f32x4 c) {
    const float av = __uint_as_float(a[s >> 2][s & 3]);
    const float bv = __uint_as_float(b[s >> 2][s & 3]);   // register selection after unrolling
    return __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, c, 0, 0, 0);
  }
};

// Each XCD takes a contiguous range of the linear (edge, tile) order; consecutive blocks are
// dispatched to the XCDs round-robin.
__device__ __forceinline__ int64_t xcd_linear(int64_t bid, int64_t nb) {
  const int64_t q = nb / 8, rr = nb % 8;
  const int64_t x = bid % 8, local = bid / 8;
  return x < rr ? x * (q + 1) + local : rr * (q + 1) + (x - rr) * q + local;
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
  return v;
}

constexpr int kFar = -(1 << 22);   // window origin of a pixel with no taps

template <int D>
__device__ __forceinline__ int window_origin(float c, int lvl) {
  const float v = floorf(c * (1.f / static_cast<float>(1 << lvl)));   // exact scaling
  // non-finite or far-away coordinates: an empty window (NaN fails both compares)
  return (v >= -1048576.f && v <= 1048576.f) ? static_cast<int>(v) - (D - 2) / 2 : kFar;
}

template <typename T, int D>
__global__ void __launch_bounds__(kApThreads)
altcorr_pyramid_kernel(ApLevels<T> lv, const float* __restrict__ coords, const int64_t* __restrict__ ii,
                       const int64_t* __restrict__ jj, T* __restrict__ out, int N, int H, int W, int M,
                       int64_t BM, int tiles_x, int tiles) {
  // taps of a pixel: D x D floats, tap (dy, dx) at dy D + dx; TS odd, so that the combine's
  // per-pixel reads hit distinct banks
  constexpr int R = D - 1, TS = D * D + 1;
  constexpr int SLICE = kApC / 4;                         // channels per lane group
  constexpr int NV = SLICE * sizeof(T) / 16;              // 16-byte pieces per lane slice
  constexpr int PP = kApC * sizeof(T) / 16;               // 16-byte pieces per map position
  constexpr int ROWB = kApC * sizeof(T) + 16;             // padded LDS bytes per staged position
  constexpr int CHUNK = ap_chunk<T>(), TPC = CHUNK / 16;  // positions and tiles per chunk
  constexpr int SPT = CHUNK * PP / kApThreads;            // staged pieces per thread
  constexpr int TAPB = ((kApPix * TS + 4 * kApPix) * 4 + 15) / 16 * 16;   // taps + four spare floats per lane
  extern __shared__ __align__(16) unsigned char smem[];
  float* s_tap = reinterpret_cast<float*>(smem);
  unsigned char* s_b = smem + TAPB;
  __shared__ float2 s_frac[kApPix];

  const int64_t nb = BM * tiles;
  const int64_t lin = xcd_linear(blockIdx.x, nb);
  const int64_t bm = lin / tiles;
  const int tile = static_cast<int>(lin - bm * tiles);
  const int m = static_cast<int>(bm % M);
  const int64_t b = bm / M;
  const int ty0 = (tile / tiles_x) * kApRows, tx0 = (tile % tiles_x) * kApCols;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int g = lane >> 4, col = lane & 15;
  const int64_t fi = ii[m], fj = jj[m];
  const bool edge_ok = fi >= 0 && fi < N && fj >= 0 && fj < N;
  const int HW = H * W;

  // A operand: row `col` of the MFMA is pixel `col` of this wave's 4 x 4 sub-patch; lane group g
  // supplies channels [g SLICE, (g + 1) SLICE)
  u32x4 a[NV];
  {
    const int ay = ty0 + (col >> 2), ax = tx0 + wave * 4 + (col & 3);
    if (edge_ok && ay < H && ax < W) {
      const u32x4* src = reinterpret_cast<const u32x4*>(
          lv.p[0] + (((b * N + fi) * H + ay) * (int64_t)W + ax) * kApC + g * SLICE);
#pragma unroll
      for (int k = 0; k < NV; ++k) a[k] = src[k];
    } else {
#pragma unroll
      for (int k = 0; k < NV; ++k) a[k] = u32x4{0u, 0u, 0u, 0u};
    }
  }
  // this lane's pixel for origins and fractions: block pixel `lane` (row lane / 16, column lane % 16)
  const int py = ty0 + (lane >> 4), px = tx0 + (lane & 15);
  const bool pin = edge_ok && py < H && px < W;
  float cx = 0.f, cy = 0.f;
  if (pin) {
    const float* c = coords + ((bm * H + py) * (int64_t)W + px) * 2;
    cx = c[0];
    cy = c[1];
  }
  const bool mine = ((lane & 15) >> 2) == wave;        // lane's block pixel lies in this wave's sub-patch

  // The MFMA runs with the staged map positions as its rows (A) and the wave's 16 pixels as its
  // columns (B), so lane (g, col) receives the results of pixel `col` (block pixel bp) at the
  // tile's positions 4g .. 4g + 3: row g / 2 of the tile, columns 4 (g % 2) .. 4 (g % 2) + 3,
  // four consecutive taps of one window row sharing one address computation.  A result outside
  // the pixel's window goes to one of this lane's four spare floats behind the taps (never
  // read), so the stores need no branch.
  const int bp = (col >> 2) * kApCols + wave * 4 + (col & 3);
  const int trash = kApPix * TS + 4 * lane;
  const int tb = bp * TS;

  for (int lvl = 0; lvl < lv.L; ++lvl) {
    const int Hl = lv.h[lvl], Wl = lv.w[lvl];
    const int ox = pin ? window_origin<D>(cx, lvl) : kFar;
    const int oy = pin ? window_origin<D>(cy, lvl) : kFar;
    const bool hit = ox != kFar && oy != kFar && ox + D > 0 && ox < Wl && oy + D > 0 && oy < Hl;
    const int y_lo = hit ? max(oy, 0) : INT_MAX, y_hi = hit ? min(oy + D - 1, Hl - 1) : INT_MIN;
    const int x_lo = hit ? max(ox, 0) : INT_MAX, x_hi = hit ? min(ox + D - 1, Wl - 1) : INT_MIN;
    // the block's box and this wave's part of it are the same in every lane: scalar registers
    const int Y0 = __builtin_amdgcn_readfirstlane(wave_min_i(y_lo));
    const int Y1 = __builtin_amdgcn_readfirstlane(wave_max_i(y_hi));
    const int X0 = __builtin_amdgcn_readfirstlane(wave_min_i(x_lo));
    const int X1 = __builtin_amdgcn_readfirstlane(wave_max_i(x_hi));
    const int sy0 = __builtin_amdgcn_readfirstlane(wave_min_i(mine ? y_lo : INT_MAX));
    const int sy1 = __builtin_amdgcn_readfirstlane(wave_max_i(mine ? y_hi : INT_MIN));
    const int sx0 = __builtin_amdgcn_readfirstlane(wave_min_i(mine ? x_lo : INT_MAX));
    const int sx1 = __builtin_amdgcn_readfirstlane(wave_max_i(mine ? x_hi : INT_MIN));
    const int ry = (g >> 1) - __shfl(oy, bp, kWave);
    const int rx = 4 * (g & 1) - __shfl(ox, bp, kWave);
    if (wave == 0) {
      const float sc = 1.f / static_cast<float>(1 << lvl);
      const float x = cx * sc, y = cy * sc;
      s_frac[lane] = make_float2(x - floorf(x), y - floorf(y));
    }
    {
      f32x4* tz = reinterpret_cast<f32x4*>(s_tap);
      for (int k = t; k < (kApPix * TS + 3) / 4; k += kApThreads) tz[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    __syncthreads();

    const bool any = Y0 <= Y1;
    const int TC = any ? (X1 - X0 + 8) >> 3 : 0;         // tile columns (8 positions each)
    const int TR = any ? (Y1 - Y0 + 2) >> 1 : 0;         // tile rows (2 positions each)
    const int nch = (TR * TC + TPC - 1) / TPC;
    const T* lp = lv.p[lvl] + (b * N + (edge_ok ? fj : 0)) * (int64_t)Hl * Wl * kApC;

    // staging: the k-th piece of thread t belongs to tile k / (SPT / TPC) of the chunk, position
    // j of that tile and 16-byte piece pc of the position's channel vector
    constexpr int PPT = SPT / TPC;                        // pieces per thread and tile
    const int pj = ((t / PP) & 15), pc = t & (PP - 1);
    int fr = 0, fc = 0;                                   // tile cursor of the next chunk to fetch
    auto fetch = [&](u32x4 (&st)[SPT]) {
#pragma unroll
      for (int k = 0; k < SPT; ++k) {
        int r = fr, c = fc + k / PPT;
        while (c >= TC) {
          c -= TC;
          ++r;
        }
        const int j = PPT == 1 ? pj : ((k % PPT) * (kApThreads / PP) + t / PP) & 15;
        const int Y = Y0 + 2 * r + (j >> 3), X = X0 + 8 * c + (j & 7);
        if (r < TR && Y < Hl && X < Wl)
          st[k] = reinterpret_cast<const u32x4*>(lp + (Y * Wl + X) * kApC)[pc];
        else
          st[k] = u32x4{0u, 0u, 0u, 0u};
      }
      fc += TPC;
      while (fc >= TC && TC > 0) {
        fc -= TC;
        ++fr;
      }
    };
    int mr = 0, mc = 0;                                   // tile cursor of the chunk in compute
    // chunk c: its staged pieces (fetched two chunks earlier) to LDS buffer c % 2, one barrier,
    // the fetch of chunk c + 2 into the freed registers, then the chunk's tiles
    auto chunk = [&](int c, u32x4 (&st)[SPT]) {
      unsigned char* buf = s_b + (c & 1) * CHUNK * ROWB;
#pragma unroll
      for (int k = 0; k < SPT; ++k) {
        const int piece = k * kApThreads + t;
        const int q = piece / PP;
        *reinterpret_cast<u32x4*>(buf + q * ROWB + pc * 16) = st[k];
      }
      __syncthreads();
      if (c + 2 < nch) fetch(st);
#pragma unroll
      for (int t4 = 0; t4 < TPC; ++t4) {
        int r = mr, cc = mc + t4;
        while (cc >= TC) {
          cc -= TC;
          ++r;
        }
        if (r >= TR) break;
        const int Ty = Y0 + 2 * r, Tx = X0 + 8 * cc;
        if (Ty > sy1 || Ty + 1 < sy0 || Tx > sx1 || Tx + 7 < sx0) continue;   // wave-uniform
        u32x4 bb[NV];
        const u32x4* src = reinterpret_cast<const u32x4*>(buf + (t4 * 16 + col) * ROWB + g * SLICE * sizeof(T));
#pragma unroll
        for (int k = 0; k < NV; ++k) bb[k] = src[k];
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < ApMma<T>::kSteps; ++s) acc = ApMma<T>::step(bb, a, s, acc);
        const int dy = Ty + ry, dx = Tx + rx;
        const bool row = static_cast<unsigned>(dy) < static_cast<unsigned>(D);
        const int at = tb + dy * D + dx;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const bool in = row && static_cast<unsigned>(dx + v) < static_cast<unsigned>(D);
          s_tap[in ? at + v : trash + v] = acc[v];
        }
      }
      mc += TPC;
      while (mc >= TC && TC > 0) {
        mc -= TC;
        ++mr;
      }
    };
    u32x4 st0[SPT], st1[SPT];
    if (nch > 0) fetch(st0);
    if (nch > 1) fetch(st1);
    for (int c = 0; c < nch; c += 2) {
      chunk(c, st0);
      if (c + 1 < nch) chunk(c + 1, st1);
    }
    __syncthreads();

    // bilinear combine: thread t takes block pixel P = t % 64 and the tap rows by = t / 64 + 4 n;
    // each pass reads two rows of taps once and forms the R outputs (ax = 0 .. R-1) of row by
    {
      const int P = t & 63;
      const int y = ty0 + (P >> 4), x = tx0 + (P & 15);
      const bool pok = y < H && x < W;
      const float2 frc = s_frac[P];
      const float w00 = (1.f - frc.x) * (1.f - frc.y) * 0.0625f, w01 = frc.x * (1.f - frc.y) * 0.0625f;
      const float w10 = (1.f - frc.x) * frc.y * 0.0625f, w11 = frc.x * frc.y * 0.0625f;
      T* ob = out + (bm * lv.L + lvl) * (int64_t)(R * R) * HW + (pok ? y * W + x : 0);
      for (int by = t >> 6; by < R; by += kApThreads / kApPix) {
        const float* tp = s_tap + P * TS + by * D;
        float r0[D], r1[D];
#pragma unroll
        for (int e = 0; e < D; ++e) {
          r0[e] = tp[e];
          r1[e] = tp[D + e];
        }
        if (pok) {
#pragma unroll
          for (int ax = 0; ax < R; ++ax) {
            const float v = w00 * r0[ax] + w01 * r0[ax + 1] + w10 * r1[ax] + w11 * r1[ax + 1];
            ob[(int64_t)(ax * R + by) * HW] = from_f<T>(v);
          }
        }
      }
    }
    __syncthreads();
  }
}

// level `lvl` of the channel-last feature pyramid: the mean of each 2^l x 2^l block of level 0
// (F.avg_pool2d applied l times with floor sizes), summed in FP32 and rounded once
template <typename T>
__global__ void __launch_bounds__(256)
pool_channels_last_kernel(const T* __restrict__ in, T* __restrict__ out, int64_t n8, int C8, int H, int W,
                          int Hl, int Wl, int s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const int c8 = static_cast<int>(i % C8);
  int64_t rest = i / C8;
  const int x = static_cast<int>(rest % Wl);
  rest /= Wl;
  const int y = static_cast<int>(rest % Hl);
  const int64_t f = rest / Hl;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int dy = 0; dy < s; ++dy)
    for (int dx = 0; dx < s; ++dx) {
      const T* p = in + ((f * H + (int64_t)y * s + dy) * W + (int64_t)x * s + dx) * (C8 * 8) + c8 * 8;
      const u32x4 v = *reinterpret_cast<const u32x4*>(p);
      if constexpr (sizeof(T) == 2) {
        const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] += to_f<T>(e[k]);
      } else {
        const u32x4 w = *reinterpret_cast<const u32x4*>(p + 4);
        const float* e = reinterpret_cast<const float*>(&v);
        const float* e2 = reinterpret_cast<const float*>(&w);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          acc[k] += e[k];
          acc[k + 4] += e2[k];
        }
      }
    }
  const float inv = 1.f / static_cast<float>(s * s);
  T* q = out + i * 8;
#pragma unroll
  for (int k = 0; k < 8; ++k) q[k] = from_f<T>(acc[k] * inv);
}

template <typename T, int D>
int launch_altcorr_pyramid(const void* const* levels, int L, const float* coords, const int64_t* ii,
                           const int64_t* jj, void* corr, int64_t B, int64_t N, int64_t H, int64_t W,
                           int64_t M, hipStream_t s) {
  ApLevels<T> lv{};
  for (int l = 0; l < L; ++l) {
    lv.p[l] = static_cast<const T*>(levels[l]);
    lv.h[l] = static_cast<int>(H >> l);
    lv.w[l] = static_cast<int>(W >> l);
  }
  lv.L = L;
  const int tiles_x = static_cast<int>((W + kApCols - 1) / kApCols);
  const int tiles = static_cast<int>((H + kApRows - 1) / kApRows) * tiles_x;
  const int64_t nb = B * M * tiles;
  HSL2_REQUIRE(nb < (1ll << 31), HSL2_ERR_SHAPE, "altcorr pyramid: too many blocks");
  constexpr int TS = D * D + 1;
  const size_t lds = ((kApPix * TS + 4 * kApPix) * 4 + 15) / 16 * 16 + 2 * ap_chunk<T>() * (kApC * sizeof(T) + 16);
  hipLaunchKernelGGL((altcorr_pyramid_kernel<T, D>), dim3(static_cast<unsigned>(nb)), dim3(kApThreads), lds, s, lv,
                     coords, ii, jj, static_cast<T*>(corr), static_cast<int>(N), static_cast<int>(H),
                     static_cast<int>(W), static_cast<int>(M), B * M, tiles_x, tiles);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

template <typename T>
int dispatch_radius(int radius, const void* const* levels, int L, const float* coords, const int64_t* ii,
                    const int64_t* jj, void* corr, int64_t B, int64_t N, int64_t H, int64_t W, int64_t M,
                    hipStream_t s) {
  if (radius == 3) return launch_altcorr_pyramid<T, 8>(levels, L, coords, ii, jj, corr, B, N, H, W, M, s);
  return launch_altcorr_pyramid<T, 10>(levels, L, coords, ii, jj, corr, B, N, H, W, M, s);
}

}  // namespace
}  // namespace hsl2

using namespace hsl2;

extern "C" int hsl2_pool_channels_last(const void* level0, void* out, int dtype, int64_t frames, int64_t h,
                                       int64_t w, int64_t channels, int level, void* stream) {
  HSL2_REQUIRE(level0 && out, HSL2_ERR_ARGUMENT, "pool_channels_last: null pointer");
  HSL2_REQUIRE(dtype == HSL2_F32 || dtype == HSL2_BF16 || dtype == HSL2_F16, HSL2_ERR_ARGUMENT,
               "pool_channels_last: dtype must be F32, F16 or BF16");
  HSL2_REQUIRE(frames >= 0 && h > 0 && w > 0 && channels > 0 && channels % 8 == 0 && level >= 0 && level < 16,
               HSL2_ERR_SHAPE, "pool_channels_last: bad sizes (channels a multiple of 8)");
  const int64_t Hl = h >> level, Wl = w >> level;
  HSL2_REQUIRE(Hl > 0 && Wl > 0, HSL2_ERR_SHAPE, "pool_channels_last: level too deep for the map");
  HSL2_REQUIRE((reinterpret_cast<uintptr_t>(level0) | reinterpret_cast<uintptr_t>(out)) % 16 == 0,
               HSL2_ERR_ARGUMENT, "pool_channels_last: buffers must be 16-byte aligned");
  const int64_t n8 = frames * Hl * Wl * (channels / 8);
  if (n8 == 0) return HSL2_OK;
  const unsigned grid = static_cast<unsigned>((n8 + 255) / 256);
  hipStream_t s = as_stream(stream);
  const int C8 = static_cast<int>(channels / 8), S = 1 << level;
#define HSL2_POOL(T)                                                                                        \
  hipLaunchKernelGGL(pool_channels_last_kernel<T>, dim3(grid), dim3(256), 0, s, static_cast<const T*>(level0), \
                     static_cast<T*>(out), n8, C8, static_cast<int>(h), static_cast<int>(w),                   \
                     static_cast<int>(Hl), static_cast<int>(Wl), S)
  if (dtype == HSL2_F32)
    HSL2_POOL(float);
  else if (dtype == HSL2_BF16)
    HSL2_POOL(__hip_bfloat16);
  else
    HSL2_POOL(__half);
#undef HSL2_POOL
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" int hsl2_altcorr_pyramid_forward(const void* const* levels, int num_levels, const float* coords,
                                            const int64_t* ii, const int64_t* jj, void* corr, int dtype,
                                            int64_t B, int64_t N, int64_t C, int64_t H, int64_t W, int64_t M,
                                            int radius, void* stream) {
  HSL2_REQUIRE(levels && coords && ii && jj && corr, HSL2_ERR_ARGUMENT, "altcorr pyramid: null pointer");
  HSL2_REQUIRE(dtype == HSL2_F32 || dtype == HSL2_BF16 || dtype == HSL2_F16, HSL2_ERR_ARGUMENT,
               "altcorr pyramid: dtype must be F32, F16 or BF16");
  HSL2_REQUIRE(num_levels >= 1 && num_levels <= kApMaxLevels, HSL2_ERR_UNSUPPORTED,
               "altcorr pyramid: 1 to 4 levels");
  HSL2_REQUIRE(radius == 3 || radius == 4, HSL2_ERR_UNSUPPORTED, "altcorr pyramid: radius 3 or 4");
  HSL2_REQUIRE(C == kApC, HSL2_ERR_UNSUPPORTED, "altcorr pyramid: 128 channels");
  HSL2_REQUIRE(B >= 0 && N > 0 && M >= 0 && H > 0 && W > 0 && (H >> (num_levels - 1)) > 0 &&
                   (W >> (num_levels - 1)) > 0 && H * W * C < (1ll << 31),
               HSL2_ERR_SHAPE, "altcorr pyramid: bad sizes");
  for (int l = 0; l < num_levels; ++l) {
    HSL2_REQUIRE(levels[l] && reinterpret_cast<uintptr_t>(levels[l]) % 16 == 0, HSL2_ERR_ARGUMENT,
                 "altcorr pyramid: levels must be non-null and 16-byte aligned");
  }
  if (B * M == 0) return HSL2_OK;
  hipStream_t s = as_stream(stream);
  if (dtype == HSL2_BF16)
    return dispatch_radius<__hip_bfloat16>(radius, levels, num_levels, coords, ii, jj, corr, B, N, H, W, M, s);
  if (dtype == HSL2_F16)
    return dispatch_radius<__half>(radius, levels, num_levels, coords, ii, jj, corr, B, N, H, W, M, s);
  return dispatch_radius<float>(radius, levels, num_levels, coords, ii, jj, corr, B, N, H, W, M, s);
}
