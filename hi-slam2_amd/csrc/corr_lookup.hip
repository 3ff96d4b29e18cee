// Radius-r bilinear lookup into a stored correlation volume, and its adjoint.
//
// Semantics of corr_index_forward_kernel / corr_index_backward_kernel
// (src/correlation_kernels.cu:19-124): for pixel (y, x) of item n, with
// (x0, y0) = coords[n, :, y, x], the (2r+2)^2 integer taps at
// (floor(x0) - r + i, floor(y0) - r + j) of the pixel's own h2 x w2 plane are
// bilinearly combined into (2r+1)^2 outputs ordered [a = x-offset][b = y-offset];
// taps outside the plane are zero.
//
// Mapping (not the reference's one-thread-per-pixel): a 256-thread block owns 64
// consecutive pixels.  Phase 1 gathers the 64 tap windows into LDS with lanes
// running along window rows (contiguous in memory); phase 2 combines taps with
// pixels on the fast lane axis, so the [a][b][pixel] output rows are written fully
// coalesced.  Taps and sums are kept in float (double for f64 volumes) and rounded
// once on store.
#include <type_traits>

#include "common.hpp"

namespace hsl2 {
namespace {

constexpr int kPix = 64;      // pixels per block
constexpr int kThreads = 256;

template <typename T> struct Acc { using type = float; };
template <> struct Acc<double> { using type = double; };

template <typename T> __device__ __forceinline__ typename Acc<T>::type load_acc(const T* p) {
  if constexpr (std::is_same<T, double>::value) return *p;
  else return to_f<T>(*p);
}

template <typename T, typename A> __device__ __forceinline__ T store_acc(A v) {
  if constexpr (std::is_same<T, double>::value) return static_cast<double>(v);
  else return from_f<T>(static_cast<float>(v));
}

// Up to four pyramid levels looked up by one launch (blockIdx.y = level).
struct LevelTable {
  const void* vol[4];
  int h2[4], w2[4];
  int layout[4];           // kPlaneRows / kPlaneTiles / kPlaneGroups (pyramid_offset)
  float scale[4];          // coordinate scale of the level (1 / 2^level)
  int64_t out_offset[4];   // element offset of the level's first output channel
};

// A block's pixels pix0 .. pix0 + 63 (pix0 a multiple of 64) are one pixel group of an
// interleaved level, so tap (p, y, x) of the block lies at pyramid_offset(layout, p, y, x, ...)
// from the block's base (n * hw1 + pix0) * plane in every layout.  In the tiled layouts a
// radius-3 window (8x8 taps) touches at most four 8x8 tiles instead of eight plane rows, and
// the pooled levels' 16-byte row segments of neighbouring pixels share 128-byte lines.
static_assert(kPix == kPlaneGroup, "lookup blocks must coincide with pixel groups");

// Coordinates of pixel p of item n: x = coords[n * cb + p * cp], y = x-address + cy.
// Output: out + n * out_batch_stride + level offset + o * hw1 + pix, o = a * (2r+1) + b.
template <typename T>
__global__ void __launch_bounds__(kThreads)
corr_lookup_fwd_kernel(LevelTable lv, const float* __restrict__ coords, int64_t cb, int cp,
                       int64_t cy_off, T* __restrict__ out, int hw1, int r,
                       int64_t out_batch_stride, int tiles_per_item) {
  using A = typename Acc<T>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  A* taps = reinterpret_cast<A*>(smem_raw);                 // [kPix][D*D + 1]
  __shared__ int s_x0[kPix], s_y0[kPix];
  __shared__ float s_dx[kPix], s_dy[kPix];

  const int level = blockIdx.y;
  const T* __restrict__ volume = static_cast<const T*>(lv.vol[level]);
  const int h2 = lv.h2[level], w2 = lv.w2[level];
  const int layout = lv.layout[level];
  const float scale = lv.scale[level];
  const int n = blockIdx.x / tiles_per_item;
  const int pix0 = (blockIdx.x % tiles_per_item) * kPix;
  const int D = 2 * r + 2;
  const int D2 = D * D;
  const int TS = D2 + 1;        // padded per-pixel stride: pixels on the fast lane axis hit distinct banks
  const int npix = min(kPix, hw1 - pix0);
  const int t = threadIdx.x;

  if (t < kPix) {
    int x0 = 0, y0 = 0;
    float fx = 0.f, fy = 0.f;
    if (t < npix) {
      const float* cp_ = coords + (int64_t)n * cb + (int64_t)(pix0 + t) * cp;
      const float cx = cp_[0] * scale;
      const float cy = cp_[cy_off] * scale;
      const float flx = floorf(cx), fly = floorf(cy);
      x0 = static_cast<int>(flx) - r;
      y0 = static_cast<int>(fly) - r;
      fx = cx - flx;
      fy = cy - fly;
    }
    s_x0[t] = x0;
    s_y0[t] = y0;
    s_dx[t] = fx;
    s_dy[t] = fy;
  }
  __syncthreads();

  const int64_t plane = (int64_t)h2 * w2;
  const T* vbase = volume + ((int64_t)n * hw1 + pix0) * plane;
  for (int idx = t; idx < npix * D2; idx += kThreads) {
    const int p = idx / D2;
    const int tap = idx - p * D2;
    const int ty = tap / D;
    const int tx = tap - ty * D;
    const int y1 = s_y0[p] + ty;
    const int x1 = s_x0[p] + tx;
    A v = 0;
    if (y1 >= 0 && y1 < h2 && x1 >= 0 && x1 < w2)
      v = load_acc<T>(vbase + pyramid_offset(layout, p, y1, x1, w2, plane));
    taps[p * TS + tap] = v;
  }
  __syncthreads();

  const int R = 2 * r + 1;
  T* obase = out + (int64_t)n * out_batch_stride + lv.out_offset[level] + pix0;
  for (int idx = t; idx < npix * R * R; idx += kThreads) {
    const int o = idx / npix;
    const int p = idx - o * npix;
    const int a = o / R;        // x offset
    const int b = o - a * R;    // y offset
    // bilinear weights in float, as the reference forms them (src/correlation_kernels.cu:41-63)
    const float fx = s_dx[p], fy = s_dy[p];
    const A w00 = (1.f - fx) * (1.f - fy), w10 = fx * (1.f - fy);
    const A w01 = (1.f - fx) * fy, w11 = fx * fy;
    const A* tp = taps + p * TS;
    const A v = w00 * tp[b * D + a] + w10 * tp[b * D + a + 1] + w01 * tp[(b + 1) * D + a] +
                w11 * tp[(b + 1) * D + a + 1];
    obase[(int64_t)o * hw1 + p] = store_acc<T>(v);
  }
}

// Radius known at compile time: 64 pixels of one level per 256-thread block.
// Phase 1: each wave gathers whole tap windows, one window per load instruction (lanes run
// along the window rows, so an instruction touches at most four 8x8 tiles of a tiled
// plane); taps land in LDS with a padded per-pixel stride.  Phase 2: four threads per
// pixel (pixels on the fast lane axis) read the tap columns they need and write the
// [a][b][pixel] output rows coalesced.  All index arithmetic is shifts or constant
// divisions.
template <typename T, int RAD>
__global__ void __launch_bounds__(256)
corr_lookup_win_kernel(LevelTable lv, const float* __restrict__ coords, int64_t cb, int cp,
                       int64_t cy_off, T* __restrict__ out, int hw1, int64_t out_batch_stride,
                       int tiles_per_item) {
  using A = typename Acc<T>::type;
  constexpr int D = 2 * RAD + 2, D2 = D * D, R = 2 * RAD + 1;
  constexpr int TS = D2 + 1;                    // padded per-pixel tap stride
  __shared__ A taps[kPix * TS];
  __shared__ int s_x0[kPix], s_y0[kPix];
  __shared__ float s_dx[kPix], s_dy[kPix];

  const int level = blockIdx.y;
  const T* __restrict__ volume = static_cast<const T*>(lv.vol[level]);
  const int h2 = lv.h2[level], w2 = lv.w2[level];
  const int layout = lv.layout[level];
  const float scale = lv.scale[level];
  const int n = blockIdx.x / tiles_per_item;
  const int pix0 = (blockIdx.x - n * tiles_per_item) * kPix;
  const int npix = min(kPix, hw1 - pix0);
  const int t = threadIdx.x;

  if (t < kPix) {
    int x0 = 0, y0 = 0;
    float fx = 0.f, fy = 0.f;
    if (t < npix) {
      const float* cp_ = coords + (int64_t)n * cb + (int64_t)(pix0 + t) * cp;
      const float cx = cp_[0] * scale;
      const float cy = cp_[cy_off] * scale;
      const float flx = floorf(cx), fly = floorf(cy);
      x0 = static_cast<int>(flx) - RAD;
      y0 = static_cast<int>(fly) - RAD;
      fx = cx - flx;
      fy = cy - fly;
    }
    s_x0[t] = x0;
    s_y0[t] = y0;
    s_dx[t] = fx;
    s_dy[t] = fy;
  }
  __syncthreads();

  const int wave = t >> 6, lane = t & 63;
  const int64_t plane = (int64_t)h2 * w2;
  const T* vbase = volume + ((int64_t)n * hw1 + pix0) * plane;
  // each wave gathers the windows of pixels wave, wave+4, ...: all of a lane's loads are
  // issued before any is consumed, so up to 16 * KT gathers are in flight per lane
  constexpr int PPW = kPix / 4;
  constexpr int KT = (D2 + 63) / 64;
  A v[PPW][KT];
  bool okv[PPW][KT];
#pragma unroll
  for (int j = 0; j < PPW; ++j) {
    const int p = wave + 4 * j;
    const int xs = s_x0[p], ys = s_y0[p];
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
      const int k = lane + 64 * kt;
      const int ty = k / D, tx = k - ty * D;    // constant divisor
      const int y = ys + ty, x = xs + tx;
      const bool ok = p < npix && k < D2 && (unsigned)y < (unsigned)h2 && (unsigned)x < (unsigned)w2;
      // out-of-plane taps read the block's first element, then zero
      v[j][kt] = load_acc<T>(vbase + (ok ? pyramid_offset(layout, p, y, x, w2, plane) : 0));
      okv[j][kt] = ok;
    }
  }
#pragma unroll
  for (int j = 0; j < PPW; ++j)
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
      const int k = lane + 64 * kt;
      if (k < D2) taps[(wave + 4 * j) * TS + k] = okv[j][kt] ? v[j][kt] : A(0);
    }
  __syncthreads();

  const int p = lane, g = wave;
  if (p < npix) {
    const float fx = s_dx[p], fy = s_dy[p];
    const A w00 = (1.f - fx) * (1.f - fy), w10 = fx * (1.f - fy);
    const A w01 = (1.f - fx) * fy, w11 = fx * fy;
    const A* tp = taps + p * TS;
    T* o = out + (int64_t)n * out_batch_stride + lv.out_offset[level] + pix0 + p;
#pragma unroll
    for (int a = g; a < R; a += 4) {
      A c0[D], c1[D];
#pragma unroll
      for (int b = 0; b < D; ++b) {
        c0[b] = tp[b * D + a];
        c1[b] = tp[b * D + a + 1];
      }
#pragma unroll
      for (int b = 0; b < R; ++b) {
        const A v = w00 * c0[b] + w10 * c1[b] + w01 * c0[b + 1] + w11 * c1[b + 1];
        o[(int64_t)(a * R + b) * hw1] = store_acc<T>(v);
      }
    }
  }
}

// corr_lookup_win_kernel at radius 3 with all four levels of 64 pixels per block (the path of
// 4-level pyramids in FP32, and of 16-bit ones that the vec kernels do not take): the window
// origins of every level come from one read of the coordinates, and every level's gathers (lane =
// tap, 16 pixels per wave) are issued before the first level is consumed, as in
// corr_lookup_vec4_kernel.  The taps and their combine are those of corr_lookup_win_kernel.
template <typename T>
__global__ void __launch_bounds__(kThreads)
corr_lookup_win4_kernel(LevelTable lv, const float* __restrict__ coords, int64_t cb, int cp,
                        int64_t cy_off, T* __restrict__ out, int hw1, int64_t out_batch_stride,
                        int tiles_per_item) {
  using A = typename Acc<T>::type;
  constexpr int RAD = 3, D = 2 * RAD + 2, D2 = D * D, R = 2 * RAD + 1;
  constexpr int TS = D2 + 1;                    // padded per-pixel tap stride
  constexpr int PPW = kPix / 4;                 // pixels per wave
  static_assert(D2 == kWave, "one tap per lane");
  __shared__ A taps[kPix * TS];
  __shared__ int s_x0[4][kPix], s_y0[4][kPix];
  __shared__ float s_dx[4][kPix], s_dy[4][kPix];
  const int n = blockIdx.x / tiles_per_item;
  const int pix0 = (blockIdx.x - n * tiles_per_item) * kPix;
  const int npix = min(kPix, hw1 - pix0);
  const int t = threadIdx.x;
  {
    const int level = t / kPix, p = t % kPix;
    const float scale = lv.scale[level];
    int x0 = 0, y0 = 0;
    float fx = 0.f, fy = 0.f;
    if (p < npix) {
      const float* cp_ = coords + (int64_t)n * cb + (int64_t)(pix0 + p) * cp;
      const float cx = cp_[0] * scale, cy = cp_[cy_off] * scale;
      const float flx = floorf(cx), fly = floorf(cy);
      x0 = static_cast<int>(flx) - RAD;
      y0 = static_cast<int>(fly) - RAD;
      fx = cx - flx;
      fy = cy - fly;
    }
    s_x0[level][p] = x0;
    s_y0[level][p] = y0;
    s_dx[level][p] = fx;
    s_dy[level][p] = fy;
  }
  __syncthreads();
  const int wave = t >> 6, lane = t & 63;
  const int ty = lane / D, tx = lane - (lane / D) * D;
  A v[4][PPW];
#pragma unroll
  for (int level = 0; level < 4; ++level) {
    const int h2 = lv.h2[level], w2 = lv.w2[level], layout = lv.layout[level];
    const int64_t plane = (int64_t)h2 * w2;
    const T* vbase = static_cast<const T*>(lv.vol[level]) + ((int64_t)n * hw1 + pix0) * plane;
#pragma unroll
    for (int j = 0; j < PPW; ++j) {
      const int p = wave + 4 * j;
      const int y = s_y0[level][p] + ty, x = s_x0[level][p] + tx;
      v[level][j] = A(0);
      if (p < npix && (unsigned)y < (unsigned)h2 && (unsigned)x < (unsigned)w2)
        v[level][j] = load_acc<T>(vbase + pyramid_offset(layout, p, y, x, w2, plane));
    }
  }
#pragma unroll
  for (int level = 0; level < 4; ++level) {
#pragma unroll
    for (int j = 0; j < PPW; ++j) taps[(wave + 4 * j) * TS + lane] = v[level][j];
    __syncthreads();
    const int p = lane, g = wave;
    if (p < npix) {
      const float fx = s_dx[level][p], fy = s_dy[level][p];
      const A w00 = (1.f - fx) * (1.f - fy), w10 = fx * (1.f - fy);
      const A w01 = (1.f - fx) * fy, w11 = fx * fy;
      const A* tp = taps + p * TS;
      T* o = out + (int64_t)n * out_batch_stride + lv.out_offset[level] + pix0 + p;
#pragma unroll
      for (int a = g; a < R; a += 4) {
        A c0[D], c1[D];
#pragma unroll
        for (int b = 0; b < D; ++b) {
          c0[b] = tp[b * D + a];
          c1[b] = tp[b * D + a + 1];
        }
#pragma unroll
        for (int b = 0; b < R; ++b) {
          const A val = w00 * c0[b] + w10 * c1[b] + w01 * c0[b + 1] + w11 * c1[b + 1];
          o[(int64_t)(a * R + b) * hw1] = store_acc<T>(val);
        }
      }
    }
    __syncthreads();                              // taps are reused by the next level
  }
}

// 16-bit volumes (bf16 / f16), radius 3, every level width a multiple of 8 and hw1 a multiple
// of 4: 64 pixels of one level per 256-thread block, with far fewer (and wider) memory
// requests than corr_lookup_win_kernel.  Each thread fetches two window rows; a row's 8 taps
// x0 .. x0+7 lie in the two aligned 8-element chunks at floor(x0 / 8) * 8 and the next one,
// fetched as two 16-byte loads (a chunk outside the plane reads as zeros; chunks never
// straddle the plane edge because w2 % 8 == 0) and the taps extracted with a funnel shift.
// The combine gives each thread four consecutive pixels of a set of output channels, so
// every store writes 8 bytes.
template <typename T> __device__ __forceinline__ float half_bits_to_f(uint32_t h);
template <> __device__ __forceinline__ float half_bits_to_f<__hip_bfloat16>(uint32_t h) {
  return __uint_as_float(h << 16);
}
template <> __device__ __forceinline__ float half_bits_to_f<__half>(uint32_t h) {
  return __half2float(__ushort_as_half(static_cast<unsigned short>(h)));
}
template <typename T> __device__ __forceinline__ uint32_t f_to_half_bits(float v) {
  const T h = from_f<T>(v);
  return static_cast<uint32_t>(*reinterpret_cast<const unsigned short*>(&h));
}

// Window gather of a kPlaneGroups24 level for the vec kernel: item (pixel p, row pair j < 5,
// column quad s < 3) loads the 2x4 segment at rows 2 (y0 / 2 + j) + {0, 1}, columns
// 4 (x0 / 4 + s) + {0, .., 3} with one 16-byte load (none when it misses the window or lies
// outside the plane: zeros) and writes the elements that fall inside the 8x8 window to its
// taps.  Every tap of a window lies in exactly one such segment.  Loads first, then the
// writes.
// A thread's loads of one level's window gather, kept in registers between the load and the
// store phase (so that a kernel can have several levels' loads in flight).
struct WinRegs {
  uint4 v[4];
  int sh[2];
};
constexpr int kSegItems = 15 * kPix, kSegIters = (kSegItems + kThreads - 1) / kThreads;
static_assert(kSegIters <= 4, "segment gather: four loads per thread at most");

template <typename T>
__device__ __forceinline__ void load_segs24(const T* __restrict__ vbase, int64_t plane, int h2, int w2, int npix,
                                            const int* s_x0, const int* s_y0, WinRegs& g) {
  constexpr int NI = kSegItems, IT = kSegIters;
  const int t = threadIdx.x;
  uint4* sv = g.v;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int item = t + kThreads * it;
    const int p = item / 15, rest = item - 15 * p, j = rest / 3, s = rest - 3 * j;
    sv[it] = make_uint4(0u, 0u, 0u, 0u);
    if (item < NI && p < npix) {
      const int y0 = s_y0[p], x0 = s_x0[p];
      const int yp = (y0 >> 1) + j, xq = (x0 >> 2) + s;
      if (2 * yp <= y0 + 7 && 4 * xq <= x0 + 7 && yp >= 0 && 2 * yp < h2 && xq >= 0 && 4 * xq < w2)
        sv[it] = *reinterpret_cast<const uint4*>(vbase + pyramid_offset(kPlaneGroups24, p, 2 * yp, 4 * xq, w2, plane));
    }
  }
}

template <typename T>
__device__ __forceinline__ void store_segs24(const WinRegs& g, const int* s_x0, const int* s_y0, float* taps) {
  constexpr int D = 8, TS = D * D + 1, NI = kSegItems, IT = kSegIters;
  const int t = threadIdx.x;
  const uint4* sv = g.v;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int item = t + kThreads * it;
    const int p = item / 15, rest = item - 15 * p, j = rest / 3, s = rest - 3 * j;
    if (item < NI) {
      const int y0 = s_y0[p], x0 = s_x0[p];
      const int ty0 = 2 * ((y0 >> 1) + j) - y0, tx0 = 4 * ((x0 >> 2) + s) - x0;
      if (ty0 <= 7 && tx0 <= 7) {
        const uint32_t wv[4] = {sv[it].x, sv[it].y, sv[it].z, sv[it].w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int ty = ty0 + (e >> 2), tx = tx0 + (e & 3);
          if (ty >= 0 && ty < D && tx >= 0 && tx < D)
            taps[p * TS + ty * D + tx] = half_bits_to_f<T>((wv[e >> 1] >> (16 * (e & 1))) & 0xffffu);
        }
      }
    }
  }
}

template <typename T>
__device__ __forceinline__ void gather_segs24(const T* __restrict__ vbase, int64_t plane, int h2, int w2,
                                              int npix, const int* s_x0, const int* s_y0, float* taps) {
  WinRegs g;
  load_segs24<T>(vbase, plane, h2, w2, npix, s_x0, s_y0, g);
  store_segs24<T>(g, s_x0, s_y0, taps);
}

// Window gather of a row-segment level (tiles, 1x8 groups or rows) for the vec kernels: each
// thread fetches two window rows as two aligned 16-byte chunks each (exec-masked: no load, so
// no line, for a chunk outside the plane or the unneeded second chunk of an aligned window)
// and extracts the 8 taps with a barrel shift; all four loads of a thread are issued first.
template <typename T>
__device__ __forceinline__ void load_rows8(const T* __restrict__ vbase, int64_t plane, int h2, int w2, int layout,
                                           int npix, const int* s_x0, const int* s_y0, WinRegs& g) {
  const int t = threadIdx.x;
  uint4* ca = g.v;
  uint4* cbv = g.v + 2;
  int* sh = g.sh;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int item = t + kThreads * it;
    const int p = item >> 3, ty = item & 7;
    const int y = s_y0[p] + ty, x0 = s_x0[p];
    const int xa = (x0 >> 3) << 3;              // floor to a multiple of 8 (also for x0 < 0)
    sh[it] = x0 - xa;
    const bool yok = p < npix && y >= 0 && y < h2;
    const bool oka = yok && xa >= 0 && xa < w2;
    // the second chunk only when the window reaches into it (x0 not a multiple of 8): an
    // aligned window would otherwise pull in a neighbouring 8x8 tile's line for nothing
    const bool okb = yok && x0 != xa && xa + 8 >= 0 && xa + 8 < w2;
    const int pp = p < npix ? p : 0;
    // lanes whose chunk lies outside the plane issue no load (exec-masked), so no line is
    // fetched for them
    uint4 la = make_uint4(0u, 0u, 0u, 0u), lb = make_uint4(0u, 0u, 0u, 0u);
    if (oka) la = *reinterpret_cast<const uint4*>(vbase + pyramid_offset(layout, pp, y, xa, w2, plane));
    if (okb) lb = *reinterpret_cast<const uint4*>(vbase + pyramid_offset(layout, pp, y, xa + 8, w2, plane));
    ca[it] = la;
    cbv[it] = lb;
  }
}

template <typename T>
__device__ __forceinline__ void store_rows8(const WinRegs& g, float* taps) {
  constexpr int D = 8, TS = D * D + 1;
  const int t = threadIdx.x;
  const uint4* ca = g.v;
  const uint4* cbv = g.v + 2;
  const int* sh = g.sh;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int item = t + kThreads * it;
    const int p = item >> 3, ty = item & 7;
    // barrel shift the 16 halves left by sh: two word stages, then a half-word align
    uint32_t w0 = ca[it].x, w1 = ca[it].y, w2 = ca[it].z, w3 = ca[it].w;
    uint32_t w4 = cbv[it].x, w5 = cbv[it].y, w6 = cbv[it].z, w7 = cbv[it].w;
    if (sh[it] & 4) {
      w0 = w2; w1 = w3; w2 = w4; w3 = w5; w4 = w6; w5 = w7;
    }
    if (sh[it] & 2) {
      w0 = w1; w1 = w2; w2 = w3; w3 = w4; w4 = w5;
    }
    const uint32_t al = (sh[it] & 1) ? 2u : 0u;
    const uint32_t d0 = __builtin_amdgcn_alignbyte(w1, w0, al);
    const uint32_t d1 = __builtin_amdgcn_alignbyte(w2, w1, al);
    const uint32_t d2 = __builtin_amdgcn_alignbyte(w3, w2, al);
    const uint32_t d3 = __builtin_amdgcn_alignbyte(w4, w3, al);
    float* dst = taps + p * TS + ty * D;
    dst[0] = half_bits_to_f<T>(d0 & 0xffffu);
    dst[1] = half_bits_to_f<T>(d0 >> 16);
    dst[2] = half_bits_to_f<T>(d1 & 0xffffu);
    dst[3] = half_bits_to_f<T>(d1 >> 16);
    dst[4] = half_bits_to_f<T>(d2 & 0xffffu);
    dst[5] = half_bits_to_f<T>(d2 >> 16);
    dst[6] = half_bits_to_f<T>(d3 & 0xffffu);
    dst[7] = half_bits_to_f<T>(d3 >> 16);
  }
}

template <typename T>
__device__ __forceinline__ void gather_rows8(const T* __restrict__ vbase, int64_t plane, int h2, int w2, int layout,
                                             int npix, const int* s_x0, const int* s_y0, float* taps) {
  WinRegs g;
  load_rows8<T>(vbase, plane, h2, w2, layout, npix, s_x0, s_y0, g);
  store_rows8<T>(g, taps);
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
corr_lookup_vec_kernel(LevelTable lv, const float* __restrict__ coords, int64_t cb, int cp,
                       int64_t cy_off, T* __restrict__ out, int hw1, int64_t out_batch_stride,
                       int tiles_per_item) {
  constexpr int RAD = 3, D = 8, R = 7;
  constexpr int TS = D * D + 1;                 // padded per-pixel tap stride
  __shared__ float taps[kPix * TS];
  __shared__ int s_x0[kPix], s_y0[kPix];
  __shared__ float s_dx[kPix], s_dy[kPix];

  const int level = blockIdx.y;
  const T* __restrict__ volume = static_cast<const T*>(lv.vol[level]);
  const int h2 = lv.h2[level], w2 = lv.w2[level];
  const int layout = lv.layout[level];
  const float scale = lv.scale[level];
  const int n = blockIdx.x / tiles_per_item;
  const int pix0 = (blockIdx.x - n * tiles_per_item) * kPix;
  const int npix = min(kPix, hw1 - pix0);
  const int t = threadIdx.x;

  if (t < kPix) {
    int x0 = 0, y0 = 0;
    float fx = 0.f, fy = 0.f;
    if (t < npix) {
      const float* cp_ = coords + (int64_t)n * cb + (int64_t)(pix0 + t) * cp;
      const float cx = cp_[0] * scale;
      const float cy = cp_[cy_off] * scale;
      const float flx = floorf(cx), fly = floorf(cy);
      x0 = static_cast<int>(flx) - RAD;
      y0 = static_cast<int>(fly) - RAD;
      fx = cx - flx;
      fy = cy - fly;
    }
    s_x0[t] = x0;
    s_y0[t] = y0;
    s_dx[t] = fx;
    s_dy[t] = fy;
  }
  __syncthreads();

  const int64_t plane = (int64_t)h2 * w2;
  const T* vbase = volume + ((int64_t)n * hw1 + pix0) * plane;
  if (layout == kPlaneGroups24) gather_segs24<T>(vbase, plane, h2, w2, npix, s_x0, s_y0, taps);
  else gather_rows8<T>(vbase, plane, h2, w2, layout, npix, s_x0, s_y0, taps);
  __syncthreads();

  // combine: thread = (pixel quad q, channel group g); channel o = g, g + 16, g + 32, g + 48
  const int q = t & 15, g = t >> 4;
  T* obase = out + (int64_t)n * out_batch_stride + lv.out_offset[level] + pix0 + 4 * q;
  float w00[4], w10[4], w01[4], w11[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const float fx = s_dx[4 * q + u], fy = s_dy[4 * q + u];
    w00[u] = (1.f - fx) * (1.f - fy);
    w10[u] = fx * (1.f - fy);
    w01[u] = (1.f - fx) * fy;
    w11[u] = fx * fy;
  }
#pragma unroll
  for (int oo = 0; oo < 4; ++oo) {
    const int o = g + 16 * oo;
    if (o < R * R) {
      const int a = o / R, b = o - (o / R) * R;
      float v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float* tp = taps + (4 * q + u) * TS;
        v[u] = w00[u] * tp[b * D + a] + w10[u] * tp[b * D + a + 1] + w01[u] * tp[(b + 1) * D + a] +
               w11[u] * tp[(b + 1) * D + a + 1];
      }
      T* dst = obase + (int64_t)o * hw1;
      if (4 * q + 3 < npix) {
        // streamed past the caches: the output is not read again here, and the window lines
        // that neighbouring pixels share stay cached
        const unsigned long long pk =
            static_cast<unsigned long long>(f_to_half_bits<T>(v[0]) | (f_to_half_bits<T>(v[1]) << 16)) |
            (static_cast<unsigned long long>(f_to_half_bits<T>(v[2]) | (f_to_half_bits<T>(v[3]) << 16)) << 32);
        __builtin_nontemporal_store(pk, reinterpret_cast<unsigned long long*>(dst));
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (4 * q + u < npix) dst[u] = from_f<T>(v[u]);
      }
    }
  }
}

// All four levels of 64 pixels per block (the vec kernel's gathers and combine; the path of
// every 4-level pyramid): the window origins of every level are formed once from one read of the
// coordinates, and every level's window loads are issued before the first is consumed, so a
// block keeps four levels' loads in flight (one memory latency per block instead of one per
// level; 116 VGPRs, four blocks per CU: 2.43 against 2.63 ms for the per-level blocks at 3000
// edges of 48x64, bit-identical results).
template <typename T>
__global__ void __launch_bounds__(kThreads)
corr_lookup_vec4_kernel(LevelTable lv, const float* __restrict__ coords, int64_t cb, int cp,
                        int64_t cy_off, T* __restrict__ out, int hw1, int64_t out_batch_stride,
                        int tiles_per_item) {
  constexpr int RAD = 3, D = 8, R = 7;
  constexpr int TS = D * D + 1;                 // padded per-pixel tap stride
  __shared__ float taps[kPix * TS];
  __shared__ int s_x0[4][kPix], s_y0[4][kPix];
  __shared__ float s_dx[4][kPix], s_dy[4][kPix];
  const int n = blockIdx.x / tiles_per_item;
  const int pix0 = (blockIdx.x - n * tiles_per_item) * kPix;
  const int npix = min(kPix, hw1 - pix0);
  const int t = threadIdx.x;
  {
    const int level = t / kPix, p = t % kPix;
    const float scale = lv.scale[level];
    int x0 = 0, y0 = 0;
    float fx = 0.f, fy = 0.f;
    if (p < npix) {
      const float* cp_ = coords + (int64_t)n * cb + (int64_t)(pix0 + p) * cp;
      const float cx = cp_[0] * scale, cy = cp_[cy_off] * scale;
      const float flx = floorf(cx), fly = floorf(cy);
      x0 = static_cast<int>(flx) - RAD;
      y0 = static_cast<int>(fly) - RAD;
      fx = cx - flx;
      fy = cy - fly;
    }
    s_x0[level][p] = x0;
    s_y0[level][p] = y0;
    s_dx[level][p] = fx;
    s_dy[level][p] = fy;
  }
  __syncthreads();
  WinRegs g[4];
#pragma unroll
  for (int level = 0; level < 4; ++level) {
    const int h2 = lv.h2[level], w2 = lv.w2[level], layout = lv.layout[level];
    const int64_t plane = (int64_t)h2 * w2;
    const T* vbase = static_cast<const T*>(lv.vol[level]) + ((int64_t)n * hw1 + pix0) * plane;
    if (layout == kPlaneGroups24) load_segs24<T>(vbase, plane, h2, w2, npix, s_x0[level], s_y0[level], g[level]);
    else load_rows8<T>(vbase, plane, h2, w2, layout, npix, s_x0[level], s_y0[level], g[level]);
  }
  const int q = t & 15, cg = t >> 4;
#pragma unroll
  for (int level = 0; level < 4; ++level) {
    if (lv.layout[level] == kPlaneGroups24) store_segs24<T>(g[level], s_x0[level], s_y0[level], taps);
    else store_rows8<T>(g[level], taps);
    __syncthreads();
    // combine: thread = (pixel quad q, channel group cg); channel o = cg, cg + 16, cg + 32, cg + 48
    T* obase = out + (int64_t)n * out_batch_stride + lv.out_offset[level] + pix0 + 4 * q;
    float w00[4], w10[4], w01[4], w11[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float fx = s_dx[level][4 * q + u], fy = s_dy[level][4 * q + u];
      w00[u] = (1.f - fx) * (1.f - fy);
      w10[u] = fx * (1.f - fy);
      w01[u] = (1.f - fx) * fy;
      w11[u] = fx * fy;
    }
#pragma unroll
    for (int oo = 0; oo < 4; ++oo) {
      const int o = cg + 16 * oo;
      if (o < R * R) {
        const int a = o / R, b = o - (o / R) * R;
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float* tp = taps + (4 * q + u) * TS;
          v[u] = w00[u] * tp[b * D + a] + w10[u] * tp[b * D + a + 1] + w01[u] * tp[(b + 1) * D + a] +
                 w11[u] * tp[(b + 1) * D + a + 1];
        }
        T* dst = obase + (int64_t)o * hw1;
        if (4 * q + 3 < npix) {
          const unsigned long long pk =
              static_cast<unsigned long long>(f_to_half_bits<T>(v[0]) | (f_to_half_bits<T>(v[1]) << 16)) |
              (static_cast<unsigned long long>(f_to_half_bits<T>(v[2]) | (f_to_half_bits<T>(v[3]) << 16)) << 32);
          __builtin_nontemporal_store(pk, reinterpret_cast<unsigned long long*>(dst));
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (4 * q + u < npix) dst[u] = from_f<T>(v[u]);
        }
      }
    }
    __syncthreads();                              // taps are reused by the next level
  }
}

// Adjoint: writes the complete volume_grad planes of the block's 64 pixels.
template <typename T>
__global__ void __launch_bounds__(kThreads)
corr_lookup_bwd_kernel(const float* __restrict__ coords, const T* __restrict__ grad,
                       T* __restrict__ vgrad, int hw1, int h2, int w2, int r,
                       int tiles_per_item) {
  using A = typename Acc<T>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  A* gt = reinterpret_cast<A*>(smem_raw);                   // [kPix][D*D]
  __shared__ int s_x0[kPix], s_y0[kPix];
  __shared__ float s_dx[kPix], s_dy[kPix];

  const int n = blockIdx.x / tiles_per_item;
  const int pix0 = (blockIdx.x % tiles_per_item) * kPix;
  const int D = 2 * r + 2;
  const int D2 = D * D;
  const int R = 2 * r + 1;
  const int npix = min(kPix, hw1 - pix0);
  const int t = threadIdx.x;

  if (t < kPix) {
    int x0 = 0, y0 = 0;
    float fx = 0.f, fy = 0.f;
    if (t < npix) {
      const float cx = coords[(int64_t)n * 2 * hw1 + pix0 + t];
      const float cy = coords[(int64_t)n * 2 * hw1 + hw1 + pix0 + t];
      const float flx = floorf(cx), fly = floorf(cy);
      x0 = static_cast<int>(flx) - r;
      y0 = static_cast<int>(fly) - r;
      fx = cx - flx;
      fy = cy - fly;
    }
    s_x0[t] = x0;
    s_y0[t] = y0;
    s_dx[t] = fx;
    s_dy[t] = fy;
  }
  __syncthreads();

  // tap gradients: tap (ty, tx) collects the up-to-four outputs that read it
  const T* gbase = grad + (int64_t)n * R * R * hw1 + pix0;
  for (int idx = t; idx < npix * D2; idx += kThreads) {
    const int p = idx % npix;
    const int tap = idx / npix;
    const int ty = tap / D;
    const int tx = tap - ty * D;
    const float fx = s_dx[p], fy = s_dy[p];
    const A w00 = (1.f - fx) * (1.f - fy), w10 = fx * (1.f - fy);
    const A w01 = (1.f - fx) * fy, w11 = fx * fy;
    A g = 0;
    // output (a, b) reads tap (a + i, b + j) with weight w_ij
    if (tx < R && ty < R) g += w00 * load_acc<T>(gbase + (int64_t)(tx * R + ty) * hw1 + p);
    if (tx > 0 && ty < R) g += w10 * load_acc<T>(gbase + (int64_t)((tx - 1) * R + ty) * hw1 + p);
    if (tx < R && ty > 0) g += w01 * load_acc<T>(gbase + (int64_t)(tx * R + ty - 1) * hw1 + p);
    if (tx > 0 && ty > 0) g += w11 * load_acc<T>(gbase + (int64_t)((tx - 1) * R + ty - 1) * hw1 + p);
    gt[p * D2 + tap] = g;
  }
  __syncthreads();

  const int64_t plane = (int64_t)h2 * w2;
  T* vbase = vgrad + ((int64_t)n * hw1 + pix0) * plane;
  for (int p = 0; p < npix; ++p) {
    const int xs = s_x0[p], ys = s_y0[p];
    T* pl = vbase + p * plane;
    for (int e = t; e < plane; e += kThreads) {
      const int y2 = e / w2;
      const int x2 = e - y2 * w2;
      const int ty = y2 - ys, tx = x2 - xs;
      A v = 0;
      if (ty >= 0 && ty < D && tx >= 0 && tx < D) v = gt[p * D2 + ty * D + tx];
      pl[e] = store_acc<T>(v);
    }
  }
}

template <typename T>
int launch_fwd(const LevelTable& lv, int levels, const float* coords, int64_t cb, int cp,
               int64_t cy_off, void* corr, int64_t batch, int64_t hw1, int r,
               int64_t out_batch_stride, hipStream_t s) {
  if (batch * hw1 == 0) return HSL2_OK;
  const int tiles = static_cast<int>((hw1 + kPix - 1) / kPix);
  if constexpr (sizeof(T) == 2) {
    // 16-byte volume chunks and 8-byte output quads must be aligned
    bool vec = r == 3 && hw1 % 4 == 0 && out_batch_stride % 4 == 0 && batch * tiles < (1ll << 31) &&
               reinterpret_cast<uintptr_t>(corr) % 8 == 0;
    for (int l = 0; l < levels; ++l)
      vec = vec && lv.w2[l] % 8 == 0 && lv.out_offset[l] % 4 == 0 &&
            reinterpret_cast<uintptr_t>(lv.vol[l]) % 16 == 0;
    if (vec && levels == 4) {
      hipLaunchKernelGGL(corr_lookup_vec4_kernel<T>, dim3(static_cast<unsigned>(batch * tiles)), dim3(256), 0, s, lv,
                         coords, cb, cp, cy_off, static_cast<T*>(corr), static_cast<int>(hw1), out_batch_stride,
                         tiles);
      HSL2_LAUNCH_CHECK();
      return HSL2_OK;
    }
    if (vec) {
      dim3 grid(static_cast<unsigned>(batch * tiles), levels);
      hipLaunchKernelGGL(corr_lookup_vec_kernel<T>, grid, dim3(256), 0, s, lv, coords, cb, cp, cy_off,
                         static_cast<T*>(corr), static_cast<int>(hw1), out_batch_stride, tiles);
      HSL2_LAUNCH_CHECK();
      return HSL2_OK;
    }
  }
  if constexpr (sizeof(T) <= 4) {
    if (r == 3 && levels == 4 && batch * tiles < (1ll << 31)) {
      hipLaunchKernelGGL(corr_lookup_win4_kernel<T>, dim3(static_cast<unsigned>(batch * tiles)), dim3(256), 0, s,
                         lv, coords, cb, cp, cy_off, static_cast<T*>(corr), static_cast<int>(hw1), out_batch_stride,
                         tiles);
      HSL2_LAUNCH_CHECK();
      return HSL2_OK;
    }
  }
  if ((r == 3 || r == 4) && batch * tiles < (1ll << 31)) {
    dim3 grid(static_cast<unsigned>(batch * tiles), levels);
    if (r == 3)
      hipLaunchKernelGGL((corr_lookup_win_kernel<T, 3>), grid, dim3(256), 0, s, lv, coords, cb, cp,
                         cy_off, static_cast<T*>(corr), static_cast<int>(hw1), out_batch_stride, tiles);
    else
      hipLaunchKernelGGL((corr_lookup_win_kernel<T, 4>), grid, dim3(256), 0, s, lv, coords, cb, cp,
                         cy_off, static_cast<T*>(corr), static_cast<int>(hw1), out_batch_stride, tiles);
    HSL2_LAUNCH_CHECK();
    return HSL2_OK;
  }
  // other radii: the same scheme with the radius as a runtime value
  const int64_t blocks = batch * tiles;
  HSL2_REQUIRE(blocks < (1ll << 31), HSL2_ERR_SHAPE, "corr lookup: grid too large");
  const int D = 2 * r + 2;
  const size_t lds = sizeof(typename Acc<T>::type) * kPix * (D * D + 1);
  HSL2_REQUIRE(lds <= 64 * 1024, HSL2_ERR_UNSUPPORTED, "corr lookup: radius too large");
  hipLaunchKernelGGL(corr_lookup_fwd_kernel<T>, dim3(static_cast<unsigned>(blocks), levels),
                     dim3(kThreads), lds, s, lv, coords, cb, cp, cy_off, static_cast<T*>(corr),
                     static_cast<int>(hw1), r, out_batch_stride, tiles);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

template <typename T>
int launch_bwd(const float* coords, const void* grad, void* vgrad, int64_t batch, int64_t hw1,
               int64_t h2, int64_t w2, int r, hipStream_t s) {
  const int tiles = static_cast<int>((hw1 + kPix - 1) / kPix);
  const int64_t blocks = batch * tiles;
  HSL2_REQUIRE(blocks < (1ll << 31), HSL2_ERR_SHAPE, "corr lookup: grid too large");
  const int D = 2 * r + 2;
  const size_t lds = sizeof(typename Acc<T>::type) * kPix * D * D;
  HSL2_REQUIRE(lds <= 64 * 1024, HSL2_ERR_UNSUPPORTED, "corr lookup: radius too large");
  if (blocks == 0) return HSL2_OK;
  hipLaunchKernelGGL(corr_lookup_bwd_kernel<T>, dim3(static_cast<unsigned>(blocks)), dim3(kThreads),
                     lds, s, coords, static_cast<const T*>(grad), static_cast<T*>(vgrad),
                     static_cast<int>(hw1), static_cast<int>(h2), static_cast<int>(w2), r, tiles);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

// ---- the lookup fused with the first 1x1 convolution of its consumer ----------------------
// corr_encoder's first layer (hislam2/modules/droid_net.py:86-90: Conv2d(196, C_out, 1) and
// ReLU) applied to the CorrBlock lookup without writing the 196 channels: out[n, c, p] =
// act(bias[c] + sum_k W[c, k] corr[n, k, p]), k in the lookup's channel order
// [level][x-offset][y-offset].  One 256-thread block per 64 pixels of an item: the vec kernel's
// window loads of all four levels are issued together (one memory latency per block), then per
// level the taps go to LDS and the bilinear combine rounds each channel to bf16 (as the
// lookup's own output) into an LDS row per pixel; then the 64 x C_out x 196 product on
// v_mfma_f32_32x32x16_bf16 (FP32 accumulation; each wave one 32-row slice of W, its fragments
// read from L2 a step group ahead, 16 bytes per lane), bias, ReLU and
// bf16 stores of 4 consecutive pixels.  bf16 pyramids of 4 levels, radius 3.
typedef __bf16 conv_bf16x8 __attribute__((ext_vector_type(8)));
typedef float conv_acc16 __attribute__((ext_vector_type(16)));
constexpr int kConvK = 196;                     // 4 levels x 49 channels
constexpr int kConvKP = 208;                    // padded to whole MFMA steps of 16
constexpr int kConvRow = kConvKP + 8;           // LDS row (bf16), 432 bytes
static_assert(kThreads == 4 * kPix, "one thread per (level, pixel) for the window origins");

template <int COUT>
__global__ void __launch_bounds__(kThreads)
corr_lookup_conv_kernel(LevelTable lv, const float* __restrict__ coords, int64_t cb, int cp, int64_t cy_off,
                        const __hip_bfloat16* __restrict__ W, const float* __restrict__ bias, int relu,
                        __hip_bfloat16* __restrict__ out, int hw1, int tiles_per_item) {
  using T = __hip_bfloat16;
  constexpr int RAD = 3, D = 8, R = 7, TS = D * D + 1;
  __shared__ float taps[kPix * TS];
  __shared__ int s_x0[4][kPix], s_y0[4][kPix];
  __shared__ float s_dx[4][kPix], s_dy[4][kPix];
  __shared__ __align__(16) unsigned short A[kPix][kConvRow];
  const int n = blockIdx.x / tiles_per_item;
  const int pix0 = (blockIdx.x - n * tiles_per_item) * kPix;
  const int npix = min(kPix, hw1 - pix0);
  const int t = threadIdx.x;
  if (t < kPix)
    for (int k = kConvK; k < kConvRow; ++k) A[t][k] = 0;
  // window origins of all four levels (thread t: level t / 64, pixel t % 64)
  {
    const int level = t / kPix, p = t % kPix;
    const float scale = lv.scale[level];
    int x0 = 0, y0 = 0;
    float fx = 0.f, fy = 0.f;
    if (p < npix) {
      const float* cp_ = coords + (int64_t)n * cb + (int64_t)(pix0 + p) * cp;
      const float cx = cp_[0] * scale, cy = cp_[cy_off] * scale;
      const float flx = floorf(cx), fly = floorf(cy);
      x0 = static_cast<int>(flx) - RAD;
      y0 = static_cast<int>(fly) - RAD;
      fx = cx - flx;
      fy = cy - fly;
    }
    s_x0[level][p] = x0;
    s_y0[level][p] = y0;
    s_dx[level][p] = fx;
    s_dy[level][p] = fy;
  }
  __syncthreads();
  // every level's loads are issued before any is consumed: one memory latency per block
  WinRegs g[4];
#pragma unroll
  for (int level = 0; level < 4; ++level) {
    const int h2 = lv.h2[level], w2 = lv.w2[level], layout = lv.layout[level];
    const int64_t plane = (int64_t)h2 * w2;
    const T* vbase = static_cast<const T*>(lv.vol[level]) + ((int64_t)n * hw1 + pix0) * plane;
    if (layout == kPlaneGroups24) load_segs24<T>(vbase, plane, h2, w2, npix, s_x0[level], s_y0[level], g[level]);
    else load_rows8<T>(vbase, plane, h2, w2, layout, npix, s_x0[level], s_y0[level], g[level]);
  }
#pragma unroll
  for (int level = 0; level < 4; ++level) {
    if (lv.layout[level] == kPlaneGroups24) store_segs24<T>(g[level], s_x0[level], s_y0[level], taps);
    else store_rows8<T>(g[level], taps);
    __syncthreads();
    // combine: thread (pixel p, channel group g) forms channels o = g, g + 4, ... < 49
    {
      const int p = t & 63, cg = t >> 6;
      const float fx = s_dx[level][p], fy = s_dy[level][p];
      const float w00 = (1.f - fx) * (1.f - fy), w10 = fx * (1.f - fy), w01 = (1.f - fx) * fy, w11 = fx * fy;
      const float* tp = taps + p * TS;
      for (int o = cg; o < R * R; o += 4) {
        const int a = o / R, b = o - (o / R) * R;
        const float v = w00 * tp[b * D + a] + w10 * tp[b * D + a + 1] + w01 * tp[(b + 1) * D + a] +
                        w11 * tp[(b + 1) * D + a + 1];
        A[p][level * R * R + o] = static_cast<unsigned short>(f_to_half_bits<T>(v));
      }
    }
    __syncthreads();                              // taps are reused by the next level
  }
  // 64 pixels x COUT outputs in 32 x 32 tiles: wave w takes output tile w % (COUT / 32) and
  // PH pixel halves (both for COUT = 128, so each W fragment is fetched once per block)
  constexpr int NOT = COUT / 32, NP = kThreads / kWave / NOT;
  constexpr int PH = 2 / NP;
  static_assert(NOT * NP == kThreads / kWave && PH * NP == 2, "tile split");
  const int lane = t % kWave, wave = t / kWave;
  const int otile = wave % NOT, ph0 = (wave / NOT) * PH;
  const int kh = 8 * (lane / 32);              // lane half: k [8 kh', 8 kh' + 8) of each step
  conv_acc16 acc[PH];
#pragma unroll
  for (int j = 0; j < PH; ++j)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[j][v] = 0.f;
  // W's fragments come from L2: steps in groups of KG (the last group partial), the next group's
  // loads issued before this group's MFMAs.  W rows hold 196 values: the piece [192, 200) is
  // read as four values and four zeros, pieces from 200 on are zeros (as A is there), so no
  // load leaves the row
  constexpr int KS = kConvKP / 16, KG = 4, NG = (KS + KG - 1) / KG;
  static_assert(kConvK % 8 == 4 && kConvKP - kConvK < 16, "W row tail handling");
  const int oc_l = otile * 32 + lane % 32;
  const __hip_bfloat16* wrow = W + (int64_t)oc_l * kConvK + kh;
  auto wload = [&](int grp, conv_bf16x8 (&b)[KG]) {
#pragma unroll
    for (int q = 0; q < KG; ++q) {
      const int k0 = (grp * KG + q) * 16;
      if (grp * KG + q >= KS) continue;
      // two 8-byte loads: a W row is 392 bytes, so a piece is 8-byte but not always 16-byte aligned
      typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
      const int kk = k0 + kh;
      bf16x4 lo = {0, 0, 0, 0}, hi = {0, 0, 0, 0};
      if (kk + 4 <= kConvK) lo = *reinterpret_cast<const bf16x4*>(wrow + k0);
      if (kk + 8 <= kConvK) hi = *reinterpret_cast<const bf16x4*>(wrow + k0 + 4);
      b[q] = conv_bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    }
  };
  conv_bf16x8 wb[2][KG];
  wload(0, wb[0]);
#pragma unroll
  for (int grp = 0; grp < NG; ++grp) {
    if (grp + 1 < NG) wload(grp + 1, wb[(grp + 1) & 1]);
#pragma unroll
    for (int q = 0; q < KG; ++q) {
      if (grp * KG + q >= KS) continue;
      const int k0 = (grp * KG + q) * 16;
#pragma unroll
      for (int j = 0; j < PH; ++j) {
        const conv_bf16x8 a = *reinterpret_cast<const conv_bf16x8*>(&A[(ph0 + j) * 32 + lane % 32][k0 + kh]);
        acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, wb[grp & 1][q], acc[j], 0, 0, 0);
      }
    }
  }
  // D value v of lane l: pixel row 8 (v / 4) + 4 (l / 32) + v % 4 of the half, column l % 32
  const float bb = bias[oc_l];
  T* obase = out + ((int64_t)n * COUT + oc_l) * hw1 + pix0;
#pragma unroll
  for (int j = 0; j < PH; ++j) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int px = (ph0 + j) * 32 + 8 * q + 4 * (lane / 32);
      uint32_t h[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float y = acc[j][4 * q + v] + bb;
        if (relu) y = fmaxf(y, 0.f);
        h[v] = f_to_half_bits<T>(y);
      }
      if (px + 3 < npix) {
        uint2 pk;
        pk.x = h[0] | (h[1] << 16);
        pk.y = h[2] | (h[3] << 16);
        *reinterpret_cast<uint2*>(obase + px) = pk;
      } else {
        for (int v = 0; v < 4; ++v)
          if (px + v < npix) reinterpret_cast<unsigned short*>(obase)[px + v] = static_cast<unsigned short>(h[v]);
      }
    }
  }
}

int check_lookup_args(const void* a, const void* b, const void* c, int64_t batch, int64_t h1,
                      int64_t w1, int64_t h2, int64_t w2, int radius) {
  HSL2_REQUIRE(a && b && c, HSL2_ERR_ARGUMENT, "corr lookup: null pointer");
  HSL2_REQUIRE(batch >= 0 && h1 >= 0 && w1 >= 0 && h2 > 0 && w2 > 0, HSL2_ERR_SHAPE,
               "corr lookup: bad sizes");
  HSL2_REQUIRE(radius >= 0 && radius <= 8, HSL2_ERR_ARGUMENT, "corr lookup: radius must be in [0, 8]");
  HSL2_REQUIRE(h1 * w1 < (1ll << 31) && h2 * w2 < (1ll << 31), HSL2_ERR_SHAPE,
               "corr lookup: plane too large");
  return HSL2_OK;
}

}  // namespace

int corr_lookup_fwd(const LevelTable& lv, int levels, const float* coords, int64_t cb, int cp,
                    int64_t cy_off, void* corr, int dtype, int64_t batch, int64_t hw1, int radius,
                    int64_t out_batch_stride, hipStream_t s) {
  switch (dtype) {
    case HSL2_F32: return launch_fwd<float>(lv, levels, coords, cb, cp, cy_off, corr, batch, hw1, radius, out_batch_stride, s);
    case HSL2_F16: return launch_fwd<__half>(lv, levels, coords, cb, cp, cy_off, corr, batch, hw1, radius, out_batch_stride, s);
    case HSL2_BF16: return launch_fwd<__hip_bfloat16>(lv, levels, coords, cb, cp, cy_off, corr, batch, hw1, radius, out_batch_stride, s);
    case HSL2_F64: return launch_fwd<double>(lv, levels, coords, cb, cp, cy_off, corr, batch, hw1, radius, out_batch_stride, s);
    default: set_error("corr lookup: unknown dtype"); return HSL2_ERR_ARGUMENT;
  }
}

}  // namespace hsl2

using namespace hsl2;

extern "C" int hsl2_corr_index_forward(const void* volume, const float* coords, void* corr,
                                       int dtype, int64_t batch, int64_t h1, int64_t w1,
                                       int64_t h2, int64_t w2, int radius, void* stream) {
  int st = check_lookup_args(volume, coords, corr, batch, h1, w1, h2, w2, radius);
  if (st) return st;
  const int64_t R = 2 * radius + 1;
  LevelTable lv{};
  lv.vol[0] = volume;
  lv.h2[0] = static_cast<int>(h2);
  lv.w2[0] = static_cast<int>(w2);
  lv.layout[0] = kPlaneRows;
  lv.scale[0] = 1.0f;
  lv.out_offset[0] = 0;
  // planar coords [batch][2][h1*w1]
  return corr_lookup_fwd(lv, 1, coords, 2 * h1 * w1, 1, h1 * w1, corr, dtype, batch, h1 * w1, radius,
                         R * R * h1 * w1, as_stream(stream));
}

extern "C" int hsl2_corr_pyramid_lookup(const void* const* levels, int num_levels,
                                        const float* coords, void* corr, int dtype, int64_t batch,
                                        int64_t h1, int64_t w1, int radius, int tiled,
                                        void* stream) {
  HSL2_REQUIRE(levels && num_levels >= 1 && num_levels <= 4, HSL2_ERR_ARGUMENT,
               "pyramid lookup: 1..4 levels");
  LevelTable lv{};
  const int64_t R = 2 * radius + 1;
  for (int i = 0; i < num_levels; ++i) {
    const int64_t h2 = h1 >> i, w2 = w1 >> i;
    int st = check_lookup_args(levels[i], coords, corr, batch, h1, w1, h2, w2, radius);
    if (st) return st;
    lv.vol[i] = levels[i];
    lv.h2[i] = static_cast<int>(h2);
    lv.w2[i] = static_cast<int>(w2);
    lv.layout[i] = pyramid_layout(i, h1 * w1, h2, w2, tiled);
    lv.scale[i] = 1.0f / static_cast<float>(1 << i);
    lv.out_offset[i] = i * R * R * h1 * w1;
  }
  // interleaved coords [batch][h1*w1][2]
  return corr_lookup_fwd(lv, num_levels, coords, 2 * h1 * w1, 2, 1, corr, dtype, batch, h1 * w1,
                         radius, num_levels * R * R * h1 * w1, as_stream(stream));
}

extern "C" int hsl2_corr_index_backward(const float* coords, const void* corr_grad,
                                        void* volume_grad, int dtype, int64_t batch, int64_t h1,
                                        int64_t w1, int64_t h2, int64_t w2, int radius,
                                        void* stream) {
  int st = check_lookup_args(coords, corr_grad, volume_grad, batch, h1, w1, h2, w2, radius);
  if (st) return st;
  const int64_t hw1 = h1 * w1;
  hipStream_t s = as_stream(stream);
  switch (dtype) {
    case HSL2_F32: return launch_bwd<float>(coords, corr_grad, volume_grad, batch, hw1, h2, w2, radius, s);
    case HSL2_F16: return launch_bwd<__half>(coords, corr_grad, volume_grad, batch, hw1, h2, w2, radius, s);
    case HSL2_BF16: return launch_bwd<__hip_bfloat16>(coords, corr_grad, volume_grad, batch, hw1, h2, w2, radius, s);
    case HSL2_F64: return launch_bwd<double>(coords, corr_grad, volume_grad, batch, hw1, h2, w2, radius, s);
    default: set_error("corr lookup: unknown dtype"); return HSL2_ERR_ARGUMENT;
  }
}

extern "C" int hsl2_corr_lookup_conv(const void* const* levels, int num_levels, const float* coords,
                                     const void* weight, const float* bias, int relu, void* out, int64_t batch,
                                     int64_t h1, int64_t w1, int radius, int tiled, int64_t cout, void* stream) {
  HSL2_REQUIRE(levels && coords && weight && bias && out, HSL2_ERR_ARGUMENT, "corr lookup conv: null pointer");
  HSL2_REQUIRE(num_levels == 4 && radius == 3, HSL2_ERR_UNSUPPORTED,
               "corr lookup conv: 4 levels of radius 3 (196 channels)");
  HSL2_REQUIRE(cout == 128 || cout == 64, HSL2_ERR_UNSUPPORTED, "corr lookup conv: 64 or 128 output channels");
  const int64_t hw1 = h1 * w1;
  HSL2_REQUIRE(batch >= 0 && h1 > 0 && w1 > 0 && hw1 % 4 == 0 && hw1 < (1ll << 31), HSL2_ERR_SHAPE,
               "corr lookup conv: h * w must be a multiple of 4");
  LevelTable lv{};
  for (int i = 0; i < 4; ++i) {
    const int64_t h2 = h1 >> i, w2 = w1 >> i;
    HSL2_REQUIRE(levels[i] && w2 % 8 == 0 && h2 > 0 && reinterpret_cast<uintptr_t>(levels[i]) % 16 == 0,
                 HSL2_ERR_UNSUPPORTED, "corr lookup conv: every level width a multiple of 8, 16-byte aligned");
    lv.vol[i] = levels[i];
    lv.h2[i] = static_cast<int>(h2);
    lv.w2[i] = static_cast<int>(w2);
    lv.layout[i] = pyramid_layout(i, hw1, h2, w2, tiled);
    lv.scale[i] = 1.0f / static_cast<float>(1 << i);
  }
  HSL2_REQUIRE(reinterpret_cast<uintptr_t>(weight) % 8 == 0 && reinterpret_cast<uintptr_t>(out) % 8 == 0,
               HSL2_ERR_ARGUMENT, "corr lookup conv: weight and out must be 8-byte aligned");
  if (batch == 0) return HSL2_OK;
  const int tiles = static_cast<int>((hw1 + kPix - 1) / kPix);
  HSL2_REQUIRE(batch * tiles < (1ll << 31), HSL2_ERR_SHAPE, "corr lookup conv: grid too large");
  dim3 grid(static_cast<unsigned>(batch * tiles));
  const auto* Wb = static_cast<const __hip_bfloat16*>(weight);
  auto* ob = static_cast<__hip_bfloat16*>(out);
  // interleaved coords [batch][h1*w1][2], as hsl2_corr_pyramid_lookup
  if (cout == 128)
    hipLaunchKernelGGL(corr_lookup_conv_kernel<128>, grid, dim3(kThreads), 0, as_stream(stream), lv, coords,
                       2 * hw1, 2, int64_t{1}, Wb, bias, relu, ob, static_cast<int>(hw1), tiles);
  else
    hipLaunchKernelGGL(corr_lookup_conv_kernel<64>, grid, dim3(kThreads), 0, as_stream(stream), lv, coords,
                       2 * hw1, 2, int64_t{1}, Wb, bias, relu, ob, static_cast<int>(hw1), tiles);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}
