// All-pairs correlation volume with a fused average-pool pyramid, on MFMA.
//
// Semantics of CorrBlock.__init__ / CorrBlock.corr (hislam2/modules/corr.py:24-38,
// 63-71): corr[n, p, q] = sum_c (f1[n, c, p] / 4) (f2[n, c, q] / 4) for pixels p of
// image 1 and q of image 2, then levels i = 1..L-1 by repeated 2x2 average pooling
// over q's (h, w) grid.
//
// One 256-thread block computes a 32 (p) x 8 image rows x WT columns (q) tile with
// v_mfma_f32_32x32x16_{bf16,f16} (f32 operands: v_mfma_f32_32x32x2f32), f32 accumulation.  The
// feature maps arrive in MFMA fragment tiles (frag_tile_offset: 32 pixels x G channels
// contiguous), so every lane fetches its operand fragment (8 consecutive channels of one pixel,
// 16 bytes) straight from memory and one wave's fetch is 1 KB of contiguous memory: no LDS
// staging and no barrier in the channel loop, which for 128 channels keeps three channel groups
// of fetches in flight.  Each wave owns four 32x32 output blocks.  Blocks are ordered so that
// each XCD works through whole edges (the operands of one edge stay in that XCD's L2).  The
// finished f32 tile is staged in LDS and written, with non-temporal stores, as level 0 plus every
// pooled level it fully contains (8 rows cover 2x2, 4x4 and 8x8 pooling), so the volume is
// written once and never read back.
#include "common.hpp"

namespace hsl2 {
namespace {

constexpr int kThreads = 256;
constexpr int kTP = 32;       // p rows per MFMA operand
constexpr int kER = 16;       // p rows staged in LDS and emitted per pass (two passes per operand)
constexpr int kRows = 8;      // image-2 rows per block

typedef float acc16 __attribute__((ext_vector_type(16)));

// Operand fragments are fetched 16 bytes per lane.  The sum over channels is order-free, so
// within a group of channels the lanes of the two MFMA k-halves take contiguous 16-byte pieces:
// for bf16 / f16 a group is 16 channels, lane half kg loads [8 kg, 8 kg + 8), one
// v_mfma_f32_32x32x16 per group; for f32 a group is 8 channels, lane half kg loads [4 kg, 4 kg + 4)
// and feeds one value to each of four v_mfma_f32_32x32x2f32.  A and B use the same assignment,
// so every channel is multiplied with its own partner exactly once.
template <typename T> struct Mfma;
template <> struct Mfma<__hip_bfloat16> {
  static constexpr int G = frag_group<__hip_bfloat16>();   // channels per group
  static constexpr int S = 1;                               // MFMAs per group
  typedef __bf16 Frag __attribute__((ext_vector_type(8)));
  static __device__ __forceinline__ Frag load(const __hip_bfloat16* p) { return *reinterpret_cast<const Frag*>(p); }
  static __device__ __forceinline__ acc16 run(const Frag& a, const Frag& b, int, acc16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<__half> {
  static constexpr int G = frag_group<__half>();
  static constexpr int S = 1;
  typedef _Float16 Frag __attribute__((ext_vector_type(8)));
  static __device__ __forceinline__ Frag load(const __half* p) { return *reinterpret_cast<const Frag*>(p); }
  static __device__ __forceinline__ acc16 run(const Frag& a, const Frag& b, int, acc16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  static constexpr int G = frag_group<float>();
  static constexpr int S = 4;
  typedef float Frag __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ Frag load(const float* p) { return *reinterpret_cast<const Frag*>(p); }
  static __device__ __forceinline__ acc16 run(const Frag& a, const Frag& b, int s, acc16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], b[s], c, 0, 0, 0);
  }
};

// Chunk width for vector stores of a row segment of `width` elements.
__host__ __device__ constexpr int chunk_of(int width) {
  return width % 8 == 0 ? 8 : (width % 4 == 0 ? 4 : (width % 2 == 0 ? 2 : 1));
}

template <typename To, int CW>
struct alignas(CW * sizeof(To)) OutV {
  To v[CW];
};

// Non-temporal store of one chunk: the volume is written once and only read later by the
// lookup, so streaming it past the L2 keeps the feature-map operands resident there.
template <typename To, int CW>
__device__ __forceinline__ void store_nt(To* dst, const OutV<To, CW>& o) {
  constexpr int B = CW * static_cast<int>(sizeof(To));
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  if constexpr (B == 32) {
    struct Two { u32x4 a, b; };
    const Two t = __builtin_bit_cast(Two, o);
    __builtin_nontemporal_store(t.a, reinterpret_cast<u32x4*>(dst));
    __builtin_nontemporal_store(t.b, reinterpret_cast<u32x4*>(dst) + 1);
  } else if constexpr (B == 16) {
    __builtin_nontemporal_store(__builtin_bit_cast(u32x4, o), reinterpret_cast<u32x4*>(dst));
  } else if constexpr (B == 8) {
    __builtin_nontemporal_store(__builtin_bit_cast(u32x2, o), reinterpret_cast<u32x2*>(dst));
  } else if constexpr (B == 4) {
    __builtin_nontemporal_store(__builtin_bit_cast(unsigned, o), reinterpret_cast<unsigned*>(dst));
  } else if constexpr (B == 2) {
    __builtin_nontemporal_store(__builtin_bit_cast(unsigned short, o), reinterpret_cast<unsigned short*>(dst));
  } else {
    *reinterpret_cast<OutV<To, CW>*>(dst) = o;
  }
}

// Pooled level (hierarchical: each level averages 2x2 of the level above it, as repeated
// avg_pool2d does).  src holds the level above for the tile ([kER][RS][WS] floats, pitch PS per
// p); the values are stored to `out` (layout `layout`, pyramid_offset) and returned in v, so
// that the caller can place them in LDS for the next level once every thread has finished
// reading src.  Thread chunk c covers CW consecutive x of one (p, y), with p the slowest index;
// for pixel-interleaved levels p is the fastest index (the p of one segment position are then
// one contiguous run in memory), and in the 2 x 4 segment layout a chunk is one whole segment
// (two rows of four x), so that every lane writes whole 16-element segments; pool_stash uses the
// same split.
template <int RS, int WS>
struct Pool {
  static constexpr int RD = RS / 2, WD = WS / 2, CW = chunk_of(WD);
  static constexpr int NCH = kER * RD * WD / CW;     // chunks of the tile
  static constexpr int NIT = (NCH + kThreads - 1) / kThreads;
  // kPlaneGroups24 levels: a chunk is one whole 2-row x 4-column segment
  static constexpr bool SEG = RD % 2 == 0 && WD % 4 == 0;
  static constexpr int NCH24 = SEG ? kER * RD * WD / 8 : 1;
  static constexpr int NIT24 = (NCH24 + kThreads - 1) / kThreads;
  static constexpr int VN = NIT * CW > NIT24 * 8 ? NIT * CW : NIT24 * 8;   // values per thread
  static constexpr int PD = RD * WD + 4;             // pitch per p of this level in LDS
};

template <int RS, int WS>
__device__ __forceinline__ void pool_split(int c, bool pfast, int& p, int& y, int& x) {
  using PL = Pool<RS, WS>;
  constexpr int WD = PL::WD, CW = PL::CW, PER = PL::RD * WD / CW;
  const int r = pfast ? c / kER : c % PER;
  p = pfast ? c % kER : c / PER;
  y = r / (WD / CW);
  x = (r % (WD / CW)) * CW;
}

template <typename To, int RS, int WS, int PS>
__device__ __forceinline__ void pool_compute(const float* __restrict__ src, To* __restrict__ out,
                                             int64_t plane0, int64_t plane_stride, int gy0, int gx0,
                                             int wl, int layout, float (&v)[Pool<RS, WS>::VN]) {
  using PL = Pool<RS, WS>;
  constexpr int CW = PL::CW;
  const int t = threadIdx.x;
  if (layout == kPlaneGroups24) {
    // one whole segment (output rows 2 yp, 2 yp + 1, columns x .. x + 3) per chunk, p fastest:
    // the 16 p of a chunk row are one contiguous run of 16-element segments in memory
    if constexpr (PL::SEG) {
      constexpr int XS = PL::WD / 4;
#pragma unroll
      for (int it = 0; it < PL::NIT24; ++it) {
        const int c = it * kThreads + t;
        if (PL::NCH24 % kThreads == 0 || c < PL::NCH24) {
          const int p = c % kER, r = c / kER;
          const int yp = r / XS, x = (r % XS) * 4;
          OutV<To, 8> o;
#pragma unroll
          for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const float* a = src + p * PS + (2 * (2 * yp + dy)) * WS + 2 * (x + e);
              const float s = 0.25f * ((a[0] + a[1]) + (a[WS] + a[WS + 1]));
              v[it * 8 + dy * 4 + e] = s;
              o.v[dy * 4 + e] = from_f<To>(s);
            }
          store_nt<To, 8>(out + pyramid_offset(layout, plane0 + p, gy0 + 2 * yp, gx0 + x, wl, plane_stride), o);
        }
      }
    }
    return;
  }
  const bool pfast = layout == kPlaneGroups;
#pragma unroll
  for (int it = 0; it < PL::NIT; ++it) {
    const int c = it * kThreads + t;
    if (PL::NCH % kThreads == 0 || c < PL::NCH) {
      int p, y, x;
      pool_split<RS, WS>(c, pfast, p, y, x);
      const float* a = src + p * PS + (2 * y) * WS + 2 * x;
      OutV<To, CW> o;
#pragma unroll
      for (int e = 0; e < CW; ++e) {
        v[it * CW + e] = 0.25f * ((a[2 * e] + a[2 * e + 1]) + (a[WS + 2 * e] + a[WS + 2 * e + 1]));
        o.v[e] = from_f<To>(v[it * CW + e]);
      }
      // a kPlaneGroups level has CW == 8 and gx a multiple of 8: the chunk is one segment
      store_nt<To, CW>(out + pyramid_offset(layout, plane0 + p, gy0 + y, gx0 + x, wl, plane_stride), o);
    }
  }
}

template <int RS, int WS>
__device__ __forceinline__ void pool_stash(float* __restrict__ dst, int layout, const float (&v)[Pool<RS, WS>::VN]) {
  using PL = Pool<RS, WS>;
  constexpr int WD = PL::WD, CW = PL::CW;
  const int t = threadIdx.x;
  if (layout == kPlaneGroups24) {
    if constexpr (PL::SEG) {
      constexpr int XS = WD / 4;
#pragma unroll
      for (int it = 0; it < PL::NIT24; ++it) {
        const int c = it * kThreads + t;
        if (PL::NCH24 % kThreads == 0 || c < PL::NCH24) {
          const int p = c % kER, r = c / kER;
          const int yp = r / XS, x = (r % XS) * 4;
#pragma unroll
          for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int e = 0; e < 4; ++e) dst[p * PL::PD + (2 * yp + dy) * WD + x + e] = v[it * 8 + dy * 4 + e];
        }
      }
    }
    return;
  }
  const bool pfast = layout == kPlaneGroups;
#pragma unroll
  for (int it = 0; it < PL::NIT; ++it) {
    const int c = it * kThreads + t;
    if (PL::NCH % kThreads == 0 || c < PL::NCH) {
      int p, y, x;
      pool_split<RS, WS>(c, pfast, p, y, x);
#pragma unroll
      for (int e = 0; e < CW; ++e) dst[p * PL::PD + y * WD + x + e] = v[it * CW + e];
    }
  }
}

// The f32 tile of kER p rows staged in LDS -> level 0 and the pooled levels of those rows.
// pl0 = index of the first p row's plane.  Ends with every thread done reading LDS.
template <typename To, int WT>
__device__ void emit_rows(const float* __restrict__ sout_in, float* __restrict__ sout, To* __restrict__ l0,
                          To* __restrict__ l1, To* __restrict__ l2, To* __restrict__ l3, int64_t pl0,
                          int h, int w, int y0, int x0, int levels, int tiled) {
  constexpr int NQ = kRows * WT;
  constexpr int SQ = NQ + 4;
  const int HW = h * w;
  const int t = threadIdx.x;
  // level 0: row segments of 8 elements (one row of an 8x8 block in the tiled layout), one
  // vector store each; consecutive threads read consecutive segments of the LDS tile
  const int lay0 = pyramid_layout(0, HW, h, w, tiled);
  constexpr int NCH0 = kER * NQ / 8;
#pragma unroll
  for (int c0 = 0; c0 < NCH0; c0 += kThreads) {
    const int c = c0 + t;
    const int p = c / (NQ / 8), r = c % (NQ / 8);
    const int y = r / (WT / 8), x = (r % (WT / 8)) * 8;
    const float* src = sout_in + p * SQ + y * WT + x;
    OutV<To, 8> o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o.v[e] = from_f<To>(src[e]);
    store_nt<To, 8>(l0 + pyramid_offset(lay0, pl0 + p, y0 + y, x0 + x, w, HW), o);
  }
  if (levels < 2) {
    __syncthreads();
    return;
  }
  // pooled levels, each from the level above; every level is kept in LDS for the next one,
  // inside the (then finished) f32 tile region: level 1 at its start, level 2 behind it
  using P1 = Pool<kRows, WT>;
  using P2 = Pool<kRows / 2, WT / 2>;
  using P3 = Pool<kRows / 4, WT / 4>;
  float* s1 = sout;
  float* s2 = s1 + kER * P1::PD;
  {
    float v[P1::VN];
    const int hl = h >> 1, wl = w >> 1, lay = pyramid_layout(1, HW, hl, wl, tiled);
    pool_compute<To, kRows, WT, SQ>(sout_in, l1, pl0, (int64_t)hl * wl, y0 >> 1, x0 >> 1, wl, lay, v);
    __syncthreads();                               // every read of the f32 tile is done
    if (levels < 3) return;
    pool_stash<kRows, WT>(s1, lay, v);
  }
  __syncthreads();
  {
    float v[P2::VN];
    const int hl = h >> 2, wl = w >> 2, lay = pyramid_layout(2, HW, hl, wl, tiled);
    pool_compute<To, kRows / 2, WT / 2, P1::PD>(s1, l2, pl0, (int64_t)hl * wl, y0 >> 2, x0 >> 2, wl, lay, v);
    if (levels < 4) {
      __syncthreads();
      return;
    }
    pool_stash<kRows / 2, WT / 2>(s2, lay, v);      // disjoint from s1: no barrier needed first
  }
  __syncthreads();
  {
    float v[P3::VN];
    const int hl = h >> 3, wl = w >> 3, lay = pyramid_layout(3, HW, hl, wl, tiled);
    pool_compute<To, kRows / 4, WT / 4, P2::PD>(s2, l3, pl0, (int64_t)hl * wl, y0 >> 3, x0 >> 3, wl, lay, v);
  }
  __syncthreads();
}

// 1-D grid over (p tile, q tile, edge), p tile fastest; see the block order note above.
// A block covers 32 p rows x 8 image rows x WT columns; WT is a compile-time constant so that
// every index split is a shift or a constant division.  NG > 0: the channel-group count C / G
// as a compile-time constant (the channel loop is unrolled with kPipe groups of fetches in
// flight); NG == 0: any multiple of G, one group at a time.
constexpr int kPipe = 3;

template <typename Ti, typename To, int WT, int NG>
__global__ void __launch_bounds__(kThreads)
corr_build_kernel(const Ti* __restrict__ f1, const Ti* __restrict__ f2, To* __restrict__ l0,
                  To* __restrict__ l1, To* __restrict__ l2, To* __restrict__ l3, int C, int h,
                  int w, int levels, int tiled, int ptiles, int qtiles, int64_t nblocks) {
  constexpr int NQ = kRows * WT;                   // columns of the tile
  constexpr int SQ = NQ + 4;                       // padded LDS row of the f32 tile
  constexpr int PW = NQ / 32 / 4;                  // 32-wide MFMA blocks per wave
  using M = Mfma<Ti>;
  using Frag = typename M::Frag;
  constexpr int G = M::G;
  constexpr int FT = kFragPixels * G;              // elements of one fragment tile
  extern __shared__ __align__(16) unsigned char smem[];
  float* sout = reinterpret_cast<float*>(smem);    // [kER][SQ]

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
  const int t = threadIdx.x;
  const int wave = t / kWave, lane = t % kWave;
  // lane (i, kg): pixel i of a 32-pixel fragment tile, channels [kg G / 2, (kg + 1) G / 2) of
  // each group; group g of the tile starting at pixel P (a multiple of 32) is at P C + g FT
  const int i = lane % 32, kg = lane / 32;
  const int lane_off = i * G + kg * (G / 2);
  const Ti* ap = f1 + ((int64_t)n * HW + p0) * C + lane_off;
  const Ti* bp[PW];
#pragma unroll
  for (int k = 0; k < PW; ++k) {
    // the 32 columns of MFMA block k are 32 consecutive pixels starting at a multiple of 32
    // (WT is 64 or the whole width)
    const int q0 = (wave * PW + k) * 32;
    const int qpix0 = (y0 + q0 / WT) * w + x0 + q0 % WT;
    bp[k] = f2 + ((int64_t)n * HW + qpix0) * C + lane_off;
  }

  acc16 acc[PW];
#pragma unroll
  for (int k = 0; k < PW; ++k)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[k][v] = 0.f;
  if constexpr (NG > 0) {
    constexpr int D = NG < kPipe ? NG : kPipe;
    Frag fa[D], fb[D][PW];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      fa[d] = M::load(ap + d * FT);
#pragma unroll
      for (int k = 0; k < PW; ++k) fb[d][k] = M::load(bp[k] + d * FT);
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const int sl = g % D;
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int st = 0; st < M::S; ++st)
#pragma unroll
        for (int k = 0; k < PW; ++k) acc[k] = M::run(fa[sl], fb[sl][k], st, acc[k]);
      __builtin_amdgcn_sched_barrier(0);
      if (g + D < NG) {
        fa[sl] = M::load(ap + (g + D) * FT);
#pragma unroll
        for (int k = 0; k < PW; ++k) fb[sl][k] = M::load(bp[k] + (g + D) * FT);
      }
    }
  } else {
    for (int g = 0; g < C / G; ++g) {
      const Frag a = M::load(ap + g * FT);
      Frag bf[PW];
#pragma unroll
      for (int k = 0; k < PW; ++k) bf[k] = M::load(bp[k] + g * FT);
#pragma unroll
      for (int st = 0; st < M::S; ++st)
#pragma unroll
        for (int k = 0; k < PW; ++k) acc[k] = M::run(a, bf[k], st, acc[k]);
    }
  }

  // the 32 p rows through the LDS tile 16 at a time; MFMA output layout: value v of lane l
  // is row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32 (rows 0-15 are values 0-7 of each
  // lane, rows 16-31 values 8-15)
#pragma unroll
  for (int part = 0; part < kTP / kER; ++part) {
#pragma unroll
    for (int k = 0; k < PW; ++k) {
      const int q = (wave * PW + k) * 32 + lane % 32;
#pragma unroll
      for (int v = 8 * part; v < 8 * part + 8; ++v) {
        const int p = 8 * (v / 4) + 4 * (lane / 32) + (v % 4) - kER * part;
        sout[p * SQ + q] = acc[k][v] * 0.0625f;
      }
    }
    __syncthreads();
    emit_rows<To, WT>(sout, sout, l0, l1, l2, l3, (int64_t)n * HW + p0 + part * kER, h, w, y0, x0, levels, tiled);
  }
}

template <typename Ti, typename To, int WT, int NG>
int launch_wt(const void* f1, const void* f2, void* const* levels, int64_t batch, int64_t C,
              int64_t h, int64_t w, int L, int tiled, hipStream_t s) {
  constexpr int NQ = kRows * WT, SQ = NQ + 4;
  const size_t lds = sizeof(float) * kER * SQ;
  HSL2_REQUIRE(lds <= 160 * 1024, HSL2_ERR_UNSUPPORTED, "corr build: LDS budget exceeded");
  const int ptiles = static_cast<int>(h * w / kTP);
  const int qtiles = static_cast<int>((h / kRows) * (w / WT));
  const int64_t nblocks = (int64_t)ptiles * qtiles * batch;
  if (nblocks == 0) return HSL2_OK;
  HSL2_REQUIRE(nblocks < (1ll << 31), HSL2_ERR_SHAPE, "corr build: grid too large");
  // pool_compute stores whole segments: 8-element chunks for kPlaneGroups levels, chunks of 4
  // or 8 for kPlaneGroups24 levels
  for (int i = 1; i < L; ++i) {
    const int lay = pyramid_layout(i, h * w, h >> i, w >> i, tiled);
    const int cw = chunk_of(WT >> i);
    HSL2_REQUIRE(!(lay == kPlaneGroups && cw != 8) && !(lay == kPlaneGroups24 && cw % 4 != 0),
                 HSL2_ERR_UNSUPPORTED, "corr build: pyramid layout and tile width do not match");
  }
  To* l[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int i = 0; i < L; ++i) l[i] = static_cast<To*>(levels[i]);
  hipLaunchKernelGGL((corr_build_kernel<Ti, To, WT, NG>), dim3(static_cast<unsigned>(nblocks)), dim3(kThreads), lds, s,
                     static_cast<const Ti*>(f1), static_cast<const Ti*>(f2), l[0], l[1], l[2], l[3],
                     static_cast<int>(C), static_cast<int>(h), static_cast<int>(w), L, tiled, ptiles, qtiles,
                     nblocks);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

// tile widths: 64 for widths that are multiples of 64, else the width itself (16, 32 or 48)
template <typename Ti, typename To, int NG>
int launch_ng(const void* f1, const void* f2, void* const* levels, int64_t batch, int64_t C,
              int64_t h, int64_t w, int L, int tiled, hipStream_t s) {
  if (w % 64 == 0) return launch_wt<Ti, To, 64, NG>(f1, f2, levels, batch, C, h, w, L, tiled, s);
  if (w == 48) return launch_wt<Ti, To, 48, NG>(f1, f2, levels, batch, C, h, w, L, tiled, s);
  if (w == 32) return launch_wt<Ti, To, 32, NG>(f1, f2, levels, batch, C, h, w, L, tiled, s);
  if (w == 16) return launch_wt<Ti, To, 16, NG>(f1, f2, levels, batch, C, h, w, L, tiled, s);
  set_error("corr build: width must be a multiple of 64, or 16, 32 or 48");
  return HSL2_ERR_UNSUPPORTED;
}

// 128 channels (the update network's feature maps) get the unrolled, pipelined channel loop
template <typename Ti, typename To>
int launch(const void* f1, const void* f2, void* const* levels, int64_t batch, int64_t C,
           int64_t h, int64_t w, int L, int tiled, hipStream_t s) {
  HSL2_REQUIRE(C % Mfma<Ti>::G == 0, HSL2_ERR_UNSUPPORTED, "corr build: channel count not a multiple of the group");
  constexpr int kNG128 = 128 / Mfma<Ti>::G;
  if (C == 128) return launch_ng<Ti, To, kNG128>(f1, f2, levels, batch, C, h, w, L, tiled, s);
  return launch_ng<Ti, To, 0>(f1, f2, levels, batch, C, h, w, L, tiled, s);
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
  HSL2_REQUIRE(channels % 16 == 0, HSL2_ERR_UNSUPPORTED, "corr build: channels must be a multiple of 16");
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
