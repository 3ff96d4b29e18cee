// Dense Cholesky factorisation and solve of the reduced camera system on the GPU.
//
// Replaces the host-side Eigen SimplicialLLT of SparseBlock::solve
// (src/droid_kernels.cu:1659-1680).  The system is held densely, column-major, padded to a
// multiple of 64 with an identity block (so every tile is full) and factored in place by a
// right-looking tiled algorithm over 64x64 tiles:
//
//   for each tile column k:
//     chol_panel   one block per tile row i >= k: in every block two waves factor the
//                  diagonal tile in registers (columns 0-31 and 32-63; redundantly in every
//                  block, so no inter-block wait is needed) and publish the finished columns
//                  of L to LDS behind a workgroup-scope counter, four at a time; a third wave
//                  consumes the columns four at a time as they arrive (no barrier): block 0 forms
//                  L_kk^-1 (kept in a separate diagonal buffer, A_kk is still being read by the
//                  other blocks) and y_k = L_kk^-1 b_k, the others solve L_ik L_kk^T = A_ik
//                  row by row
//     chol_update  one block per lower trailing tile (i, j), k < j <= i: A_ij -= L_ik L_jk^T,
//                  plus one block per tile row i > k: b_i -= L_ik y_k
//   back-substitution, one chol_back_all launch: one block per tile row j applies
//                  y_j -= L_kj^T x_k as each x_k (k > j) is published through an agent-scope
//                  counter, then writes x_j = (L_jj^-1)^T y_j and publishes it (a single-tile
//                  system: block 0's consumer wave forms x = (L_00^-1)^T y itself, one launch)
//
// A pivot that is not positive (or NaN) sets info = 1, as Eigen's LLT reports a
// NumericalIssue; the caller then treats the solution as zero.  Templated on the element
// type: double for the split-precision path, float for the plain FP32 path.
#include <utility>

#include "common.hpp"

namespace hsl2 {

constexpr int kTile = 64;
// info[0]: factorisation status; info[kBackFlags + j]: x_j published (back-substitution)
constexpr int kBackFlags = 4;
constexpr int kLd = kTile + 1;      // padded LDS row

namespace {

// Copy a 64x64 column-major tile (column stride ld) into LDS T[r][c]: every thread issues
// its 16 loads before any LDS store, so they are in flight together.
template <typename R>
__device__ __forceinline__ void tile_to_lds(const R* __restrict__ src, int64_t ld, R (*T)[kLd]) {
  const int t = threadIdx.x;
  R v[16];
#pragma unroll
  for (int it = 0; it < 16; ++it) {
    const int idx = t + 256 * it, c = idx >> 6, r = idx & 63;
    v[it] = src[r + c * ld];
  }
#pragma unroll
  for (int it = 0; it < 16; ++it) {
    const int idx = t + 256 * it, c = idx >> 6, r = idx & 63;
    T[r][c] = v[it];
  }
}

// 1/x and 1/sqrt(x) from the hardware estimates refined by Newton steps (two for double,
// reaching full precision from the ~2^-26 estimate; one for float), instead of the IEEE
// division / square-root sequences, which sit on the factorisation's critical path.
__device__ __forceinline__ double fast_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = r * (2.0 - x * r);
  r = r * (2.0 - x * r);
  return r;
}
__device__ __forceinline__ float fast_rcp(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  return r * (2.f - x * r);
}
__device__ __forceinline__ double fast_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * (1.5 - 0.5 * x * y * y);
  y = y * (1.5 - 0.5 * x * y * y);
  return y;
}
// one Newton step from the hardware estimate: ~1.5 ulp, for the factor's critical path, where
// each pivot's latency adds to every column
__device__ __forceinline__ double pivot_rsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return y * (1.5 - 0.5 * x * y * y);
}
__device__ __forceinline__ float pivot_rsqrt(float x) { return fast_rsqrt(x); }
__device__ __forceinline__ float fast_rsqrt(float x) {
  float y = __builtin_amdgcn_rsqf(x);
  return y * (1.5f - 0.5f * x * y * y);
}

// Broadcast lane `lane`'s value to the whole wave (v_readlane into SGPRs; `lane` must be
// wave-uniform -- here always a compile-time constant after unrolling).
template <typename R> __device__ __forceinline__ R lane_bcast(R v, int lane);
template <> __device__ __forceinline__ double lane_bcast<double>(double v, int lane) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
template <> __device__ __forceinline__ float lane_bcast<float>(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Wait until the LDS word *flag (written by another wave of the block) reaches v.  The poll
// loop is a single asm block so the caller's fully unrolled code stays straight-line for the
// register allocator; the load is waited for inside the block.  Bounded: after 2^22 polls it
// gives up and returns false (the caller then reports a failed factorisation).
__device__ __forceinline__ bool wait_flag(const int* flag, int v) {
  const unsigned addr = static_cast<unsigned>(reinterpret_cast<uintptr_t>(flag));
  int val;
  unsigned left;
  asm volatile(
      "s_mov_b32 %[left], 0x400000\n"
      "1:\n"
      "ds_read_b32 %[val], %[addr]\n"
      "s_waitcnt lgkmcnt(0)\n"
      "v_cmp_lt_i32 vcc, %[val], %[want]\n"
      "s_cbranch_vccz 2f\n"
      "s_sub_u32 %[left], %[left], 1\n"
      "s_cmp_eq_u32 %[left], 0\n"
      "s_cbranch_scc1 2f\n"
      "s_sleep 1\n"
      "s_branch 1b\n"
      "2:\n"
      : [val] "=&v"(val), [left] "=&s"(left)
      : [addr] "v"(addr), [want] "v"(v)
      : "vcc", "scc", "memory");
  return left != 0;
}

// LDS shared by the waves of a panel block.
template <typename R> struct PanelShared {
  R Lc[kTile][kLd];        // Lc[c][r] = L_kk[r][c]
  R dinv[kTile];           // 1 / L_kk[c][c]
  R bk[kTile];             // b_k (block 0), staged before the factor for y_k = L_kk^-1 b_k
  int published;           // columns of L_kk published so far
};

// ---- the diagonal factor on kFactorWaves waves -------------------------------------------
// Factor wave w owns the W columns [OWN0, OWN0 + W) = [kFactorCol[w], kFactorCol[w + 1]) of the
// diagonal tile (lane r: row r, a[s] = A_kk[r][OWN0 + s]); the ranges start at multiples of four,
// the publication granule.  It first applies the rank-1 updates of the columns before its own,
// four published columns at a time, then factors its columns four at a time:
//   * the 4 x 4 diagonal block is read from its four lanes by v_readlane (ten values) and
//     factored in every lane alike: pivots e_j, their reciprocal square roots, and the block's
//     off-diagonal entries l_ij;
//   * each lane forms its row's four entries of L by the same updates the column-by-column
//     algorithm applies (a[c+j] -= L_m l_jm for m < j, in column order, then times 1 / L_jj), so
//     every value, the pivots included, is the one that algorithm computes, operation for
//     operation;
//   * the four columns and 1 / L_jj go to LDS and are published together (workgroup-scope
//     release of the counter), and the wave's later columns take the rank-4 update, reading the
//     published columns back from LDS as broadcasts four columns at a time, the next group's
//     reads in flight; the first group is the next block's four columns.
// So the cross-lane round trips and the pivot latencies are paid once per four columns.  The
// waves publish in column order, so the consumers are unchanged.  LDS addresses are taken
// relative to an opaque zero so that each group's reads share one base register.
// Two waves of 32 columns: three (24 / 20 / 20) measured no faster with the column-by-column
// step, the per-column cost being the pivot chain and the LDS round trip.
constexpr int kFactorWaves = 2;
constexpr int kFactorCol[kFactorWaves + 1] = {0, 32, kTile};
// chol_step blocks: the factor waves and one consumer wave, at most one wave per SIMD so that
// each may use the whole register file; the update tasks use all 256 threads
constexpr int kStepThreads = 256;
static_assert((kFactorWaves + 1) * kWave <= kStepThreads, "factor and consumer waves fit the block");

__device__ __forceinline__ int opaque_zero() {
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// A non-positive (or NaN) pivot is flagged (info = 1, as Eigen's LLT reports a NumericalIssue)
// and left in place: its reciprocal square root is then NaN or infinite, which only poisons
// values that every consumer of the solve discards once info is set (the solution counts as
// zero), and the pivot chain carries no compare-and-select.
template <typename R>
__device__ __forceinline__ R checked_pivot(R e, bool& bad) {
  bad |= !(e > R(0));
  return e;
}

// Rank-4 update of local columns s = c + 4 + 4 G .. + 3 by the block's columns C4 .. C4 + 3:
// a[s] -= L_j L[OWN0 + s][C4 + j] for j = 0..3 in order; v[j][i] = L[OWN0 + s0 + i][C4 + j].
template <typename R, int C4, int OWN0, int W, int G>
__device__ __forceinline__ void fb_read(const PanelShared<R>& sh, int z, R (&v)[4][4]) {
  constexpr int s0 = C4 - OWN0 + 4 + 4 * G;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (s0 + i < W) v[j][i] = sh.Lc[C4 + j][z + OWN0 + s0 + i];
}

template <typename R, int C4, int OWN0, int W, int G>
__device__ __forceinline__ void fb_fma(R (&a)[W], const R (&L)[4], const R (&v)[4][4]) {
  constexpr int s0 = C4 - OWN0 + 4 + 4 * G;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (s0 + i < W) {
#pragma unroll
      for (int j = 0; j < 4; ++j) a[s0 + i] -= L[j] * v[j][i];
    }
}

template <typename R, int C4, int OWN0, int W, int... Gs>
__device__ __forceinline__ void fb_bulk(R (&a)[W], const R (&L)[4], const PanelShared<R>& sh, int z,
                                        std::integer_sequence<int, Gs...>) {
  R v[2][4][4];
  if constexpr (sizeof...(Gs) > 0) fb_read<R, C4, OWN0, W, 0>(sh, z, v[0]);
  (([&] {
     if constexpr (Gs + 1 < static_cast<int>(sizeof...(Gs))) fb_read<R, C4, OWN0, W, Gs + 1>(sh, z, v[(Gs + 1) & 1]);
     fb_fma<R, C4, OWN0, W, Gs>(a, L, v[Gs & 1]);
   }()),
   ...);
}

// Columns C4 .. C4 + 3 (local c .. c + 3).  Every expression below repeats, in the same order,
// the operation that the column-by-column step performs on the lane that owns the value.
template <typename R, int C4, int OWN0, int W>
__device__ __forceinline__ void fw_block4(R (&a)[W], PanelShared<R>& sh, int z, bool& bad) {
  static_assert((C4 - OWN0) % 4 == 0 && C4 - OWN0 + 4 <= W, "blocks of four columns inside the wave's range");
  const int r = threadIdx.x % kWave;
  constexpr int c = C4 - OWN0;
  // the diagonal block, lower triangle
  const R d00 = lane_bcast<R>(a[c], C4);
  const R d10 = lane_bcast<R>(a[c], C4 + 1), d11 = lane_bcast<R>(a[c + 1], C4 + 1);
  const R d20 = lane_bcast<R>(a[c], C4 + 2), d21 = lane_bcast<R>(a[c + 1], C4 + 2);
  const R d22 = lane_bcast<R>(a[c + 2], C4 + 2);
  const R d30 = lane_bcast<R>(a[c], C4 + 3), d31 = lane_bcast<R>(a[c + 1], C4 + 3);
  const R d32 = lane_bcast<R>(a[c + 2], C4 + 3), d33 = lane_bcast<R>(a[c + 3], C4 + 3);
  R L[4];
  // column C4
  const R e0 = checked_pivot<R>(d00, bad);
  const R is0 = pivot_rsqrt(e0);
  const R l10 = d10 * is0, l20 = d20 * is0, l30 = d30 * is0;
  L[0] = (r == C4) ? e0 * is0 : (r > C4 ? a[c] * is0 : R(0));
  a[c + 1] -= L[0] * l10;
  a[c + 2] -= L[0] * l20;
  a[c + 3] -= L[0] * l30;
  // column C4 + 1
  R t11 = d11;
  t11 -= l10 * l10;
  const R e1 = checked_pivot<R>(t11, bad);
  const R is1 = pivot_rsqrt(e1);
  R t21 = d21;
  t21 -= l20 * l10;
  R t31 = d31;
  t31 -= l30 * l10;
  const R l21 = t21 * is1, l31 = t31 * is1;
  L[1] = (r == C4 + 1) ? e1 * is1 : (r > C4 + 1 ? a[c + 1] * is1 : R(0));
  a[c + 2] -= L[1] * l21;
  a[c + 3] -= L[1] * l31;
  // column C4 + 2
  R t22 = d22;
  t22 -= l20 * l20;
  t22 -= l21 * l21;
  const R e2 = checked_pivot<R>(t22, bad);
  const R is2 = pivot_rsqrt(e2);
  R t32 = d32;
  t32 -= l30 * l20;
  t32 -= l31 * l21;
  const R l32 = t32 * is2;
  L[2] = (r == C4 + 2) ? e2 * is2 : (r > C4 + 2 ? a[c + 2] * is2 : R(0));
  a[c + 3] -= L[2] * l32;
  // column C4 + 3
  R t33 = d33;
  t33 -= l30 * l30;
  t33 -= l31 * l31;
  t33 -= l32 * l32;
  const R e3 = checked_pivot<R>(t33, bad);
  const R is3 = pivot_rsqrt(e3);
  L[3] = (r == C4 + 3) ? e3 * is3 : (r > C4 + 3 ? a[c + 3] * is3 : R(0));
  // the four columns and their 1 / L_jj, published together (the release store waits for
  // every LDS access in flight)
#pragma unroll
  for (int j = 0; j < 4; ++j) sh.Lc[C4 + j][r] = L[j];
  sh.dinv[C4] = is0;
  sh.dinv[C4 + 1] = is1;
  sh.dinv[C4 + 2] = is2;
  sh.dinv[C4 + 3] = is3;
  __hip_atomic_store(&sh.published, C4 + 4, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  constexpr int rest = W - (c + 4);
  fb_bulk<R, C4, OWN0, W>(a, L, sh, z, std::make_integer_sequence<int, (rest + 3) / 4>{});
}

template <typename R, int OWN0, int W, int... Bs>
__device__ __forceinline__ void fw_all(R (&a)[W], PanelShared<R>& sh, int z, bool& bad,
                                       std::integer_sequence<int, Bs...>) {
  (fw_block4<R, OWN0 + 4 * Bs, OWN0, W>(a, sh, z, bad), ...);
}

// a[s] -= sum_{c < OWN0} L[r][c] L[OWN0 + s][c], four published columns per step; false if a
// wait gave up.
template <typename R, int OWN0, int W>
__device__ __forceinline__ bool fw_apply_earlier(R (&a)[W], const PanelShared<R>& sh, int z) {
  const int r = threadIdx.x % kWave;
  unsigned late = 0;
  for (int c0 = 0; c0 < OWN0; c0 += 4) {
    late |= wait_flag(&sh.published, c0 + 4) ? 0u : 1u;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const R l = sh.Lc[c0 + m][r];
      R v[W];
#pragma unroll
      for (int q = 0; q < W; ++q) v[q] = sh.Lc[c0 + m][z + OWN0 + q];
#pragma unroll
      for (int q = 0; q < W; ++q) a[q] -= l * v[q];
    }
  }
  return late == 0;
}

template <typename R, int OWN0, int W>
__device__ void factor_part(const R* __restrict__ A, int64_t lda, int k, PanelShared<R>& sh, bool report,
                            int* __restrict__ info) {
  static_assert(OWN0 % 4 == 0 && W % 4 == 0 && W > 0, "factor ranges are whole blocks of four columns");
  const int r = threadIdx.x % kWave;
  const int64_t base = (int64_t)k * kTile;
  R a[W];
#pragma unroll
  for (int c = 0; c < W; ++c) a[c] = A[(base + r) + (base + OWN0 + c) * lda];
  const int z = opaque_zero();
  bool bad = false;
  if constexpr (OWN0 > 0) bad = !fw_apply_earlier<R, OWN0, W>(a, sh, z);
  fw_all<R, OWN0, W>(a, sh, z, bad, std::make_integer_sequence<int, W / 4>{});
  if (bad && report && r == 0) info[0] = 1;
}

template <typename R, int... Ws>
__device__ __forceinline__ void factor_waves(const R* __restrict__ A, int64_t lda, int k, PanelShared<R>& sh,
                                             bool report, int* __restrict__ info, std::integer_sequence<int, Ws...>) {
  const int w = threadIdx.x / kWave;
  ((w == Ws ? factor_part<R, kFactorCol[Ws], kFactorCol[Ws + 1] - kFactorCol[Ws]>(A, lda, k, sh, report, info)
            : void()),
   ...);
}

// Consumer step over columns M .. M+3 of L_kk once they are published: the 4x4 triangle
// x[M+m] = (x[M+m] - sum_{q<m} x[M+q] L[M+m][M+q]) / L[M+m][M+m] -- final values, handed to
// `out` at once so their registers are free for the rest of the sweep -- then
// x[j] -= sum_m x[M+m] L[j][M+m] for every j > M+3, in groups of kConsumeGroup j whose LDS
// reads are issued one group ahead of the FMAs.
constexpr int kConsumeGroup = 8;

template <typename R, int M, int G>
__device__ __forceinline__ void cons_read(const PanelShared<R>& sh, R (&v)[4][kConsumeGroup]) {
  constexpr int j0 = M + 4 + G * kConsumeGroup;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int i = 0; i < kConsumeGroup; ++i)
      if (j0 + i < kTile) v[m][i] = sh.Lc[M + m][j0 + i];
}

template <typename R, int M, int G>
__device__ __forceinline__ void cons_fma(R (&x)[kTile], const R (&v)[4][kConsumeGroup]) {
  constexpr int j0 = M + 4 + G * kConsumeGroup;
#pragma unroll
  for (int i = 0; i < kConsumeGroup; ++i)
    if (j0 + i < kTile)
      x[j0 + i] -= x[M] * v[0][i] + x[M + 1] * v[1][i] + x[M + 2] * v[2][i] + x[M + 3] * v[3][i];
}

// the trailing update of a consumer step, group by group, the next group's reads in flight
template <typename R, int M, int... Gs>
__device__ __forceinline__ void cons_update(R (&x)[kTile], const PanelShared<R>& sh,
                                            std::integer_sequence<int, Gs...>) {
  R v[2][4][kConsumeGroup];
  cons_read<R, M, 0>(sh, v[0]);
  (([&] {
     if constexpr (Gs + 1 < static_cast<int>(sizeof...(Gs))) cons_read<R, M, Gs + 1>(sh, v[(Gs + 1) & 1]);
     __builtin_amdgcn_sched_barrier(0);
     cons_fma<R, M, Gs>(x, v[Gs & 1]);
     __builtin_amdgcn_sched_barrier(0);
   }()),
   ...);
}

template <typename R, int M, typename Out>
__device__ __forceinline__ void consume_block4(R (&x)[kTile], PanelShared<R>& sh, unsigned& late, Out& out) {
  late |= wait_flag(&sh.published, M + 4) ? 0u : 1u;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    R v = x[M + m];
#pragma unroll
    for (int q = 0; q < m; ++q) v -= x[M + q] * sh.Lc[M + q][M + m];
    x[M + m] = v * sh.dinv[M + m];
    out(M + m, x[M + m]);
  }
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (M + 4 < kTile)
    cons_update<R, M>(x, sh, std::make_integer_sequence<int, (kTile - 4 - M + kConsumeGroup - 1) / kConsumeGroup>{});
}

// All 64 columns; returns false if a wait gave up.
template <typename R, typename Out, int... Ms>
__device__ __forceinline__ bool consume_all(R (&x)[kTile], PanelShared<R>& sh, Out& out,
                                            std::integer_sequence<int, Ms...>) {
  unsigned late = 0;
  (consume_block4<R, 4 * Ms>(x, sh, late, out), ...);
  return late == 0;
}

// Consumer, off-diagonal block: L_ik = A_ik L_kk^-T, i.e. x L_kk^T = a for each of the 64 rows
// (lane q = row q), right-looking as the columns of L arrive: x_m = acc_m / L[m][m], then
// acc_j -= x_m L[j][m] for j > m.
template <typename R>
__device__ bool solve_rows(R* __restrict__ A, int64_t lda, int64_t bi, int64_t bk, PanelShared<R>& sh) {
  const int q = threadIdx.x - kFactorWaves * kWave;
  R x[kTile];
#pragma unroll
  for (int c = 0; c < kTile; ++c) x[c] = A[(bi + q) + (bk + c) * lda];
  auto out = [&](int c, R v) { A[(bi + q) + (bk + c) * lda] = v; };
  return consume_all<R>(x, sh, out, std::make_integer_sequence<int, kTile / 4>{});
}

// Consumer, diagonal block: X = L_kk^-1 by columns (lane j = column j), right-looking as the
// columns of L arrive; X goes to Xs (row-major) for the caller.
template <typename R>
__device__ bool invert_columns(PanelShared<R>& sh, R (*Xs)[kLd]) {
  const int j = threadIdx.x - kFactorWaves * kWave;
  R x[kTile];
#pragma unroll
  for (int m = 0; m < kTile; ++m) x[m] = m == j ? R(1) : R(0);
  auto out = [&](int m, R v) { Xs[m][j] = v; };
  return consume_all<R>(x, sh, out, std::make_integer_sequence<int, kTile / 4>{});
}

// Panel block for tile row i = k + blk (threads 0..191; the launch's other threads have
// returned).  Waves 0 and 1 factor the diagonal tile and publish L four columns at a time; wave 2
// consumes the columns as they arrive: block 0 inverts L_kk (kept in Ld for the
// back-substitution, and y_k = L_kk^-1 b_k), every other block solves its tile L_ik.  The only
// block barrier is the one after the counter is cleared.
template <typename R>
__device__ void panel_block(R* __restrict__ A, int64_t lda, R* __restrict__ Ld, R* __restrict__ b, int k,
                            int blk, int* __restrict__ info, PanelShared<R>& sh, R (*Xs)[kLd],
                            R* __restrict__ xsol) {
  const int t = threadIdx.x;
  const int i = k + blk;
  const int64_t bk = (int64_t)k * kTile, bi = (int64_t)i * kTile;
  if (t == 0) sh.published = 0;
  __syncthreads();
  if (t < kFactorWaves * kWave) {
    factor_waves<R>(A, lda, k, sh, blk == 0, info, std::make_integer_sequence<int, kFactorWaves>{});
    return;
  }
  if (blk > 0) {
    if (!solve_rows<R>(A, lda, bi, bk, sh) && t == kFactorWaves * kWave) info[0] = 1;
    return;
  }
  // the consumer wave of block 0 (its own LDS writes and reads, in order): b_k staged in LDS
  // now, its load hidden behind the factor (b_k is final: the previous launch updated it, this
  // one does not touch it); then L_kk^-1 by columns, written column-major into Ld, and
  // y_k = L_kk^-1 b_k
  const int u = t - kFactorWaves * kWave;
  sh.bk[u] = b[bk + u];
  const bool ok = invert_columns<R>(sh, Xs);
  if (!ok && t == kFactorWaves * kWave) info[0] = 1;
  R* Lk = Ld + (int64_t)k * kTile * kTile;
#pragma unroll 8
  for (int c = 0; c < kTile; ++c) Lk[c * kTile + u] = Xs[u][c];
  R acc = 0;
#pragma unroll 8
  for (int c = 0; c < kTile; ++c) acc += Xs[u][c] * sh.bk[c];
  b[bk + u] = acc;
  if (xsol) {
    // a single-tile system: the back-substitution x = L_kk^-T y here, in this wave (lane u:
    // x_u = sum_c (L_kk^-1)[c][u] y_c), instead of a chol_back_all launch.  y replaces b_k in
    // LDS: every read of b_k above has returned (its value was used), and the fence orders the
    // stores before the loads of the other lanes
    sh.bk[u] = acc;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    R xv = 0;
#pragma unroll 8
    for (int c = 0; c < kTile; ++c) xv += Xs[c][u] * sh.bk[c];
    xsol[u] = xv;
  }
}

// Trailing-update work of step k (256 threads): tile (k+1+ii, k+1+jj) -= L_{k+1+ii,k} L_{k+1+jj,k}^T.
template <typename R>
__device__ void update_tile(R* __restrict__ A, int64_t lda, int k, int ii, int jj, R (*Li)[kLd], R (*Lj)[kLd]) {
  const int t = threadIdx.x;
  const int64_t bk = (int64_t)k * kTile;
  const int64_t bi = (int64_t)(k + 1 + ii) * kTile, bj = (int64_t)(k + 1 + jj) * kTile;
  tile_to_lds<R>(A + bi + bk * lda, lda, Li);
  tile_to_lds<R>(A + bj + bk * lda, lda, Lj);
  __syncthreads();
  // thread (tc, tr) owns rows tc + 16 p and columns tr + 16 q of the tile: within a wave the
  // Li reads hit 16 consecutive padded rows (distinct banks) and the Lj reads 4 broadcast
  // addresses; the global read-modify-write covers 16 consecutive rows per column
  const int tc = t & 15, tr = t >> 4;
  R acc[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[p][q] = 0;
#pragma unroll 8
  for (int kk = 0; kk < kTile; ++kk) {
    R a[4], bb[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      a[p] = Li[tc + 16 * p][kk];
      bb[p] = Lj[tr + 16 * p][kk];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * bb[q];
  }
  R old[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int p = 0; p < 4; ++p) old[p][q] = A[(bi + tc + 16 * p) + (bj + tr + 16 * q) * lda];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int p = 0; p < 4; ++p) A[(bi + tc + 16 * p) + (bj + tr + 16 * q) * lda] = old[p][q] - acc[p][q];
}

// ... and its rhs part: b_{k+1+ii} -= L_{k+1+ii,k} y_k.
template <typename R>
__device__ void update_rhs(const R* __restrict__ A, int64_t lda, R* __restrict__ b, int k, int ii,
                           R (*part)[kTile]) {
  const int t = threadIdx.x;
  const int64_t bk = (int64_t)k * kTile, bi = (int64_t)(k + 1 + ii) * kTile;
  const int r = t & 63, cg = t >> 6;
  R acc = 0;
#pragma unroll
  for (int c = cg; c < kTile; c += 4) acc += A[(bi + r) + (bk + c) * lda] * b[bk + c];
  part[cg][r] = acc;
  __syncthreads();
  if (t < kTile) b[bi + t] -= part[0][t] + part[1][t] + part[2][t] + part[3][t];
}

// LDS of a chol_step block: a panel block's shared state and L_kk^-1, or an update block's
// two L tiles and rhs partial sums.
template <typename R> union StepShared {
  struct {
    PanelShared<R> sh;
    R Xs[kTile][kLd];
  } panel;
  struct {
    R Li[kTile][kLd];
    R Lj[kTile][kLd];
    R part[4][kTile];
  } update;
};

// Step k of the factorisation (256-thread blocks).  Blocks [0, T - k) are the panel blocks of
// tile column k.  For k >= 1 the further blocks apply the rest of step k - 1's trailing update
// (column k and rhs row k were updated by chol_update_next before this launch): with
// m = T - k, tiles (ii, jj) of step k - 1 with 1 <= jj <= ii < m, then rhs rows 1 .. m - 1.
// Neither part touches what the other reads or writes.
template <typename R>
__global__ void __launch_bounds__(kStepThreads)
chol_step(R* __restrict__ A, int64_t lda, R* __restrict__ Ld, R* __restrict__ b, int k, int T,
          int* __restrict__ info, R* __restrict__ xsol) {
  __shared__ StepShared<R> smem;
  if (k == 0 && blockIdx.x == 0)                 // clear the back-substitution's tile counters
    for (int i = threadIdx.x; i < T; i += blockDim.x) info[kBackFlags + i] = 0;
  const int npanel = T - k;
  const int blk = blockIdx.x;
  if (blk < npanel) {
    if (threadIdx.x >= (kFactorWaves + 1) * kWave) return;
    panel_block<R>(A, lda, Ld, b, k, blk, info, smem.panel.sh, smem.panel.Xs, xsol);
    return;
  }
  const int kp = k - 1, m1 = T - k - 1;           // sub-triangle size of the remaining tiles
  const int task = blk - npanel;
  const int ntiles = m1 * (m1 + 1) / 2;
  if (task >= ntiles) {
    update_rhs<R>(A, lda, b, kp, 1 + (task - ntiles), smem.update.part);
    return;
  }
  // lower-triangular task index -> (ii, jj), jj <= ii, of the sub-triangle
  int ii = static_cast<int>((sqrt(8.0 * task + 1.0) - 1.0) * 0.5);
  while ((ii + 1) * (ii + 2) / 2 <= task) ++ii;
  while (ii * (ii + 1) / 2 > task) --ii;
  const int jj = task - ii * (ii + 1) / 2;
  update_tile<R>(A, lda, kp, 1 + ii, 1 + jj, smem.update.Li, smem.update.Lj);
}

// The part of step k's trailing update that the next panel needs: tiles (ii, 0) of column
// k + 1 for ii < m = T - k - 1, each cut into four 32 x 32 quarters (blocks 4 ii .. 4 ii + 3,
// so four CUs share a tile's latency), and rhs row k + 1 (block 4 m).  A quarter stages the 32
// rows of L_{k+1+ii,k} and of L_{k+1,k} it needs; thread (tc, tr) owns rows tc + 16 p and
// columns tr + 16 q of the quarter.
template <typename R>
__global__ void __launch_bounds__(256)
chol_update_next(R* __restrict__ A, int64_t lda, R* __restrict__ b, int k, int m) {
  __shared__ R Li[kTile / 2][kLd];
  __shared__ R Lj[kTile / 2][kLd];
  __shared__ R part[4][kTile];
  const int task = blockIdx.x;
  if (task >= 4 * m) {
    update_rhs<R>(A, lda, b, k, 0, part);
    return;
  }
  const int ii = task >> 2, rh = task & 1, ch = (task >> 1) & 1;
  const int t = threadIdx.x;
  const int64_t bk = (int64_t)k * kTile;
  const int64_t bi = (int64_t)(k + 1 + ii) * kTile + 32 * rh, bj = (int64_t)(k + 1) * kTile + 32 * ch;
  // stage 32 rows x 64 columns of each operand (rows along the lanes: coalesced)
  {
    R vi[8], vj[8];
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const int idx = t + 256 * it, c = idx >> 5, r = idx & 31;
      vi[it] = A[(bi + r) + (bk + c) * lda];
      vj[it] = A[(bj + r) + (bk + c) * lda];
    }
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const int idx = t + 256 * it, c = idx >> 5, r = idx & 31;
      Li[r][c] = vi[it];
      Lj[r][c] = vj[it];
    }
  }
  __syncthreads();
  const int tc = t & 15, tr = t >> 4;             // rows tc, tc + 16; columns tr, tr + 16
  R acc[2][2] = {{0, 0}, {0, 0}};
#pragma unroll 8
  for (int kk = 0; kk < kTile; ++kk) {
    const R a0 = Li[tc][kk], a1 = Li[tc + 16][kk];
    const R b0 = Lj[tr][kk], b1 = Lj[tr + 16][kk];
    acc[0][0] += a0 * b0;
    acc[0][1] += a0 * b1;
    acc[1][0] += a1 * b0;
    acc[1][1] += a1 * b1;
  }
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int p = 0; p < 2; ++p) A[(bi + tc + 16 * p) + (bj + tr + 16 * q) * lda] -= acc[p][q];
}

// Back-substitution x = L^-T y in one launch, one block per tile row j (block b takes
// j = T - 1 - b, so the blocks without dependencies start first).  Forward progress: block b
// waits only for x_k with k > j, which blocks of lower index publish, so the lowest-index block
// still running never waits for one that has not started, as long as blocks are dispatched in
// index order (as this hardware does; the bounded wait turns a violation into info = 1, not a
// hang).  Up to T = 512 all blocks are resident together as well (68 KB of LDS each in FP64: two
// per CU on 256 CUs), which is why hsl2_spd_solve accepts n <= 32768; the BA's reduced systems
// (6 P unknowns) stay within that for P <= 5461 poses.  Block j keeps y_j in
// LDS and, for k = T-1 down to j+1, stages tile (k, j) of L, waits until x_k is published
// (agent-scope acquire of flags[k]; bounded, a timeout sets info[0] = 1) and applies
// y_j -= L_kj^T x_k; then x_j = (L_jj^-1)^T y_j is written and flags[j] raised with an
// agent-scope release.  The tile loads of later steps overlap the waits for earlier ones.
constexpr int kBackSpin = 1 << 22;

template <typename R>
__global__ void __launch_bounds__(256)
chol_back_all(const R* __restrict__ A, int64_t lda, const R* __restrict__ Ld, const R* __restrict__ y,
              R* __restrict__ x, int T, int* __restrict__ info) {
  __shared__ R Lt[kTile][kLd];
  __shared__ R Xj[kTile][kLd];
  __shared__ R ys[kTile];
  __shared__ R xs[kTile];
  __shared__ R part[4][kTile];
  int* flags = info + kBackFlags;
  const int j = T - 1 - static_cast<int>(blockIdx.x);
  const int t = threadIdx.x, c = t & 63, rg = t >> 6;
  const int64_t bj = (int64_t)j * kTile;
  tile_to_lds<R>(Ld + (int64_t)j * kTile * kTile, kTile, Xj);   // L_jj^-1, Xj[r][c]
  if (t < kTile) ys[t] = y[bj + t];
  bool late = false;
  for (int k = T - 1; k > j; --k) {
    const int64_t bk = (int64_t)k * kTile;
    tile_to_lds<R>(A + bk + bj * lda, lda, Lt);   // Lt[r][c] = L_kj[r][c]
    if (t < kTile) {
      int spins = 0;
      while (__hip_atomic_load(&flags[k], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (++spins > kBackSpin) {
          late = true;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      xs[t] = x[bk + t];
    }
    __syncthreads();
    R acc = 0;
#pragma unroll 8
    for (int r = rg; r < kTile; r += 4) acc += Lt[r][c] * xs[r];
    part[rg][c] = acc;
    __syncthreads();
    if (t < kTile) ys[t] -= part[0][t] + part[1][t] + part[2][t] + part[3][t];
    __syncthreads();
  }
  __syncthreads();                               // Xj and ys complete (also when no k > j)
  {
    R acc = 0;
#pragma unroll 8
    for (int r = rg; r < kTile; r += 4) acc += Xj[r][c] * ys[r];
    part[rg][c] = acc;
  }
  __syncthreads();
  if (t < kTile) {
    x[bj + t] = part[0][t] + part[1][t] + part[2][t] + part[3][t];
    if (late) info[0] = 1;
  }
  __syncthreads();
  if (t == 0) __hip_atomic_store(&flags[j], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

// Factor the padded n_pad x n_pad matrix A (column-major, lda = n_pad, n_pad % 64 == 0) and
// solve A x = b.  A's strictly-lower tiles receive L, Ld (n_pad * 64 elements) the inverses
// of the diagonal factors; b is overwritten by the forward-substitution result; x receives the solution;
// info[0] is set to 1 on a non-positive pivot (it must be cleared by the caller).
template <typename R>
int cholesky_solve(R* A, int64_t n_pad, R* Ld, R* b, R* x, int* info, hipStream_t s) {
  HSL2_REQUIRE(n_pad % kTile == 0, HSL2_ERR_SHAPE, "cholesky: size must be a multiple of 64");
  const int T = static_cast<int>(n_pad / kTile);
  if (T == 1) {
    // one tile: the panel block also back-substitutes (no chol_back_all launch)
    hipLaunchKernelGGL(chol_step<R>, dim3(1), dim3(kStepThreads), 0, s, A, n_pad, Ld, b, 0, T, info, x);
    HSL2_LAUNCH_CHECK();
    return HSL2_OK;
  }
  for (int k = 0; k < T; ++k) {
    // panel k, with the rest of step k - 1's trailing update alongside
    const int m1 = k > 0 ? T - k - 1 : 0;
    const int rest = k > 0 ? m1 * (m1 + 1) / 2 + m1 : 0;
    hipLaunchKernelGGL(chol_step<R>, dim3(T - k + rest), dim3(kStepThreads), 0, s, A, n_pad, Ld, b, k, T, info,
                       static_cast<R*>(nullptr));
    HSL2_LAUNCH_CHECK();
    const int m = T - k - 1;
    if (m > 0) {
      hipLaunchKernelGGL(chol_update_next<R>, dim3(4 * m + 1), dim3(256), 0, s, A, n_pad, b, k, m);
      HSL2_LAUNCH_CHECK();
    }
  }
  hipLaunchKernelGGL(chol_back_all<R>, dim3(T), dim3(256), 0, s, A, n_pad, Ld, b, x, T, info);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

template int cholesky_solve<double>(double*, int64_t, double*, double*, double*, int*, hipStream_t);
template int cholesky_solve<float>(float*, int64_t, float*, float*, float*, int*, hipStream_t);

namespace {

template <typename R>
__global__ void spd_pad(const R* __restrict__ A, const R* __restrict__ b, int64_t n, int64_t n_pad,
                        R* __restrict__ Ap, R* __restrict__ bp, int* __restrict__ info) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx == 0) info[0] = 0;
  if (idx < n_pad) bp[idx] = idx < n ? b[idx] : R(0);
  if (idx >= n_pad * n_pad) return;
  const int64_t c = idx / n_pad, r = idx - c * n_pad;
  Ap[idx] = (r < n && c < n) ? A[r + c * n] : (r == c ? R(1) : R(0));
}

template <typename R>
__global__ void spd_unpad(const R* __restrict__ xp, const int* __restrict__ info, int64_t n,
                          R* __restrict__ x) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) x[idx] = info[0] == 0 ? xp[idx] : R(0);
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace
}  // namespace hsl2

using namespace hsl2;

extern "C" size_t hsl2_spd_solve_workspace_bytes(int dtype, int64_t n) {
  const size_t es = dtype == HSL2_F64 ? 8 : 4;
  const size_t np = static_cast<size_t>((n + 63) / 64) * 64;
  return align256(np * np * es) + 3 * align256(np * es) + align256(np * 64 * es) +
         align256((4 + np / 64) * sizeof(int));
}

extern "C" int hsl2_spd_solve(int dtype, const void* A, const void* b, void* x, int64_t n,
                              void* workspace, size_t workspace_bytes, int* ok, void* stream) {
  HSL2_REQUIRE(A && b && x && workspace, HSL2_ERR_ARGUMENT, "spd_solve: null pointer");
  HSL2_REQUIRE(n > 0 && n <= (1 << 15), HSL2_ERR_SHAPE, "spd_solve: n must be in [1, 32768]");
  HSL2_REQUIRE(dtype == HSL2_F64 || dtype == HSL2_F32, HSL2_ERR_ARGUMENT, "spd_solve: dtype must be F32 or F64");
  HSL2_REQUIRE(workspace_bytes >= hsl2_spd_solve_workspace_bytes(dtype, n), HSL2_ERR_WORKSPACE,
               "spd_solve: workspace too small");
  HSL2_REQUIRE(reinterpret_cast<uintptr_t>(workspace) % 256 == 0, HSL2_ERR_ARGUMENT,
               "spd_solve: workspace must be 256-byte aligned");
  hipStream_t s = as_stream(stream);
  const int64_t np = ((n + 63) / 64) * 64;
  const size_t es = dtype == HSL2_F64 ? 8 : 4;
  char* w = static_cast<char*>(workspace);
  char* Ap = w;
  char* bp = Ap + align256(np * np * es);
  char* xp = bp + align256(np * es);
  char* Ld = xp + align256(np * es);
  int* info = reinterpret_cast<int*>(Ld + align256(np * 64 * es));
  const unsigned g1 = static_cast<unsigned>((np * np + 255) / 256), g2 = static_cast<unsigned>((n + 255) / 256);
  int st;
  if (dtype == HSL2_F64) {
    hipLaunchKernelGGL(spd_pad<double>, dim3(g1), dim3(256), 0, s, static_cast<const double*>(A),
                       static_cast<const double*>(b), n, np, reinterpret_cast<double*>(Ap),
                       reinterpret_cast<double*>(bp), info);
    HSL2_LAUNCH_CHECK();
    st = cholesky_solve<double>(reinterpret_cast<double*>(Ap), np, reinterpret_cast<double*>(Ld),
                                reinterpret_cast<double*>(bp), reinterpret_cast<double*>(xp), info, s);
    if (st) return st;
    hipLaunchKernelGGL(spd_unpad<double>, dim3(g2), dim3(256), 0, s, reinterpret_cast<double*>(xp), info, n,
                       static_cast<double*>(x));
  } else {
    hipLaunchKernelGGL(spd_pad<float>, dim3(g1), dim3(256), 0, s, static_cast<const float*>(A),
                       static_cast<const float*>(b), n, np, reinterpret_cast<float*>(Ap),
                       reinterpret_cast<float*>(bp), info);
    HSL2_LAUNCH_CHECK();
    st = cholesky_solve<float>(reinterpret_cast<float*>(Ap), np, reinterpret_cast<float*>(Ld),
                               reinterpret_cast<float*>(bp), reinterpret_cast<float*>(xp), info, s);
    if (st) return st;
    hipLaunchKernelGGL(spd_unpad<float>, dim3(g2), dim3(256), 0, s, reinterpret_cast<float*>(xp), info, n,
                       static_cast<float*>(x));
  }
  HSL2_LAUNCH_CHECK();
  if (ok) {
    int h = 0;
    HSL2_HIP(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, s));
    HSL2_HIP(hipStreamSynchronize(s));
    *ok = h == 0 ? 1 : 0;
  }
  return HSL2_OK;
}
