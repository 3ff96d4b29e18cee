// Sim3 pose-graph bundle adjustment: droid_backends.pgba (src/droid.cpp:171-195 ->
// src/droid_kernels.cu:1940-2044, pixel terms 608-858).
//
// One call is one Gauss-Newton step over the poses of [t0, t1) (7 parameters each: translation,
// rotation, log-scale) and the disparities of the frames of [t0, t1) ++ ii, with the pose-graph
// blocks Hsp / vsp of the caller added to the pose system.  Device pipeline:
//   pgba_linearize   one block per edge: Sim3 pixel terms (FP32 as the reference), the edge's
//                    14x14 pose block and 14-vector reduced in FP64, Eij / Eii rows, C, bz;
//   pgba_frame       per (256 pixels, depth frame): C, w, Q = 1 / (C + eta) and the Ei rows of
//                    the frames of [t0, t1) (the sums of their edges' Eii);
//   pgba_clear       the padded FP64 matrix (identity on the padding) and the rhs;
//   pgba_assemble    one block per nonzero lower 7x7 block (a >= b) of the reduced system:
//                    edge and pose-graph blocks minus the Schur term sum_x,y E_x Q E_y^T over
//                    the row pairs of a shared depth frame, damping on the diagonal, written
//                    with its mirror; diagonal blocks also form the rhs;
//   cholesky_solve   the dense FP64 LLT shared with ba (chol.hip);
//   pgba_backsub     per (256 pixels, frame): dz = Q (w - sum E^T dx) without pose slot 0,
//                    disps += dz; pgba_dx writes dx (0 when the factorisation failed).
// The lower triangle of each block is what enters the factorisation, as with the reference's
// SimplicialLLT (src/droid_kernels.cu:1659-1680), which reads the lower triangle only.
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "hsl2.h"
#include "se3.hpp"

namespace hsl2 {
template <typename R>
int cholesky_solve(R* A, int64_t n_pad, R* Ld, R* b, R* x, int* info, hipStream_t s);
}

struct hsl2_pgba_plan {
  int64_t E = 0, Np = 0, N = 0, ht = 0, wd = 0, HW = 0;
  int t0 = 0, t1 = 0, P = 0;
  int64_t F = 0, R = 0, NB = 0, n = 0, n_pad = 0;
  std::vector<int64_t> kx;
  std::vector<int32_t> blob;
  int64_t o_edge_ii = 0, o_edge_jj = 0, o_frame_kx = 0, o_frame_slot = 0, o_frame_eptr = 0, o_frame_edges = 0;
  int64_t o_blk_ab = 0, o_blk_tptr = 0, o_blk_trip = 0, o_blk_hptr = 0, o_blk_h = 0, o_blk_pptr = 0, o_blk_p = 0;
  int64_t o_slot_vptr = 0, o_slot_v = 0, o_slot_qptr = 0, o_slot_q = 0, o_slot_rptr = 0, o_slot_rows = 0;
  int64_t o_frame_rptr = 0, o_frame_rows = 0;
  size_t w_blob = 0, w_hv = 0, w_eall = 0, w_eii = 0, w_cii = 0, w_bz = 0, w_q = 0, w_w = 0;
  size_t w_mat = 0, w_rhs = 0, w_x = 0, w_ld = 0, w_info = 0, total = 0;
};

namespace hsl2 {
namespace {

constexpr int kD = 7;                          // Sim3 tangent: translation, rotation, scale
constexpr int kPacked = 14 * 15 / 2;           // lower triangle of the 14x14 edge block
constexpr int kHv = 120;                       // doubles per edge: packed block, vi, vj, pad
constexpr float kPgMinDepth = 0.25f;           // src/droid_kernels.cu:26

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Y = adjSim3(t, q, s, X) exactly as src/droid_kernels.cu:121-146 writes it
__device__ __forceinline__ void adj_sim3(const float t[3], const float q[4], float s, const float X[7], float Y[7]) {
  const float qinv[4] = {-q[0], -q[1], -q[2], q[3]};
  float a[3], b[3], u[3], v[3];
  quat_rotate(qinv, X, a);
  quat_rotate(qinv, X + 3, b);
  u[0] = t[2] * X[1] - t[1] * X[2];
  u[1] = t[0] * X[2] - t[2] * X[0];
  u[2] = t[1] * X[0] - t[0] * X[1];
  quat_rotate(qinv, u, v);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Y[c] = s * a[c];
    Y[3 + c] = s * (b[c] + v[c]) + X[6] * t[c];
  }
  Y[6] = X[6];
}

// One block per edge: the pixel loop of projective_transform_sim3_kernel
// (src/droid_kernels.cu:697-825) with the block sums in FP64.
__global__ void __launch_bounds__(256)
pgba_linearize(const float* __restrict__ targets, const float* __restrict__ weights, const float* __restrict__ poses,
               const float* __restrict__ disps, const float* __restrict__ intr, const int* __restrict__ edge_ii,
               const int* __restrict__ edge_jj, int ht, int wd, int P, double* __restrict__ Hv,
               float* __restrict__ Eall, float* __restrict__ Eii, float* __restrict__ Cii, float* __restrict__ bz) {
  __shared__ double lds[16 * 256];
  const int e = blockIdx.x;
  const int HW = ht * wd;
  const int ix = edge_ii[e], jx = edge_jj[e];
  const float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
  const float* pi = poses + (int64_t)ix * 8;
  const float* pj = poses + (int64_t)jx * 8;
  // relSim3 (src/droid_kernels.cu:161-177)
  float tij[3], qij[4], rti[3];
  relative_pose(pi, pj, tij, qij);
  const float sij = pj[7] / pi[7];
  quat_rotate(qij, pi, rti);
#pragma unroll
  for (int c = 0; c < 3; ++c) tij[c] = pj[c] - sij * rti[c];
  const bool self = ix == jx;

  float h[kPacked + 14];
#pragma unroll
  for (int l = 0; l < kPacked + 14; ++l) h[l] = 0.f;

  const float* tg = targets + (int64_t)e * 2 * HW;
  const float* wt = weights + (int64_t)e * 2 * HW;
  const float* dp = disps + (int64_t)ix * HW;
  float* eij = Eall + (int64_t)(P + e) * kD * HW;
  float* eii = Eii + (int64_t)e * kD * HW;
  for (int k = threadIdx.x; k < HW; k += blockDim.x) {
    const int row = k / wd, col = k - row * wd;
    const float d = dp[k];
    const float X[3] = {(static_cast<float>(col) - cx) / fx, (static_cast<float>(row) - cy) / fy, 1.f};
    float Y[3];
    quat_rotate(qij, X, Y);                                     // actSim3 (src/droid_kernels.cu:95-102)
    const float x = sij * Y[0] + d * tij[0];
    const float y = sij * Y[1] + d * tij[1];
    const float zr = sij * Y[2] + d * tij[2];
    const bool valid = zr >= kPgMinDepth;
    const float z = valid ? zr : 0.f;
    const float dd = valid ? 1.f / zr : 0.f;
    const float d2 = dd * dd;
    const float Jp[6] = {fx * dd, 0.f, -fx * x * d2, 0.f, fy * dd, -fy * y * d2};
    const float res[2] = {valid ? tg[k] - (fx * x * dd + cx) : 0.f, valid ? tg[HW + k] - (fy * y * dd + cy) : 0.f};
    float c_acc = 0.f, b_acc = 0.f, ei[kD] = {}, ej[kD] = {};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const float* J = Jp + 3 * r;
      const float w = valid ? 0.001f * wt[r * HW + k] : 0.f;
      float Jx[14];
      float* Ji = Jx;
      float* Jj = Jx + 7;
      Jj[0] = J[0] * d;
      Jj[1] = J[1] * d;
      Jj[2] = J[2] * d;
      Jj[3] = -J[1] * z + J[2] * y;
      Jj[4] = J[0] * z - J[2] * x;
      Jj[5] = -J[0] * y + J[1] * x;
      Jj[6] = J[0] * x + J[1] * y + J[2] * z;
      const float Jz = J[0] * tij[0] + J[1] * tij[1] + J[2] * tij[2];
      c_acc += w * Jz * Jz;
      b_acc += w * res[r] * Jz;
      const float wp = self ? 0.f : w;
      adj_sim3(tij, qij, sij, Jj, Ji);
#pragma unroll
      for (int n = 0; n < kD; ++n) Ji[n] = -Ji[n];
      int l = 0;
#pragma unroll
      for (int n = 0; n < 14; ++n) {
        const float wj = wp * Jx[n];
#pragma unroll
        for (int m = 0; m <= n; ++m) h[l++] += wj * Jx[m];
      }
#pragma unroll
      for (int n = 0; n < 14; ++n) h[kPacked + n] += wp * res[r] * Jx[n];
#pragma unroll
      for (int n = 0; n < kD; ++n) {
        ei[n] += wp * Jz * Ji[n];
        ej[n] += wp * Jz * Jj[n];
      }
    }
    Cii[(int64_t)e * HW + k] = c_acc;
    bz[(int64_t)e * HW + k] = b_acc;
#pragma unroll
    for (int n = 0; n < kD; ++n) {
      eij[(int64_t)n * HW + k] = ej[n];
      eii[(int64_t)n * HW + k] = ei[n];
    }
  }
  block_reduce_store<kPacked + 14, 256>(h, lds, Hv + (int64_t)e * kHv);
}

// Per (256 pixels, depth frame f): C and w summed over the edges leaving f, Q = 1 / (C + eta),
// and, for a frame of [t0, t1), its Ei row (src/droid_kernels.cu:2001-2006).
__global__ void __launch_bounds__(256)
pgba_frame(const float* __restrict__ Cii, const float* __restrict__ bz, const float* __restrict__ Eii,
           const float* __restrict__ eta, int eta_rows, const int* __restrict__ frame_slot,
           const int* __restrict__ frame_eptr, const int* __restrict__ frame_edges, int HW,
           float* __restrict__ Q, float* __restrict__ W, float* __restrict__ Eall) {
  const int f = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= HW) return;
  const int e0 = frame_eptr[f], e1 = frame_eptr[f + 1];
  const int slot = frame_slot[f];
  float c = 0.f, w = 0.f, ei[kD] = {};
  for (int q = e0; q < e1; ++q) {
    const int e = frame_edges[q];
    c += Cii[(int64_t)e * HW + k];
    w += bz[(int64_t)e * HW + k];
    if (slot >= 0) {
#pragma unroll
      for (int n = 0; n < kD; ++n) ei[n] += Eii[((int64_t)e * kD + n) * HW + k];
    }
  }
  const float et = eta[(int64_t)(eta_rows == 1 ? 0 : f) * HW + k];
  Q[(int64_t)f * HW + k] = 1.f / (c + et);
  W[(int64_t)f * HW + k] = w;
  if (slot >= 0) {
#pragma unroll
    for (int n = 0; n < kD; ++n) Eall[((int64_t)slot * kD + n) * HW + k] = ei[n];
  }
}

// Column-major padded matrix: zero, identity on the padding; rhs zero; factorisation status clear.
__global__ void pgba_clear(double* __restrict__ A, double* __restrict__ b, int64_t n, int64_t n_pad,
                           int* __restrict__ info) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx == 0) info[0] = 0;
  if (idx < n_pad) b[idx] = 0.0;
  if (idx >= n_pad * n_pad) return;
  const int64_t c = idx / n_pad, r = idx - c * n_pad;
  A[idx] = (r == c && r >= n) ? 1.0 : 0.0;
}

// Entry (r, c) of block type k (0: ii-ii, 1: ii-jj, 2: jj-ii, 3: jj-jj) of an edge's packed
// 14x14 lower triangle.
__device__ __forceinline__ double edge_block(const double* __restrict__ hv, int k, int r, int c) {
  const int R = r + ((k & 2) ? 7 : 0);
  const int C = c + ((k & 1) ? 7 : 0);
  const int n = R > C ? R : C, m = R > C ? C : R;
  return hv[n * (n + 1) / 2 + m];
}

// One block per nonzero lower 7x7 block (a >= b) of A - S (src/droid_kernels.cu:1982-2009 with
// schur_block 1689-1779 and SparseBlock 1598-1680).
__global__ void __launch_bounds__(256)
pgba_assemble(const float* __restrict__ Eall, const float* __restrict__ Q, const float* __restrict__ W,
              const double* __restrict__ Hv, const float* __restrict__ Hsp, const float* __restrict__ vsp,
              int64_t Np, const int* __restrict__ blk_ab, const int* __restrict__ blk_tptr,
              const int* __restrict__ blk_trip, const int* __restrict__ blk_hptr, const int* __restrict__ blk_h,
              const int* __restrict__ blk_pptr, const int* __restrict__ blk_p, const int* __restrict__ slot_vptr,
              const int* __restrict__ slot_v, const int* __restrict__ slot_qptr, const int* __restrict__ slot_q,
              const int* __restrict__ slot_rptr, const int* __restrict__ slot_rows, int HW, int64_t n_pad,
              float lm, float ep, double* __restrict__ A, double* __restrict__ rhs) {
  __shared__ double lds[16 * 256];
  __shared__ double red[kD * kD + kD];
  const int blk = blockIdx.x;
  const int a = blk_ab[2 * blk], b = blk_ab[2 * blk + 1];
  const bool diag = a == b;
  // per-thread sums: FP32 over a triple's pixels, FP64 across triples (A - S cancels heavily
  // in the scale / depth direction, so the Schur term is kept at FP64 accuracy)
  double acc[kD * kD + kD];
#pragma unroll
  for (int l = 0; l < kD * kD + kD; ++l) acc[l] = 0.0;
  for (int t = blk_tptr[blk]; t < blk_tptr[blk + 1]; ++t) {
    const int x = blk_trip[3 * t], y = blk_trip[3 * t + 1], f = blk_trip[3 * t + 2];
    const float* ex = Eall + (int64_t)x * kD * HW;
    const float* ey = Eall + (int64_t)y * kD * HW;
    const float* qf = Q + (int64_t)f * HW;
    float part[kD * kD];
#pragma unroll
    for (int l = 0; l < kD * kD; ++l) part[l] = 0.f;
    for (int k = threadIdx.x; k < HW; k += blockDim.x) {
      const float q = qf[k];
      float u[kD], v[kD];
#pragma unroll
      for (int n = 0; n < kD; ++n) {
        u[n] = ex[(int64_t)n * HW + k] * q;
        v[n] = ey[(int64_t)n * HW + k];
      }
#pragma unroll
      for (int n = 0; n < kD; ++n)
#pragma unroll
        for (int m = 0; m < kD; ++m) part[n * kD + m] += u[n] * v[m];
    }
#pragma unroll
    for (int l = 0; l < kD * kD; ++l) acc[l] += static_cast<double>(part[l]);
  }
  if (diag) {
    // bS[a] = sum over the rows of slot a of E_x (Q w) of the row's frame
    for (int t = slot_rptr[a]; t < slot_rptr[a + 1]; ++t) {
      const int x = slot_rows[2 * t], f = slot_rows[2 * t + 1];
      const float* ex = Eall + (int64_t)x * kD * HW;
      float part[kD] = {};
      for (int k = threadIdx.x; k < HW; k += blockDim.x) {
        const float qw = Q[(int64_t)f * HW + k] * W[(int64_t)f * HW + k];
#pragma unroll
        for (int n = 0; n < kD; ++n) part[n] += ex[(int64_t)n * HW + k] * qw;
      }
#pragma unroll
      for (int n = 0; n < kD; ++n) acc[kD * kD + n] += static_cast<double>(part[n]);
    }
  }
  block_reduce_store<kD * kD + kD, 256>(acc, lds, red);
  __syncthreads();
  const int t = threadIdx.x;
  if (t < kD * kD) {
    const int r = t / kD, c = t % kD;
    if (!diag || r >= c) {
      double v = -red[t];
      for (int q = blk_hptr[blk]; q < blk_hptr[blk + 1]; ++q) {
        const int code = blk_h[q];
        v += edge_block(Hv + (int64_t)(code >> 2) * kHv, code & 3, r, c);
      }
      for (int q = blk_pptr[blk]; q < blk_pptr[blk + 1]; ++q) {
        const int code = blk_p[q];
        v += static_cast<double>(Hsp[(((int64_t)(code & 3) * Np + (code >> 2)) * kD + r) * kD + c]);
      }
      if (diag && r == c) v += static_cast<double>(ep) + static_cast<double>(lm) * v;
      const int64_t gr = (int64_t)a * kD + r, gc = (int64_t)b * kD + c;
      A[gr + gc * n_pad] = v;
      A[gc + gr * n_pad] = v;
    }
  } else if (diag && t < kD * kD + kD) {
    const int r = t - kD * kD;
    double v = -red[t];
    for (int q = slot_vptr[a]; q < slot_vptr[a + 1]; ++q) {
      const int code = slot_v[q];
      v += Hv[(int64_t)(code >> 1) * kHv + kPacked + (code & 1) * kD + r];
    }
    for (int q = slot_qptr[a]; q < slot_qptr[a + 1]; ++q) {
      const int code = slot_q[q];
      v += static_cast<double>(vsp[((int64_t)(code & 1) * Np + (code >> 1)) * kD + r]);
    }
    rhs[(int64_t)a * kD + r] = v;
  }
}

// dz = Q (w - sum over the frame's rows with pose slot >= 1 of E_x^T dx) (src/droid_kernels.cu:2011-2041
// with EvT6x1_kernel 1562-1582), dx = 0 when the factorisation failed; disps[kx[f]] += dz.
__global__ void __launch_bounds__(256)
pgba_backsub(const float* __restrict__ Eall, const float* __restrict__ Q, const float* __restrict__ W,
             const double* __restrict__ x, const int* __restrict__ info, const int* __restrict__ frame_rptr,
             const int* __restrict__ frame_rows, const int* __restrict__ frame_kx, int HW,
             float* __restrict__ dz, float* __restrict__ disps) {
  const int f = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= HW) return;
  const bool ok = info[0] == 0;
  float acc = 0.f;
  if (ok) {
    for (int q = frame_rptr[f]; q < frame_rptr[f + 1]; ++q) {
      const int row = frame_rows[2 * q], slot = frame_rows[2 * q + 1];
      const float* ex = Eall + (int64_t)row * kD * HW;
#pragma unroll
      for (int n = 0; n < kD; ++n) acc += ex[(int64_t)n * HW + k] * static_cast<float>(x[(int64_t)slot * kD + n]);
    }
  }
  const float v = Q[(int64_t)f * HW + k] * (W[(int64_t)f * HW + k] - acc);
  dz[(int64_t)f * HW + k] = v;
  disps[(int64_t)frame_kx[f] * HW + k] += v;
}

__global__ void pgba_dx(const double* __restrict__ x, const int* __restrict__ info, int64_t n,
                        float* __restrict__ dx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dx[i] = info[0] == 0 ? static_cast<float>(x[i]) : 0.f;
}

}  // namespace
}  // namespace hsl2

using namespace hsl2;

extern "C" int hsl2_pgba_plan_create(const int64_t* ii, const int64_t* jj, int64_t E, const int64_t* iip,
                                     const int64_t* jjp, int64_t Np, int64_t num_poses, int64_t ht, int64_t wd,
                                     int t0, int t1, hsl2_pgba_plan** out) {
  HSL2_REQUIRE(out, HSL2_ERR_ARGUMENT, "pgba: null plan pointer");
  *out = nullptr;
  HSL2_REQUIRE(E >= 0 && Np >= 0 && (E == 0 || (ii && jj)) && (Np == 0 || (iip && jjp)), HSL2_ERR_ARGUMENT,
               "pgba: null index array");
  HSL2_REQUIRE(ht > 0 && wd > 0 && ht * wd < (int64_t(1) << 31), HSL2_ERR_SHAPE, "pgba: bad frame size");
  HSL2_REQUIRE(0 <= t0 && t0 < t1 && t1 <= num_poses, HSL2_ERR_ARGUMENT, "pgba: need 0 <= t0 < t1 <= number of poses");
  const int P = t1 - t0;
  HSL2_REQUIRE(int64_t(P) * 7 < (1 << 16), HSL2_ERR_SHAPE, "pgba: at most 9362 poses in [t0, t1)");
  HSL2_REQUIRE(E < (int64_t(1) << 28) && Np < (int64_t(1) << 28), HSL2_ERR_SHAPE, "pgba: too many edges");
  for (int64_t e = 0; e < E; ++e)
    HSL2_REQUIRE(ii[e] >= 0 && jj[e] >= 0 && ii[e] < t1 && jj[e] < t1, HSL2_ERR_INDEX,
                 "pgba: edge index outside [0, t1)");
  for (int64_t e = 0; e < Np; ++e)
    HSL2_REQUIRE(iip[e] >= 0 && jjp[e] >= 0 && iip[e] < t1 && jjp[e] < t1, HSL2_ERR_INDEX,
                 "pgba: pose-graph index outside [0, t1)");
  std::unique_ptr<hsl2_pgba_plan> p(new (std::nothrow) hsl2_pgba_plan());
  HSL2_REQUIRE(p, HSL2_ERR_ARGUMENT, "pgba: out of host memory");
  p->E = E;
  p->Np = Np;
  p->N = num_poses;
  p->ht = ht;
  p->wd = wd;
  p->HW = ht * wd;
  p->t0 = t0;
  p->t1 = t1;
  p->P = P;
  const int64_t R = P + E;
  p->R = R;
  // expanded rows: [t0, t1) then the edges (src/droid_kernels.cu:1960-1968)
  std::vector<int64_t> ii_exp(R), jj_exp(R);
  for (int s = 0; s < P; ++s) ii_exp[s] = jj_exp[s] = t0 + s;
  for (int64_t e = 0; e < E; ++e) {
    ii_exp[P + e] = ii[e];
    jj_exp[P + e] = jj[e];
  }
  p->kx = ii_exp;
  std::sort(p->kx.begin(), p->kx.end());
  p->kx.erase(std::unique(p->kx.begin(), p->kx.end()), p->kx.end());
  const int64_t F = static_cast<int64_t>(p->kx.size());
  p->F = F;
  std::vector<int32_t> kk(R), slot(R);
  for (int64_t n = 0; n < R; ++n) {
    kk[n] = static_cast<int32_t>(std::lower_bound(p->kx.begin(), p->kx.end(), ii_exp[n]) - p->kx.begin());
    slot[n] = static_cast<int32_t>(jj_exp[n] - t0);      // >= P cannot occur (jj < t1)
  }
  std::vector<int32_t> edge_ii(E), edge_jj(E), frame_kx(F), frame_slot(F), frame_eptr(F + 1, 0), frame_edges(E);
  for (int64_t e = 0; e < E; ++e) {
    edge_ii[e] = static_cast<int32_t>(ii[e]);
    edge_jj[e] = static_cast<int32_t>(jj[e]);
    ++frame_eptr[kk[P + e] + 1];
  }
  for (int64_t f = 0; f < F; ++f) {
    frame_kx[f] = static_cast<int32_t>(p->kx[f]);
    frame_slot[f] = (p->kx[f] >= t0 && p->kx[f] < t1) ? static_cast<int32_t>(p->kx[f] - t0) : -1;
    frame_eptr[f + 1] += frame_eptr[f];
  }
  {
    std::vector<int32_t> fill(frame_eptr.begin(), frame_eptr.end() - 1);
    for (int64_t e = 0; e < E; ++e) frame_edges[fill[kk[P + e]]++] = static_cast<int32_t>(e);
  }
  // rows of each frame with a pose slot in [0, P)
  std::vector<std::vector<int32_t>> frows(F);
  for (int64_t n = 0; n < R; ++n)
    if (slot[n] >= 0) frows[kk[n]].push_back(static_cast<int32_t>(n));
  // lower blocks: every diagonal block first, then the rest as they appear
  std::unordered_map<int64_t, int32_t> bid;
  std::vector<int32_t> blk_ab;
  auto block_of = [&](int a, int b) {
    const int64_t key = int64_t(a) * P + b;
    auto it = bid.find(key);
    if (it != bid.end()) return it->second;
    const int32_t id = static_cast<int32_t>(blk_ab.size() / 2);
    bid.emplace(key, id);
    blk_ab.push_back(a);
    blk_ab.push_back(b);
    return id;
  };
  for (int a = 0; a < P; ++a) block_of(a, a);
  std::vector<std::vector<int32_t>> trip, hl, pl;
  auto grow = [&]() {
    const size_t nb = blk_ab.size() / 2;
    if (trip.size() < nb) {
      trip.resize(nb);
      hl.resize(nb);
      pl.resize(nb);
    }
  };
  grow();
  for (int64_t f = 0; f < F; ++f) {
    for (int32_t x : frows[f])
      for (int32_t y : frows[f]) {
        if (slot[x] < slot[y]) continue;
        const int32_t id = block_of(slot[x], slot[y]);
        grow();
        trip[id].push_back(x);
        trip[id].push_back(y);
        trip[id].push_back(static_cast<int32_t>(f));
      }
  }
  auto add_pose_blocks = [&](const int64_t* a_idx, const int64_t* b_idx, int64_t cnt, bool edges) {
    for (int64_t e = 0; e < cnt; ++e) {
      const int64_t i = a_idx[e] - t0, j = b_idx[e] - t0;
      const int64_t rows[4] = {i, i, j, j}, cols[4] = {i, j, i, j};
      for (int k = 0; k < 4; ++k) {
        if (rows[k] < 0 || cols[k] < 0 || rows[k] < cols[k]) continue;
        const int32_t id = block_of(static_cast<int>(rows[k]), static_cast<int>(cols[k]));
        grow();
        (edges ? hl : pl)[id].push_back(static_cast<int32_t>(e * 4 + k));
      }
    }
  };
  add_pose_blocks(ii, jj, E, true);
  add_pose_blocks(iip, jjp, Np, false);
  const int64_t NB = static_cast<int64_t>(blk_ab.size() / 2);
  p->NB = NB;
  // the blocks with most Schur work are dispatched first
  std::vector<int32_t> order(NB);
  for (int64_t b = 0; b < NB; ++b) order[b] = static_cast<int32_t>(b);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t u, int32_t v) { return trip[u].size() > trip[v].size(); });
  std::vector<int32_t> o_ab(2 * NB), tptr(NB + 1, 0), hptr(NB + 1, 0), pptr(NB + 1, 0), tv, hv, pv;
  for (int64_t q = 0; q < NB; ++q) {
    const int32_t b = order[q];
    o_ab[2 * q] = blk_ab[2 * b];
    o_ab[2 * q + 1] = blk_ab[2 * b + 1];
    tv.insert(tv.end(), trip[b].begin(), trip[b].end());
    hv.insert(hv.end(), hl[b].begin(), hl[b].end());
    pv.insert(pv.end(), pl[b].begin(), pl[b].end());
    tptr[q + 1] = static_cast<int32_t>(tv.size() / 3);
    hptr[q + 1] = static_cast<int32_t>(hv.size());
    pptr[q + 1] = static_cast<int32_t>(pv.size());
  }
  // rhs lists per slot: edge gradients (e * 2 + side), pose-graph gradients, rows for bS
  std::vector<std::vector<int32_t>> sv(P), sq(P), sr(P);
  for (int64_t e = 0; e < E; ++e) {
    if (ii[e] - t0 >= 0) sv[ii[e] - t0].push_back(static_cast<int32_t>(e * 2));
    if (jj[e] - t0 >= 0) sv[jj[e] - t0].push_back(static_cast<int32_t>(e * 2 + 1));
  }
  for (int64_t e = 0; e < Np; ++e) {
    if (iip[e] - t0 >= 0) sq[iip[e] - t0].push_back(static_cast<int32_t>(e * 2));
    if (jjp[e] - t0 >= 0) sq[jjp[e] - t0].push_back(static_cast<int32_t>(e * 2 + 1));
  }
  for (int64_t n = 0; n < R; ++n)
    if (slot[n] >= 0) {
      sr[slot[n]].push_back(static_cast<int32_t>(n));
      sr[slot[n]].push_back(kk[n]);
    }
  auto csr = [](const std::vector<std::vector<int32_t>>& lists, int per, std::vector<int32_t>& ptr,
                std::vector<int32_t>& val) {
    ptr.assign(lists.size() + 1, 0);
    for (size_t i = 0; i < lists.size(); ++i) {
      val.insert(val.end(), lists[i].begin(), lists[i].end());
      ptr[i + 1] = static_cast<int32_t>(val.size() / per);
    }
  };
  std::vector<int32_t> svp, svv, sqp, sqv, srp, srv, frp, frv;
  csr(sv, 1, svp, svv);
  csr(sq, 1, sqp, sqv);
  csr(sr, 2, srp, srv);
  // back-substitution rows per frame: slot >= 1 (EvT6x1_kernel skips slot 0)
  std::vector<std::vector<int32_t>> fr(F);
  for (int64_t n = 0; n < R; ++n)
    if (slot[n] >= 1) {
      fr[kk[n]].push_back(static_cast<int32_t>(n));
      fr[kk[n]].push_back(slot[n]);
    }
  csr(fr, 2, frp, frv);

  auto& B = p->blob;
  auto put = [&B](const std::vector<int32_t>& v) {
    const int64_t off = static_cast<int64_t>(B.size());
    B.insert(B.end(), v.begin(), v.end());
    B.push_back(0);                                    // keeps every table non-empty
    return off;
  };
  p->o_edge_ii = put(edge_ii);
  p->o_edge_jj = put(edge_jj);
  p->o_frame_kx = put(frame_kx);
  p->o_frame_slot = put(frame_slot);
  p->o_frame_eptr = put(frame_eptr);
  p->o_frame_edges = put(frame_edges);
  p->o_blk_ab = put(o_ab);
  p->o_blk_tptr = put(tptr);
  p->o_blk_trip = put(tv);
  p->o_blk_hptr = put(hptr);
  p->o_blk_h = put(hv);
  p->o_blk_pptr = put(pptr);
  p->o_blk_p = put(pv);
  p->o_slot_vptr = put(svp);
  p->o_slot_v = put(svv);
  p->o_slot_qptr = put(sqp);
  p->o_slot_q = put(sqv);
  p->o_slot_rptr = put(srp);
  p->o_slot_rows = put(srv);
  p->o_frame_rptr = put(frp);
  p->o_frame_rows = put(frv);

  const int64_t HW = p->HW;
  p->n = int64_t(P) * kD;
  p->n_pad = (p->n + 63) / 64 * 64;
  size_t off = 0;
  auto take = [&off](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  p->w_blob = take(B.size() * sizeof(int32_t));
  p->w_hv = take(size_t(std::max<int64_t>(E, 1)) * kHv * sizeof(double));
  p->w_eall = take(size_t(R) * kD * HW * sizeof(float));
  p->w_eii = take(size_t(std::max<int64_t>(E, 1)) * kD * HW * sizeof(float));
  p->w_cii = take(size_t(std::max<int64_t>(E, 1)) * HW * sizeof(float));
  p->w_bz = take(size_t(std::max<int64_t>(E, 1)) * HW * sizeof(float));
  p->w_q = take(size_t(F) * HW * sizeof(float));
  p->w_w = take(size_t(F) * HW * sizeof(float));
  p->w_mat = take(size_t(p->n_pad) * p->n_pad * sizeof(double));
  p->w_rhs = take(size_t(p->n_pad) * sizeof(double));
  p->w_x = take(size_t(p->n_pad) * sizeof(double));
  p->w_ld = take(size_t(p->n_pad) * 64 * sizeof(double));
  p->w_info = take((4 + p->n_pad / 64) * sizeof(int));
  p->total = off;
  *out = p.release();
  return HSL2_OK;
}

extern "C" void hsl2_pgba_plan_destroy(hsl2_pgba_plan* plan) { delete plan; }

extern "C" int64_t hsl2_pgba_plan_num_frames(const hsl2_pgba_plan* plan) { return plan ? plan->F : 0; }

extern "C" int hsl2_pgba_plan_frames(const hsl2_pgba_plan* plan, int64_t* kx) {
  HSL2_REQUIRE(plan && kx, HSL2_ERR_ARGUMENT, "pgba: null pointer");
  std::memcpy(kx, plan->kx.data(), plan->kx.size() * sizeof(int64_t));
  return HSL2_OK;
}

extern "C" size_t hsl2_pgba_workspace_bytes(const hsl2_pgba_plan* plan) { return plan ? plan->total : 0; }

extern "C" int hsl2_pgba_run(const hsl2_pgba_plan* p, const float* poses, float* disps, const float* intrinsics,
                             const float* targets, const float* weights, const float* eta, int64_t eta_rows,
                             const float* Hsp, const float* vsp, float lm, float ep, float* dx, float* dz,
                             void* workspace, size_t workspace_bytes, int* solve_ok, void* stream) {
  HSL2_REQUIRE(p, HSL2_ERR_ARGUMENT, "pgba: null plan");
  HSL2_REQUIRE(poses && disps && intrinsics && eta && dx && dz && workspace, HSL2_ERR_ARGUMENT,
               "pgba: null pointer");
  HSL2_REQUIRE(p->E == 0 || (targets && weights), HSL2_ERR_ARGUMENT, "pgba: null targets / weights");
  HSL2_REQUIRE(p->Np == 0 || (Hsp && vsp), HSL2_ERR_ARGUMENT, "pgba: null Hsp / vsp");
  HSL2_REQUIRE(eta_rows == 1 || eta_rows == p->F, HSL2_ERR_SHAPE,
               "pgba: eta must have one row per unique frame of [t0, t1) ++ ii, or one row");
  HSL2_REQUIRE(workspace_bytes >= p->total, HSL2_ERR_WORKSPACE, "pgba: workspace too small");
  HSL2_REQUIRE(reinterpret_cast<uintptr_t>(workspace) % 256 == 0, HSL2_ERR_ARGUMENT,
               "pgba: workspace must be 256-byte aligned");
  hipStream_t s = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  const int* blob = reinterpret_cast<const int*>(ws + p->w_blob);
  double* Hv = reinterpret_cast<double*>(ws + p->w_hv);
  float* Eall = reinterpret_cast<float*>(ws + p->w_eall);
  float* Eii = reinterpret_cast<float*>(ws + p->w_eii);
  float* Cii = reinterpret_cast<float*>(ws + p->w_cii);
  float* bz = reinterpret_cast<float*>(ws + p->w_bz);
  float* Q = reinterpret_cast<float*>(ws + p->w_q);
  float* W = reinterpret_cast<float*>(ws + p->w_w);
  double* A = reinterpret_cast<double*>(ws + p->w_mat);
  double* rhs = reinterpret_cast<double*>(ws + p->w_rhs);
  double* x = reinterpret_cast<double*>(ws + p->w_x);
  double* Ld = reinterpret_cast<double*>(ws + p->w_ld);
  int* info = reinterpret_cast<int*>(ws + p->w_info);
  const int HW = static_cast<int>(p->HW);
  const unsigned gx = static_cast<unsigned>((HW + 255) / 256);

  HSL2_HIP(hipMemcpyAsync(ws + p->w_blob, p->blob.data(), p->blob.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, s));
  if (p->E > 0) {
    hipLaunchKernelGGL(pgba_linearize, dim3(static_cast<unsigned>(p->E)), dim3(256), 0, s, targets, weights, poses,
                       disps, intrinsics, blob + p->o_edge_ii, blob + p->o_edge_jj, static_cast<int>(p->ht),
                       static_cast<int>(p->wd), p->P, Hv, Eall, Eii, Cii, bz);
    HSL2_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(pgba_frame, dim3(gx, static_cast<unsigned>(p->F)), dim3(256), 0, s, Cii, bz, Eii, eta,
                     static_cast<int>(eta_rows), blob + p->o_frame_slot, blob + p->o_frame_eptr,
                     blob + p->o_frame_edges, HW, Q, W, Eall);
  HSL2_LAUNCH_CHECK();
  const int64_t nn = p->n_pad * p->n_pad;
  hipLaunchKernelGGL(pgba_clear, dim3(static_cast<unsigned>((nn + 255) / 256)), dim3(256), 0, s, A, rhs, p->n,
                     p->n_pad, info);
  HSL2_LAUNCH_CHECK();
  hipLaunchKernelGGL(pgba_assemble, dim3(static_cast<unsigned>(p->NB)), dim3(256), 0, s, Eall, Q, W, Hv, Hsp, vsp,
                     p->Np, blob + p->o_blk_ab, blob + p->o_blk_tptr, blob + p->o_blk_trip, blob + p->o_blk_hptr,
                     blob + p->o_blk_h, blob + p->o_blk_pptr, blob + p->o_blk_p, blob + p->o_slot_vptr,
                     blob + p->o_slot_v, blob + p->o_slot_qptr, blob + p->o_slot_q, blob + p->o_slot_rptr,
                     blob + p->o_slot_rows, HW, p->n_pad, lm, ep, A, rhs);
  HSL2_LAUNCH_CHECK();
  const int st = cholesky_solve<double>(A, p->n_pad, Ld, rhs, x, info, s);
  if (st) return st;
  hipLaunchKernelGGL(pgba_dx, dim3(static_cast<unsigned>((p->n + 255) / 256)), dim3(256), 0, s, x, info, p->n, dx);
  HSL2_LAUNCH_CHECK();
  hipLaunchKernelGGL(pgba_backsub, dim3(gx, static_cast<unsigned>(p->F)), dim3(256), 0, s, Eall, Q, W, x, info,
                     blob + p->o_frame_rptr, blob + p->o_frame_rows, blob + p->o_frame_kx, HW, dz, disps);
  HSL2_LAUNCH_CHECK();
  // the plan's host tables must outlive the upload, and solve_ok needs the status
  int h = 0;
  HSL2_HIP(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, s));
  HSL2_HIP(hipStreamSynchronize(s));
  if (solve_ok) *solve_ok = h == 0 ? 1 : 0;
  return HSL2_OK;
}
