// Kernels of the dense-BA iteration shared by the ba_cuda driver (ba.hip) and the
// torch-semantics Gauss-Newton step with its adjoint (gn_step.hip).  Both semantics run the
// same kernels; the differences are switches:
//   ba_cuda (src/droid_kernels.cu)        torch BA (hislam2/geom/ba.py, chol.py)
//   one intrinsics vector                 per-frame intrinsics (iproj with frame ii, proj with jj)
//   valid: X1.z >= 0.25                   valid: X1.z > 0.2 (projective_ops.py:6, 61)
//   Q = 1/(C + eta)                       Q = 1/(C + eta + 1e-7) (ba.py:92)
//   damping on the reduced diagonal       damping on H before the Schur complement (chol.py:57)
//   back-substitution skips slot 0        full back-substitution (chol.py:70)
//   Exp drops V for |phi| <= 1e-4         lietorch Exp with the full left Jacobian V
//   no clamp                              disps > 10 -> 0, clamp(min = 0.001) (ba.py:104-105)
#pragma once
#include <type_traits>

#include "ba_plan.hpp"
#include "common.hpp"
#include "se3.hpp"

namespace hsl2 {

template <typename R>
int cholesky_solve(R* A, int64_t n_pad, R* Ld, R* b, R* x, int* info, hipStream_t s);

enum { kSemCuda = 0, kSemTorch = 1 };

namespace {

constexpr float kMinDepth = 0.25f;       // src/droid_kernels.cu:26
constexpr float kMinDepthTorch = 0.2f;   // hislam2/geom/projective_ops.py:6
constexpr int kLinThreads = 256;

__device__ __forceinline__ void quat_to_rot(const float q[4], float R[9]) {
  const float x = q[0], y = q[1], z = q[2], w = q[3];
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - z * w);     R[2] = 2 * (x * z + y * w);
  R[3] = 2 * (x * y + z * w);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - x * w);
  R[6] = 2 * (x * z - y * w);     R[7] = 2 * (y * z + x * w);     R[8] = 1 - 2 * (x * x + y * y);
}

struct EdgeGeom {
  float R[9];
  float t[3];
  float fxi, fyi, cxi, cyi;     // intrinsics of frame i (inverse projection)
  float fx, fy, cx, cy;         // intrinsics of frame j (projection)
};

// intr_stride = 0: one intrinsics vector for every frame (ba_cuda); 4: one row per frame.
__device__ __forceinline__ void load_edge_geom(const float* __restrict__ poses,
                                               const float* __restrict__ intr, int intr_stride,
                                               int ix, int jx, EdgeGeom& g) {
  float pi[7], pj[7], q[4];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    pi[k] = poses[ix * 7 + k];
    pj[k] = poses[jx * 7 + k];
  }
  relative_pose(pi, pj, g.t, q);
  quat_to_rot(q, g.R);
  const float* ki = intr + (int64_t)ix * intr_stride;
  const float* kj = intr + (int64_t)jx * intr_stride;
  g.fxi = ki[0];
  g.fyi = ki[1];
  g.cxi = ki[2];
  g.cyi = ki[3];
  g.fx = kj[0];
  g.fy = kj[1];
  g.cx = kj[2];
  g.cy = kj[3];
}

// One residual row (u or v) of one pixel: weight, residual, Jacobians.
struct RowJac {
  float w, r, Jz;
  float Ji[6], Jj[6];
};

// Per-pixel geometry of projective_transform_kernel (src/droid_kernels.cu:451-561), or of the
// torch projective_transform (hislam2/geom/projective_ops.py:44-74, pinhole.py:6-53).
// Returns false for an invalid pixel (ba_cuda: transformed depth < 0.25; torch: not > 0.2):
// then every weight is 0.  The torch projection's Z < 0.1 -> 1 only ever applies to pixels
// that are invalid, so it needs no code here.
template <int kSem>
__device__ __forceinline__ bool pixel_rows(const EdgeGeom& g, float u, float v, float d,
                                           float tu, float tv, float wtu, float wtv,
                                           RowJac& ru, RowJac& rv) {
  // explicit multiply-adds and no implicit contraction: the rows are formed alike wherever they
  // are inlined
#pragma clang fp contract(off)
  const float X0 = (u - g.cxi) / g.fxi;
  const float Y0 = (v - g.cyi) / g.fyi;
  const float x = fmaf(d, g.t[0], fmaf(g.R[1], Y0, fmaf(g.R[0], X0, g.R[2])));
  const float y = fmaf(d, g.t[1], fmaf(g.R[4], Y0, fmaf(g.R[3], X0, g.R[5])));
  const float zr = fmaf(d, g.t[2], fmaf(g.R[7], Y0, fmaf(g.R[6], X0, g.R[8])));
  if (kSem == kSemCuda) {
    if (!(zr >= kMinDepth)) return false;
  } else {
    if (!(zr > kMinDepthTorch)) return false;
  }
  const float z = zr;
  const float dinv = 1.f / zr;
  const float d2 = dinv * dinv;
  ru.w = 0.001f * wtu;
  rv.w = 0.001f * wtv;
  ru.r = tu - fmaf(g.fx, x * dinv, g.cx);
  rv.r = tv - fmaf(g.fy, y * dinv, g.cy);
  const float P0 = g.fx * dinv, P2 = -g.fx * x * d2;   // u row: [P0, 0, P2]
  const float Q1 = g.fy * dinv, Q2 = -g.fy * y * d2;   // v row: [0, Q1, Q2]
  ru.Jj[0] = P0 * d;  ru.Jj[1] = 0.f;    ru.Jj[2] = P2 * d;
  ru.Jj[3] = P2 * y;  ru.Jj[4] = fmaf(P0, z, -(P2 * x));  ru.Jj[5] = -P0 * y;
  rv.Jj[0] = 0.f;     rv.Jj[1] = Q1 * d; rv.Jj[2] = Q2 * d;
  rv.Jj[3] = fmaf(-Q1, z, Q2 * y);  rv.Jj[4] = -Q2 * x;  rv.Jj[5] = Q1 * x;
  ru.Jz = fmaf(P0, g.t[0], P2 * g.t[2]);
  rv.Jz = fmaf(Q1, g.t[1], Q2 * g.t[2]);
  // Ji = -Adj(G_ij)^T Jj: tau part -R^T Jt, phi part -R^T (Jr - t x Jt) (src/droid_kernels.cu:104-119)
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    RowJac& J = k == 0 ? ru : rv;
    const float a0 = J.Jj[0], a1 = J.Jj[1], a2 = J.Jj[2];
    const float c0 = J.Jj[3] - fmaf(g.t[1], a2, -(g.t[2] * a1));
    const float c1 = J.Jj[4] - fmaf(g.t[2], a0, -(g.t[0] * a2));
    const float c2 = J.Jj[5] - fmaf(g.t[0], a1, -(g.t[1] * a0));
    J.Ji[0] = -fmaf(g.R[6], a2, fmaf(g.R[3], a1, g.R[0] * a0));
    J.Ji[1] = -fmaf(g.R[7], a2, fmaf(g.R[4], a1, g.R[1] * a0));
    J.Ji[2] = -fmaf(g.R[8], a2, fmaf(g.R[5], a1, g.R[2] * a0));
    J.Ji[3] = -fmaf(g.R[6], c2, fmaf(g.R[3], c1, g.R[0] * c0));
    J.Ji[4] = -fmaf(g.R[7], c2, fmaf(g.R[4], c1, g.R[1] * c0));
    J.Ji[5] = -fmaf(g.R[8], c2, fmaf(g.R[5], c1, g.R[2] * c0));
  }
  return true;
}

__device__ __forceinline__ int tri(int n, int m) { return n * (n + 1) / 2 + m; }

// A valid pixel's terms of one edge in its depth frame's sums (FP32): the depth diagonal
// c = w Jz^2, the depth gradient b = w r Jz and the pose-i coupling ei = w Jz Ji (zero for an
// edge from a frame to itself).  The one definition serves ba_linearize's per-edge planes and
// frame_sums_block, so both form the same values.
__device__ __forceinline__ void depth_terms(const RowJac& ru, const RowJac& rv, bool self, float& c, float& b,
                                            float (&ei)[6]) {
#pragma clang fp contract(off)
  const float wzu = ru.w * ru.Jz, wzv = rv.w * rv.Jz;
  c = fmaf(wzv, rv.Jz, wzu * ru.Jz);
  b = fmaf(wzv, rv.r, wzu * ru.r);
  if (!self) {
#pragma unroll
    for (int n = 0; n < 6; ++n) ei[n] = fmaf(wzv, rv.Ji[n], wzu * ru.Ji[n]);
  }
}

// The pose-j coupling ej = w Jz Jj of a valid pixel of an edge between two frames, formed as
// depth_terms forms ei.
__device__ __forceinline__ void pose_j_terms(const RowJac& ru, const RowJac& rv, float (&ej)[6]) {
#pragma clang fp contract(off)
  const float wzu = ru.w * ru.Jz, wzv = rv.w * rv.Jz;
#pragma unroll
  for (int n = 0; n < 6; ++n) ej[n] = fmaf(wzv, rv.Jj[n], wzu * ru.Jj[n]);
}

// ---- 1. per-edge linearization ----------------------------------------------
// With Cii == nullptr the per-edge planes Cii, bz and Eii are not written (frame_sums_block
// forms the frame sums from the pixels instead); E_ij goes to Eij either way.
// Block of edge e and pixel slice z (of `slices`): with several slices per edge the block
// writes its 90 partial sums to Hpart [slices][E][90], which lin_reduce adds up in slice
// order.
template <bool kSchur, int kSem, typename Acc>
__device__ __forceinline__ void linearize_block(
    const float* __restrict__ poses, const float* __restrict__ disps, const float* __restrict__ intr,
    int intr_stride, const float* __restrict__ targets, const float* __restrict__ weights,
    const int* __restrict__ edge_ii, const int* __restrict__ edge_jj, int HW, int wd, int E, Acc* __restrict__ Hs,
    Acc* __restrict__ vs, float* __restrict__ Eij, float* __restrict__ Eii, float* __restrict__ Cii,
    float* __restrict__ bz, int slices, Acc* __restrict__ Hpart, int e, int z) {
  __shared__ Acc lds[16 * kLinThreads];
  __shared__ Acc red[90];
  const int span = (HW + slices - 1) / slices;
  const int k_begin = z * span, k_end = min(HW, k_begin + span);
  const int ix = edge_ii[e], jx = edge_jj[e];
  const bool self = ix == jx;
  EdgeGeom g;
  load_edge_geom(poses, intr, intr_stride, ix, jx, g);

  // Per-pixel products and each thread's partial sums (HW / 256 pixels, 12 at 48x64) are
  // formed in FP32; the block reduction of the partial sums runs in Acc (FP64 in the split
  // mode, FP32 in the all-FP32 mode).
  float acc[90];
#pragma unroll
  for (int k = 0; k < 90; ++k) acc[k] = 0.f;

  const float* dsp = disps + (int64_t)ix * HW;
  const float* tg = targets + (int64_t)e * 2 * HW;
  const float* wt = weights + (int64_t)e * 2 * HW;
  // the next pixel's five loads are issued before this pixel's arithmetic
  float nd = 0.f, ntu = 0.f, ntv = 0.f, nwu = 0.f, nwv = 0.f;
  if (k_begin + (int)threadIdx.x < k_end) {
    const int k = k_begin + threadIdx.x;
    nd = dsp[k], ntu = tg[k], ntv = tg[HW + k], nwu = wt[k], nwv = wt[HW + k];
  }
  for (int k = k_begin + threadIdx.x; k < k_end; k += kLinThreads) {
    const float d = nd, tu = ntu, tv = ntv, wu = nwu, wv = nwv;
    const int kn = k + kLinThreads;
    if (kn < k_end) nd = dsp[kn], ntu = tg[kn], ntv = tg[HW + kn], nwu = wt[kn], nwv = wt[HW + kn];
    __builtin_amdgcn_sched_barrier(0);
    const int row = k / wd;
    const int col = k - row * wd;
    RowJac ru, rv;
    const bool ok = pixel_rows<kSem>(g, static_cast<float>(col), static_cast<float>(row), d, tu, tv, wu,
                               wv, ru, rv);
    if (kSchur) {
      float ej[6] = {0, 0, 0, 0, 0, 0};
      if (ok && !self) pose_j_terms(ru, rv, ej);
#pragma unroll
      for (int n = 0; n < 6; ++n) Eij[((int64_t)e * 6 + n) * HW + k] = ej[n];
      if (Cii) {                                  // the per-edge planes of ba_frame_reduce
        float c = 0.f, b = 0.f, ei[6] = {0, 0, 0, 0, 0, 0};
        if (ok) depth_terms(ru, rv, self, c, b, ei);
        Cii[(int64_t)e * HW + k] = c;
        bz[(int64_t)e * HW + k] = b;
#pragma unroll
        for (int n = 0; n < 6; ++n) Eii[((int64_t)e * 6 + n) * HW + k] = ei[n];
      }
    }
    if (ok && !self) {
      float Ju[12], Jv[12];
#pragma unroll
      for (int n = 0; n < 6; ++n) {
        Ju[n] = ru.Ji[n];
        Ju[6 + n] = ru.Jj[n];
        Jv[n] = rv.Ji[n];
        Jv[6 + n] = rv.Jj[n];
      }
      // Ju[7] (the u row's Jj[1]) and Jv[6] (the v row's Jj[0]) are zero by construction
      // (pixel_rows): the products with them only add zeros, so they are left out at compile
      // time (26 of the 180 multiply-adds; every sum is unchanged for finite values)
      constexpr int kZu = 7, kZv = 6;
#pragma unroll
      for (int n = 0; n < 12; ++n) {
        const float wun = ru.w * Ju[n], wvn = rv.w * Jv[n];
#pragma unroll
        for (int m = 0; m <= n; ++m) {
          float v = acc[tri(n, m)];
          if (n != kZu && m != kZu) v = fmaf(wun, Ju[m], v);
          if (n != kZv && m != kZv) v = fmaf(wvn, Jv[m], v);
          acc[tri(n, m)] = v;
        }
      }
      const float wru = ru.w * ru.r, wrv = rv.w * rv.r;
#pragma unroll
      for (int n = 0; n < 12; ++n) {
        float v = acc[78 + n];
        if (n != kZu) v = fmaf(wru, Ju[n], v);
        if (n != kZv) v = fmaf(wrv, Jv[n], v);
        acc[78 + n] = v;
      }
    }
  }

  block_reduce_store<90, kLinThreads, Acc>(acc, lds, red);
  __syncthreads();
  const int t = threadIdx.x;
  if (slices > 1) {
    if (t < 90) Hpart[((int64_t)z * E + e) * 90 + t] = red[t];
    return;
  }
  if (t < 144) {
    const int q = t / 36, i = (t % 36) / 6, j = t % 6;
    int l;
    if (q == 0) l = tri(max(i, j), min(i, j));
    else if (q == 1) l = tri(6 + j, i);
    else if (q == 2) l = tri(6 + i, j);
    else l = tri(6 + max(i, j), 6 + min(i, j));
    Hs[((int64_t)q * E + e) * 36 + i * 6 + j] = red[l];
  } else if (t < 156) {
    const int k = t - 144;
    vs[((int64_t)(k / 6) * E + e) * 6 + (k % 6)] = red[78 + k];
  }
}

template <bool kSchur, int kSem, typename Acc = double>
__global__ void __launch_bounds__(kLinThreads)
ba_linearize(const float* __restrict__ poses, const float* __restrict__ disps,
             const float* __restrict__ intr, int intr_stride, const float* __restrict__ targets,
             const float* __restrict__ weights, const int* __restrict__ edge_ii,
             const int* __restrict__ edge_jj, int HW, int wd, int E,
             Acc* __restrict__ Hs, Acc* __restrict__ vs, float* __restrict__ Eij,
             float* __restrict__ Eii, float* __restrict__ Cii, float* __restrict__ bz,
             int slices = 1, Acc* __restrict__ Hpart = nullptr) {
  linearize_block<kSchur, kSem, Acc>(poses, disps, intr, intr_stride, targets, weights, edge_ii, edge_jj, HW, wd, E,
                                     Hs, vs, Eij, Eii, Cii, bz, slices, Hpart, static_cast<int>(blockIdx.x),
                                     static_cast<int>(blockIdx.y));
}

// Hs [4, E, 6, 6] and vs [2, E, 6] from the slices' partial sums Hpart [slices][E][90] (the
// packed lower triangle of the edge's 12 x 12 block, then its 12-vector), summed in slice order.
template <typename Acc>
__device__ __forceinline__ void lin_reduce_elem(const Acc* __restrict__ Hpart, int slices, int E, Acc* __restrict__ Hs,
                                                Acc* __restrict__ vs, int64_t idx) {
  if (idx >= (int64_t)E * 156) return;
  const int e = static_cast<int>(idx / 156), t = static_cast<int>(idx % 156);
  int l;
  if (t < 144) {
    const int q = t / 36, i = (t % 36) / 6, j = t % 6;
    if (q == 0) l = tri(max(i, j), min(i, j));
    else if (q == 1) l = tri(6 + j, i);
    else if (q == 2) l = tri(6 + i, j);
    else l = tri(6 + max(i, j), 6 + min(i, j));
  } else {
    l = 78 + (t - 144);
  }
  Acc v = Hpart[(int64_t)e * 90 + l];
  for (int z = 1; z < slices; ++z) v += Hpart[((int64_t)z * E + e) * 90 + l];
  if (t < 144) {
    const int q = t / 36;
    Hs[((int64_t)q * E + e) * 36 + (t % 36)] = v;
  } else {
    const int k = t - 144;
    vs[((int64_t)(k / 6) * E + e) * 6 + (k % 6)] = v;
  }
}

template <typename Acc>
__global__ void __launch_bounds__(256)
lin_reduce(const Acc* __restrict__ Hpart, int slices, int E, Acc* __restrict__ Hs, Acc* __restrict__ vs) {
  lin_reduce_elem(Hpart, slices, E, Hs, vs, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- 2. per-frame sums: C, w, Q and the pose-i Schur block Ei ----------------
// Sums in Acc (FP64 in the split mode, FP32 in the all-FP32 mode).  With disps_sens (the
// sensor-disparity form of upstream DROID-SLAM's ba): a pixel whose sensor disparity is positive
// takes the prior alpha (d + dz - d_sens)^2 / 2 in place of eta, C + alpha and w - alpha (d - d_sens);
// the other pixels keep exactly the arithmetic of the call without it.
template <typename Acc>
__global__ void __launch_bounds__(256)
ba_frame_reduce(const float* __restrict__ Cii, const float* __restrict__ bz,
                const float* __restrict__ Eii, const float* __restrict__ eta, int eta_rows,
                const int* __restrict__ frame_eptr, const int* __restrict__ frame_edges,
                const int* __restrict__ frame_pose, const int* __restrict__ frame_order, int HW, float c_eps,
                float* __restrict__ Q, float* __restrict__ w, float* __restrict__ Eall,
                const float* __restrict__ disps, const float* __restrict__ disps_sens,
                const int* __restrict__ frame_kx, float alpha) {
  const int f = frame_order[blockIdx.y];          // the frames with most edges first
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= HW) return;
  const int e0 = frame_eptr[f], e1 = frame_eptr[f + 1];
  Acc c = 0, b = 0;
  int q = e0;
  for (; q + 4 <= e1; q += 4) {                  // four edges' loads in flight together
    float cv[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = frame_edges[q + u];
      cv[u] = Cii[(int64_t)e * HW + k];
      bv[u] = bz[(int64_t)e * HW + k];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c += cv[u];
      b += bv[u];
    }
  }
  for (; q < e1; ++q) {
    const int e = frame_edges[q];
    c += Cii[(int64_t)e * HW + k];
    b += bz[(int64_t)e * HW + k];
  }
  float et = eta[(int64_t)(eta_rows == 1 ? 0 : f) * HW + k];
  if (disps_sens) {
    const int64_t px = (int64_t)frame_kx[f] * HW + k;
    const float ds = disps_sens[px];
    if (ds > 0.f) {
      et = alpha;
      b -= static_cast<Acc>(alpha) * (static_cast<Acc>(disps[px]) - static_cast<Acc>(ds));
    }
  }
  Q[(int64_t)f * HW + k] = static_cast<float>(Acc(1) / (c + et + c_eps));
  w[(int64_t)f * HW + k] = static_cast<float>(b);
  const int p = frame_pose[f];
  if (p >= 0) {
    Acc ei[6] = {0, 0, 0, 0, 0, 0};
    int q2 = e0;
    for (; q2 + 2 <= e1; q2 += 2) {
      float v[2][6];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = frame_edges[q2 + u];
#pragma unroll
        for (int n = 0; n < 6; ++n) v[u][n] = Eii[((int64_t)e * 6 + n) * HW + k];
      }
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int n = 0; n < 6; ++n) ei[n] += v[u][n];
    }
    for (; q2 < e1; ++q2) {
      const int e = frame_edges[q2];
#pragma unroll
      for (int n = 0; n < 6; ++n) ei[n] += Eii[((int64_t)e * 6 + n) * HW + k];
    }
#pragma unroll
    for (int n = 0; n < 6; ++n) Eall[((int64_t)p * 6 + n) * HW + k] = static_cast<float>(ei[n]);
  }
}

template <typename Acc = double>
inline void launch_frame_reduce(const float* Cii, const float* bz, const float* Eii, const float* eta,
                                int eta_rows, const int* frame_eptr, const int* frame_edges,
                                const int* frame_pose, const int* frame_order, int HW, int F, float c_eps,
                                float* Q, float* w, float* Eall, hipStream_t s, const float* disps = nullptr,
                                const float* disps_sens = nullptr, const int* frame_kx = nullptr,
                                float alpha = 0.f) {
  hipLaunchKernelGGL(ba_frame_reduce<Acc>, dim3((HW + 255) / 256, F), dim3(256), 0, s, Cii, bz, Eii, eta, eta_rows,
                     frame_eptr, frame_edges, frame_pose, frame_order, HW, c_eps, Q, w, Eall, disps, disps_sens,
                     frame_kx, alpha);
}

// ---- 2'. per-frame sums straight from the pixels --------------------------------
// The results of ba_frame_reduce (Q, w and the pose-i block Ei of each depth frame) without the
// per-edge planes: thread k takes pixel k of frame f and, edge by edge in the frame's edge order,
// recomputes the pixel's rows (pixel_rows, as ba_linearize does), forms the edge's depth_terms
// in FP32 and adds them in Acc in the order in which ba_frame_reduce adds the stored planes
// (an invalid pixel's or a self edge's zeros leave a sum unchanged), so the sums are the same.
// Per pixel and edge it reads the targets and weights, 16 B, where ba_linearize wrote and
// ba_frame_reduce read back Cii, bz and Eii, 2 x 32 B.  The edges' geometry is staged in LDS
// kFsEdges at a time; the targets and weights of the next four edges are in flight while four
// are summed.  The sensor-disparity prior is applied as in ba_frame_reduce.
constexpr int kFsEdges = 64;

// Block (fb, bx): pixels 256 bx .. of the fb-th frame of frame_order.
template <int kSem, typename Acc>
__device__ __forceinline__ void frame_sums_block(
    const float* __restrict__ poses, const float* __restrict__ disps, const float* __restrict__ intr,
    int intr_stride, const float* __restrict__ targets, const float* __restrict__ weights,
    const int* __restrict__ edge_ii, const int* __restrict__ edge_jj, const float* __restrict__ eta, int eta_rows,
    const int* __restrict__ frame_eptr, const int* __restrict__ frame_edges, const int* __restrict__ frame_pose,
    const int* __restrict__ frame_order, const int* __restrict__ frame_kx, int HW, int wd, float c_eps,
    float* __restrict__ Q, float* __restrict__ w, float* __restrict__ Eall, const float* __restrict__ disps_sens,
    float alpha, int fb, int bx) {
  __shared__ EdgeGeom geo[kFsEdges];
  __shared__ int eid[kFsEdges];
  __shared__ int eself[kFsEdges];
  const int f = frame_order[fb];                  // the frames with most edges first
  const int k = bx * blockDim.x + threadIdx.x;
  const bool in = k < HW;
  const int e0 = frame_eptr[f], e1 = frame_eptr[f + 1];
  const int p = frame_pose[f];
  const int64_t kx = frame_kx[f];
  const int row = k / wd;
  const float pu = static_cast<float>(k - row * wd), pv = static_cast<float>(row);
  const float d = in ? disps[kx * HW + k] : 0.f;
  Acc c = 0, b = 0, ei[6] = {0, 0, 0, 0, 0, 0};
  for (int base = e0; base < e1; base += kFsEdges) {
    const int cnt = min(kFsEdges, e1 - base);
    __syncthreads();                              // the previous chunk's geometry is consumed
    if (static_cast<int>(threadIdx.x) < cnt) {
      const int e = frame_edges[base + threadIdx.x];
      const int ix = edge_ii[e], jx = edge_jj[e];
      eid[threadIdx.x] = e;
      eself[threadIdx.x] = ix == jx;
      load_edge_geom(poses, intr, intr_stride, ix, jx, geo[threadIdx.x]);
    }
    __syncthreads();
    if (!in) continue;
    // targets and weights (tu, tv, wu, wv) of edges q0 .. q0 + 3 of the chunk
    auto fetch = [&](int q0, float (&o)[4][4]) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (q0 + j < cnt) {
          const int64_t e = eid[q0 + j];
          const float* tg = targets + e * 2 * HW;
          const float* wt = weights + e * 2 * HW;
          o[j][0] = tg[k];
          o[j][1] = tg[HW + k];
          o[j][2] = wt[k];
          o[j][3] = wt[HW + k];
        }
      }
    };
    float cur[4][4], nxt[4][4];
    fetch(0, cur);
    for (int q0 = 0; q0 < cnt; q0 += 4) {
      if (q0 + 4 < cnt) fetch(q0 + 4, nxt);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (q0 + j < cnt) {
          RowJac ru, rv;
          if (pixel_rows<kSem>(geo[q0 + j], pu, pv, d, cur[j][0], cur[j][1], cur[j][2], cur[j][3], ru, rv)) {
            float cc, bb, ee[6] = {0, 0, 0, 0, 0, 0};
            depth_terms(ru, rv, eself[q0 + j] != 0, cc, bb, ee);
            c += cc;
            b += bb;
            if (p >= 0) {
#pragma unroll
              for (int n = 0; n < 6; ++n) ei[n] += ee[n];
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int n = 0; n < 4; ++n) cur[j][n] = nxt[j][n];
    }
  }
  if (!in) return;
  float et = eta[(int64_t)(eta_rows == 1 ? 0 : f) * HW + k];
  if (disps_sens) {
    const int64_t px = kx * HW + k;
    const float ds = disps_sens[px];
    if (ds > 0.f) {
      et = alpha;
      b -= static_cast<Acc>(alpha) * (static_cast<Acc>(disps[px]) - static_cast<Acc>(ds));
    }
  }
  Q[(int64_t)f * HW + k] = static_cast<float>(Acc(1) / (c + et + c_eps));
  w[(int64_t)f * HW + k] = static_cast<float>(b);
  if (p >= 0) {
#pragma unroll
    for (int n = 0; n < 6; ++n) Eall[((int64_t)p * 6 + n) * HW + k] = static_cast<float>(ei[n]);
  }
}

// The linearisation (blocks [0, nlin): edge b % E, pixel slice b / E) and the frame sums (the
// blocks after them: frame fb = (b - nlin) / nx, pixel block (b - nlin) % nx) in one launch.  The
// frame sums recompute their rows from the pixels, so they need nothing from the linearisation;
// the two write disjoint outputs (E_ij rows, Hessian blocks / the frames' Q, w and E_i rows), and
// the frame sums use the memory bandwidth that the linearisation's arithmetic leaves free.
template <int kSem, typename Acc>
__global__ void __launch_bounds__(256)
ba_lin_frame(const float* __restrict__ poses, const float* __restrict__ disps, const float* __restrict__ intr,
             const float* __restrict__ targets, const float* __restrict__ weights, const int* __restrict__ edge_ii,
             const int* __restrict__ edge_jj, int HW, int wd, int E, Acc* __restrict__ Hs, Acc* __restrict__ vs,
             float* __restrict__ Eij, int slices, Acc* __restrict__ Hpart, int nlin, const float* __restrict__ eta,
             int eta_rows, const int* __restrict__ frame_eptr, const int* __restrict__ frame_edges,
             const int* __restrict__ frame_pose, const int* __restrict__ frame_order,
             const int* __restrict__ frame_kx, float* __restrict__ Q, float* __restrict__ w, float* __restrict__ Eall,
             const float* __restrict__ disps_sens, float alpha, int nx) {
  static_assert(kLinThreads == 256, "one block size for both parts");
  static_assert(kSem == kSemCuda, "the ba_cuda loop: one intrinsics vector, Q = 1 / (C + eta)");
  const int b = blockIdx.x;
  if (b < nlin) {
    linearize_block<true, kSem, Acc>(poses, disps, intr, 0, targets, weights, edge_ii, edge_jj, HW, wd, E, Hs, vs,
                                     Eij, nullptr, nullptr, nullptr, slices, Hpart, b % E, b / E);
    return;
  }
  const int fb = b - nlin;
  frame_sums_block<kSem, Acc>(poses, disps, intr, 0, targets, weights, edge_ii, edge_jj, eta, eta_rows, frame_eptr,
                              frame_edges, frame_pose, frame_order, frame_kx, HW, wd, 0.f, Q, w, Eall, disps_sens,
                              alpha, fb / nx, fb % nx);
}

// The frame sums alone (grid: pixel blocks x frames), for graphs large enough to fill the GPU
// with the linearisation by itself: a launch of their own keeps their occupancy instead of the
// linearisation's (212 VGPRs).  With FP64 sums the kernel is held to five waves per SIMD (93
// VGPRs, no spills; 108 and four waves unconstrained), which measured slightly faster; the
// FP32-sum instantiation would spill at five and keeps the compiler's choice.
template <int kSem, typename Acc>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(Acc) == 8 ? 5 : 1)))
ba_frame_sums(const float* __restrict__ poses, const float* __restrict__ disps, const float* __restrict__ intr,
              const float* __restrict__ targets, const float* __restrict__ weights, const int* __restrict__ edge_ii,
              const int* __restrict__ edge_jj, int HW, int wd, const float* __restrict__ eta, int eta_rows,
              const int* __restrict__ frame_eptr, const int* __restrict__ frame_edges,
              const int* __restrict__ frame_pose, const int* __restrict__ frame_order,
              const int* __restrict__ frame_kx, float* __restrict__ Q, float* __restrict__ w, float* __restrict__ Eall,
              const float* __restrict__ disps_sens, float alpha) {
  static_assert(kSem == kSemCuda, "the ba_cuda loop: one intrinsics vector, Q = 1 / (C + eta)");
  frame_sums_block<kSem, Acc>(poses, disps, intr, 0, targets, weights, edge_ii, edge_jj, eta, eta_rows, frame_eptr,
                              frame_edges, frame_pose, frame_order, frame_kx, HW, wd, 0.f, Q, w, Eall, disps_sens,
                              alpha, static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.x));
}

// ---- 3. Ev rows, one wave per slot -------------------------------------------
// Ev[s] = sum_px Q w E_s for the Ev tasks [task0, NT) of the plan's task list (task0 = NPAIR:
// the pair blocks come from ba_schur_gram_lds).
template <typename Acc>
__global__ void __launch_bounds__(256)
ba_schur_ev(const float* __restrict__ Eall, const float* __restrict__ Q,
            const float* __restrict__ w, const int* __restrict__ tasks,
            const int* __restrict__ slot_row, int task0, int NT, int HW, int task_blocks8,
            Acc* __restrict__ Ev) {
  // XCD-aware order: the 8 XCDs receive blocks round-robin, so give each XCD a
  // contiguous run of tasks (tasks of one depth frame share Q and w in L2).
  const int b = blockIdx.x;
  const int tb = (b % 8) * task_blocks8 + b / 8;
  const int lane = threadIdx.x % kWave;
  const int task = task0 + tb * 4 + threadIdx.x / kWave;
  if (task >= NT) return;
  const int sa = tasks[task * 4 + 1];
  const int f = tasks[task * 4 + 3];
  const float* Qf = Q + (int64_t)f * HW;
  const float* Ea = Eall + (int64_t)slot_row[sa] * 6 * HW;
  const float* wf = w + (int64_t)f * HW;
  Acc acc[6] = {0, 0, 0, 0, 0, 0};
  for (int k = lane; k < HW; k += kWave) {
    const Acc qw = static_cast<Acc>(Qf[k]) * wf[k];
#pragma unroll
    for (int n = 0; n < 6; ++n) acc[n] += qw * Ea[n * HW + k];
  }
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    const Acc v = wave_sum(acc[n]);
    if (lane == n) Ev[(int64_t)sa * 6 + n] = v;
  }
}

// ---- 3b. Schur Gram blocks of the slot pairs on the matrix cores ----------------
// For depth frame f with m slots, G = E_f diag(Q_f) E_f^T is the 6m x 6m Gram matrix of the
// slots' E rows (row 6a + n = component n of slot a); the Schur pair (a, b >= a) is its 6 x 6
// block at (6a, 6b), computed in 16 x 16 tiles with v_mfma_f32_16x16x4f32.
// MFMA operand layout (16x16x4, f32): lane l holds A[l % 16][l / 16] and B[l / 16][l % 16];
// accumulator element i of lane l is D[4 (l / 16) + i][l % 16].
typedef float gram_f4 __attribute__((ext_vector_type(4)));

// row[k .. k+3] with elements at or beyond `end` read as zero
__device__ __forceinline__ float4 gram_load4(const float* __restrict__ row, int k, int end, bool vec) {
  if (vec) return *reinterpret_cast<const float4*>(row + k);
  return make_float4(k < end ? row[k] : 0.f, k + 1 < end ? row[k + 1] : 0.f, k + 2 < end ? row[k + 2] : 0.f,
                     k + 3 < end ? row[k + 3] : 0.f);
}

// One block per item (f, ga, gb): the 16-row tiles of groups ga and gb (kGramGroup tiles each)
// of frame f's Gram matrix, i.e. the tile pairs (ti <= tj) inside group ga when ga == gb and
// all pairs (ti in ga, tj in gb) otherwise.  Each chunk of kGramKC pixels of the item's rows
// (at most 2 x 64) and of Q is staged in LDS (pixel-major, rows padded so that fragment reads
// and staging writes are free of bank conflicts) while the previous chunk is consumed; the
// next chunk's loads are in flight meanwhile (one barrier per chunk).  Tile pair p goes to
// wave p % 4; step s of a chunk feeds lane l pixel 4 s + l / 16 of its A row (scaled by Q) and
// B row.  FP32 tile sums are folded into FP64 every 256 pixels.  Only entries with a <= b are
// written; a diagonal block (a == b) is written from its entries r <= c, mirrored, so it is
// exactly symmetric.  (Acc = float, the all-FP32 mode, folds into FP32.)  A diagonal item (ga == gb) also forms the Ev rows of its group's rows,
// Ev = sum_px Q w E (EEt6x6 / Ev6x1's Ev, src/droid_kernels.cu:1466-1560), from the staged
// chunks: each thread 16 pixels of one row per chunk in FP32, FP64 across chunks and threads.
// When the plan has few items (small problems: one item per frame for a frontend window) the
// pixel range is cut into `slices` runs of whole 256-pixel folds, one block per (item, slice):
// each block writes its partial S and Ev to its slice of Spart ([NPAIR 36 | NS 6] per slice) and
// gram_reduce sums the slices in order (so the result does not depend on the schedule).
constexpr int kGramKC = 32;                          // pixels per staged chunk
constexpr int kGramPT = kGramKC / 2;                 // pixels of one row staged per thread
constexpr int kGramFold = 256 / kGramKC;             // chunks per FP64 fold
constexpr int kGramRS = 2 * 16 * kGramGroup + 16;   // LDS row stride (floats), 16 mod 64

template <typename Acc>
__global__ void __launch_bounds__(256, 3)
ba_schur_gram_lds(const float* __restrict__ Eall, const float* __restrict__ Q,
                  const int* __restrict__ items, int n_items, int item_blocks8,
                  const int* __restrict__ slot_ptr, const int* __restrict__ slot_row,
                  const int* __restrict__ frame_pair0, int HW, const float* __restrict__ wvec,
                  Acc* __restrict__ S, Acc* __restrict__ Ev, int slices, int64_t npair, int64_t part_stride,
                  Acc* __restrict__ Spart, const Acc* __restrict__ Hpart, int lin_slices, int E,
                  Acc* __restrict__ Hs, Acc* __restrict__ vs) {
  __shared__ float Ls[2][kGramKC][kGramRS];
  __shared__ float qs[2][kGramKC], qws[2][kGramKC];
  __shared__ Acc evs[4][16 * kGramGroup];
  const int blk = blockIdx.x;
  if (blk >= 8 * item_blocks8) {                  // lin_reduce of a sliced linearisation, alongside
    lin_reduce_elem(Hpart, lin_slices, E, Hs, vs, (int64_t)(blk - 8 * item_blocks8) * blockDim.x + threadIdx.x);
    return;
  }
  const int idx = (blk % 8) * item_blocks8 + blk / 8;   // XCD-aware: runs of consecutive (item, slice)
  if (idx >= n_items * slices) return;
  const int item = idx / slices, slice = idx - item * slices;
  const int t = threadIdx.x, lane = t % kWave;
  const int wv = __builtin_amdgcn_readfirstlane(t / kWave);
  const int f = items[3 * item], ga = items[3 * item + 1], gb = items[3 * item + 2];
  const int s0 = slot_ptr[f], m = slot_ptr[f + 1] - s0, R = 6 * m;
  const int nt = (R + 15) / 16;
  const int a0 = ga * kGramGroup, a1 = min(nt, a0 + kGramGroup);
  const int b0 = gb * kGramGroup, b1 = min(nt, b0 + kGramGroup);
  const bool diag = ga == gb;
  const int na = 16 * (a1 - a0), nr = na + (diag ? 0 : 16 * (b1 - b0));
  // this wave's tile pairs: staged row offsets of the A and B tiles, and the tile indices
  int rowA[4] = {0, 0, 0, 0}, rowB[4] = {0, 0, 0, 0}, tiA[4] = {0, 0, 0, 0}, tjB[4] = {0, 0, 0, 0};
  int np = 0;
  {
    int p = 0;
    for (int ti = a0; ti < a1; ++ti)
      for (int tj = diag ? ti : b0; tj < (diag ? a1 : b1); ++tj, ++p)
        if (p % 4 == wv && np < 4) {
          rowA[np] = 16 * (ti - a0);
          rowB[np] = diag ? 16 * (tj - a0) : na + 16 * (tj - b0);
          tiA[np] = ti;
          tjB[np] = tj;
          ++np;
        }
    np = __builtin_amdgcn_readfirstlane(np);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      rowA[u] = __builtin_amdgcn_readfirstlane(rowA[u]);
      rowB[u] = __builtin_amdgcn_readfirstlane(rowB[u]);
    }
  }
  // staging role: thread t stages pixels [kGramPT h, kGramPT (h + 1)) of staged row r; rows
  // beyond the frame's 6m stage zeros
  const int r = t & 127, h = t >> 7;
  const float* src = nullptr;
  if (r < nr) {
    const int gr = r < na ? 16 * a0 + r : 16 * b0 + (r - na);
    if (gr < R) src = Eall + ((int64_t)slot_row[s0 + gr / 6] * 6 + gr % 6) * HW;
  }
  const float* qf = Q + (int64_t)f * HW;
  const float* wf = wvec + (int64_t)f * HW;
  const bool vec = (HW % 4) == 0;
  // staging registers of one chunk: kGramPT pixels of one row, and Q, w of 4 pixels
  struct Stage {
    float4 x[kGramPT / 4], q, w;
  };
  // Ev rows of a diagonal item's rows: thread t sums row t % 64 over a quarter of every chunk
  // (pixels kGramKC / 4 * (t / 64) + [0, kGramKC / 4))
  const int er = t % 64, eq = t / 64;
  const bool ev_row = diag && er < na && 16 * a0 + er < R;
  Acc evd = 0;
  auto load = [&](Stage& g, int k0) {
    const int kk = k0 + kGramPT * h;
#pragma unroll
    for (int j = 0; j < kGramPT / 4; ++j) g.x[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    g.q = make_float4(0.f, 0.f, 0.f, 0.f);
    g.w = g.q;
    if (vec && k0 + kGramKC <= HW) {               // whole chunk inside the row: 16-byte loads
      if (src) {
#pragma unroll
        for (int j = 0; j < kGramPT / 4; ++j) g.x[j] = *reinterpret_cast<const float4*>(src + kk + 4 * j);
      }
      if (t < kGramKC / 4) {
        g.q = *reinterpret_cast<const float4*>(qf + k0 + 4 * t);
        if (diag) g.w = *reinterpret_cast<const float4*>(wf + k0 + 4 * t);
      }
    } else {
      if (src) {
#pragma unroll
        for (int j = 0; j < kGramPT / 4; ++j) g.x[j] = gram_load4(src, kk + 4 * j, HW, false);
      }
      if (t < kGramKC / 4) {
        g.q = gram_load4(qf, k0 + 4 * t, HW, false);
        if (diag) g.w = gram_load4(wf, k0 + 4 * t, HW, false);
      }
    }
  };
  auto store = [&](const Stage& g, int buf) {
    if (r < nr) {
#pragma unroll
      for (int j = 0; j < kGramPT / 4; ++j) {
        const int k = kGramPT * h + 4 * j;
        Ls[buf][k][r] = g.x[j].x;
        Ls[buf][k + 1][r] = g.x[j].y;
        Ls[buf][k + 2][r] = g.x[j].z;
        Ls[buf][k + 3][r] = g.x[j].w;
      }
    }
    if (t < kGramKC / 4) {
      qs[buf][4 * t] = g.q.x;
      qs[buf][4 * t + 1] = g.q.y;
      qs[buf][4 * t + 2] = g.q.z;
      qs[buf][4 * t + 3] = g.q.w;
      qws[buf][4 * t] = g.q.x * g.w.x;
      qws[buf][4 * t + 1] = g.q.y * g.w.y;
      qws[buf][4 * t + 2] = g.q.z * g.w.z;
      qws[buf][4 * t + 3] = g.q.w * g.w.w;
    }
  };
  Acc dacc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int i = 0; i < 4; ++i) dacc[u][i] = 0.0;
  const int nch_all = (HW + kGramKC - 1) / kGramKC;
  const int nfold = (nch_all + kGramFold - 1) / kGramFold;
  const int c_begin = (slice * nfold / slices) * kGramFold;
  const int nch = min(nch_all, ((slice + 1) * nfold / slices) * kGramFold) - c_begin;   // chunks of this block
  // NP tile pairs per wave, the same for the whole block: waves with fewer pairs run
  // copies of a staged tile whose sums are not written
  auto run = [&](auto np_c) {
    constexpr int NP = decltype(np_c)::value;
    gram_f4 acc[NP];
#pragma unroll
    for (int u = 0; u < NP; ++u) acc[u] = gram_f4{0.f, 0.f, 0.f, 0.f};
    // chunk c + 2 is loaded while chunk c is consumed and chunk c + 1 waits in registers
    Stage g0, g1;
    load(g0, c_begin * kGramKC);
    if (nch > 1) load(g1, (c_begin + 1) * kGramKC);
    store(g0, 0);
    __syncthreads();
    // cur held chunk c (already in LDS), nxt holds chunk c + 1
    auto chunk = [&](int c, Stage& cur, const Stage& nxt) {
      const int buf = c & 1;
      if (c + 2 < nch) load(cur, (c_begin + c + 2) * kGramKC);
#pragma unroll
      for (int st = 0; st < kGramKC / 4; ++st) {
        const int kr = 4 * st + lane / 16;
        const float q = qs[buf][kr];
#pragma unroll
        for (int u = 0; u < NP; ++u) {
          const float a = Ls[buf][kr][rowA[u] + lane % 16] * q;
          const float b = Ls[buf][kr][rowB[u] + lane % 16];
          acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[u], 0, 0, 0);
        }
      }
      if (ev_row) {
        float e = 0.f;
#pragma unroll
        for (int j = 0; j < kGramKC / 4; ++j)
          e = fmaf(Ls[buf][kGramKC / 4 * eq + j][er], qws[buf][kGramKC / 4 * eq + j], e);
        evd += static_cast<Acc>(e);
      }
      if (c % kGramFold == kGramFold - 1 || c + 1 == nch) {   // every 256 pixels (c_begin: whole folds)
#pragma unroll
        for (int u = 0; u < NP; ++u) {
#pragma unroll
          for (int i = 0; i < 4; ++i) dacc[u][i] += static_cast<Acc>(acc[u][i]);
          acc[u] = gram_f4{0.f, 0.f, 0.f, 0.f};
        }
      }
      if (c + 1 < nch) store(nxt, buf ^ 1);
      __syncthreads();
    };
    for (int c = 0; c < nch; c += 2) {
      chunk(c, g0, g1);
      if (c + 1 < nch) chunk(c + 1, g1, g0);
    }
  };
  const int npairs = diag ? (a1 - a0) * (a1 - a0 + 1) / 2 : (a1 - a0) * (b1 - b0);
  switch ((npairs + 3) / 4) {
    case 1: run(std::integral_constant<int, 1>{}); break;
    case 2: run(std::integral_constant<int, 2>{}); break;
    case 3: run(std::integral_constant<int, 3>{}); break;
    default: run(std::integral_constant<int, 4>{}); break;
  }

  if (slices > 1) {                                      // partial sums of this slice
    Acc* part = Spart + (int64_t)slice * part_stride;
    S = part;
    Ev = part + npair * 36;
  }
  if (diag) {
    evs[eq][er] = evd;
    __syncthreads();
    if (ev_row && eq == 0) {
      const int gr = 16 * a0 + er;
      Ev[(int64_t)(s0 + gr / 6) * 6 + gr % 6] = (evs[0][er] + evs[1][er]) + (evs[2][er] + evs[3][er]);
    }
  }
  const int64_t pair0 = frame_pair0[f];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (u >= np) continue;
    const int cidx = 16 * tjB[u] + lane % 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ridx = 16 * tiA[u] + 4 * (lane / 16) + i;
      if (ridx >= R || cidx >= R || ridx > cidx) continue;
      const int a = ridx / 6, n = ridx - 6 * a, b = cidx / 6, n2 = cidx - 6 * b;
      const int64_t pair = pair0 + (int64_t)a * m - (int64_t)a * (a - 1) / 2 + (b - a);
      S[pair * 36 + n * 6 + n2] = dacc[u][i];
      if (a == b) S[pair * 36 + n2 * 6 + n] = dacc[u][i];
    }
  }
}

// Sum of the slices' partial S and Ev (in slice order): part [slices][npair 36 + ns 6].
template <typename Acc>
__global__ void __launch_bounds__(256)
gram_reduce(const Acc* __restrict__ part, int slices, int64_t npair, int64_t stride, Acc* __restrict__ S,
            Acc* __restrict__ Ev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= stride) return;
  Acc v = part[i];
  for (int z = 1; z < slices; ++z) v += part[(int64_t)z * stride + i];
  if (i < npair * 36) S[i] = v;
  else Ev[i - npair * 36] = v;
}

// The Ev rows with the weight vector wvec (one wave per slot).
template <typename Acc>
inline void launch_schur_ev(const hsl2_ba_plan* p, const int* blob, const float* Eall, const float* Q,
                            const float* wvec, Acc* Ev, hipStream_t s) {
  const int task0 = static_cast<int>(p->NPAIR);
  const int nt = static_cast<int>(p->NT) - task0;
  if (nt <= 0) return;
  const int tb8 = ((nt + 3) / 4 + 7) / 8;
  hipLaunchKernelGGL(ba_schur_ev<Acc>, dim3(tb8 * 8), dim3(256), 0, s, Eall, Q, wvec, blob + p->o_task,
                     blob + p->o_slot_row, task0, static_cast<int>(p->NT), static_cast<int>(p->HW), tb8, Ev);
}

// Schur pair blocks and Ev rows (both from the Gram kernel; Ev alone when there are no
// Gram items, i.e. motion-only).
// With Hpart != nullptr (a linearisation cut into lin_slices > 1 pixel slices) the Gram launch
// also carries lin_reduce (Hpart -> Hs, vs) in blocks after its own; without Gram items it runs
// on its own.
template <typename Acc>
inline void launch_schur_gram(const hsl2_ba_plan* p, const int* blob, const float* Eall, const float* Q,
                              const float* wvec, Acc* S, Acc* Ev, Acc* Spart, hipStream_t s,
                              const Acc* Hpart = nullptr, Acc* Hs = nullptr, Acc* vs = nullptr) {
  const int HW = static_cast<int>(p->HW);
  const int E = static_cast<int>(p->E);
  const int nlr = Hpart ? static_cast<int>((E * 156 + 255) / 256) : 0;
  if (p->NGGI > 0) {
    const int ni = static_cast<int>(p->NGGI);
    const int z = p->gram_slices;
    const int ib8 = (ni * z + 7) / 8;
    const int64_t stride = p->NPAIR * 36 + p->NS * 6;
    hipLaunchKernelGGL(ba_schur_gram_lds<Acc>, dim3(ib8 * 8 + nlr), dim3(256), 0, s, Eall, Q,
                       blob + p->o_gram_group_item, ni, ib8, blob + p->o_slot_ptr, blob + p->o_slot_row,
                       blob + p->o_frame_pair0, HW, wvec, S, Ev, z, p->NPAIR, stride, Spart, Hpart, p->lin_slices, E,
                       Hs, vs);
    if (z > 1)
      hipLaunchKernelGGL(gram_reduce<Acc>, dim3(static_cast<unsigned>((stride + 255) / 256)), dim3(256), 0, s, Spart,
                         z, p->NPAIR, stride, S, Ev);
  } else {
    launch_schur_ev(p, blob, Eall, Q, wvec, Ev, s);
    if (nlr > 0)
      hipLaunchKernelGGL(lin_reduce<Acc>, dim3(static_cast<unsigned>(nlr)), dim3(256), 0, s, Hpart, p->lin_slices, E,
                         Hs, vs);
  }
}

// ---- 4. reduced camera system, dense column-major ----------------------------
template <typename R>
__device__ __forceinline__ void assemble_elem(const R* __restrict__ Hs, const R* __restrict__ S,
                                              const int* __restrict__ blk_ptr, const int* __restrict__ blk_src,
                                              int E, int P, int64_t n_pad, float lm, float ep, int damp_on_H,
                                              R* __restrict__ A, int* __restrict__ info, int64_t idx) {
  const int64_t n = 6 * (int64_t)P;
  if (idx == 0) info[0] = 0;                     // this iteration's factorisation status
  if (idx >= n_pad * n_pad) return;
  const int64_t col = idx / n_pad;
  const int64_t row = idx - col * n_pad;
  if (row >= n || col >= n) {                    // identity padding up to a multiple of 64
    A[idx] = row == col ? R(1) : R(0);
    return;
  }
  const int a = static_cast<int>(row / 6), i = static_cast<int>(row % 6);
  const int b = static_cast<int>(col / 6), j = static_cast<int>(col % 6);
  const int64_t blk = (int64_t)a * P + b;
  R vh = 0, vs = 0;
  // contributions four at a time (their code and value loads in flight together), summed in
  // list order
  auto src = [&](int code) {
    const int kind = code & 7;
    const int64_t id = code >> 3;
    if (kind < 4) return Hs + ((int64_t)kind * E + id) * 36 + i * 6 + j;
    return S + id * 36 + (kind == kBlkS ? i * 6 + j : j * 6 + i);
  };
  const int q1 = blk_ptr[blk + 1];
  int q = blk_ptr[blk];
  for (; q + 4 <= q1; q += 4) {
    int code[4];
    R val[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) code[u] = blk_src[q + u];
#pragma unroll
    for (int u = 0; u < 4; ++u) val[u] = *src(code[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if ((code[u] & 7) < 4) vh += val[u];
      else vs += val[u];
    }
  }
  for (; q < q1; ++q) {
    const int code = blk_src[q];
    const R val = *src(code);
    if ((code & 7) < 4) vh += val;
    else vs += val;
  }
  R v = vh - vs;
  // damp_on_H = 0: ba_cuda damps the reduced diagonal (src/droid_kernels.cu:1663-1664);
  // 1: the torch formulation's schur_solve damps H's diagonal before the Schur complement
  // (hislam2/geom/chol.py:56-57); 2: its block_solve (chol.py:36-37) adds (ep + lm H) * eye(6)
  // to every 6x6 block [P, P, 6, 6], off-diagonal blocks included.
  const bool damped = damp_on_H == 2 ? (i == j) : (row == col);
  if (damped) v += static_cast<R>(ep) + static_cast<R>(lm) * (damp_on_H ? vh : v);
  A[idx] = v;
}

template <typename R>
__global__ void __launch_bounds__(256)
ba_assemble(const R* __restrict__ Hs, const R* __restrict__ S,
            const int* __restrict__ blk_ptr, const int* __restrict__ blk_src, int E, int P,
            int64_t n_pad, float lm, float ep, int damp_on_H, R* __restrict__ A,
            int* __restrict__ info) {
  assemble_elem(Hs, S, blk_ptr, blk_src, E, P, n_pad, lm, ep, damp_on_H, A, info,
                (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// one wave per pose (six rows): lanes take the pose's contributions in turn, each lane sums its
// share of the six components, then the wave reduces them (fixed order).  Wave-uniform `a`.
template <typename R>
__device__ __forceinline__ void assemble_rhs_wave(const R* __restrict__ vs, const R* __restrict__ Ev,
                                                  const int* __restrict__ rhs_ptr, const int* __restrict__ rhs_src,
                                                  int E, int P, int n_pad, R* __restrict__ bvec, int a) {
  const int lane = threadIdx.x % kWave;
  if (6 * a >= n_pad) return;
  R v[6] = {0, 0, 0, 0, 0, 0};
  if (a < P) {
    for (int q = rhs_ptr[a] + lane; q < rhs_ptr[a + 1]; q += kWave) {
      const int code = rhs_src[q];
      const int kind = code & 3;
      const int64_t id = code >> 2;
      const R* src = kind == kRhsVi ? vs + id * 6 : (kind == kRhsVj ? vs + ((int64_t)E + id) * 6 : Ev + id * 6);
      const R sg = kind == kRhsEv ? R(-1) : R(1);
#pragma unroll
      for (int i = 0; i < 6; ++i) v[i] += sg * src[i];
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const R tot = wave_sum(v[i]);
    const int row = 6 * a + i;
    if (lane == i && row < n_pad) bvec[row] = row < 6 * P ? tot : R(0);
  }
}

template <typename R>
__global__ void __launch_bounds__(256)
ba_assemble_rhs(const R* __restrict__ vs, const R* __restrict__ Ev,
                const int* __restrict__ rhs_ptr, const int* __restrict__ rhs_src, int E, int P,
                int n_pad, R* __restrict__ bvec) {
  assemble_rhs_wave(vs, Ev, rhs_ptr, rhs_src, E, P, n_pad, bvec,
                    static_cast<int>(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
}

// The reduced system and its right-hand side in one launch: blocks [0, nA) assemble A (one
// element per thread), the blocks after them the rhs (one pose per wave).
template <typename R>
__global__ void __launch_bounds__(256)
ba_assemble_sys(const R* __restrict__ Hs, const R* __restrict__ S, const int* __restrict__ blk_ptr,
                const int* __restrict__ blk_src, const R* __restrict__ vs, const R* __restrict__ Ev,
                const int* __restrict__ rhs_ptr, const int* __restrict__ rhs_src, int E, int P, int64_t n_pad,
                float lm, float ep, R* __restrict__ A, R* __restrict__ bvec, int* __restrict__ info, int nA) {
  const int b = blockIdx.x;
  if (b < nA) {
    assemble_elem(Hs, S, blk_ptr, blk_src, E, P, n_pad, lm, ep, 0, A, info, (int64_t)b * blockDim.x + threadIdx.x);
    return;
  }
  assemble_rhs_wave(vs, Ev, rhs_ptr, rhs_src, E, P, static_cast<int>(n_pad), bvec,
                    (b - nA) * static_cast<int>(blockDim.x / kWave) + static_cast<int>(threadIdx.x / kWave));
}

// ---- 6. back-substitution and disparity update -------------------------------
template <typename R>
__device__ __forceinline__ void backsub_px(const float* __restrict__ Eall, const float* __restrict__ Q,
                                           const float* __restrict__ w, const R* __restrict__ x,
                                           const int* __restrict__ info, const int* __restrict__ slot_ptr,
                                           const int* __restrict__ slot_row, const int* __restrict__ slot_pose,
                                           const int* __restrict__ frame_kx, int HW, int t0_skip,
                                           float* __restrict__ disps, float* __restrict__ dz, int f, int k) {
  if (k >= HW) return;
  const bool solved = info[0] == 0;
  R acc = 0;
  if (solved) {
    // slots two at a time, their loads in flight together; summed in slot order
    const int s1 = slot_ptr[f + 1];
    auto add = [&](int a, const float (&e)[6]) {
      if (t0_skip && a <= 0) return;              // t0 skip, src/droid_kernels.cu:1572
#pragma unroll
      for (int n = 0; n < 6; ++n) acc += static_cast<R>(e[n]) * x[6 * a + n];
    };
    int s = slot_ptr[f];
    for (; s + 2 <= s1; s += 2) {
      float e[2][6];
      int a[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        a[u] = slot_pose[s + u];
        const float* Er = Eall + (int64_t)slot_row[s + u] * 6 * HW + k;
#pragma unroll
        for (int n = 0; n < 6; ++n) e[u][n] = Er[(int64_t)n * HW];
      }
      add(a[0], e[0]);
      add(a[1], e[1]);
    }
    if (s < s1) {
      float e[6];
      const float* Er = Eall + (int64_t)slot_row[s] * 6 * HW + k;
#pragma unroll
      for (int n = 0; n < 6; ++n) e[n] = Er[(int64_t)n * HW];
      add(slot_pose[s], e);
    }
  }
  const float qv = Q[(int64_t)f * HW + k];
  const float d = static_cast<float>(static_cast<R>(qv) * (static_cast<R>(w[(int64_t)f * HW + k]) - acc));
  if (dz) dz[(int64_t)f * HW + k] = d;
  if (disps) disps[(int64_t)frame_kx[f] * HW + k] += d;
}

template <typename R>
__global__ void __launch_bounds__(256)
ba_backsub(const float* __restrict__ Eall, const float* __restrict__ Q,
           const float* __restrict__ w, const R* __restrict__ x, const int* __restrict__ info,
           const int* __restrict__ slot_ptr, const int* __restrict__ slot_row,
           const int* __restrict__ slot_pose, const int* __restrict__ frame_kx, int HW,
           int t0_skip, float* __restrict__ disps, float* __restrict__ dz) {
  backsub_px(Eall, Q, w, x, info, slot_ptr, slot_row, slot_pose, frame_kx, HW, t0_skip, disps, dz,
             static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x));
}

// ---- 7. pose retraction -------------------------------------------------------
// full_v = 0: ba_cuda's expSE3, which applies V only for |phi| > 1e-4; full_v = 1: lietorch's
// Exp with the SO3 left Jacobian V at every angle, its coefficients from their Taylor series
// below 0.1 rad (no cancellation in float).
__device__ void retr_se3(const float xi[6], float* pose, int full_v) {
  // Exp(xi) * T (src/droid_kernels.cu:179-244, 1342-1360)
  const float* tau = xi;
  const float* phi = xi + 3;
  const float th2 = phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2];
  const float th = sqrtf(th2);
  float imag, real;
  if (th2 < 1e-8f) {
    imag = 0.5f - (1.f / 48.f) * th2 + (1.f / 3840.f) * th2 * th2;
    real = 1.f - (1.f / 8.f) * th2 + (1.f / 384.f) * th2 * th2;
  } else {
    imag = sinf(0.5f * th) / th;
    real = cosf(0.5f * th);
  }
  const float dq[4] = {imag * phi[0], imag * phi[1], imag * phi[2], real};
  float dt[3] = {tau[0], tau[1], tau[2]};
  if (full_v || th > 1e-4f) {
    float a, b;
    if (!full_v) {
      a = (1.f - cosf(th)) / th2;
      b = (th - sinf(th)) / (th * th2);
    } else if (th < 0.1f) {
      a = 0.5f - th2 * (1.f / 24.f) + th2 * th2 * (1.f / 720.f);
      b = (1.f / 6.f) - th2 * (1.f / 120.f) + th2 * th2 * (1.f / 5040.f);
    } else {
      const float sh = sinf(0.5f * th);
      a = 2.f * sh * sh / th2;
      b = (th - sinf(th)) / (th * th2);
    }
    const float c1[3] = {phi[1] * tau[2] - phi[2] * tau[1], phi[2] * tau[0] - phi[0] * tau[2],
                         phi[0] * tau[1] - phi[1] * tau[0]};
    const float c2[3] = {phi[1] * c1[2] - phi[2] * c1[1], phi[2] * c1[0] - phi[0] * c1[2],
                         phi[0] * c1[1] - phi[1] * c1[0]};
#pragma unroll
    for (int k = 0; k < 3; ++k) dt[k] += a * c1[k] + b * c2[k];
  }
  const float* q = pose + 3;
  float q1[4];
  q1[0] = dq[3] * q[0] + dq[0] * q[3] + dq[1] * q[2] - dq[2] * q[1];
  q1[1] = dq[3] * q[1] + dq[1] * q[3] + dq[2] * q[0] - dq[0] * q[2];
  q1[2] = dq[3] * q[2] + dq[2] * q[3] + dq[0] * q[1] - dq[1] * q[0];
  q1[3] = dq[3] * q[3] - dq[0] * q[0] - dq[1] * q[1] - dq[2] * q[2];
  float t1[3];
  quat_rotate(dq, pose, t1);
  pose[0] = t1[0] + dt[0];
  pose[1] = t1[1] + dt[1];
  pose[2] = t1[2] + dt[2];
  pose[3] = q1[0];
  pose[4] = q1[1];
  pose[5] = q1[2];
  pose[6] = q1[3];
}

template <typename R>
__device__ __forceinline__ void pose_retr_one(float* __restrict__ poses, const R* __restrict__ x,
                                              int* __restrict__ info, int t0, int P, int full_v,
                                              float* __restrict__ dx, int a) {
  const bool solved = info[0] == 0;
  if (a == 0 && !solved) info[1] = 1;            // remember a failed solve
  if (a >= P) return;
  float xi[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) xi[n] = solved ? static_cast<float>(x[6 * a + n]) : 0.f;
  if (dx) {
#pragma unroll
    for (int n = 0; n < 6; ++n) dx[6 * a + n] = xi[n];
  }
  float pose[7];
#pragma unroll
  for (int n = 0; n < 7; ++n) pose[n] = poses[(int64_t)(t0 + a) * 7 + n];
  retr_se3(xi, pose, full_v);
#pragma unroll
  for (int n = 0; n < 7; ++n) poses[(int64_t)(t0 + a) * 7 + n] = pose[n];
}

template <typename R>
__global__ void ba_pose_retr(float* __restrict__ poses, const R* __restrict__ x,
                             int* __restrict__ info, int t0, int P, int full_v,
                             float* __restrict__ dx) {
  pose_retr_one(poses, x, info, t0, P, full_v, dx, static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x));
}

// The depth back-substitution with the disparity update (blocks [0, nbs): frame b / nx, pixel
// block b % nx) and the pose retraction (the blocks after them, one pose per thread) in one
// launch: both only read the solution and write disjoint outputs (disparities, poses).
template <typename R>
__global__ void __launch_bounds__(256)
ba_update(const float* __restrict__ Eall, const float* __restrict__ Q, const float* __restrict__ w,
          const R* __restrict__ x, int* __restrict__ info, const int* __restrict__ slot_ptr,
          const int* __restrict__ slot_row, const int* __restrict__ slot_pose, const int* __restrict__ frame_kx,
          int HW, float* __restrict__ disps, float* __restrict__ dz, int nx, int nbs, float* __restrict__ poses,
          int t0, int P, float* __restrict__ dx) {
  const int b = blockIdx.x;
  if (b < nbs) {
    backsub_px(Eall, Q, w, x, info, slot_ptr, slot_row, slot_pose, frame_kx, HW, 1, disps, dz, b / nx,
               (b % nx) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x));
    return;
  }
  pose_retr_one(poses, x, info, t0, P, 0, dx, (b - nbs) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x));
}

__global__ void ba_clear_flags(int* info) {
  if (threadIdx.x < 4) info[threadIdx.x] = 0;
}

}  // namespace
}  // namespace hsl2
