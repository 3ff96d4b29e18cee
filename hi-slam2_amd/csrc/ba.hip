// Dense bundle adjustment on gfx950: the iteration loop of ba_cuda
// (src/droid_kernels.cu:1820-1937) with every step on the device.
//
// Per Gauss-Newton iteration:
//   1. ba_lin_frame      the linearisation and the frame sums in one launch (small graphs;
//                        ba_linearize then ba_frame_sums for large ones):
//                        linearize_block, one block per edge (and pixel slice): relative pose,
//                        per-pixel residuals and Jacobians, the 12x12 Hessian and gradient of
//                        the edge reduced through LDS, and the pose-j Schur rows E_ij
//                        (src/droid_kernels.cu:364-606); frame_sums_block, per depth frame
//                        from the pixels (each edge's rows recomputed): C, w, Q = 1/(C+eta)
//                        and the pose-i block Ei of the frame's edges (accum_cuda,
//                        1413-1463, 1901-1906).
//   2. ba_schur_gram_lds per depth frame, 16x16 tiles of the slots' Gram matrix
//                        S = sum_px E_a Q E_b^T on the matrix cores, one block per group
//                        pair of row tiles staged in LDS by 32-pixel chunks (FP32 tile
//                        sums per 256 pixels, totals in the solve type); the blocks of diagonal
//                        groups also form Ev = sum_px Q w E (EEt6x6 / Ev6x1, 1466-1560)
//                        (+ gram_reduce for sliced items; a sliced linearisation's
//                        lin_reduce runs in the same launch).
//   3. ba_assemble_sys   the damped reduced camera system A - S and its rhs in FP64
//                        (FP32 in the all-FP32 mode, whose per-edge Hessian blocks, frame
//                        sums and Gram totals are FP32 as well), written densely in column-major
//                        order, padded to a multiple of 64 with an identity block
//                        (SparseBlock, 1584-1686).
//   4. cholesky_solve    tiled Cholesky factorisation and solve (chol.hip); a non-positive
//                        pivot gives dx = 0 (1669-1677).
//   5. ba_update         dz = Q (w - sum E^T dx) with the reference's t0 skip and the
//                        disparity update (EvT6x1, disp_retr, 1562-1582, 1398-1411), and
//                        T <- Exp(dx) T (1342-1396), in one launch.
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "ba_kernels.hpp"

namespace hsl2 {
namespace {

// Sensor-disparity prior weight of the 15-argument ba form (upstream DROID-SLAM's RGB-D mode).
constexpr float kSensAlpha = 0.05f;
// Graphs whose linearisation and frame sums take at most this many blocks together run them in
// one launch (a frontend window: 70 edges x 4 slices + 12 frames x 12 pixel blocks = 424); larger
// ones launch them apart, each at its own occupancy.
constexpr int kFuseBlocks = 1024;

template <typename R>
int run_iterations(const hsl2_ba_plan* p, float* poses, float* disps, const float* intr,
                   const float* disps_sens, const float* targets, const float* weights, const float* eta,
                   int eta_rows, int iterations, float lm, float ep, float* dx, float* dz, char* ws,
                   const int* blob, hipStream_t s) {
  float* Eall = reinterpret_cast<float*>(ws + p->w_Eall);
  float* Q = reinterpret_cast<float*>(ws + p->w_Q);
  float* wv = reinterpret_cast<float*>(ws + p->w_w);
  // sums in the solve type: FP64 in the split mode, FP32 in the all-FP32 mode
  R* Hs = reinterpret_cast<R*>(ws + p->w_Hs);
  R* vs = reinterpret_cast<R*>(ws + p->w_vs);
  R* S = reinterpret_cast<R*>(ws + p->w_S);
  R* Ev = reinterpret_cast<R*>(ws + p->w_Ev);
  R* Spart = reinterpret_cast<R*>(ws + p->w_Spart);
  R* Hpart = reinterpret_cast<R*>(ws + p->w_Hpart);
  const int zl = p->lin_slices;
  R* A = reinterpret_cast<R*>(ws + p->w_A);
  R* bvec = reinterpret_cast<R*>(ws + p->w_b);
  R* xvec = reinterpret_cast<R*>(ws + p->w_x);
  R* Ld = reinterpret_cast<R*>(ws + p->w_Ld);
  int* info = reinterpret_cast<int*>(ws + p->w_info);
  const int E = static_cast<int>(p->E), P = p->P, HW = static_cast<int>(p->HW);
  const int F = static_cast<int>(p->F);
  const int wd = static_cast<int>(p->wd);
  const bool schur = !p->motion_only;
  const int64_t n_pad = p->n_pad;

  hipLaunchKernelGGL(ba_clear_flags, dim3(1), dim3(64), 0, s, info);
  HSL2_LAUNCH_CHECK();
  const int nx = (HW + 255) / 256;                // pixel blocks per frame
  for (int it = 0; it < iterations; ++it) {
    const bool last = it == iterations - 1;
    if (E > 0) {
      const int nlin = E * zl;
      if (schur && nlin + nx * F <= kFuseBlocks) {
        // a small graph: linearisation and frame sums (independent) in one launch
        hipLaunchKernelGGL((ba_lin_frame<kSemCuda, R>), dim3(static_cast<unsigned>(nlin + nx * F)), dim3(256), 0, s,
                           poses, disps, intr, targets, weights, blob + p->o_edge_ii, blob + p->o_edge_jj, HW, wd, E,
                           Hs, vs, Eall + (int64_t)P * 6 * HW, zl, Hpart, nlin, eta, eta_rows,
                           blob + p->o_frame_eptr, blob + p->o_frame_edges, blob + p->o_frame_pose,
                           blob + p->o_frame_order, blob + p->o_frame_kx, Q, wv, Eall, disps_sens, kSensAlpha, nx);
        HSL2_LAUNCH_CHECK();
      } else if (schur) {
        hipLaunchKernelGGL((ba_linearize<true, kSemCuda, R>), dim3(E, zl), dim3(kLinThreads), 0, s, poses, disps,
                           intr, 0, targets, weights, blob + p->o_edge_ii, blob + p->o_edge_jj, HW, wd, E,
                           Hs, vs, Eall + (int64_t)P * 6 * HW, nullptr, nullptr, nullptr, zl, Hpart);
        HSL2_LAUNCH_CHECK();
        hipLaunchKernelGGL((ba_frame_sums<kSemCuda, R>), dim3(nx, F), dim3(256), 0, s, poses, disps, intr, targets,
                           weights, blob + p->o_edge_ii, blob + p->o_edge_jj, HW, wd, eta, eta_rows,
                           blob + p->o_frame_eptr, blob + p->o_frame_edges, blob + p->o_frame_pose,
                           blob + p->o_frame_order, blob + p->o_frame_kx, Q, wv, Eall, disps_sens, kSensAlpha);
        HSL2_LAUNCH_CHECK();
      } else {
        hipLaunchKernelGGL((ba_linearize<false, kSemCuda, R>), dim3(E, zl), dim3(kLinThreads), 0, s, poses, disps,
                           intr, 0, targets, weights, blob + p->o_edge_ii, blob + p->o_edge_jj, HW, wd, E,
                           Hs, vs, nullptr, nullptr, nullptr, nullptr, zl, Hpart);
        HSL2_LAUNCH_CHECK();
        if (zl > 1) {
          hipLaunchKernelGGL(lin_reduce<R>, dim3(static_cast<unsigned>((E * 156 + 255) / 256)), dim3(256), 0, s,
                             Hpart, zl, E, Hs, vs);
          HSL2_LAUNCH_CHECK();
        }
      }
    }
    if (schur) {
      // the Gram (with the slices' lin_reduce alongside when the linearisation was sliced)
      launch_schur_gram(p, blob, Eall, Q, wv, S, Ev, Spart, s, E > 0 && zl > 1 ? Hpart : nullptr, Hs, vs);
      HSL2_LAUNCH_CHECK();
    }
    // the reduced system and its rhs
    const int nA = static_cast<int>((n_pad * n_pad + 255) / 256);
    const int nR = static_cast<int>((n_pad / 6 + 4) / 4);
    hipLaunchKernelGGL(ba_assemble_sys<R>, dim3(static_cast<unsigned>(nA + nR)), dim3(256), 0, s, Hs, S,
                       blob + p->o_blk_ptr, blob + p->o_blk_src, vs, Ev, blob + p->o_rhs_ptr, blob + p->o_rhs_src, E,
                       P, n_pad, lm, ep, A, bvec, info, nA);
    HSL2_LAUNCH_CHECK();
    {
      const int st = cholesky_solve<R>(A, n_pad, Ld, bvec, xvec, info, s);
      if (st) return st;
    }
    // depth back-substitution and pose retraction in one launch
    const int nbs = schur ? nx * F : 0;
    hipLaunchKernelGGL(ba_update<R>, dim3(static_cast<unsigned>(nbs + (P + 255) / 256)), dim3(256), 0, s, Eall, Q,
                       wv, xvec, info, blob + p->o_slot_ptr, blob + p->o_slot_row, blob + p->o_slot_pose,
                       blob + p->o_frame_kx, HW, disps, last ? dz : nullptr, nx, nbs, poses, p->t0, P,
                       last ? dx : nullptr);
    HSL2_LAUNCH_CHECK();
  }
  return HSL2_OK;
}

// ---- mono-prior step pieces: proj_trans, bi_inter ------------------------------------------
// Depth-only terms of the edges summed per ii frame (projective_transform2_kernel + accum_cuda,
// src/droid_kernels.cu:246-362, 1781-1817): for pixel k of frame m (its edges frame_edges
// [frame_ptr[m], frame_ptr[m+1]), all with ii = frame m's index),
//   C[m][k] = sum_e sum_r w_r Jz_r^2,   w[m][k] = sum_e sum_r w_r r_r Jz_r
// with the BA linearisation's pixel geometry (weights x 0.001, no contribution when the
// transformed depth is below 0.25).  One block per (256 pixels, frame); the frame's edges are
// taken 64 at a time, their relative poses formed once per block in LDS; sums in FP64.
constexpr int kPtEdges = 64;

__global__ void __launch_bounds__(256)
proj_trans_kernel(const float* __restrict__ poses, const float* __restrict__ disps,
                  const float* __restrict__ intr, const float* __restrict__ targets,
                  const float* __restrict__ weights, const int64_t* __restrict__ ii,
                  const int64_t* __restrict__ jj, const int* __restrict__ frame_ptr,
                  const int* __restrict__ frame_edges, int ht, int wd, float* __restrict__ C,
                  float* __restrict__ w) {
  __shared__ EdgeGeom geo[kPtEdges];
  __shared__ int eid[kPtEdges];
  const int HW = ht * wd;
  const int m = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int e0 = frame_ptr[m], e1 = frame_ptr[m + 1];
  if (e1 <= e0) {                                 // a frame always has an edge; kept for safety
    if (k < HW) C[(int64_t)m * HW + k] = w[(int64_t)m * HW + k] = 0.f;
    return;
  }
  const int64_t ix = ii[frame_edges[e0]];
  const float u = static_cast<float>(k % wd), v = static_cast<float>(k / wd);
  const float d = k < HW ? disps[ix * HW + k] : 0.f;
  double c = 0.0, b = 0.0;
  for (int base = e0; base < e1; base += kPtEdges) {
    const int cnt = min(kPtEdges, e1 - base);
    __syncthreads();                              // the previous chunk's geometry is consumed
    if ((int)threadIdx.x < cnt) {
      const int e = frame_edges[base + threadIdx.x];
      eid[threadIdx.x] = e;
      load_edge_geom(poses, intr, 0, static_cast<int>(ii[e]), static_cast<int>(jj[e]), geo[threadIdx.x]);
    }
    __syncthreads();
    if (k < HW) {
      for (int q = 0; q < cnt; ++q) {
        const int64_t e = eid[q];
        const float* tg = targets + e * 2 * HW;
        const float* wt = weights + e * 2 * HW;
        RowJac ru, rv;
        if (pixel_rows<kSemCuda>(geo[q], u, v, d, tg[k], tg[HW + k], wt[k], wt[HW + k], ru, rv)) {
          c += static_cast<double>(ru.w * ru.Jz * ru.Jz) + static_cast<double>(rv.w * rv.Jz * rv.Jz);
          b += static_cast<double>(ru.w * ru.r * ru.Jz) + static_cast<double>(rv.w * rv.r * rv.Jz);
        }
      }
    }
  }
  if (k < HW) {
    C[(int64_t)m * HW + k] = static_cast<float>(c);
    w[(int64_t)m * HW + k] = static_cast<float>(b);
  }
}

// Bilinear interpolation of per-frame scale grids and its Jacobian (bi_inter_kernel,
// src/droid_kernels.cu:1275-1315): pixel k of frame m samples scales[m] at column
// grids[m][k][0], row grids[m][k][1]; weights wa (y0, x0), wb (y1, x0), wc (y0, x1), wd (y1, x1)
// formed as the reference forms them.  The pixel's Jacobian row (hs * ws values) is written
// whole: zeros except the four taps.  Taps outside the grid contribute nothing and are not
// written (the reference indexes outside its tensors there).
__global__ void __launch_bounds__(256)
bi_inter_kernel(const float* __restrict__ scales, const float* __restrict__ grids, int HW, int hs,
                int ws, float* __restrict__ out, float* __restrict__ jac) {
  const int m = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= HW) return;
  const int64_t pix = (int64_t)m * HW + k;
  const float xind = grids[pix * 2], yind = grids[pix * 2 + 1];
  const int x0 = static_cast<int>(floorf(xind)), y0 = static_cast<int>(floorf(yind));
  const int x1 = x0 + 1, y1 = y0 + 1;
  const float wa = (x1 - xind) * (y1 - yind);
  const float wb = (x1 - xind) * (yind - y0);
  const float wc = (xind - x0) * (y1 - yind);
  const float wdd = (xind - x0) * (yind - y0);
  const float* sc = scales + (int64_t)m * hs * ws;
  float* jr = jac + pix * hs * ws;
  for (int q = 0; q < hs * ws; ++q) jr[q] = 0.f;
  const int ty[4] = {y0, y1, y0, y1}, tx[4] = {x0, x0, x1, x1};
  const float tw[4] = {wa, wb, wc, wdd};
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (ty[j] >= 0 && ty[j] < hs && tx[j] >= 0 && tx[j] < ws) {
      acc += tw[j] * sc[ty[j] * ws + tx[j]];
      jr[ty[j] * ws + tx[j]] = tw[j];
    }
  }
  out[pix] = acc;
}

// One JDSA step's reduced system, per slot (hislam2/geom/ba.py:171-219 with
// schur_solve_mono_prior, chol.py:80-108), for scale grids of D = hs * ws <= 4 parameters.
// The M*D reduced system is block-diagonal: slot s holds the D scale parameters of the frame
// src[s] (the rank m with kx[m] - kx[0] == s; -1: none, all terms zero) and couples them with
// the pixels of rank s.  Per pixel k, with the prior p of frame src[s], its mask q = p > 0, the
// bilinear weights J (bi_inter on the linspace grid gx, gy) and rd = disp - p * (J . scales):
//   Jso = -q p J,  H = a sum Jso Jso^T,  E_k = a Jso,  v = -a sum Jso rd
//   C' = C + q_s a + (1 - q_s) eta,  w' = w - q_s a rd_s  (rank s),  Q = 1 / C'
//   S = H + (ep + lm H) I - sum E_k Q E_k^T,  rhs = v - sum E_k Q w'.
// jdsa_reduce: one block per slot sums these in FP64 and factors S (Cholesky, D x D); the
// factor, dso and a failure flag go to the workspace.  jdsa_apply: one block per (256 pixels,
// slot): dz = Q (w' - E_k . dso), dzcov = Q^2 |L^-1 E_k|^2 + Q, disps_out[kx[s]] = disps + dz
// and dscales_out[kx[s]] = dscales + dso (the inputs are read unmodified, so slots never see
// each other's updates); if any slot's factorisation failed every dso is taken as zero and
// dzcov = Q.
constexpr int kJdsaMaxD = 4;
struct JdsaSlot {
  double L[kJdsaMaxD][kJdsaMaxD];
  double dso[kJdsaMaxD];
  int fail;
};

struct JdsaPixel {
  float jso[kJdsaMaxD];       // Jso of the slot's source frame at this pixel
  float rd;                   // rd of the source frame
};

__device__ __forceinline__ void jdsa_pixel(const float* __restrict__ prior_f, const float* __restrict__ disp_f,
                                           const float* __restrict__ scales_f, const float* __restrict__ gx,
                                           const float* __restrict__ gy, int k, int wd, int hs, int ws,
                                           JdsaPixel& px) {
  const int y = k / wd, x = k - (k / wd) * wd;
  const float xind = gx[x], yind = gy[y];
  const int x0 = static_cast<int>(floorf(xind)), y0 = static_cast<int>(floorf(yind));
  const int x1 = x0 + 1, y1 = y0 + 1;
  const float wts[4] = {(x1 - xind) * (y1 - yind), (x1 - xind) * (yind - y0), (xind - x0) * (y1 - yind),
                        (xind - x0) * (yind - y0)};
  const int ty[4] = {y0, y1, y0, y1}, tx[4] = {x0, x0, x1, x1};
  float J[kJdsaMaxD] = {0.f, 0.f, 0.f, 0.f};
  float sc = 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (ty[j] >= 0 && ty[j] < hs && tx[j] >= 0 && tx[j] < ws) {
      const int d = ty[j] * ws + tx[j];
      sc += wts[j] * scales_f[d];
#pragma unroll
      for (int q = 0; q < kJdsaMaxD; ++q)
        if (q == d) J[q] = wts[j];
    }
  }
  const float p = prior_f[k];
  const float msk = p > 0.f ? 1.f : 0.f;
  px.rd = disp_f[k] - p * sc;
#pragma unroll
  for (int q = 0; q < kJdsaMaxD; ++q) px.jso[q] = -msk * p * J[q];
}

__global__ void __launch_bounds__(256)
jdsa_reduce(const float* __restrict__ C, const float* __restrict__ w, const float* __restrict__ disps,
            const float* __restrict__ prior, const float* __restrict__ dscales, const float* __restrict__ eta,
            const int64_t* __restrict__ kx, const int* __restrict__ src, const float* __restrict__ gx,
            const float* __restrict__ gy, int HW, int wd, int hs, int ws, float alpha, float ep, float lm,
            JdsaSlot* __restrict__ slots) {
  // sums: H (10, upper triangle of 4 x 4), EQE (10), v (4), EQw (4)
  constexpr int NV = 28;
  __shared__ double lds[16 * 256];
  __shared__ double red[NV];
  const int s = blockIdx.x;
  const int D = hs * ws;
  const int m = src[s];
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0.0;
  if (m >= 0) {
    const int64_t fm = kx[m], fs = kx[s];
    const double a = alpha;
    for (int k = threadIdx.x; k < HW; k += blockDim.x) {
      JdsaPixel px;
      jdsa_pixel(prior + fm * HW, disps + fm * HW, dscales + fm * D, gx, gy, k, wd, hs, ws, px);
      // rank s's pixel: C', w' (its own prior mask and rd)
      JdsaPixel ps;
      jdsa_pixel(prior + fs * HW, disps + fs * HW, dscales + fs * D, gx, gy, k, wd, hs, ws, ps);
      const double qs = prior[fs * HW + k] > 0.f ? 1.0 : 0.0;
      const double Cp = static_cast<double>(C[(int64_t)s * HW + k]) + qs * a +
                        (1.0 - qs) * static_cast<double>(eta[(int64_t)s * HW + k]);
      const double wp = static_cast<double>(w[(int64_t)s * HW + k]) - qs * a * static_cast<double>(ps.rd);
      const double Q = 1.0 / Cp;
      int t = 0;
#pragma unroll
      for (int i = 0; i < kJdsaMaxD; ++i)
#pragma unroll
        for (int j = i; j < kJdsaMaxD; ++j, ++t) {
          const double jj2 = static_cast<double>(px.jso[i]) * px.jso[j];
          acc[t] += a * jj2;
          acc[10 + t] += a * a * Q * jj2;
        }
#pragma unroll
      for (int i = 0; i < kJdsaMaxD; ++i) {
        acc[20 + i] += -a * static_cast<double>(px.jso[i]) * px.rd;
        acc[24 + i] += a * static_cast<double>(px.jso[i]) * Q * wp;
      }
    }
  }
  block_reduce_store<NV, 256, double>(acc, lds, red);
  __syncthreads();
  if (threadIdx.x == 0) {
    double S[kJdsaMaxD][kJdsaMaxD], b[kJdsaMaxD];
    int t = 0;
    for (int i = 0; i < kJdsaMaxD; ++i)
      for (int j = i; j < kJdsaMaxD; ++j, ++t) {
        const double h = red[t];
        double v = h - red[10 + t];
        if (i == j) v += ep + lm * h;
        S[i][j] = S[j][i] = v;
      }
    for (int i = 0; i < kJdsaMaxD; ++i) b[i] = red[20 + i] - red[24 + i];
    // Cholesky of the leading D x D block (padding rows of a smaller grid are unused).  The loops
    // run to the constant kJdsaMaxD with D as a guard, so every array index is a constant after
    // unrolling and the arrays stay in registers (a loop bound of D would index them through
    // scratch memory); the operations and their order are those of the plain D loops.
    double L[kJdsaMaxD][kJdsaMaxD] = {};
    int fail = 0;
#pragma unroll
    for (int j = 0; j < kJdsaMaxD; ++j) {
      if (j < D && !fail) {
        double d = S[j][j];
#pragma unroll
        for (int q = 0; q < j; ++q) d -= L[j][q] * L[j][q];
        if (!(d > 0.0)) {
          fail = 1;
        } else {
          L[j][j] = sqrt(d);
#pragma unroll
          for (int i = j + 1; i < kJdsaMaxD; ++i) {
            if (i < D) {
              double v = S[i][j];
#pragma unroll
              for (int q = 0; q < j; ++q) v -= L[i][q] * L[j][q];
              L[i][j] = v / L[j][j];
            }
          }
        }
      }
    }
    double x[kJdsaMaxD] = {0.0, 0.0, 0.0, 0.0};
    if (!fail) {
      double y[kJdsaMaxD] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < kJdsaMaxD; ++i) {
        if (i < D) {
          double v = b[i];
#pragma unroll
          for (int q = 0; q < i; ++q) v -= L[i][q] * y[q];
          y[i] = v / L[i][i];
        }
      }
#pragma unroll
      for (int i = kJdsaMaxD - 1; i >= 0; --i) {
        if (i < D) {
          double v = y[i];
#pragma unroll
          for (int q = i + 1; q < kJdsaMaxD; ++q)
            if (q < D) v -= L[q][i] * x[q];
          x[i] = v / L[i][i];
        }
      }
    }
    JdsaSlot& out = slots[s];
    for (int i = 0; i < kJdsaMaxD; ++i) {
      out.dso[i] = x[i];
      for (int j = 0; j < kJdsaMaxD; ++j) out.L[i][j] = L[i][j];
    }
    out.fail = fail;
  }
}

__global__ void __launch_bounds__(256)
jdsa_apply(const float* __restrict__ C, const float* __restrict__ w, const float* __restrict__ disps,
           const float* __restrict__ prior, const float* __restrict__ dscales, const float* __restrict__ eta,
           const int64_t* __restrict__ kx, const int* __restrict__ src, const float* __restrict__ gx,
           const float* __restrict__ gy, const JdsaSlot* __restrict__ slots, int M, int HW, int wd, int hs,
           int ws, float alpha, float* __restrict__ disps_out, float* __restrict__ dscales_out,
           float* __restrict__ dzcov) {
  const int s = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int D = hs * ws;
  __shared__ int any_fail;
  if (threadIdx.x == 0) any_fail = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < M; i += blockDim.x)
    if (slots[i].fail) any_fail = 1;             // every writer stores the same value
  __syncthreads();
  const bool ok = any_fail == 0;
  const int m = src[s];
  const JdsaSlot& sl = slots[s];
  const int64_t fs = kx[s];
  if (k == 0 && ok)                               // dscales[kx[s]] += dso[s]
    for (int i = 0; i < D; ++i) dscales_out[fs * D + i] = dscales[fs * D + i] + static_cast<float>(sl.dso[i]);
  if (k >= HW) return;
  const double a = alpha;
  JdsaPixel ps;
  jdsa_pixel(prior + fs * HW, disps + fs * HW, dscales + fs * D, gx, gy, k, wd, hs, ws, ps);
  const double qs = prior[fs * HW + k] > 0.f ? 1.0 : 0.0;
  const double Cp = static_cast<double>(C[(int64_t)s * HW + k]) + qs * a +
                    (1.0 - qs) * static_cast<double>(eta[(int64_t)s * HW + k]);
  const double wp = static_cast<double>(w[(int64_t)s * HW + k]) - qs * a * static_cast<double>(ps.rd);
  const double Q = 1.0 / Cp;
  double dz = Q * wp, cov = Q;
  if (m >= 0 && ok) {
    const int64_t fm = kx[m];
    JdsaPixel px;
    jdsa_pixel(prior + fm * HW, disps + fm * HW, dscales + fm * D, gx, gy, k, wd, hs, ws, px);
    double e[kJdsaMaxD], z[kJdsaMaxD];
    double ed = 0.0;
    for (int i = 0; i < kJdsaMaxD; ++i) {
      e[i] = a * px.jso[i] * Q;                   // (E Q)_k
      ed += a * px.jso[i] * sl.dso[i];
    }
    dz = Q * (wp - ed);
    double f2 = 0.0;
    for (int i = 0; i < D; ++i) {                 // z = L^-1 (E Q)_k
      double v = e[i];
      for (int q = 0; q < i; ++q) v -= sl.L[i][q] * z[q];
      z[i] = v / sl.L[i][i];
      f2 += z[i] * z[i];
    }
    cov = f2 + Q;
  }
  dzcov[(int64_t)s * HW + k] = static_cast<float>(cov);
  disps_out[fs * HW + k] = disps[fs * HW + k] + static_cast<float>(dz);
}

}  // namespace
}  // namespace hsl2

using namespace hsl2;

extern "C" int hsl2_ba_run_sens(const hsl2_ba_plan* plan, float* poses, float* disps,
                                const float* intrinsics, const float* disps_sens, const float* targets,
                                const float* weights, const float* eta, int64_t eta_rows, int iterations,
                                float lm, float ep, float* dx, float* dz, void* workspace,
                                size_t workspace_bytes, int* solve_ok, void* stream) {
  HSL2_REQUIRE(plan, HSL2_ERR_ARGUMENT, "ba: null plan");
  HSL2_REQUIRE(poses && disps && intrinsics && dx, HSL2_ERR_ARGUMENT, "ba: null pointer");
  HSL2_REQUIRE(plan->E == 0 || (targets && weights), HSL2_ERR_ARGUMENT, "ba: null targets/weights");
  HSL2_REQUIRE(iterations >= 0, HSL2_ERR_ARGUMENT, "ba: negative iteration count");
  HSL2_REQUIRE(workspace_bytes >= plan->total, HSL2_ERR_WORKSPACE, "ba: workspace too small");
  HSL2_REQUIRE(workspace && (reinterpret_cast<uintptr_t>(workspace) % 256) == 0,
               HSL2_ERR_ARGUMENT, "ba: workspace must be 256-byte aligned");
  if (!plan->motion_only) {
    HSL2_REQUIRE(eta, HSL2_ERR_ARGUMENT, "ba: null eta");
    HSL2_REQUIRE(eta_rows == plan->F || eta_rows == 1, HSL2_ERR_SHAPE,
                 "ba: eta must have one row per unique depth frame (or a single row)");
  }
  hipStream_t s = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  // the plan's tables: the registered device copy when it lives on this device, else uploaded
  // into the workspace
  const int* blob = reinterpret_cast<const int*>(ws + plan->w_blob);
  int dev = -1;
  if (plan->dev_blob) HSL2_HIP(hipGetDevice(&dev));
  if (plan->dev_blob && dev == plan->dev_blob_device)
    blob = plan->dev_blob;
  else
    HSL2_HIP(hipMemcpyAsync(ws + plan->w_blob, plan->blob.data(), plan->blob.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice, s));
  int st;
  if (plan->precision == HSL2_BA_SPLIT)
    st = run_iterations<double>(plan, poses, disps, intrinsics, disps_sens, targets, weights, eta,
                                static_cast<int>(eta_rows), iterations, lm, ep, dx, dz, ws, blob, s);
  else
    st = run_iterations<float>(plan, poses, disps, intrinsics, disps_sens, targets, weights, eta,
                               static_cast<int>(eta_rows), iterations, lm, ep, dx, dz, ws, blob, s);
  if (st) return st;
  if (iterations == 0) HSL2_HIP(hipMemsetAsync(dx, 0, sizeof(float) * 6 * plan->P, s));
  if (solve_ok) {
    int flags[2] = {0, 0};
    HSL2_HIP(hipMemcpyAsync(flags, ws + plan->w_info, sizeof(flags), hipMemcpyDeviceToHost, s));
    HSL2_HIP(hipStreamSynchronize(s));
    *solve_ok = flags[1] == 0 ? 1 : 0;
  }
  return HSL2_OK;
}

// ---- hsl2_ba_run_static: the same run with a library-owned static workspace ------------------
// For callers that cannot allocate device memory.  The workspace is a file-scope device array:
// each GPU the library runs on holds its own copy (HSL2_BA_STATIC_SCRATCH_BYTES) until the
// process ends.  Host code never passes the array's name to a kernel or a copy (on the host it
// is a stand-in, not a device address); the device address comes from hipGetSymbolAddress,
// taken once per device with that device current and kept in scratch_ptr[].  Every call on one
// GPU shares the array, so calls that overlap on one GPU race.  The kernels need no zeroed
// workspace: ba_clear_flags resets the solver flags and the Cholesky resets its own counters,
// every other buffer is written before it is read within an iteration.
namespace hsl2 {
namespace {
static __device__ __align__(256) unsigned char scratch[HSL2_BA_STATIC_SCRATCH_BYTES];

constexpr int kMaxStaticDevices = 64;
std::mutex scratch_mu;
void* scratch_ptr[kMaxStaticDevices] = {};

int static_scratch(void** out) {
  int dev = -1;
  HSL2_HIP(hipGetDevice(&dev));
  HSL2_REQUIRE(dev >= 0 && dev < kMaxStaticDevices, HSL2_ERR_UNSUPPORTED,
               "ba_run_static: device index beyond the static workspace table");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    HSL2_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    HSL2_REQUIRE(p && (reinterpret_cast<uintptr_t>(p) % 256) == 0, HSL2_ERR_HIP,
                 "ba_run_static: static workspace has no aligned device address");
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return HSL2_OK;
}
}  // namespace
}  // namespace hsl2

extern "C" int hsl2_ba_run_static(const hsl2_ba_plan* plan, float* poses, float* disps,
                                  const float* intrinsics, const float* targets, const float* weights,
                                  const float* eta, int64_t eta_rows, int iterations, float lm, float ep,
                                  float* dx, float* dz, int* solve_ok, void* stream) {
  HSL2_REQUIRE(plan, HSL2_ERR_ARGUMENT, "ba_run_static: null plan");
  HSL2_REQUIRE(plan->total <= HSL2_BA_STATIC_SCRATCH_BYTES, HSL2_ERR_WORKSPACE,
               "ba_run_static: the plan needs " + std::to_string(plan->total) +
                   " bytes of workspace, more than the static " +
                   std::to_string(static_cast<size_t>(HSL2_BA_STATIC_SCRATCH_BYTES)) + " bytes");
  void* ws = nullptr;
  const int st = static_scratch(&ws);
  if (st) return st;
  return hsl2_ba_run(plan, poses, disps, intrinsics, targets, weights, eta, eta_rows, iterations, lm, ep, dx,
                     dz, ws, HSL2_BA_STATIC_SCRATCH_BYTES, solve_ok, stream);
}

extern "C" int hsl2_ba_run(const hsl2_ba_plan* plan, float* poses, float* disps,
                           const float* intrinsics, const float* targets, const float* weights,
                           const float* eta, int64_t eta_rows, int iterations, float lm, float ep,
                           float* dx, float* dz, void* workspace, size_t workspace_bytes,
                           int* solve_ok, void* stream) {
  return hsl2_ba_run_sens(plan, poses, disps, intrinsics, nullptr, targets, weights, eta, eta_rows, iterations, lm,
                          ep, dx, dz, workspace, workspace_bytes, solve_ok, stream);
}

extern "C" int hsl2_ba_debug_views(const hsl2_ba_plan* plan, void* workspace, double** Hs,
                                   double** vs, void** reduced_matrix, void** reduced_rhs) {
  HSL2_REQUIRE(plan && workspace, HSL2_ERR_ARGUMENT, "ba: null pointer");
  char* ws = static_cast<char*>(workspace);
  if (Hs) *Hs = reinterpret_cast<double*>(ws + plan->w_Hs);
  if (vs) *vs = reinterpret_cast<double*>(ws + plan->w_vs);
  if (reduced_matrix) *reduced_matrix = ws + plan->w_A;
  if (reduced_rhs) *reduced_rhs = ws + plan->w_b;
  return HSL2_OK;
}

// ---- mono-prior step pieces (src/droid.cpp:121-138, 244-250) --------------------------------

extern "C" int hsl2_edge_frames(const int64_t* ii, int64_t num_edges, int64_t* kx, int32_t* frame_ptr,
                                int32_t* frame_edges, int64_t* num_frames) {
  HSL2_REQUIRE(num_edges >= 0 && num_frames && (num_edges == 0 || (ii && kx && frame_ptr && frame_edges)),
               HSL2_ERR_ARGUMENT, "edge frames: null pointer");
  std::vector<int64_t> u(ii, ii + num_edges);
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  const int64_t M = static_cast<int64_t>(u.size());
  std::vector<int32_t> cnt(M + 1, 0);
  std::vector<int32_t> rank(num_edges);
  for (int64_t e = 0; e < num_edges; ++e) {
    rank[e] = static_cast<int32_t>(std::lower_bound(u.begin(), u.end(), ii[e]) - u.begin());
    ++cnt[rank[e] + 1];
  }
  for (int64_t m = 0; m < M; ++m) cnt[m + 1] += cnt[m];
  for (int64_t m = 0; m <= M; ++m) frame_ptr[m] = cnt[m];
  std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
  for (int64_t e = 0; e < num_edges; ++e) frame_edges[fill[rank[e]]++] = static_cast<int32_t>(e);
  for (int64_t m = 0; m < M; ++m) kx[m] = u[m];
  *num_frames = M;
  return HSL2_OK;
}

extern "C" int hsl2_proj_trans(const float* poses, const float* disps, const float* intrinsics,
                               const float* targets, const float* weights, const int64_t* ii,
                               const int64_t* jj, const int32_t* frame_ptr, const int32_t* frame_edges,
                               int64_t num_frames, int64_t ht, int64_t wd, float* C, float* w, void* stream) {
  using namespace hsl2;
  HSL2_REQUIRE(num_frames >= 0 && ht > 0 && wd > 0 && ht * wd < (1ll << 31), HSL2_ERR_SHAPE,
               "proj_trans: bad sizes");
  if (num_frames == 0) return HSL2_OK;
  HSL2_REQUIRE(poses && disps && intrinsics && targets && weights && ii && jj && frame_ptr && frame_edges && C && w,
               HSL2_ERR_ARGUMENT, "proj_trans: null pointer");
  HSL2_REQUIRE(num_frames < 65536, HSL2_ERR_SHAPE, "proj_trans: too many frames");
  const int HW = static_cast<int>(ht * wd);
  hipLaunchKernelGGL(proj_trans_kernel, dim3((HW + 255) / 256, static_cast<unsigned>(num_frames)), dim3(256), 0,
                     as_stream(stream), poses, disps, intrinsics, targets, weights, ii, jj, frame_ptr, frame_edges,
                     static_cast<int>(ht), static_cast<int>(wd), C, w);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" int hsl2_bi_inter(const float* scales, const float* grids, int64_t num, int64_t ht, int64_t wd,
                             int64_t hs, int64_t ws, float* outputs, float* jacobis, void* stream) {
  using namespace hsl2;
  HSL2_REQUIRE(num >= 0 && ht > 0 && wd > 0 && hs > 0 && ws > 0 && ht * wd < (1ll << 31) && hs * ws < (1 << 20),
               HSL2_ERR_SHAPE, "bi_inter: bad sizes");
  if (num == 0) return HSL2_OK;
  HSL2_REQUIRE(scales && grids && outputs && jacobis, HSL2_ERR_ARGUMENT, "bi_inter: null pointer");
  HSL2_REQUIRE(num < 65536, HSL2_ERR_SHAPE, "bi_inter: too many frames");
  const int HW = static_cast<int>(ht * wd);
  hipLaunchKernelGGL(bi_inter_kernel, dim3((HW + 255) / 256, static_cast<unsigned>(num)), dim3(256), 0,
                     as_stream(stream), scales, grids, HW, static_cast<int>(hs), static_cast<int>(ws), outputs,
                     jacobis);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" size_t hsl2_jdsa_workspace_bytes(int64_t num_frames) {
  return static_cast<size_t>(num_frames > 0 ? num_frames : 0) * sizeof(hsl2::JdsaSlot);
}

extern "C" int hsl2_jdsa_solve(const float* C, const float* w, const float* disps, const float* prior,
                               const float* dscales, const float* eta, const int64_t* kx, const int32_t* slot_src,
                               const float* gx, const float* gy, int64_t num_frames, int64_t ht, int64_t wd,
                               int64_t hs, int64_t ws, float alpha, float ep, float lm, void* workspace,
                               size_t workspace_bytes, float* disps_out, float* dscales_out, float* dzcov,
                               void* stream) {
  using namespace hsl2;
  HSL2_REQUIRE(num_frames >= 0 && ht > 0 && wd > 0 && ht * wd < (1ll << 31), HSL2_ERR_SHAPE,
               "jdsa: bad sizes");
  HSL2_REQUIRE(hs > 0 && ws > 0 && hs * ws <= kJdsaMaxD, HSL2_ERR_UNSUPPORTED,
               "jdsa: scale grids of at most 4 parameters");
  if (num_frames == 0) return HSL2_OK;
  HSL2_REQUIRE(C && w && disps && prior && dscales && eta && kx && slot_src && gx && gy && disps_out && dscales_out &&
                   dzcov && workspace,
               HSL2_ERR_ARGUMENT, "jdsa: null pointer");
  HSL2_REQUIRE(workspace_bytes >= hsl2_jdsa_workspace_bytes(num_frames), HSL2_ERR_WORKSPACE,
               "jdsa: workspace too small");
  HSL2_REQUIRE(num_frames < 65536, HSL2_ERR_SHAPE, "jdsa: too many frames");
  const int HW = static_cast<int>(ht * wd);
  hipStream_t s = as_stream(stream);
  JdsaSlot* slots = static_cast<JdsaSlot*>(workspace);
  hipLaunchKernelGGL(jdsa_reduce, dim3(static_cast<unsigned>(num_frames)), dim3(256), 0, s, C, w, disps, prior,
                     dscales, eta, kx, slot_src, gx, gy, HW, static_cast<int>(wd), static_cast<int>(hs),
                     static_cast<int>(ws), alpha, ep, lm, slots);
  HSL2_LAUNCH_CHECK();
  hipLaunchKernelGGL(jdsa_apply, dim3((HW + 255) / 256, static_cast<unsigned>(num_frames)), dim3(256), 0, s, C, w,
                     disps, prior, dscales, eta, kx, slot_src, gx, gy, slots, static_cast<int>(num_frames), HW,
                     static_cast<int>(wd), static_cast<int>(hs), static_cast<int>(ws), alpha, disps_out, dscales_out,
                     dzcov);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" size_t hsl2_ba_plan_table_bytes(const hsl2_ba_plan* plan) {
  return plan ? plan->blob.size() * sizeof(int32_t) : 0;
}

extern "C" int hsl2_ba_plan_upload_tables(hsl2_ba_plan* plan, void* device_buffer, void* stream) {
  HSL2_REQUIRE(plan && device_buffer, HSL2_ERR_ARGUMENT, "ba_plan_upload_tables: null pointer");
  HSL2_REQUIRE(reinterpret_cast<uintptr_t>(device_buffer) % 256 == 0, HSL2_ERR_ARGUMENT,
               "ba_plan_upload_tables: the buffer must be 256-byte aligned");
  int dev = -1;
  HSL2_HIP(hipGetDevice(&dev));
  hipStream_t s = as_stream(stream);
  HSL2_HIP(hipMemcpyAsync(device_buffer, plan->blob.data(), plan->blob.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, s));
  plan->dev_blob = static_cast<const int32_t*>(device_buffer);
  plan->dev_blob_device = dev;
  return HSL2_OK;
}
