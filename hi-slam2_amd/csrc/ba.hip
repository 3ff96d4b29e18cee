// Dense bundle adjustment on gfx950: the iteration loop of ba_cuda
// (src/droid_kernels.cu:1820-1937) with every step on the device.
//
// Per Gauss-Newton iteration:
//   1. ba_linearize      one block per edge: relative pose, per-pixel residuals and
//                        Jacobians, the 12x12 Hessian and gradient of the edge
//                        reduced through LDS, and the per-pixel Schur terms
//                        (src/droid_kernels.cu:364-606).
//   2. ba_frame_reduce   per depth frame: C, w, Q = 1/(C+eta) and the pose-i block
//                        Ei of the frame's edges (accum_cuda, 1413-1463, 1901-1906).
//   3. ba_schur_gram_lds per depth frame, 16x16 tiles of the slots' Gram matrix
//                        S = sum_px E_a Q E_b^T on the matrix cores, one block per group
//                        pair of row tiles staged in LDS by 64-pixel chunks (FP32 tile
//                        sums per 256 pixels, FP64 totals); the blocks of diagonal
//                        groups also form Ev = sum_px Q w E (EEt6x6 / Ev6x1, 1466-1560)
//                        (ba_schur_ev, one wave per slot, when motion-only).
//   4. ba_assemble       the damped reduced camera system A - S, b - bS in FP64
//                        (FP32 in the plain mode), written densely in column-major
//                        order, padded to a multiple of 64 with an identity block
//                        (SparseBlock, 1584-1686).
//   5. cholesky_solve    tiled Cholesky factorisation and solve (chol.hip); a non-positive
//                        pivot gives dx = 0 (1669-1677).
//   6. ba_backsub        dz = Q (w - sum E^T dx) with the reference's t0 skip, and
//                        the disparity update (EvT6x1, disp_retr, 1562-1582, 1398-1411).
//   7. ba_pose_retr      T <- Exp(dx) T (1342-1396).
#include "ba_kernels.hpp"

namespace hsl2 {
namespace {

template <typename R>
int run_iterations(const hsl2_ba_plan* p, float* poses, float* disps, const float* intr,
                   const float* targets, const float* weights, const float* eta, int eta_rows,
                   int iterations, float lm, float ep, float* dx, float* dz, char* ws,
                   hipStream_t s) {
  const int* blob = reinterpret_cast<const int*>(ws + p->w_blob);
  float* Eall = reinterpret_cast<float*>(ws + p->w_Eall);
  float* Eii = reinterpret_cast<float*>(ws + p->w_Eii);
  float* Cii = reinterpret_cast<float*>(ws + p->w_Cii);
  float* bz = reinterpret_cast<float*>(ws + p->w_bz);
  float* Q = reinterpret_cast<float*>(ws + p->w_Q);
  float* wv = reinterpret_cast<float*>(ws + p->w_w);
  double* Hs = reinterpret_cast<double*>(ws + p->w_Hs);
  double* vs = reinterpret_cast<double*>(ws + p->w_vs);
  double* S = reinterpret_cast<double*>(ws + p->w_S);
  double* Ev = reinterpret_cast<double*>(ws + p->w_Ev);
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
  for (int it = 0; it < iterations; ++it) {
    const bool last = it == iterations - 1;
    if (E > 0) {
      if (schur)
        hipLaunchKernelGGL((ba_linearize<true, kSemCuda>), dim3(E), dim3(kLinThreads), 0, s, poses, disps,
                           intr, 0, targets, weights, blob + p->o_edge_ii, blob + p->o_edge_jj, HW, wd, E,
                           Hs, vs, Eall + (int64_t)P * 6 * HW, Eii, Cii, bz);
      else
        hipLaunchKernelGGL((ba_linearize<false, kSemCuda>), dim3(E), dim3(kLinThreads), 0, s, poses, disps,
                           intr, 0, targets, weights, blob + p->o_edge_ii, blob + p->o_edge_jj, HW, wd, E,
                           Hs, vs, nullptr, nullptr, nullptr, nullptr);
      HSL2_LAUNCH_CHECK();
    }
    if (schur) {
      launch_frame_reduce(Cii, bz, Eii, eta, eta_rows, blob + p->o_frame_eptr, blob + p->o_frame_edges,
                          blob + p->o_frame_pose, HW, static_cast<int>(F), 0.f, Q, wv, Eall, s);
      HSL2_LAUNCH_CHECK();
      launch_schur_gram(p, blob, Eall, Q, wv, S, Ev, s);
      HSL2_LAUNCH_CHECK();
    }
    const int64_t nn = n_pad * n_pad;
    hipLaunchKernelGGL(ba_assemble<R>, dim3(static_cast<unsigned>((nn + 255) / 256)), dim3(256), 0, s,
                       Hs, S, blob + p->o_blk_ptr, blob + p->o_blk_src, E, P, n_pad, lm, ep, 0,
                       A, info);
    HSL2_LAUNCH_CHECK();
    hipLaunchKernelGGL(ba_assemble_rhs<R>, dim3(static_cast<unsigned>((n_pad / 6 + 4) / 4)), dim3(256), 0, s,
                       vs, Ev, blob + p->o_rhs_ptr, blob + p->o_rhs_src, E, P,
                       static_cast<int>(n_pad), bvec);
    HSL2_LAUNCH_CHECK();
    {
      const int st = cholesky_solve<R>(A, n_pad, Ld, bvec, xvec, info, s);
      if (st) return st;
    }
    if (schur) {
      hipLaunchKernelGGL(ba_backsub<R>, dim3((HW + 255) / 256, F), dim3(256), 0, s, Eall, Q, wv,
                         xvec, info, blob + p->o_slot_ptr, blob + p->o_slot_row,
                         blob + p->o_slot_pose, blob + p->o_frame_kx, HW, 1, disps,
                         last ? dz : nullptr);
      HSL2_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(ba_pose_retr<R>, dim3((P + 63) / 64), dim3(64), 0, s, poses, xvec, info,
                       p->t0, P, 0, last ? dx : nullptr);
    HSL2_LAUNCH_CHECK();
  }
  return HSL2_OK;
}

}  // namespace
}  // namespace hsl2

using namespace hsl2;

extern "C" int hsl2_ba_run(const hsl2_ba_plan* plan, float* poses, float* disps,
                           const float* intrinsics, const float* targets, const float* weights,
                           const float* eta, int64_t eta_rows, int iterations, float lm, float ep,
                           float* dx, float* dz, void* workspace, size_t workspace_bytes,
                           int* solve_ok, void* stream) {
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
  HSL2_HIP(hipMemcpyAsync(ws + plan->w_blob, plan->blob.data(), plan->blob.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, s));
  if (plan->NPAIR == 0 && plan->NS == 0) {
    // motion-only (or empty) plans read S/Ev through the assembly lists only when present
  }
  int st;
  if (plan->precision == HSL2_BA_SPLIT)
    st = run_iterations<double>(plan, poses, disps, intrinsics, targets, weights, eta,
                                static_cast<int>(eta_rows), iterations, lm, ep, dx, dz, ws, s);
  else
    st = run_iterations<float>(plan, poses, disps, intrinsics, targets, weights, eta,
                               static_cast<int>(eta_rows), iterations, lm, ep, dx, dz, ws, s);
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
