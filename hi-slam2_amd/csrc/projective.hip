// The torch formulation's projective transform (hislam2/geom/projective_ops.py:44-74 with
// iproj_pinhole / proj_pinhole of hislam2/geom/pinhole.py:6-53) and its adjoint.
//
// Per edge e and pixel (u, v) of frame ii[e]:
//   X0 = [(u - cx_i) / fx_i, (v - cy_i) / fy_i, 1, d]           intrinsics of frame ii
//   X1 = G X0 = [R X0[:3] + t d, d],  G = T_jj T_ii^-1
//   Z' = (Z < 0.1 ? 1 : Z);  x = fx_j X / Z' + cx_j,  y = fy_j Y / Z' + cy_j  (intrinsics of jj)
//   [optionally a third channel d / Z' (return_depth)]
//   valid = X1.z > 0.2 && X0.z > 0.2
//   Jj = Jp Ja, Ji = -Jj Adj(G), Jz = Jp [t; 1]   (jacobian)
// Output layouts are the reference's: coords [E, h, w, 2 or 3], valid [E, h, w],
// Ji / Jj [E, h, w, 2, 6], Jz [E, h, w, 2] (one batch element per call).
//
// Adjoint: the gradient of coords with respect to the left tangents of T_ii, T_jj and the
// disparity, as autograd forms it through the reference's code: the Z < 0.1 branch is a
// constant, so its derivative with respect to Z is 0 there.  Per-edge sums in FP64, then a
// deterministic per-frame reduction.
#include "common.hpp"

namespace hsl2 {
namespace {

constexpr int kThreads = 256;

struct Geo {
  double R[9], t[3];
  double fxi, fyi, cxi, cyi, fx, fy, cx, cy;
};

__device__ void edge_geo(const float* __restrict__ poses, const float* __restrict__ intr, int ix,
                         int jx, Geo& g) {
  double qi[4], qj[4], pi[3], pj[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pi[k] = poses[ix * 7 + k];
    pj[k] = poses[jx * 7 + k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    qi[k] = poses[ix * 7 + 3 + k];
    qj[k] = poses[jx * 7 + 3 + k];
  }
  // q = q_j * conj(q_i)
  const double x = -qj[3] * qi[0] + qj[0] * qi[3] - qj[1] * qi[2] + qj[2] * qi[1];
  const double y = -qj[3] * qi[1] + qj[1] * qi[3] - qj[2] * qi[0] + qj[0] * qi[2];
  const double z = -qj[3] * qi[2] + qj[2] * qi[3] - qj[0] * qi[1] + qj[1] * qi[0];
  const double w = qj[3] * qi[3] + qj[0] * qi[0] + qj[1] * qi[1] + qj[2] * qi[2];
  double* R = g.R;
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - z * w);     R[2] = 2 * (x * z + y * w);
  R[3] = 2 * (x * y + z * w);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - x * w);
  R[6] = 2 * (x * z - y * w);     R[7] = 2 * (y * z + x * w);     R[8] = 1 - 2 * (x * x + y * y);
#pragma unroll
  for (int k = 0; k < 3; ++k) g.t[k] = pj[k] - (R[3 * k] * pi[0] + R[3 * k + 1] * pi[1] + R[3 * k + 2] * pi[2]);
  g.fxi = intr[ix * 4 + 0]; g.fyi = intr[ix * 4 + 1]; g.cxi = intr[ix * 4 + 2]; g.cyi = intr[ix * 4 + 3];
  g.fx = intr[jx * 4 + 0];  g.fy = intr[jx * 4 + 1];  g.cx = intr[jx * 4 + 2];  g.cy = intr[jx * 4 + 3];
}

__global__ void __launch_bounds__(kThreads)
proj_fwd(const float* __restrict__ poses, const float* __restrict__ disps, const float* __restrict__ intr,
         const int64_t* __restrict__ ii, const int64_t* __restrict__ jj, int HW, int wd, int nc,
         float* __restrict__ coords, float* __restrict__ valid, float* __restrict__ Ji,
         float* __restrict__ Jj, float* __restrict__ Jz) {
  const int e = blockIdx.y;
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= HW) return;
  const int ix = static_cast<int>(ii[e]), jx = static_cast<int>(jj[e]);
  Geo g;
  edge_geo(poses, intr, ix, jx, g);
  const int row = k / wd, col = k - (k / wd) * wd;
  const double d = disps[(int64_t)ix * HW + k];
  const double X0 = (col - g.cxi) / g.fxi, Y0 = (row - g.cyi) / g.fyi;
  const double X = g.R[0] * X0 + g.R[1] * Y0 + g.R[2] + g.t[0] * d;
  const double Y = g.R[3] * X0 + g.R[4] * Y0 + g.R[5] + g.t[1] * d;
  const double Z = g.R[6] * X0 + g.R[7] * Y0 + g.R[8] + g.t[2] * d;
  const double Zs = Z < 0.1 ? 1.0 : Z;                 // pinhole.py:32
  const double iz = 1.0 / Zs;
  const int64_t o = (int64_t)e * HW + k;
  coords[o * nc + 0] = static_cast<float>(g.fx * (X * iz) + g.cx);
  coords[o * nc + 1] = static_cast<float>(g.fy * (Y * iz) + g.cy);
  if (nc == 3) coords[o * nc + 2] = static_cast<float>(d * iz);
  if (valid) valid[o] = (Z > 0.2 ? 1.f : 0.f);         // X0.z = 1 > 0.2 always
  if (Ji) {
    // Jp = [[fx/Z', 0, -fx X/Z'^2], [0, fy/Z', -fy Y/Z'^2]]; Ja = [[d I, -[X1]x]]
    const double p0 = g.fx * iz, p2 = -g.fx * X * iz * iz;
    const double q1 = g.fy * iz, q2 = -g.fy * Y * iz * iz;
    double ju[6] = {p0 * d, 0.0, p2 * d, p2 * Y, p0 * Z - p2 * X, -p0 * Y};
    double jv[6] = {0.0, q1 * d, q2 * d, -q1 * Z + q2 * Y, -q2 * X, q1 * X};
    float* oi = Ji + o * 12;
    float* oj = Jj + o * 12;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const double* J = r == 0 ? ju : jv;
      // Ji = -J Adj(G): tau part -R^T J_tau, phi part -R^T (J_phi - t x J_tau)
      const double c0 = J[3] - (g.t[1] * J[2] - g.t[2] * J[1]);
      const double c1 = J[4] - (g.t[2] * J[0] - g.t[0] * J[2]);
      const double c2 = J[5] - (g.t[0] * J[1] - g.t[1] * J[0]);
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        oi[r * 6 + n] = static_cast<float>(-(g.R[n] * J[0] + g.R[3 + n] * J[1] + g.R[6 + n] * J[2]));
        oi[r * 6 + 3 + n] = static_cast<float>(-(g.R[n] * c0 + g.R[3 + n] * c1 + g.R[6 + n] * c2));
      }
#pragma unroll
      for (int n = 0; n < 6; ++n) oj[r * 6 + n] = static_cast<float>(J[n]);
    }
    Jz[o * 2 + 0] = static_cast<float>(p0 * g.t[0] + p2 * g.t[2]);
    Jz[o * 2 + 1] = static_cast<float>(q1 * g.t[1] + q2 * g.t[2]);
  }
}

// one block per edge: per-edge tangent gradients [12] (FP64) and per-pixel disparity gradients
__global__ void __launch_bounds__(kThreads)
proj_bwd_edges(const float* __restrict__ poses, const float* __restrict__ disps,
               const float* __restrict__ intr, const int64_t* __restrict__ ii,
               const int64_t* __restrict__ jj, int HW, int wd, int nc,
               const float* __restrict__ gcoords, double* __restrict__ pose_part,
               float* __restrict__ disp_part) {
  __shared__ double lds[12 * kThreads];
  __shared__ double red[12];
  const int e = blockIdx.x;
  const int ix = static_cast<int>(ii[e]), jx = static_cast<int>(jj[e]);
  Geo g;
  edge_geo(poses, intr, ix, jx, g);
  double acc[12];
#pragma unroll
  for (int n = 0; n < 12; ++n) acc[n] = 0.0;
  for (int k = threadIdx.x; k < HW; k += kThreads) {
    const int row = k / wd, col = k - (k / wd) * wd;
    const double d = disps[(int64_t)ix * HW + k];
    const double X0 = (col - g.cxi) / g.fxi, Y0 = (row - g.cyi) / g.fyi;
    const double X = g.R[0] * X0 + g.R[1] * Y0 + g.R[2] + g.t[0] * d;
    const double Y = g.R[3] * X0 + g.R[4] * Y0 + g.R[5] + g.t[1] * d;
    const double Z = g.R[6] * X0 + g.R[7] * Y0 + g.R[8] + g.t[2] * d;
    const bool clamp = Z < 0.1;
    const double iz = clamp ? 1.0 : 1.0 / Z;
    const int64_t o = (int64_t)e * HW + k;
    const double gx = gcoords[o * nc + 0], gy = gcoords[o * nc + 1];
    const double gd = nc == 3 ? static_cast<double>(gcoords[o * nc + 2]) : 0.0;
    // gradient with respect to the point P = [X, Y, Z, W = d]
    const double GX = gx * g.fx * iz, GY = gy * g.fy * iz;
    const double GZ = clamp ? 0.0 : -(gx * g.fx * X + gy * g.fy * Y + gd * d) * iz * iz;
    const double GW = gd * iz;
    // T_j tangent: G_P . (e_s^ P), P = [X, Y, Z, d]
    const double tj[6] = {GX * d, GY * d, GZ * d, GZ * Y - GY * Z, GX * Z - GZ * X, GY * X - GX * Y};
    // T_i tangent: -G_P . (Adj(G) e_s)^ P = -(Adj(G)^T tj)_s
    const double c0 = tj[3] - (g.t[1] * tj[2] - g.t[2] * tj[1]);
    const double c1 = tj[4] - (g.t[2] * tj[0] - g.t[0] * tj[2]);
    const double c2 = tj[5] - (g.t[0] * tj[1] - g.t[1] * tj[0]);
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      acc[n] -= g.R[n] * tj[0] + g.R[3 + n] * tj[1] + g.R[6 + n] * tj[2];
      acc[3 + n] -= g.R[n] * c0 + g.R[3 + n] * c1 + g.R[6 + n] * c2;
    }
#pragma unroll
    for (int n = 0; n < 6; ++n) acc[6 + n] += tj[n];
    // disparity: dP/dd = [t; 1]
    disp_part[o] = static_cast<float>(GX * g.t[0] + GY * g.t[1] + GZ * g.t[2] + GW);
  }
  block_reduce_store<12, kThreads, double>(acc, lds, red);
  __syncthreads();
  if (threadIdx.x < 12) pose_part[(int64_t)e * 12 + threadIdx.x] = red[threadIdx.x];
}

__global__ void proj_bwd_frames(const float* __restrict__ disp_part, const int64_t* __restrict__ ii,
                                int E, int HW, float* __restrict__ grad_disps) {
  const int n = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= HW) return;
  double s = 0.0;
  for (int e = 0; e < E; ++e)
    if (ii[e] == n) s += disp_part[(int64_t)e * HW + k];
  grad_disps[(int64_t)n * HW + k] = static_cast<float>(s);
}

__global__ void proj_bwd_poses(const double* __restrict__ pose_part, const int64_t* __restrict__ ii,
                               const int64_t* __restrict__ jj, int E, int N, float* __restrict__ grad_poses) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 6 * N) return;
  const int k = idx / 6, n = idx % 6;
  double s = 0.0;
  for (int e = 0; e < E; ++e) {
    if (ii[e] == k) s += pose_part[(int64_t)e * 12 + n];
    if (jj[e] == k) s += pose_part[(int64_t)e * 12 + 6 + n];
  }
  grad_poses[idx] = static_cast<float>(s);
}

}  // namespace
}  // namespace hsl2

using namespace hsl2;

extern "C" int hsl2_projective_transform(const float* poses, const float* disps, const float* intrinsics,
                                         const int64_t* ii, const int64_t* jj, int64_t num_edges,
                                         int64_t ht, int64_t wd, int return_depth, float* coords,
                                         float* valid, float* Ji, float* Jj, float* Jz, void* stream) {
  HSL2_REQUIRE(num_edges >= 0 && ht > 0 && wd > 0 && ht * wd < (1ll << 31), HSL2_ERR_SHAPE,
               "projective_transform: bad sizes");
  HSL2_REQUIRE(poses && disps && intrinsics && coords, HSL2_ERR_ARGUMENT, "projective_transform: null pointer");
  HSL2_REQUIRE(num_edges == 0 || (ii && jj), HSL2_ERR_ARGUMENT, "projective_transform: null edge arrays");
  HSL2_REQUIRE((Ji == nullptr) == (Jj == nullptr) && (Ji == nullptr) == (Jz == nullptr), HSL2_ERR_ARGUMENT,
               "projective_transform: Ji, Jj, Jz must all be given or all be null");
  HSL2_REQUIRE(num_edges < 65536, HSL2_ERR_SHAPE, "projective_transform: at most 65535 edges per call");
  if (num_edges == 0) return HSL2_OK;
  const int HW = static_cast<int>(ht * wd);
  hipLaunchKernelGGL(proj_fwd, dim3((HW + kThreads - 1) / kThreads, static_cast<unsigned>(num_edges)),
                     dim3(kThreads), 0, as_stream(stream), poses, disps, intrinsics, ii, jj, HW,
                     static_cast<int>(wd), return_depth ? 3 : 2, coords, valid, Ji, Jj, Jz);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" size_t hsl2_projective_transform_workspace_bytes(int64_t num_edges, int64_t ht, int64_t wd) {
  const size_t a = static_cast<size_t>(num_edges) * 12 * sizeof(double);
  return ((a + 255) & ~size_t(255)) + static_cast<size_t>(num_edges) * ht * wd * sizeof(float);
}

extern "C" int hsl2_projective_transform_backward(const float* poses, const float* disps,
                                                  const float* intrinsics, const int64_t* ii,
                                                  const int64_t* jj, int64_t num_edges, int64_t num_poses,
                                                  int64_t ht, int64_t wd, int return_depth,
                                                  const float* grad_coords, float* grad_poses,
                                                  float* grad_disps, void* workspace,
                                                  size_t workspace_bytes, void* stream) {
  HSL2_REQUIRE(num_edges >= 0 && num_poses > 0 && ht > 0 && wd > 0 && ht * wd < (1ll << 31),
               HSL2_ERR_SHAPE, "projective_transform backward: bad sizes");
  HSL2_REQUIRE(poses && disps && intrinsics && grad_poses && grad_disps && (num_edges == 0 || grad_coords),
               HSL2_ERR_ARGUMENT, "projective_transform backward: null pointer");
  HSL2_REQUIRE(num_edges == 0 || (ii && jj), HSL2_ERR_ARGUMENT, "projective_transform backward: null edges");
  HSL2_REQUIRE(workspace_bytes >= hsl2_projective_transform_workspace_bytes(num_edges, ht, wd) &&
                   (reinterpret_cast<uintptr_t>(workspace) % 256) == 0,
               HSL2_ERR_WORKSPACE, "projective_transform backward: workspace too small or misaligned");
  hipStream_t s = as_stream(stream);
  const int HW = static_cast<int>(ht * wd), E = static_cast<int>(num_edges), N = static_cast<int>(num_poses);
  char* ws = static_cast<char*>(workspace);
  double* pose_part = reinterpret_cast<double*>(ws);
  float* disp_part = reinterpret_cast<float*>(ws + ((static_cast<size_t>(E) * 12 * sizeof(double) + 255) & ~size_t(255)));
  if (E > 0) {
    hipLaunchKernelGGL(proj_bwd_edges, dim3(E), dim3(kThreads), 0, s, poses, disps, intrinsics, ii, jj, HW,
                       static_cast<int>(wd), return_depth ? 3 : 2, grad_coords, pose_part, disp_part);
    HSL2_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(proj_bwd_frames, dim3((HW + 255) / 256, N), dim3(256), 0, s, disp_part, ii, E, HW, grad_disps);
  HSL2_LAUNCH_CHECK();
  hipLaunchKernelGGL(proj_bwd_poses, dim3((6 * N + 255) / 256), dim3(256), 0, s, pose_part, ii, jj, E, N, grad_poses);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}
