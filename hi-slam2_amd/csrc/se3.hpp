// SE3 helpers of the pixel-geometry kernels (xyzw quaternions, G_ij = G_j G_i^-1).
#pragma once
#include "common.hpp"

namespace hsl2 {
namespace {

// ---- SE3 helpers (xyzw quaternions, tangent [tau, phi]) ----------------------
__device__ __forceinline__ void quat_rotate(const float q[4], const float X[3], float Y[3]) {
  // Y = X + 2 w (v x X) + 2 v x (v x X)   (src/droid_kernels.cu:74-84)
  const float uvx = 2.f * (q[1] * X[2] - q[2] * X[1]);
  const float uvy = 2.f * (q[2] * X[0] - q[0] * X[2]);
  const float uvz = 2.f * (q[0] * X[1] - q[1] * X[0]);
  Y[0] = X[0] + q[3] * uvx + (q[1] * uvz - q[2] * uvy);
  Y[1] = X[1] + q[3] * uvy + (q[2] * uvx - q[0] * uvz);
  Y[2] = X[2] + q[3] * uvz + (q[0] * uvy - q[1] * uvx);
}

// G_ij = G_j G_i^-1 (src/droid_kernels.cu:148-159)
__device__ __forceinline__ void relative_pose(const float* pi, const float* pj, float tij[3],
                                              float qij[4]) {
  const float* qi = pi + 3;
  const float* qj = pj + 3;
  qij[0] = -qj[3] * qi[0] + qj[0] * qi[3] - qj[1] * qi[2] + qj[2] * qi[1];
  qij[1] = -qj[3] * qi[1] + qj[1] * qi[3] - qj[2] * qi[0] + qj[0] * qi[2];
  qij[2] = -qj[3] * qi[2] + qj[2] * qi[3] - qj[0] * qi[1] + qj[1] * qi[0];
  qij[3] = qj[3] * qi[3] + qj[0] * qi[0] + qj[1] * qi[1] + qj[2] * qi[2];
  float r[3];
  quat_rotate(qij, pi, r);
  tij[0] = pj[0] - r[0];
  tij[1] = pj[1] - r[1];
  tij[2] = pj[2] - r[2];
}

}  // namespace
}  // namespace hsl2
