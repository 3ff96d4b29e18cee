// Edge-selection and point-cloud geometry: frame_distance, covis_distance, depth_filter,
// iproj (src/droid_kernels.cu:860-1273, 2047-2143; bindings src/droid.cpp:170-241).
//
// All four project each pixel of a frame i with its disparity, (u - cx) / fx, (v - cy) / fy,
// 1, d, through a pose (relative G_ij = G_j G_i^-1, or the frame's own pose for iproj) and the
// pinhole model of one camera.  The pixel arithmetic is FP32 as in the reference; per-pair
// sums are FP64.  One block per pair / frame (frame_distance, covis_distance) or per
// (256 pixels, frame) (depth_filter, iproj); no atomics: depth_filter visits a pixel's six
// neighbours in one thread and writes its count once.
#include "common.hpp"
#include "se3.hpp"

namespace hsl2 {
namespace {

constexpr float kGeoMinDepth = 0.25f;   // src/droid_kernels.cu:26

struct Cam {
  float fx, fy, cx, cy;
};

__device__ __forceinline__ Cam load_cam(const float* __restrict__ intr) {
  return Cam{intr[0], intr[1], intr[2], intr[3]};
}

// G X for a homogeneous point X = (x, y, 1, d): rotation of (x, y, 1) plus d t (actSE3,
// src/droid_kernels.cu:86-95)
__device__ __forceinline__ void act(const float t[3], const float q[4], const float X[3], float d, float Y[3]) {
  quat_rotate(q, X, Y);
  Y[0] += d * t[0];
  Y[1] += d * t[1];
  Y[2] += d * t[2];
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) lds[wv] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NT / kWave; ++i) s += lds[i];
  return s;
}

// frame_distance_kernel (src/droid_kernels.cu:860-992): for pair (ii, jj) the beta-weighted
// mean flow magnitude of the full relative motion and of its translation alone, over the
// pixels whose transformed depth exceeds 0.25; 1000 when fewer than 75 % of the weight is valid.
__global__ void __launch_bounds__(256)
frame_distance_kernel(const float* __restrict__ poses, const float* __restrict__ disps, const float* __restrict__ intr,
                      const int64_t* __restrict__ ii, const int64_t* __restrict__ jj, int ht, int wd, float beta,
                      float* __restrict__ dist) {
  __shared__ double lds[256 / kWave];
  const int e = blockIdx.x;
  const int64_t ix = ii[e], jx = jj[e];
  const Cam c = load_cam(intr);
  float t[3], q[4];
  relative_pose(poses + ix * 7, poses + jx * 7, t, q);
  const int HW = ht * wd;
  const float* di = disps + ix * HW;
  double acc = 0.0, valid = 0.0, total = 0.0;
  for (int k = threadIdx.x; k < HW; k += blockDim.x) {
    const float u = static_cast<float>(k % wd), v = static_cast<float>(k / wd);
    const float X[3] = {(u - c.cx) / c.fx, (v - c.cy) / c.fy, 1.f};
    const float d = di[k];
    float Y[3];
    act(t, q, X, d, Y);
    float du = c.fx * (Y[0] / Y[2]) + c.cx - u, dv = c.fy * (Y[1] / Y[2]) + c.cy - v;
    total += beta;
    if (Y[2] > kGeoMinDepth) {
      acc += static_cast<double>(beta) * sqrtf(du * du + dv * dv);
      valid += beta;
    }
    const float Z[3] = {X[0] + d * t[0], X[1] + d * t[1], X[2] + d * t[2]};   // translation only
    du = c.fx * (Z[0] / Z[2]) + c.cx - u;
    dv = c.fy * (Z[1] / Z[2]) + c.cy - v;
    total += 1.f - beta;
    if (Z[2] > kGeoMinDepth) {
      acc += static_cast<double>(1.f - beta) * sqrtf(du * du + dv * dv);
      valid += 1.f - beta;
    }
  }
  acc = block_sum<256>(acc, lds);
  valid = block_sum<256>(valid, lds);
  total = block_sum<256>(total, lds);
  if (threadIdx.x == 0) dist[e] = (valid / (total + 1e-8) < 0.75) ? 1000.f : static_cast<float>(acc / valid);
}

// covis_distance_kernel (src/droid_kernels.cu:994-1092): the fraction of frame ii's pixels that
// project outside the image in both neighbours ii - 1 and ii + 1.
__global__ void __launch_bounds__(256)
covis_distance_kernel(const float* __restrict__ poses, const float* __restrict__ disps, const float* __restrict__ intr,
                      const int64_t* __restrict__ ii, int ht, int wd, float* __restrict__ dist) {
  __shared__ double lds[256 / kWave];
  const int e = blockIdx.x;
  const int64_t ix = ii[e];
  const Cam c = load_cam(intr);
  float t1[3], q1[4], t2[3], q2[4];
  relative_pose(poses + ix * 7, poses + (ix - 1) * 7, t1, q1);
  relative_pose(poses + ix * 7, poses + (ix + 1) * 7, t2, q2);
  const int HW = ht * wd;
  const float* di = disps + ix * HW;
  double cnt = 0.0;
  for (int k = threadIdx.x; k < HW; k += blockDim.x) {
    const float u = static_cast<float>(k % wd), v = static_cast<float>(k / wd);
    const float X[3] = {(u - c.cx) / c.fx, (v - c.cy) / c.fy, 1.f};
    float Y1[3], Y2[3];
    act(t1, q1, X, di[k], Y1);
    act(t2, q2, X, di[k], Y2);
    const float x1 = c.fx * (Y1[0] / Y1[2]) + c.cx, y1 = c.fy * (Y1[1] / Y1[2]) + c.cy;
    const float x2 = c.fx * (Y2[0] / Y2[2]) + c.cx, y2 = c.fy * (Y2[1] / Y2[2]) + c.cy;
    const bool out1 = x1 < 0 || x1 >= wd || y1 < 0 || y1 >= ht;
    const bool out2 = x2 < 0 || x2 >= wd || y2 < 0 || y2 >= ht;
    if (out1 && out2) cnt += 1.0;
  }
  cnt = block_sum<256>(cnt, lds);
  if (threadIdx.x == 0) dist[e] = static_cast<float>(cnt / HW);
}

// depth_filter_kernel (src/droid_kernels.cu:1094-1200): for pixel k of frame inds[b], the number
// of the neighbours jx = ix - 1, ix - 2, ix - 3, ix + 3, ix + 4, ix + 5 inside [0, num) (the
// reference's offsets) in which the projected pixel has an integer corner (u0, v0) inside the
// image minus its last row and column, and one of the four corner disparities d of frame jx
// agrees with the projected depth: |1/dj - 1/d| < thresh[b] (tested in double, corners in the
// reference's order; at most one count per neighbour).
__global__ void __launch_bounds__(256)
depth_filter_kernel(const float* __restrict__ poses, const float* __restrict__ disps, const float* __restrict__ intr,
                    const int64_t* __restrict__ inds, const float* __restrict__ thresh, int num, int ht, int wd,
                    float* __restrict__ counter) {
  const int b = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int HW = ht * wd;
  if (k >= HW) return;
  const int64_t ix = inds[b];
  const Cam c = load_cam(intr);
  const float th = thresh[b];
  const float u = static_cast<float>(k % wd), v = static_cast<float>(k / wd);
  const float X[3] = {(u - c.cx) / c.fx, (v - c.cy) / c.fy, 1.f};
  const float d = disps[ix * HW + k];
  float count = 0.f;
  for (int n = 0; n < 6; ++n) {
    const int64_t jx = n < 3 ? ix - n - 1 : ix + n;
    if (jx < 0 || jx >= num) continue;
    float t[3], q[4], Y[3];
    relative_pose(poses + ix * 7, poses + jx * 7, t, q);
    act(t, q, X, d, Y);
    const float uj = c.fx * (Y[0] / Y[2]) + c.cx, vj = c.fy * (Y[1] / Y[2]) + c.cy;
    const float dj = d / Y[2];
    const int u0 = static_cast<int>(floorf(uj)), v0 = static_cast<int>(floorf(vj));
    if (u0 >= 0 && v0 >= 0 && u0 < wd - 1 && v0 < ht - 1) {
      const float* dn = disps + jx * HW;
      const float corner[4] = {dn[v0 * wd + u0], dn[v0 * wd + u0 + 1], dn[(v0 + 1) * wd + u0],
                               dn[(v0 + 1) * wd + u0 + 1]};
      const double zj = 1.0 / dj;
      bool hit = false;
#pragma unroll
      for (int m = 0; m < 4; ++m) hit = hit || fabs(zj - 1.0 / corner[m]) < th;
      if (hit) count += 1.f;
    }
  }
  counter[(int64_t)b * HW + k] = count;
}

// iproj_kernel (src/droid_kernels.cu:1204-1273): frame n's pixels lifted with their disparity
// and moved by the frame's pose, as 3-D points (x, y, z) / d.
__global__ void __launch_bounds__(256)
iproj_kernel(const float* __restrict__ poses, const float* __restrict__ disps, const float* __restrict__ intr, int ht,
             int wd, float* __restrict__ points) {
  const int n = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int HW = ht * wd;
  if (k >= HW) return;
  const Cam c = load_cam(intr);
  const float* pose = poses + (int64_t)n * 7;
  const float t[3] = {pose[0], pose[1], pose[2]}, q[4] = {pose[3], pose[4], pose[5], pose[6]};
  const float u = static_cast<float>(k % wd), v = static_cast<float>(k / wd);
  const float X[3] = {(u - c.cx) / c.fx, (v - c.cy) / c.fy, 1.f};
  const float d = disps[(int64_t)n * HW + k];
  float Y[3];
  act(t, q, X, d, Y);
  float* p = points + ((int64_t)n * HW + k) * 3;
  p[0] = Y[0] / d;
  p[1] = Y[1] / d;
  p[2] = Y[2] / d;
}

int check_geo(const void* poses, const void* disps, const void* intr, int64_t ht, int64_t wd, const char* what) {
  HSL2_REQUIRE(poses && disps && intr, HSL2_ERR_ARGUMENT, std::string(what) + ": null pointer");
  HSL2_REQUIRE(ht > 0 && wd > 0 && ht * wd < (1ll << 31), HSL2_ERR_SHAPE, std::string(what) + ": bad sizes");
  return HSL2_OK;
}

}  // namespace
}  // namespace hsl2

using namespace hsl2;

extern "C" int hsl2_frame_distance(const float* poses, const float* disps, const float* intrinsics, const int64_t* ii,
                                   const int64_t* jj, int64_t num_edges, int64_t ht, int64_t wd, float beta,
                                   float* dist, void* stream) {
  if (int st = check_geo(poses, disps, intrinsics, ht, wd, "frame_distance")) return st;
  if (num_edges == 0) return HSL2_OK;
  HSL2_REQUIRE(ii && jj && dist && num_edges > 0 && num_edges < (1ll << 31), HSL2_ERR_ARGUMENT,
               "frame_distance: bad edge arrays");
  hipLaunchKernelGGL(frame_distance_kernel, dim3(static_cast<unsigned>(num_edges)), dim3(256), 0, as_stream(stream),
                     poses, disps, intrinsics, ii, jj, static_cast<int>(ht), static_cast<int>(wd), beta, dist);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" int hsl2_covis_distance(const float* poses, const float* disps, const float* intrinsics, const int64_t* ii,
                                   int64_t num, int64_t ht, int64_t wd, float* dist, void* stream) {
  if (int st = check_geo(poses, disps, intrinsics, ht, wd, "covis_distance")) return st;
  if (num == 0) return HSL2_OK;
  HSL2_REQUIRE(ii && dist && num > 0 && num < (1ll << 31), HSL2_ERR_ARGUMENT, "covis_distance: bad index array");
  hipLaunchKernelGGL(covis_distance_kernel, dim3(static_cast<unsigned>(num)), dim3(256), 0, as_stream(stream), poses,
                     disps, intrinsics, ii, static_cast<int>(ht), static_cast<int>(wd), dist);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" int hsl2_depth_filter(const float* poses, const float* disps, const float* intrinsics, const int64_t* inds,
                                 const float* thresh, int64_t count, int64_t num_frames, int64_t ht, int64_t wd,
                                 float* counter, void* stream) {
  if (int st = check_geo(poses, disps, intrinsics, ht, wd, "depth_filter")) return st;
  if (count == 0) return HSL2_OK;
  HSL2_REQUIRE(inds && thresh && counter && count > 0 && count < 65536, HSL2_ERR_ARGUMENT,
               "depth_filter: bad index arrays");
  const int HW = static_cast<int>(ht * wd);
  hipLaunchKernelGGL(depth_filter_kernel, dim3((HW + 255) / 256, static_cast<unsigned>(count)), dim3(256), 0,
                     as_stream(stream), poses, disps, intrinsics, inds, thresh, static_cast<int>(num_frames),
                     static_cast<int>(ht), static_cast<int>(wd), counter);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}

extern "C" int hsl2_iproj(const float* poses, const float* disps, const float* intrinsics, int64_t num_frames,
                          int64_t ht, int64_t wd, float* points, void* stream) {
  if (int st = check_geo(poses, disps, intrinsics, ht, wd, "iproj")) return st;
  if (num_frames == 0) return HSL2_OK;
  HSL2_REQUIRE(points && num_frames > 0 && num_frames < 65536, HSL2_ERR_ARGUMENT, "iproj: bad sizes");
  const int HW = static_cast<int>(ht * wd);
  hipLaunchKernelGGL(iproj_kernel, dim3((HW + 255) / 256, static_cast<unsigned>(num_frames)), dim3(256), 0,
                     as_stream(stream), poses, disps, intrinsics, static_cast<int>(ht), static_cast<int>(wd), points);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}
