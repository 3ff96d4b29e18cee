// One Gauss-Newton step in the semantics of the reference's torch formulation
// (hislam2/geom/ba.py:32-158 with hislam2/geom/chol.py:5-78), and its hand-written adjoint.
//
// Forward (hsl2_gn_step): the kernels of the ba_cuda driver with the torch switches of
// ba_kernels.hpp (per-frame intrinsics, validity Z > 0.2, C + eta + 1e-7, damping of H before
// the Schur complement, full back-substitution, lietorch Exp, disparity clamp), plus the
// marginal depth variance dzcov = Q + Q^2 e^T S^-1 e per pixel (chol.py:74-76, where it is
// sum(square(L^-1 E Q)) + Q), through W = L^-1 of the tiled Cholesky factor and S^-1 = W^T W.
//
// Backward (hsl2_gn_step_backward).  The step solves K y = g for y = [dx; dz] with
//   K = [[H + diag(ep + lm diag H), E], [E^T, diag C]],  g = [v; w],
//   H, E, C, v, w = sums over edges, pixels and residual rows c of w_c J_c J_c^T, w_c J_c r_c.
// With upstream gradients (dx_bar, dz_bar) and lambda = K^-1 [dx_bar; dz_bar] (the implicit
// adjoint through the Schur solve, CholeskySolver.backward of chol.py:22-31, solved with the
// same factorisation machinery), every parameter theta receives
//   theta_bar = lambda^T (dg/dtheta - (dK/dtheta) y).
// Per pixel and row c, with a_c = J_c lambda, b_c = J_c y, r_c = target_c - pi_c:
//   target_bar_c = w_c a_c
//   weight_bar_c = 0.001 valid (a_c r_c - a_c b_c - lm sum_q lambda_q y_q J_cq^2)   (q: pose coords;
//                  for block_solve's damping of every 6x6 block the lm term is sum_i Lam_i Y_i,
//                  Lam_i = lambda_i,i J_c,i + lambda_j,i J_c,6+i, Y_i likewise with y)
//   eta_bar      = -lambda_z y_z
// and for the state (the 6+6 tangent coordinates of T_i, T_j and the disparity d):
//   state_bar_s += w_c (sum_q beta_q d^2 pi_c / dxi_q dtheta_s - a_c J_cs),
//   beta_q = lambda_q (r_c - b_c) - a_c y_q - 2 lm [q pose] lambda_q y_q J_cq.
// The mixed second derivatives are evaluated in closed form by recomputing the projection of
// each pixel (nothing per pixel is stored): with P = G X0 (homogeneous, G = T_j T_i^-1),
// generators e_s^ and A_s = (Adj_G e_s)^, the point derivatives are dP/dT_j,s = e_s^ P,
// dP/dT_i,s = -A_s P, dP/dd = [t; 1], and for any direction u = (u_i, u_j, u_z)
//   sum_q u_q d2P/dxi_q dT_j,s = u_j^ e_s^ P + e_s^ (-U_i P + u_z [t;1])
//   sum_q u_q d2P/dxi_q dT_i,s = -u_j^ A_s P - A_s (-U_i P + u_z [t;1])
//   sum_q u_q d2P/dxi_q dd     = u_j^ [t;1] - U_i [t;1],          U_i = (Adj_G u_i)^,
// where xi is a left perturbation applied after theta (J evaluated at the perturbed state).
// The result is exact (the same quantity autograd forms through the torch formulation).
#include "ba_kernels.hpp"

namespace hsl2 {
namespace {

constexpr int kAdjThreads = 256;

__global__ void ba_clamp_disps(float* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = d[i];
  v = v > 10.f ? 0.f : v;                     // hislam2/geom/ba.py:104
  d[i] = fmaxf(v, 0.001f);                    // hislam2/geom/ba.py:105
}

__global__ void fill_zero(float* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// ---- dzcov ---------------------------------------------------------------------------------
// W = L^-1, tile by tile: diagonal tiles are the stored inverses Ld_k (column-major 64x64),
// W_ik = -Ld_i sum_{m=k}^{i-1} L_im W_mk for i > k (one launch per tile row i), upper tiles 0.
__global__ void tri_inv_init(const double* __restrict__ Ld, int64_t n_pad, double* __restrict__ W) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_pad * n_pad) return;
  const int64_t col = idx / n_pad, row = idx - col * n_pad;
  const int64_t tr = row / 64, tc = col / 64;
  W[idx] = tr == tc ? Ld[tc * 4096 + (col % 64) * 64 + (row % 64)] : 0.0;
}

__global__ void __launch_bounds__(256)
tri_inv_row(const double* __restrict__ A, const double* __restrict__ Ld, int64_t n_pad, int i,
            double* __restrict__ W) {
  __shared__ double La[64][65];
  __shared__ double Wb[64][65];
  const int k = blockIdx.x;                    // k < i
  const int t = threadIdx.x, tr = t >> 4, ts = t & 15;
  double acc[4][4] = {};
  for (int m = k; m < i; ++m) {
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
      const int idx = t + 256 * it, c = idx >> 6, r = idx & 63;
      La[r][c] = A[((int64_t)i * 64 + r) + ((int64_t)m * 64 + c) * n_pad];
      Wb[r][c] = W[((int64_t)m * 64 + r) + ((int64_t)k * 64 + c) * n_pad];
    }
    __syncthreads();
    for (int q = 0; q < 64; ++q) {
      double a[4], b[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        a[p] = La[4 * tr + p][q];
        b[p] = Wb[q][4 * ts + p];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[p][u] += a[p] * b[u];
    }
    __syncthreads();
  }
  // W_ik = -Ld_i acc
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int u = 0; u < 4; ++u) Wb[4 * tr + p][4 * ts + u] = acc[p][u];
#pragma unroll 4
  for (int it = 0; it < 16; ++it) {
    const int idx = t + 256 * it, c = idx >> 6, r = idx & 63;
    La[r][c] = Ld[(int64_t)i * 4096 + c * 64 + r];
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double v = 0;
      for (int q = 0; q < 64; ++q) v += La[4 * tr + p][q] * Wb[q][4 * ts + u];
      W[((int64_t)i * 64 + 4 * tr + p) + ((int64_t)k * 64 + 4 * ts + u) * n_pad] = -v;
    }
}

// Sinv = W^T W, one block per 64x64 output tile (I, J): sum over K >= max(I, J) of W_KI^T W_KJ.
__global__ void __launch_bounds__(256)
gram_wtw(const double* __restrict__ W, int64_t n_pad, int T, double* __restrict__ Sinv) {
  __shared__ double Wa[64][65];
  __shared__ double Wb[64][65];
  const int I = blockIdx.x % T, J = blockIdx.x / T;
  const int t = threadIdx.x, tr = t >> 4, ts = t & 15;
  double acc[4][4] = {};
  for (int K = max(I, J); K < T; ++K) {
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
      const int idx = t + 256 * it, c = idx >> 6, r = idx & 63;
      Wa[r][c] = W[((int64_t)K * 64 + r) + ((int64_t)I * 64 + c) * n_pad];
      Wb[r][c] = W[((int64_t)K * 64 + r) + ((int64_t)J * 64 + c) * n_pad];
    }
    __syncthreads();
    for (int q = 0; q < 64; ++q) {
      double a[4], b[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        a[p] = Wa[q][4 * tr + p];
        b[p] = Wb[q][4 * ts + p];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[p][u] += a[p] * b[u];
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int u = 0; u < 4; ++u)
      Sinv[((int64_t)I * 64 + 4 * tr + p) + ((int64_t)J * 64 + 4 * ts + u) * n_pad] = acc[p][u];
}

// dzcov[f][k] = Q + Q^2 sum_{s, s' slots of f} E_s(k)^T Sinv[a_s, a_s'] E_s'(k)
__global__ void __launch_bounds__(256)
dzcov_kernel(const float* __restrict__ Eall, const float* __restrict__ Q,
             const double* __restrict__ Sinv, int64_t n_pad, const int* __restrict__ slot_ptr,
             const int* __restrict__ slot_row, const int* __restrict__ slot_pose, int HW,
             const int* __restrict__ info, float* __restrict__ dzcov) {
  const int f = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= HW) return;
  double acc = 0;
  if (info[0] == 0) {
    for (int s = slot_ptr[f]; s < slot_ptr[f + 1]; ++s) {
      const float* Er = Eall + (int64_t)slot_row[s] * 6 * HW + k;
      double ea[6];
#pragma unroll
      for (int n = 0; n < 6; ++n) ea[n] = Er[(int64_t)n * HW];
      const int64_t a = 6 * (int64_t)slot_pose[s];
      for (int s2 = slot_ptr[f]; s2 < slot_ptr[f + 1]; ++s2) {
        const float* Eq = Eall + (int64_t)slot_row[s2] * 6 * HW + k;
        const int64_t b = 6 * (int64_t)slot_pose[s2];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          const double eb = Eq[(int64_t)m * HW];
          double row = 0;
#pragma unroll
          for (int n = 0; n < 6; ++n) row += ea[n] * Sinv[(a + n) + (b + m) * n_pad];
          acc += row * eb;
        }
      }
    }
  }
  const double q = Q[(int64_t)f * HW + k];
  dzcov[(int64_t)f * HW + k] = static_cast<float>(q + q * q * acc);
}

// ---- adjoint -------------------------------------------------------------------------------
// rhs of the adjoint solve: dx_bar - sum over the pose's Schur slots of E Q dz_bar
__global__ void adj_rhs(const float* __restrict__ gdx, const double* __restrict__ Ev,
                        const int* __restrict__ rhs_ptr, const int* __restrict__ rhs_src, int P,
                        int n_pad, double* __restrict__ bvec) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_pad) return;
  if (idx >= 6 * P) {
    bvec[idx] = 0.0;
    return;
  }
  const int a = idx / 6, i = idx % 6;
  double v = gdx[idx];
  for (int q = rhs_ptr[a]; q < rhs_ptr[a + 1]; ++q) {
    const int code = rhs_src[q];
    if ((code & 3) == kRhsEv) v -= Ev[(int64_t)(code >> 2) * 6 + i];
  }
  bvec[idx] = v;
}

struct D4 {
  double x, y, z, w;
};

// u^ P for a twist u = [tau, phi] and a homogeneous point P: [phi x P + tau P.w, 0]
__device__ __forceinline__ D4 twist(const double u[6], const D4& P) {
  return {u[4] * P.z - u[5] * P.y + u[0] * P.w, u[5] * P.x - u[3] * P.z + u[1] * P.w,
          u[3] * P.y - u[4] * P.x + u[2] * P.w, 0.0};
}
// e_s^ P for the s-th generator
__device__ __forceinline__ D4 gen(int s, const D4& P) {
  switch (s) {
    case 0: return {P.w, 0.0, 0.0, 0.0};
    case 1: return {0.0, P.w, 0.0, 0.0};
    case 2: return {0.0, 0.0, P.w, 0.0};
    case 3: return {0.0, -P.z, P.y, 0.0};
    case 4: return {P.z, 0.0, -P.x, 0.0};
    default: return {-P.y, P.x, 0.0, 0.0};
  }
}
// Adj(G) u = [R tau + t x (R phi), R phi]
__device__ __forceinline__ void adj(const double R[9], const double t[3], const double u[6],
                                    double o[6]) {
  double rt[3], rp[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    rt[i] = R[3 * i] * u[0] + R[3 * i + 1] * u[1] + R[3 * i + 2] * u[2];
    rp[i] = R[3 * i] * u[3] + R[3 * i + 1] * u[4] + R[3 * i + 2] * u[5];
  }
  o[0] = rt[0] + t[1] * rp[2] - t[2] * rp[1];
  o[1] = rt[1] + t[2] * rp[0] - t[0] * rp[2];
  o[2] = rt[2] + t[0] * rp[1] - t[1] * rp[0];
  o[3] = rp[0];
  o[4] = rp[1];
  o[5] = rp[2];
}
__device__ __forceinline__ double dot4(const D4& a, const D4& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}
__device__ __forceinline__ D4 axpy(double a, const D4& x, const D4& y) {
  return {a * x.x + y.x, a * x.y + y.y, a * x.z + y.z, a * x.w + y.w};
}
__device__ __forceinline__ D4 sub(const D4& a, const D4& b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w};
}
__device__ __forceinline__ D4 add(const D4& a, const D4& b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
}

// One block per edge.  Outputs per pixel: target and weight gradients, the disparity-state
// gradient of frame ii (disp_part[e]); per edge: the 12 tangent-state gradients (pose_part[e]).
__global__ void __launch_bounds__(kAdjThreads)
gn_adjoint_edges(const float* __restrict__ poses, const float* __restrict__ disps,
                 const float* __restrict__ intr, const float* __restrict__ targets,
                 const float* __restrict__ weights, const int* __restrict__ edge_ii,
                 const int* __restrict__ edge_jj, const int* __restrict__ edge_frame, int t0,
                 int P, int HW, int wd, int E, float lm, int block_damp, const double* __restrict__ lam_x,
                 const int* __restrict__ info, const float* __restrict__ dx,
                 const float* __restrict__ dz, const float* __restrict__ lz,
                 float* __restrict__ grad_targets, float* __restrict__ grad_weights,
                 double* __restrict__ pose_part, float* __restrict__ disp_part) {
  __shared__ double lds[12 * kAdjThreads];
  __shared__ double red[12];
  const int e = blockIdx.x;
  const int ix = edge_ii[e], jx = edge_jj[e], fr = edge_frame[e];
  const bool self = ix == jx;
  // G = T_j T_i^-1 in double
  double R[9], tg[3];
  {
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
    double q[4];   // q_j * conj(q_i)
    q[0] = -qj[3] * qi[0] + qj[0] * qi[3] - qj[1] * qi[2] + qj[2] * qi[1];
    q[1] = -qj[3] * qi[1] + qj[1] * qi[3] - qj[2] * qi[0] + qj[0] * qi[2];
    q[2] = -qj[3] * qi[2] + qj[2] * qi[3] - qj[0] * qi[1] + qj[1] * qi[0];
    q[3] = qj[3] * qi[3] + qj[0] * qi[0] + qj[1] * qi[1] + qj[2] * qi[2];
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - z * w);     R[2] = 2 * (x * z + y * w);
    R[3] = 2 * (x * y + z * w);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - x * w);
    R[6] = 2 * (x * z - y * w);     R[7] = 2 * (y * z + x * w);     R[8] = 1 - 2 * (x * x + y * y);
#pragma unroll
    for (int k = 0; k < 3; ++k)
      tg[k] = pj[k] - (R[3 * k] * pi[0] + R[3 * k + 1] * pi[1] + R[3 * k + 2] * pi[2]);
  }
  // A_s = Adj(G) e_s
  double As[6][6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    double es[6] = {0, 0, 0, 0, 0, 0};
    es[s] = 1.0;
    adj(R, tg, es, As[s]);
  }
  const double fxi = intr[ix * 4 + 0], fyi = intr[ix * 4 + 1], cxi = intr[ix * 4 + 2],
               cyi = intr[ix * 4 + 3];
  const double fx = intr[jx * 4 + 0], fy = intr[jx * 4 + 1], cx = intr[jx * 4 + 2],
               cy = intr[jx * 4 + 3];
  // lambda and y restricted to this edge's pose slots
  const bool solved = info[0] == 0;
  const int si = ix - t0, sj = jx - t0;
  const bool vi = si >= 0 && si < P, vj = sj >= 0 && sj < P;
  double li[6], lj[6], yi[6], yj[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    li[n] = (vi && solved) ? lam_x[6 * si + n] : 0.0;
    lj[n] = (vj && solved) ? lam_x[6 * sj + n] : 0.0;
    yi[n] = vi ? static_cast<double>(dx[6 * si + n]) : 0.0;
    yj[n] = vj ? static_cast<double>(dx[6 * sj + n]) : 0.0;
  }
  double Uli[6], Uyi[6];
  adj(R, tg, li, Uli);
  adj(R, tg, yi, Uyi);
  const D4 tt = {tg[0], tg[1], tg[2], 1.0};

  double acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
  const float* dsp = disps + (int64_t)ix * HW;
  const float* tgt = targets + (int64_t)e * 2 * HW;
  const float* wgt = weights + (int64_t)e * 2 * HW;
  float* gtg = grad_targets + (int64_t)e * 2 * HW;
  float* gwt = grad_weights + (int64_t)e * 2 * HW;
  for (int k = threadIdx.x; k < HW; k += kAdjThreads) {
    const int row = k / wd, col = k - (k / wd) * wd;
    const double d = dsp[k];
    const double X0 = (col - cxi) / fxi, Y0 = (row - cyi) / fyi;
    const D4 Pt = {R[0] * X0 + R[1] * Y0 + R[2] + tg[0] * d, R[3] * X0 + R[4] * Y0 + R[5] + tg[1] * d,
                   R[6] * X0 + R[7] * Y0 + R[8] + tg[2] * d, d};
    double sd = 0.0;
    if (self || !(static_cast<float>(Pt.z) > kMinDepthTorch)) {
      gtg[k] = 0.f; gtg[HW + k] = 0.f;
      gwt[k] = 0.f; gwt[HW + k] = 0.f;
      disp_part[(int64_t)e * HW + k] = 0.f;
      continue;
    }
    const double lzk = lz ? static_cast<double>(lz[(int64_t)fr * HW + k]) : 0.0;
    const double yzk = dz ? static_cast<double>(dz[(int64_t)fr * HW + k]) : 0.0;
    const double iz = 1.0 / Pt.z;
    // first derivatives of the point for the 13 local coordinates
    D4 Ps[13];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const D4 v = twist(As[s], Pt);
      Ps[s] = {-v.x, -v.y, -v.z, 0.0};
      Ps[6 + s] = gen(s, Pt);
    }
    Ps[12] = tt;
    const D4 Pl = axpy(lzk, tt, sub(twist(lj, Pt), twist(Uli, Pt)));
    const D4 Py = axpy(yzk, tt, sub(twist(yj, Pt), twist(Uyi, Pt)));
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const double f = c == 0 ? fx : fy;
      const double Xc = c == 0 ? Pt.x : Pt.y;
      const D4 g = c == 0 ? D4{fx * iz, 0.0, -fx * Pt.x * iz * iz, 0.0}
                          : D4{0.0, fy * iz, -fy * Pt.y * iz * iz, 0.0};
      const double pix = f * Xc * iz + (c == 0 ? cx : cy);
      const double wc = 0.001 * static_cast<double>(wgt[(int64_t)c * HW + k]);
      const double r = static_cast<double>(tgt[(int64_t)c * HW + k]) - pix;
      double J[13];
#pragma unroll
      for (int s = 0; s < 13; ++s) J[s] = dot4(g, Ps[s]);
      const double a = dot4(g, Pl), b = dot4(g, Py);
      // damping term lambda^T D(w) y of this row: diagonal of H (schur_solve) or the
      // diagonal of every 6x6 block (block_solve), i.e. sum_i Lam_i Y_i with
      // Lam_i = li_i Ji_i + lj_i Jj_i and Y_i = yi_i Ji_i + yj_i Jj_i
      double beta[13], damp = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        beta[q] = li[q] * (r - b) - a * yi[q];
        beta[6 + q] = lj[q] * (r - b) - a * yj[q];
        if (block_damp) {
          const double Lq = li[q] * J[q] + lj[q] * J[6 + q], Yq = yi[q] * J[q] + yj[q] * J[6 + q];
          damp += Lq * Yq;
          beta[q] -= lm * (li[q] * Yq + yi[q] * Lq);
          beta[6 + q] -= lm * (lj[q] * Yq + yj[q] * Lq);
        } else {
          const double ki = li[q] * yi[q], kj = lj[q] * yj[q];
          damp += ki * J[q] * J[q] + kj * J[6 + q] * J[6 + q];
          beta[q] -= 2.0 * lm * ki * J[q];
          beta[6 + q] -= 2.0 * lm * kj * J[6 + q];
        }
      }
      beta[12] = lzk * (r - b) - a * yzk;
      gtg[(int64_t)c * HW + k] = static_cast<float>(wc * a);
      gwt[(int64_t)c * HW + k] = static_cast<float>(0.001 * (a * r - a * b - lm * damp));
      double Ub[6];
      adj(R, tg, beta, Ub);
      const D4 UbP = twist(Ub, Pt);
      const D4 bjP = twist(beta + 6, Pt);
      const D4 Pb = axpy(beta[12], tt, sub(bjP, UbP));
      const D4 inner = axpy(beta[12], tt, D4{-UbP.x, -UbP.y, -UbP.z, -UbP.w});
      // A^T (d2 pi_c) B
      auto hess = [&](const D4& A, const D4& B) {
        const double Acomp = c == 0 ? A.x : A.y, Bcomp = c == 0 ? B.x : B.y;
        return f * (-(Acomp * B.z + A.z * Bcomp) * iz * iz + 2.0 * Xc * A.z * B.z * iz * iz * iz);
      };
#pragma unroll
      for (int s = 0; s < 13; ++s) {
        D4 Pbs;
        if (s < 6) Pbs = sub(twist(beta + 6, Ps[s]), twist(As[s], inner));
        else if (s < 12) Pbs = add(twist(beta + 6, Ps[s]), gen(s - 6, inner));
        else Pbs = sub(twist(beta + 6, tt), twist(Ub, tt));
        const double v = wc * (dot4(g, Pbs) + hess(Pb, Ps[s]) - a * J[s]);
        if (s < 12) acc[s] += v;
        else sd += v;
      }
    }
    disp_part[(int64_t)e * HW + k] = static_cast<float>(sd);
  }
  block_reduce_store<12, kAdjThreads, double>(acc, lds, red);
  __syncthreads();
  if (threadIdx.x < 12) pose_part[(int64_t)e * 12 + threadIdx.x] = red[threadIdx.x];
}

// eta_bar = -lambda_z y_z; disparity-state gradients summed over the edges of each frame.
__global__ void __launch_bounds__(256)
gn_adjoint_frames(const float* __restrict__ disp_part, const float* __restrict__ lz,
                  const float* __restrict__ dz, const int* __restrict__ frame_eptr,
                  const int* __restrict__ frame_edges, const int* __restrict__ frame_kx, int HW,
                  float* __restrict__ grad_eta, float* __restrict__ grad_disps) {
  const int f = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= HW) return;
  if (grad_eta)
    grad_eta[(int64_t)f * HW + k] = -lz[(int64_t)f * HW + k] * dz[(int64_t)f * HW + k];
  double s = 0.0;
  for (int q = frame_eptr[f]; q < frame_eptr[f + 1]; ++q)
    s += disp_part[(int64_t)frame_edges[q] * HW + k];
  grad_disps[(int64_t)frame_kx[f] * HW + k] = static_cast<float>(s);
}

// tangent-state gradient of every pose: sum of the i-parts of its outgoing edges and the
// j-parts of its incoming edges (deterministic order)
__global__ void gn_adjoint_poses(const double* __restrict__ pose_part, const int* __restrict__ edge_ii,
                                 const int* __restrict__ edge_jj, int E, int N,
                                 float* __restrict__ grad_poses) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 6 * N) return;
  const int k = idx / 6, n = idx % 6;
  double s = 0.0;
  for (int e = 0; e < E; ++e) {
    if (edge_ii[e] == k) s += pose_part[(int64_t)e * 12 + n];
    if (edge_jj[e] == k) s += pose_part[(int64_t)e * 12 + 6 + n];
  }
  grad_poses[idx] = static_cast<float>(s);
}

struct Views {
  const int* blob;
  float *Eall, *Eii, *Cii, *bz, *Q, *w, *lz;
  double *Hs, *vs, *S, *Ev, *Spart, *Hpart, *A, *b, *x, *Ld, *W, *Sinv;
  int* info;
};

Views views(const hsl2_ba_plan* p, char* ws) {
  Views v;
  v.blob = reinterpret_cast<const int*>(ws + p->w_blob);
  v.Eall = reinterpret_cast<float*>(ws + p->w_Eall);
  v.Eii = reinterpret_cast<float*>(ws + p->w_Eii);
  v.Cii = reinterpret_cast<float*>(ws + p->w_Cii);
  v.bz = reinterpret_cast<float*>(ws + p->w_bz);
  v.Q = reinterpret_cast<float*>(ws + p->w_Q);
  v.w = reinterpret_cast<float*>(ws + p->w_w);
  v.lz = reinterpret_cast<float*>(ws + p->w_lz);
  v.Hs = reinterpret_cast<double*>(ws + p->w_Hs);
  v.vs = reinterpret_cast<double*>(ws + p->w_vs);
  v.S = reinterpret_cast<double*>(ws + p->w_S);
  v.Ev = reinterpret_cast<double*>(ws + p->w_Ev);
  v.Spart = reinterpret_cast<double*>(ws + p->w_Spart);
  v.Hpart = reinterpret_cast<double*>(ws + p->w_Hpart);
  v.A = reinterpret_cast<double*>(ws + p->w_A);
  v.b = reinterpret_cast<double*>(ws + p->w_b);
  v.x = reinterpret_cast<double*>(ws + p->w_x);
  v.Ld = reinterpret_cast<double*>(ws + p->w_Ld);
  v.W = reinterpret_cast<double*>(ws + p->w_W);
  v.Sinv = reinterpret_cast<double*>(ws + p->w_Sinv);
  v.info = reinterpret_cast<int*>(ws + p->w_info);
  return v;
}

unsigned blocks_for(int64_t n, int t = 256) { return static_cast<unsigned>((n + t - 1) / t); }


int check_common(const hsl2_ba_plan* plan, const void* workspace, size_t bytes) {
  HSL2_REQUIRE(plan, HSL2_ERR_ARGUMENT, "gn_step: null plan");
  HSL2_REQUIRE(plan->semantics == HSL2_SEM_TORCH, HSL2_ERR_ARGUMENT,
               "gn_step: the plan was not created with torch semantics");
  HSL2_REQUIRE(bytes >= plan->total, HSL2_ERR_WORKSPACE, "gn_step: workspace too small");
  HSL2_REQUIRE(workspace && (reinterpret_cast<uintptr_t>(workspace) % 256) == 0, HSL2_ERR_ARGUMENT,
               "gn_step: workspace must be 256-byte aligned");
  return HSL2_OK;
}

}  // namespace
}  // namespace hsl2

using namespace hsl2;

extern "C" int hsl2_gn_step(const hsl2_ba_plan* p, float* poses, float* disps,
                            const float* intrinsics, const float* targets, const float* weights,
                            const float* eta, float lm, float ep, float* dx, float* dz,
                            float* dzcov, void* workspace, size_t workspace_bytes, int* solve_ok,
                            void* stream) {
  if (int st = check_common(p, workspace, workspace_bytes)) return st;
  HSL2_REQUIRE(poses && disps && intrinsics && dx, HSL2_ERR_ARGUMENT, "gn_step: null pointer");
  HSL2_REQUIRE(p->E == 0 || (targets && weights), HSL2_ERR_ARGUMENT, "gn_step: null targets/weights");
  const bool schur = !p->motion_only;
  HSL2_REQUIRE(!schur || (eta && dz), HSL2_ERR_ARGUMENT, "gn_step: null eta or dz");
  hipStream_t s = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  HSL2_HIP(hipMemcpyAsync(ws + p->w_blob, p->blob.data(), p->blob.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, s));
  const Views v = views(p, ws);
  const int E = static_cast<int>(p->E), P = p->P, HW = static_cast<int>(p->HW);
  const int F = static_cast<int>(p->F), wd = static_cast<int>(p->wd);
  const int64_t n_pad = p->n_pad;
  hipLaunchKernelGGL(ba_clear_flags, dim3(1), dim3(64), 0, s, v.info);
  HSL2_LAUNCH_CHECK();
  if (E > 0) {
    const int zl = p->lin_slices;
    if (schur)
      hipLaunchKernelGGL((ba_linearize<true, kSemTorch>), dim3(E, zl), dim3(kLinThreads), 0, s, poses, disps,
                         intrinsics, 4, targets, weights, v.blob + p->o_edge_ii, v.blob + p->o_edge_jj,
                         HW, wd, E, v.Hs, v.vs, v.Eall + (int64_t)P * 6 * HW, v.Eii, v.Cii, v.bz, zl, v.Hpart);
    else
      hipLaunchKernelGGL((ba_linearize<false, kSemTorch>), dim3(E, zl), dim3(kLinThreads), 0, s, poses, disps,
                         intrinsics, 4, targets, weights, v.blob + p->o_edge_ii, v.blob + p->o_edge_jj,
                         HW, wd, E, v.Hs, v.vs, nullptr, nullptr, nullptr, nullptr, zl, v.Hpart);
    HSL2_LAUNCH_CHECK();
    if (zl > 1) {
      hipLaunchKernelGGL(lin_reduce<double>, dim3(blocks_for(int64_t(E) * 156)), dim3(256), 0, s, v.Hpart, zl, E,
                         v.Hs, v.vs);
      HSL2_LAUNCH_CHECK();
    }
  }
  if (schur) {
    launch_frame_reduce(v.Cii, v.bz, v.Eii, eta, static_cast<int>(F), v.blob + p->o_frame_eptr,
                        v.blob + p->o_frame_edges, v.blob + p->o_frame_pose, v.blob + p->o_frame_order, HW,
                        static_cast<int>(F), 1e-7f, v.Q, v.w, v.Eall, s);
    HSL2_LAUNCH_CHECK();
    launch_schur_gram(p, v.blob, v.Eall, v.Q, v.w, v.S, v.Ev, v.Spart, s);
    HSL2_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(ba_assemble<double>, dim3(blocks_for(n_pad * n_pad)), dim3(256), 0, s, v.Hs, v.S,
                     v.blob + p->o_blk_ptr, v.blob + p->o_blk_src, E, P, n_pad, lm, ep, schur ? 1 : 2, v.A,
                     v.info);
  HSL2_LAUNCH_CHECK();
  hipLaunchKernelGGL(ba_assemble_rhs<double>, dim3(static_cast<unsigned>((n_pad / 6 + 4) / 4)), dim3(256), 0, s,
                     v.vs, v.Ev, v.blob + p->o_rhs_ptr, v.blob + p->o_rhs_src, E, P, static_cast<int>(n_pad), v.b);
  HSL2_LAUNCH_CHECK();
  if (int st = cholesky_solve<double>(v.A, n_pad, v.Ld, v.b, v.x, v.info, s)) return st;
  if (schur) {
    hipLaunchKernelGGL(ba_backsub<double>, dim3(blocks_for(HW), F), dim3(256), 0, s, v.Eall, v.Q, v.w,
                       v.x, v.info, v.blob + p->o_slot_ptr, v.blob + p->o_slot_row,
                       v.blob + p->o_slot_pose, v.blob + p->o_frame_kx, HW, 0, disps, dz);
    HSL2_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(ba_pose_retr<double>, dim3(blocks_for(P, 64)), dim3(64), 0, s, poses, v.x, v.info,
                     p->t0, P, 1, dx);
  HSL2_LAUNCH_CHECK();
  if (schur) {
    const int64_t nd = p->num_poses * HW;
    hipLaunchKernelGGL(ba_clamp_disps, dim3(blocks_for(nd)), dim3(256), 0, s, disps, nd);
    HSL2_LAUNCH_CHECK();
  }
  if (schur && dzcov) {
    const int T = static_cast<int>(n_pad / 64);
    hipLaunchKernelGGL(tri_inv_init, dim3(blocks_for(n_pad * n_pad)), dim3(256), 0, s, v.Ld, n_pad, v.W);
    HSL2_LAUNCH_CHECK();
    for (int i = 1; i < T; ++i) {
      hipLaunchKernelGGL(tri_inv_row, dim3(i), dim3(256), 0, s, v.A, v.Ld, n_pad, i, v.W);
      HSL2_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(gram_wtw, dim3(T * T), dim3(256), 0, s, v.W, n_pad, T, v.Sinv);
    HSL2_LAUNCH_CHECK();
    hipLaunchKernelGGL(dzcov_kernel, dim3(blocks_for(HW), F), dim3(256), 0, s, v.Eall, v.Q, v.Sinv, n_pad,
                       v.blob + p->o_slot_ptr, v.blob + p->o_slot_row, v.blob + p->o_slot_pose, HW,
                       v.info, dzcov);
    HSL2_LAUNCH_CHECK();
  }
  if (solve_ok) {
    int flag = 0;
    HSL2_HIP(hipMemcpyAsync(&flag, ws + p->w_info, sizeof(int), hipMemcpyDeviceToHost, s));
    HSL2_HIP(hipStreamSynchronize(s));
    *solve_ok = flag == 0 ? 1 : 0;
  }
  return HSL2_OK;
}

extern "C" int hsl2_gn_step_backward(const hsl2_ba_plan* p, const float* poses, const float* disps,
                                     const float* intrinsics, const float* targets,
                                     const float* weights, float lm, float ep, const float* dx,
                                     const float* dz, const float* grad_dx, const float* grad_dz,
                                     float* grad_targets, float* grad_weights, float* grad_eta,
                                     float* grad_poses, float* grad_disps, void* workspace,
                                     size_t workspace_bytes, void* stream) {
  if (int st = check_common(p, workspace, workspace_bytes)) return st;
  HSL2_REQUIRE(poses && disps && intrinsics && dx && grad_dx && grad_poses && grad_disps,
               HSL2_ERR_ARGUMENT, "gn_step backward: null pointer");
  HSL2_REQUIRE(p->E == 0 || (targets && weights && grad_targets && grad_weights), HSL2_ERR_ARGUMENT,
               "gn_step backward: null targets/weights or their gradients");
  const bool schur = !p->motion_only;
  HSL2_REQUIRE(!schur || (dz && grad_dz && grad_eta), HSL2_ERR_ARGUMENT,
               "gn_step backward: null dz, its gradient or the eta gradient");
  hipStream_t s = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  const Views v = views(p, ws);
  const int E = static_cast<int>(p->E), P = p->P, HW = static_cast<int>(p->HW);
  const int F = static_cast<int>(p->F), wd = static_cast<int>(p->wd);
  const int N = static_cast<int>(p->num_poses);
  const int64_t n_pad = p->n_pad;
  // lambda_x = S^-1 (dx_bar - E Q dz_bar): the Ev tasks of the Gram kernel with w := dz_bar,
  // the reduced system re-assembled from the forward's Hs and S, the same tiled Cholesky
  if (schur) {
    launch_schur_ev(p, v.blob, v.Eall, v.Q, grad_dz, v.Ev, s);
    HSL2_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(ba_assemble<double>, dim3(blocks_for(n_pad * n_pad)), dim3(256), 0, s, v.Hs, v.S,
                     v.blob + p->o_blk_ptr, v.blob + p->o_blk_src, E, P, n_pad, lm, ep, schur ? 1 : 2, v.A,
                     v.info);
  HSL2_LAUNCH_CHECK();
  hipLaunchKernelGGL(adj_rhs, dim3(blocks_for(n_pad)), dim3(256), 0, s, grad_dx, v.Ev,
                     v.blob + p->o_rhs_ptr, v.blob + p->o_rhs_src, P, static_cast<int>(n_pad), v.b);
  HSL2_LAUNCH_CHECK();
  if (int st = cholesky_solve<double>(v.A, n_pad, v.Ld, v.b, v.x, v.info, s)) return st;
  // lambda_z = Q (dz_bar - E^T lambda_x)
  if (schur) {
    hipLaunchKernelGGL(ba_backsub<double>, dim3(blocks_for(HW), F), dim3(256), 0, s, v.Eall, v.Q, grad_dz,
                       v.x, v.info, v.blob + p->o_slot_ptr, v.blob + p->o_slot_row,
                       v.blob + p->o_slot_pose, v.blob + p->o_frame_kx, HW, 0, nullptr, v.lz);
    HSL2_LAUNCH_CHECK();
  }
  // per-pixel adjoint; the per-edge pose parts reuse the vs region (the forward's gradient
  // blocks are not needed again), the per-edge disparity parts have their own region
  double* pose_part = v.vs;                                                   // [E][12]
  float* disp_part = reinterpret_cast<float*>(ws + p->w_dpart);               // [E][HW]
  if (E > 0) {
    hipLaunchKernelGGL(gn_adjoint_edges, dim3(E), dim3(kAdjThreads), 0, s, poses, disps, intrinsics,
                       targets, weights, v.blob + p->o_edge_ii, v.blob + p->o_edge_jj,
                       v.blob + p->o_edge_frame, p->t0, P, HW, wd, E, lm, schur ? 0 : 1, v.x, v.info, dx,
                       schur ? dz : nullptr, schur ? v.lz : nullptr, grad_targets, grad_weights,
                       pose_part, disp_part);
    HSL2_LAUNCH_CHECK();
  }
  const int64_t nd = p->num_poses * HW;
  hipLaunchKernelGGL(fill_zero, dim3(blocks_for(nd)), dim3(256), 0, s, grad_disps, nd);
  HSL2_LAUNCH_CHECK();
  hipLaunchKernelGGL(gn_adjoint_frames, dim3(blocks_for(HW), F), dim3(256), 0, s, disp_part,
                     schur ? v.lz : nullptr, schur ? dz : nullptr, v.blob + p->o_frame_eptr,
                     v.blob + p->o_frame_edges, v.blob + p->o_frame_kx, HW, schur ? grad_eta : nullptr,
                     grad_disps);
  HSL2_LAUNCH_CHECK();
  hipLaunchKernelGGL(gn_adjoint_poses, dim3(blocks_for(6 * (int64_t)N)), dim3(256), 0, s, pose_part,
                     v.blob + p->o_edge_ii, v.blob + p->o_edge_jj, E, N, grad_poses);
  HSL2_LAUNCH_CHECK();
  return HSL2_OK;
}
