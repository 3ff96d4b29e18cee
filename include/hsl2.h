/*
 * hsl2.h -- C ABI of the MI355X (gfx950) dense-tracking library.
 *
 * Plain pointers and sizes only: no torch types cross this boundary.  Every
 * entry point returns an int status (HSL2_OK == 0); a human-readable detail of
 * the last failure on the calling thread is available from hsl2_last_error().
 * Device pointers must be allocated on the device that is current when the
 * call is made; `stream` is a hipStream_t passed as void* (NULL = the null
 * stream).  No entry point allocates device memory: scratch is supplied by the caller,
 * except for hsl2_ba_run_static, whose workspace is a static device array of the library.
 *
 * Each function names the reference interface it replaces (paths relative to
 * the HI-SLAM2 source tree).
 */
#ifndef HSL2_H_
#define HSL2_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum {
  HSL2_OK = 0,
  HSL2_ERR_ARGUMENT = 1,   /* null pointer, unknown dtype, bad scalar argument */
  HSL2_ERR_SHAPE = 2,      /* sizes inconsistent with each other */
  HSL2_ERR_INDEX = 3,      /* an edge / frame index outside its valid range */
  HSL2_ERR_HIP = 4,        /* a HIP runtime call failed */
  HSL2_ERR_WORKSPACE = 5,  /* caller-supplied workspace too small */
  HSL2_ERR_UNSUPPORTED = 6 /* valid request this build does not implement */
};

/* ---- element types -------------------------------------------------------- */
enum {
  HSL2_F32 = 0,
  HSL2_F16 = 1,
  HSL2_BF16 = 2,
  HSL2_F64 = 3
};

/* ---- BA precision modes ---------------------------------------------------- */
enum {
  HSL2_BA_SPLIT = 0,  /* FP32 linearization + Schur Gram, FP64 reduced Cholesky and back-substitution */
  HSL2_BA_FP32 = 1    /* the same kernels all in FP32: per-edge Hessian blocks, frame sums, Schur Gram
                         totals, reduced Cholesky and back-substitution */
};

/* ---- BA semantics ------------------------------------------------------------- */
enum {
  HSL2_SEM_CUDA = 0,  /* ba_cuda of src/droid_kernels.cu:1820-1937 (the droid_backends.ba binding) */
  HSL2_SEM_TORCH = 1  /* one GN step of the torch formulation, hislam2/geom/ba.py:32-158 + chol.py */
};

int hsl2_version(void);
const char* hsl2_status_string(int status);
const char* hsl2_last_error(void);

/* ---- correlation lookup ---------------------------------------------------- */

/* Replaces corr_index_forward (src/droid.cpp:254-262 -> src/correlation_kernels.cu:126-155).
 * volume [batch, h1, w1, h2, w2] of `dtype` (F32/F16/BF16/F64), coords [batch, 2, h1, w1] f32
 * (x then y), corr [batch, 2r+1, 2r+1, h1, w1] of `dtype`, fully written (no pre-zeroing). */
int hsl2_corr_index_forward(const void* volume, const float* coords, void* corr, int dtype,
                            int64_t batch, int64_t h1, int64_t w1, int64_t h2, int64_t w2,
                            int radius, void* stream);

/* Replaces corr_index_backward (src/droid.cpp:264-275 -> src/correlation_kernels.cu:157-185).
 * corr_grad [batch, 2r+1, 2r+1, h1, w1] of `dtype`; volume_grad [batch, h1, w1, h2, w2] of
 * `dtype` is fully written (zeros outside the taps). */
int hsl2_corr_index_backward(const float* coords, const void* corr_grad, void* volume_grad,
                             int dtype, int64_t batch, int64_t h1, int64_t w1, int64_t h2,
                             int64_t w2, int radius, void* stream);

/* Replaces CorrBlock.__call__ (hislam2/modules/corr.py:40-50: one corr_index_forward per
 * level on coords / 2^level, then torch.cat) with one launch.  levels[i] is the level-i volume
 * [batch, h1, w1, h1 >> i, w1 >> i] of `dtype`; coords [batch, h1, w1, 2] f32 (x, y
 * interleaved, the CorrBlock call layout); corr [batch, L*(2r+1)^2, h1, w1] of `dtype`, channel
 * order [level][x-offset][y-offset], fully written.  tiled != 0: the levels are in the tiled
 * pyramid layout that hsl2_corr_build writes with tiled != 0. */
int hsl2_corr_pyramid_lookup(const void* const* levels, int num_levels, const float* coords,
                             void* corr, int dtype, int64_t batch, int64_t h1, int64_t w1,
                             int radius, int tiled, void* stream);

/* No reference counterpart (the next step of the path: the lookup fused into its consumer).
 * The CorrBlock lookup of hsl2_corr_pyramid_lookup (4 bf16 levels, radius 3, 196 channels in
 * its channel order) followed by corr_encoder's first layer (hislam2/modules/droid_net.py:86-90:
 * a 1x1 convolution to cout channels, bias, ReLU when relu != 0), without storing the 196
 * channels: out [batch, cout, h1, w1] bf16 = act(bias + weight . corr), weight [cout, 196] bf16,
 * bias [cout] f32; each lookup channel is rounded to bf16 (as the lookup's own output) before the
 * product, which accumulates in FP32.  cout 64 or 128; h1 * w1 a multiple of 4; every level
 * width a multiple of 8. */
int hsl2_corr_lookup_conv(const void* const* levels, int num_levels, const float* coords, const void* weight,
                          const float* bias, int relu, void* out, int64_t batch, int64_t h1, int64_t w1,
                          int radius, int tiled, int64_t cout, void* stream);

/* ---- correlation build ------------------------------------------------------ */

/* Replaces CorrBlock.__init__ + CorrBlock.corr (hislam2/modules/corr.py:24-38, 63-71):
 * all-pairs correlation (f1/4)^T (f2/4) of fmap1/fmap2 followed by `num_levels` 2x2
 * average-pool levels.  fmap1/fmap2 are in MFMA FRAGMENT TILES: per batch item
 * [h w / 32][C / G][32 pixels][G channels], G = 16 for BF16/F16 and 8 for F32 (`in_dtype`;
 * hsl2_fmap_tiles converts the reference's [batch, C, h, w]).  levels[i]
 * points at [batch, h, w, h >> i, w >> i] of `out_dtype` (F32/BF16).  Requires h, w multiples
 * of 8 and C a multiple of 16.  tiled == 0 gives the reference's row-major planes
 * (element (pix, y, x) at pix * h2 w2 + y w2 + x, pix = item * h w + pixel).  tiled != 0 (read
 * by hsl2_corr_pyramid_lookup with tiled != 0; the same element count per level):
 *   level 0, h2 and w2 multiples of 8: each plane as row-major 8x8 tiles, each tile row-major;
 *   levels 1 and 2, h2 even, w2 a multiple of 4 and h w a multiple of 64: pixel groups
 *     g = pix / 64 of 64 * h2 w2 elements, laid out [y / 2][x / 4][pix % 64][y % 2][x % 4];
 *   levels >= 1 otherwise, w2 a multiple of 8 and h w a multiple of 64: pixel groups laid out
 *     [y][x / 8][pix % 64][x % 8];
 *   any other level: 8x8 tiles when h2 and w2 are multiples of 8, else row-major planes. */
int hsl2_corr_build(const void* fmap1, const void* fmap2, int in_dtype, void* const* levels,
                    int out_dtype, int64_t batch, int64_t channels, int64_t h, int64_t w,
                    int num_levels, int tiled, void* stream);

/* ---- on-the-fly correlation ------------------------------------------------- */

/* Layout helper: [rows, C, P] -> [rows, P, C] for `dtype` (F32/F16/BF16).  The altcorr
 * kernels read channel-last feature maps so that a tap's channel vector is contiguous. */
int hsl2_channels_last(const void* in, void* out, int dtype, int64_t rows, int64_t channels,
                       int64_t pixels, void* stream);
/* Layout helper for hsl2_corr_build: [rows, C, P] -> MFMA fragment tiles
 * [rows, P / 32, C / G, 32, G] (G = 32 / element size: 16 for F16/BF16, 8 for F32), so that one
 * MFMA operand fragment (32 pixels x G channels) is contiguous.  P a multiple of 32, C of G. */
int hsl2_fmap_tiles(const void* in, void* out, int dtype, int64_t rows, int64_t channels,
                    int64_t pixels, void* stream);
/* Layout helper: [rows, P, C] -> [rows, C, P], f32. */
int hsl2_channels_first_f32(const float* in, float* out, int64_t rows, int64_t channels,
                            int64_t pixels, void* stream);

/* Replaces altcorr_forward (src/droid.cpp:277-289 -> src/altcorr_kernel.cu:132-172).
 * fmap1 [B, N1, H, W, C], fmap2 [B, N2, H2, W2, C] CHANNEL-LAST of `dtype` (F32/F16/BF16;
 * hsl2_channels_last converts the reference's [B, N, C, H, W]), coords [B, M, 2, H, W] f32,
 * ii/jj [M] int64 (device) index fmap1/fmap2 (an out-of-range edge yields zeros),
 * corr [B, M, 2r+1 (x-offset), 2r+1 (y-offset), H, W] of `dtype`, fully written.
 * C must be a multiple of 8. */
int hsl2_altcorr_forward(const void* fmap1, const void* fmap2, const float* coords,
                         const int64_t* ii, const int64_t* jj, void* corr, int dtype,
                         int64_t B, int64_t N1, int64_t N2, int64_t C, int64_t H, int64_t W,
                         int64_t H2, int64_t W2, int64_t M, int radius, void* stream);

/* Replaces altcorr_backward (src/droid.cpp:291-305 -> src/altcorr_kernel.cu:175-225) with
 * the argument order the binding declares and the exact adjoint of the forward (including
 * its 1/16 scaling).  Same layouts as the forward; corr_grad is f32; fmap1_grad
 * [B, N1, H, W, C] and fmap2_grad [B, N2, H2, W2, C] are f32 channel-last, zeroed and then
 * accumulated by this call. */
int hsl2_altcorr_backward(const void* fmap1, const void* fmap2, const float* coords,
                          const float* corr_grad, const int64_t* ii, const int64_t* jj,
                          float* fmap1_grad, float* fmap2_grad, int dtype,
                          int64_t B, int64_t N1, int64_t N2, int64_t C, int64_t H, int64_t W,
                          int64_t H2, int64_t W2, int64_t M, int radius, void* stream);

/* Layout helper for the on-the-fly correlation pyramid (AltCorrBlock.__init__,
 * hislam2/modules/corr.py:92-104, which pools the feature maps with F.avg_pool2d level by
 * level): out [frames, h >> level, w >> level, channels] = the mean of each 2^level x 2^level
 * block of the channel-last level-0 maps level0 [frames, h, w, channels] (F32/F16/BF16),
 * summed in FP32 and rounded once.  channels a multiple of 8, buffers 16-byte aligned. */
int hsl2_pool_channels_last(const void* level0, void* out, int dtype, int64_t frames, int64_t h,
                            int64_t w, int64_t channels, int level, void* stream);

/* Replaces AltCorrBlock.__call__ (hislam2/modules/corr.py:107-117: one altcorr_forward per
 * level, src/altcorr_kernel.cu:24-75 + 132-172, on coords / 2^level, then stack + flatten)
 * with one launch on the matrix cores.  levels[l] = level l of the CHANNEL-LAST feature
 * pyramid [B, N, H >> l, W >> l, C] (level 0 from hsl2_channels_last, the others from
 * hsl2_pool_channels_last) of `dtype` (BF16, F16 or F32; the dot products run on MFMA with FP32
 * accumulation); fmap1 is level 0 of frame ii[m], fmap2 level l of frame jj[m] (an edge naming
 * a frame outside [0, N) yields zeros).  coords [B, M, H, W, 2] f32 (x, y interleaved: the
 * AltCorrBlock call layout, without its permute), ii / jj [M] int64 on the device, corr
 * [B, M, num_levels (2r+1)^2, H, W] of `dtype`, channel order [level][x-offset][y-offset],
 * fully written.  C must be 128, radius 3 or 4, num_levels 1 to 4. */
int hsl2_altcorr_pyramid_forward(const void* const* levels, int num_levels, const float* coords,
                                 const int64_t* ii, const int64_t* jj, void* corr, int dtype,
                                 int64_t B, int64_t N, int64_t C, int64_t H, int64_t W, int64_t M,
                                 int radius, void* stream);

/* FP8 (OCP e4m3fn) feature maps with per-channel FP32 scales, for the on-the-fly correlation
 * of high-resolution frames (no counterpart in the reference, which runs altcorr in the fmap
 * dtype).  hsl2_fp8_quantize: x is `vectors` channel-last vectors of `channels` values (F32 or
 * BF16); scales[c] = max_v |x[v][c]| / 448 (1 for an all-zero channel) and q[v][c] =
 * e4m3(x[v][c] / scales[c]), round to nearest even.  hsl2_altcorr_forward_fp8: the lookup of
 * hsl2_altcorr_forward on such maps (channel-last, C a multiple of 16), dequantised on the fly
 * (s1_c s2_c folded into fmap1); out_dtype F32 or BF16. */
int hsl2_fp8_quantize(const void* x, int dtype, int64_t vectors, int64_t channels, uint8_t* q,
                      float* scales, void* stream);
int hsl2_altcorr_forward_fp8(const uint8_t* fmap1, const float* scale1, const uint8_t* fmap2,
                             const float* scale2, const float* coords, const int64_t* ii,
                             const int64_t* jj, void* corr, int out_dtype, int64_t B, int64_t N1,
                             int64_t N2, int64_t C, int64_t H, int64_t W, int64_t H2, int64_t W2,
                             int64_t M, int radius, void* stream);

/* ---- dense bundle adjustment ------------------------------------------------- */

typedef struct hsl2_ba_plan hsl2_ba_plan;

/* Host-side preparation of one ba call (the index work of src/droid_kernels.cu:1837-1849,
 * 1413-1463 and 1689-1746, done once per call instead of per iteration).
 * ii, jj: HOST arrays of num_edges frame indices.  Requires 0 <= t0 < t1 <= num_poses and
 * 0 <= ii, jj < t1.  num_poses = rows of the poses buffer, ht x wd = disparity map size. */
int hsl2_ba_plan_create(const int64_t* ii, const int64_t* jj, int64_t num_edges,
                        int64_t num_poses, int64_t ht, int64_t wd, int t0, int t1,
                        int motion_only, int precision, hsl2_ba_plan** plan);
/* hsl2_ba_plan_create with a choice of semantics.  HSL2_SEM_TORCH plans one step of the torch
 * formulation (BA / MoBA of hislam2/geom/ba.py:32-158): t0 plays `fixedp` (poses before it are
 * held fixed), t1 = num_poses, the depth frames are unique(ii) (ba.py:70) and the precision
 * must be HSL2_BA_SPLIT.  motion_only = 1 plans MoBA (ba.py:110-158, rig = 1). */
int hsl2_ba_plan_create_ex(const int64_t* ii, const int64_t* jj, int64_t num_edges,
                           int64_t num_poses, int64_t ht, int64_t wd, int t0, int t1,
                           int motion_only, int precision, int semantics, hsl2_ba_plan** plan);
void hsl2_ba_plan_destroy(hsl2_ba_plan* plan);
/* number of unique depth frames |kx| (rows of eta and of dz) */
int64_t hsl2_ba_plan_num_frames(const hsl2_ba_plan* plan);
/* copies kx (sorted unique of [t0,t1) ++ ii) into a host array of num_frames entries */
int hsl2_ba_plan_frames(const hsl2_ba_plan* plan, int64_t* kx);
/* copies the inverse index kk_exp (one entry per expanded edge, t1-t0+num_edges) */
int hsl2_ba_plan_inverse(const hsl2_ba_plan* plan, int64_t* kk_exp);
/* bytes of device workspace hsl2_ba_run needs (256-byte aligned base required) */
size_t hsl2_ba_plan_workspace_bytes(const hsl2_ba_plan* plan);
/* No counterpart in the reference (its ba re-derives the index structures on the host every
 * iteration, src/droid_kernels.cu:1837-1849).  A plan's int32 tables are constant: their size in
 * bytes, and registering a device copy.  hsl2_ba_plan_upload_tables copies them on `stream` into
 * device_buffer (hsl2_ba_plan_table_bytes bytes, 256-byte aligned, on the current device) and
 * records it in the plan; later runs on that device read the tables there instead of uploading
 * them into the workspace on every call.  The buffer must outlive every run of the plan. */
size_t hsl2_ba_plan_table_bytes(const hsl2_ba_plan* plan);
int hsl2_ba_plan_upload_tables(hsl2_ba_plan* plan, void* device_buffer, void* stream);

/* Replaces ba (src/droid.cpp:141-168 -> src/droid_kernels.cu:1820-1937).
 * Device arrays: poses [num_poses, 7] f32 and disps [num_poses, ht, wd] f32 (updated in
 * place), intrinsics [4] f32, targets/weights [num_edges, 2, ht, wd] f32, eta
 * [eta_rows, ht, wd] f32 with eta_rows == num_frames or 1 (broadcast).
 * Outputs: dx [t1-t0, 6] f32 (last iteration's pose update), dz [num_frames, ht*wd] f32
 * (last iteration's depth update; may be NULL; ignored when motion_only).
 * *solve_ok (host, may be NULL) receives 1 if every reduced Cholesky succeeded; reading it
 * synchronizes the stream.  A failed Cholesky gives dx = 0 as in the reference. */
int hsl2_ba_run(const hsl2_ba_plan* plan, float* poses, float* disps, const float* intrinsics,
                const float* targets, const float* weights, const float* eta, int64_t eta_rows,
                int iterations, float lm, float ep, float* dx, float* dz,
                void* workspace, size_t workspace_bytes, int* solve_ok, void* stream);

/* hsl2_ba_run with the sensor-disparity argument of the 15-argument ba form named for this path,
 * ba(poses, disps, intrinsics, disps_sens, targets, weights, eta, ii, jj, t0, t1, iterations,
 * lm, ep, motion_only) (upstream DROID-SLAM's RGB-D binding; this reference's ba,
 * src/droid.cpp:141-168, has no such argument).  disps_sens [num_poses, ht, wd] f32 on the
 * device, or NULL (then exactly hsl2_ba_run).  A pixel of depth frame kx[f] whose sensor disparity
 * is positive takes the prior alpha (d + dz - d_sens)^2 / 2, alpha = 0.05, in place of eta:
 * C = sum Cii + alpha, w = sum bz - alpha (d - d_sens); every other pixel's arithmetic is that of
 * hsl2_ba_run, so an all-zero disps_sens gives bit-identical results. */
int hsl2_ba_run_sens(const hsl2_ba_plan* plan, float* poses, float* disps, const float* intrinsics,
                     const float* disps_sens, const float* targets, const float* weights, const float* eta,
                     int64_t eta_rows, int iterations, float lm, float ep, float* dx, float* dz,
                     void* workspace, size_t workspace_bytes, int* solve_ok, void* stream);

/* No counterpart in the reference.  hsl2_ba_run for callers that cannot allocate device
 * memory (code under HIP graph capture, tight tracking loops, harnesses that forbid
 * hipMalloc): the workspace is a static device array of HSL2_BA_STATIC_SCRATCH_BYTES held by
 * the library on each GPU it runs on until the process ends, so the caller passes inputs and
 * outputs only.  Plans needing more workspace (hsl2_ba_plan_workspace_bytes) are rejected with
 * HSL2_ERR_WORKSPACE; frontend windows fit (12 keyframes, 70 edges at 48x64 need
 * 7.2 MiB).  All calls on one GPU share that array and keep the last call's contents: calls
 * that overlap on one GPU race.  Allocates nothing; otherwise identical to hsl2_ba_run. */
#define HSL2_BA_STATIC_SCRATCH_BYTES (64u << 20)
int hsl2_ba_run_static(const hsl2_ba_plan* plan, float* poses, float* disps, const float* intrinsics,
                       const float* targets, const float* weights, const float* eta, int64_t eta_rows,
                       int iterations, float lm, float ep, float* dx, float* dz, int* solve_ok,
                       void* stream);

/* One Gauss-Newton step in the torch semantics: replaces BA (hislam2/geom/ba.py:32-107, with
 * schur_solve and CholeskySolver of hislam2/geom/chol.py:5-78) and, for a motion-only plan,
 * MoBA (ba.py:110-158 with block_solve).  Device arrays: poses [num_poses, 7] and disps
 * [num_poses, ht, wd] f32 updated in place (disps > 10 -> 0, then clamped at 0.001 as in
 * ba.py:104-105), intrinsics [num_poses, 4] f32 (one row per frame), targets / weights
 * [num_edges, 2, ht, wd] f32, eta [num_frames, ht, wd] f32 (num_frames = |unique(ii)|).
 * Outputs: dx [num_poses - t0, 6] f32; dz [num_frames, ht * wd] f32 (the raw Schur solution,
 * before the clamp); dzcov [num_frames, ht * wd] f32 (may be NULL) = sum(square(L^-1 E Q)) + Q
 * (chol.py:74-76).  A failed factorisation gives dx = 0 (chol.py:14-18) and dzcov = Q.
 * The workspace must stay untouched until hsl2_gn_step_backward of the same step. */
int hsl2_gn_step(const hsl2_ba_plan* plan, float* poses, float* disps, const float* intrinsics,
                 const float* targets, const float* weights, const float* eta, float lm, float ep,
                 float* dx, float* dz, float* dzcov, void* workspace, size_t workspace_bytes,
                 int* solve_ok, void* stream);

/* Adjoint of hsl2_gn_step: the backward that autograd forms through BA with CholeskySolver's
 * implicit adjoint (hislam2/geom/chol.py:22-31), written out and evaluated per pixel with the
 * Jacobians recomputed.  Inputs: the step's INPUT state (poses, disps), intrinsics, targets,
 * weights, lm, ep, the step's outputs dx and dz, and the gradients grad_dx [num_poses-t0, 6] and
 * grad_dz [num_frames, ht*wd] (grad_dz / dz unused for a motion-only plan).  Outputs:
 * grad_targets / grad_weights [num_edges, 2, ht, wd], grad_eta [num_frames, ht, wd],
 * grad_poses [num_poses, 6] (tangent: left perturbation Exp(xi) T of each input pose) and
 * grad_disps [num_poses, ht, wd] (through the linearization only; the identity path of
 * disps + dz is the caller's).  Uses the workspace the forward step left. */
int hsl2_gn_step_backward(const hsl2_ba_plan* plan, const float* poses, const float* disps,
                          const float* intrinsics, const float* targets, const float* weights,
                          float lm, float ep, const float* dx, const float* dz,
                          const float* grad_dx, const float* grad_dz, float* grad_targets,
                          float* grad_weights, float* grad_eta, float* grad_poses,
                          float* grad_disps, void* workspace, size_t workspace_bytes, void* stream);

/* Replaces projective_transform of the torch formulation (hislam2/geom/projective_ops.py:44-74
 * with pinhole.py:6-53) for one batch element.  Device arrays: poses [N, 7], disps [N, ht, wd],
 * intrinsics [N, 4] f32, ii / jj [num_edges] int64 (must index [0, N)).  Outputs: coords
 * [num_edges, ht, wd, 2] (3 with return_depth: x, y, d / Z), valid [num_edges, ht, wd] (may be
 * NULL), and, all or none, Ji / Jj [num_edges, ht, wd, 2, 6], Jz [num_edges, ht, wd, 2]. */
int hsl2_projective_transform(const float* poses, const float* disps, const float* intrinsics,
                              const int64_t* ii, const int64_t* jj, int64_t num_edges, int64_t ht,
                              int64_t wd, int return_depth, float* coords, float* valid, float* Ji,
                              float* Jj, float* Jz, void* stream);
/* Adjoint of hsl2_projective_transform with respect to the poses (left tangent, grad_poses
 * [N, 6]) and the disparities (grad_disps [N, ht, wd]), given grad_coords shaped like coords;
 * the Z < 0.1 branch of the projection has zero derivative in Z, as in the reference's code.
 * The workspace (256-byte aligned) holds per-edge partial sums. */
size_t hsl2_projective_transform_workspace_bytes(int64_t num_edges, int64_t ht, int64_t wd);
int hsl2_projective_transform_backward(const float* poses, const float* disps, const float* intrinsics,
                                       const int64_t* ii, const int64_t* jj, int64_t num_edges,
                                       int64_t num_poses, int64_t ht, int64_t wd, int return_depth,
                                       const float* grad_coords, float* grad_poses, float* grad_disps,
                                       void* workspace, size_t workspace_bytes, void* stream);

/* ---- edge-selection and point-cloud geometry (src/droid.cpp:170-241) ------------------------
 * poses [N, 7] (tx ty tz qx qy qz qw), disps [N, ht, wd], intrinsics [4] (fx fy cx cy), all f32 on
 * the device; index arrays int64.  Pixel arithmetic in FP32 as the reference, per-pair sums in
 * FP64. */

/* Replaces frame_distance_cuda (src/droid_kernels.cu:860-992, 2047-2069): dist[e] = the
 * beta-weighted mean flow magnitude of edge (ii[e], jj[e]) under the full relative motion and
 * under its translation alone, over pixels with transformed depth > 0.25; 1000 when the valid
 * weight is below 75 %. */
int hsl2_frame_distance(const float* poses, const float* disps, const float* intrinsics, const int64_t* ii,
                        const int64_t* jj, int64_t num_edges, int64_t ht, int64_t wd, float beta, float* dist,
                        void* stream);

/* Replaces covis_distance_cuda (src/droid_kernels.cu:994-1092, 2071-2090): dist[e] = the fraction
 * of frame ii[e]'s pixels that project outside the image in both ii[e] - 1 and ii[e] + 1 (both
 * must be frames). */
int hsl2_covis_distance(const float* poses, const float* disps, const float* intrinsics, const int64_t* ii,
                        int64_t num, int64_t ht, int64_t wd, float* dist, void* stream);

/* Replaces depth_filter_cuda (src/droid_kernels.cu:1094-1200, 2093-2117): counter[b][pixel] (f32,
 * every entry written) = the number of neighbours inds[b] - 1, - 2, - 3, + 3, + 4, + 5 in
 * [0, num_frames) whose disparity at one of the four integer corners around the projected pixel
 * agrees with the projected depth within thresh[b] (|1/dj - 1/d| in double). */
int hsl2_depth_filter(const float* poses, const float* disps, const float* intrinsics, const int64_t* inds,
                      const float* thresh, int64_t count, int64_t num_frames, int64_t ht, int64_t wd,
                      float* counter, void* stream);

/* Replaces iproj_cuda (src/droid_kernels.cu:1204-1273, 2120-2143): points [N, ht, wd, 3] = each
 * pixel lifted with its disparity and moved by its frame's pose, divided by the disparity. */
int hsl2_iproj(const float* poses, const float* disps, const float* intrinsics, int64_t num_frames, int64_t ht,
               int64_t wd, float* points, void* stream);

/* ---- mono-prior depth/scale step pieces --------------------------------------------------- */

/* Host only.  Groups edges by their ii frame: kx[num_frames] = sorted unique ii (torch.unique,
 * as proj_trans_cuda and JDSA use it), frame_ptr[num_frames + 1] / frame_edges[num_edges] the
 * edges of each frame in ascending edge order.  kx, frame_ptr, frame_edges need room for
 * num_edges (+1) entries. */
int hsl2_edge_frames(const int64_t* ii, int64_t num_edges, int64_t* kx, int32_t* frame_ptr,
                     int32_t* frame_edges, int64_t* num_frames);

/* Replaces proj_trans_cuda (src/droid_kernels.cu:246-362, 1781-1817; binding
 * src/droid.cpp:121-138): the depth-only terms of every edge summed per ii frame,
 * C[m][k] = sum w Jz^2, w[m][k] = sum w r Jz over the edges of frame kx[m] (ba's pixel geometry:
 * weights x 0.001, no contribution for a transformed depth below 0.25).  poses [N, 7], disps
 * [N, ht, wd], intrinsics [4] (one camera), targets / weights [E, 2, ht, wd], ii / jj int64 [E],
 * frame_ptr / frame_edges from hsl2_edge_frames, all on the device; C, w [num_frames, ht * wd]
 * f32 (FP64 sums). */
int hsl2_proj_trans(const float* poses, const float* disps, const float* intrinsics,
                    const float* targets, const float* weights, const int64_t* ii, const int64_t* jj,
                    const int32_t* frame_ptr, const int32_t* frame_edges, int64_t num_frames,
                    int64_t ht, int64_t wd, float* C, float* w, void* stream);

/* Replaces bi_inter_cuda (src/droid_kernels.cu:1275-1315, 2145-2170; binding
 * src/droid.cpp:244-250): scales [num, hs, ws] sampled bilinearly at grids [num, ht, wd, 2]
 * (x = column, y = row) -> outputs [num, ht, wd]; jacobis [num, ht, wd, hs, ws] receives the
 * interpolation weights (every entry written: zeros except the four taps).  Taps outside the
 * grid contribute nothing (the reference reads and writes outside its tensors there). */
int hsl2_bi_inter(const float* scales, const float* grids, int64_t num, int64_t ht, int64_t wd,
                  int64_t hs, int64_t ws, float* outputs, float* jacobis, void* stream);

/* Replaces the rest of JDSA after proj_trans (hislam2/geom/ba.py:171-219 with
 * schur_solve_mono_prior, hislam2/geom/chol.py:80-108) for scale grids of hs * ws <= 4: the
 * reduced system of the scale parameters is block-diagonal, so one block per frame slot forms
 * it in FP64 from C, w [M, ht*wd] (proj_trans), disps [N, ht*wd], prior [*, ht*wd], dscales
 * [N, hs*ws], eta [M, ht*wd], kx [M] (int64, the sorted unique ii) and slot_src [M] (int32: the
 * rank whose frame fills slot s = kx[rank] - kx[0], or -1), gx [wd] / gy [ht] (the
 * interpolation grid), factors it and applies the update: disps_out[kx[s]] = disps + dz,
 * dscales_out[kx[s]] = dscales + dso, dzcov [M, ht*wd].  Any failed block zeroes every dso
 * (the reference's single Cholesky fails as a whole); dzcov is then Q.  The inputs are only
 * read (disps_out / dscales_out must not alias disps / dscales); workspace of
 * hsl2_jdsa_workspace_bytes(M) bytes. */
size_t hsl2_jdsa_workspace_bytes(int64_t num_frames);
int hsl2_jdsa_solve(const float* C, const float* w, const float* disps, const float* prior, const float* dscales,
                    const float* eta, const int64_t* kx, const int32_t* slot_src, const float* gx, const float* gy,
                    int64_t num_frames, int64_t ht, int64_t wd, int64_t hs, int64_t ws, float alpha, float ep,
                    float lm, void* workspace, size_t workspace_bytes, float* disps_out, float* dscales_out,
                    float* dzcov, void* stream);

/* Dense symmetric positive-definite solve A x = b (column-major n x n, F64 or F32) with the
 * tiled Cholesky that ba uses for its reduced camera system (replaces the host Eigen
 * SimplicialLLT of SparseBlock::solve, src/droid_kernels.cu:1659-1680).  A is not modified;
 * x = 0 and *ok = 0 (host, may be NULL; reading it synchronizes) when a pivot is not
 * positive.  1 <= n <= 32768.  The workspace (256-byte aligned) holds the padded copy and the
 * factors. */
size_t hsl2_spd_solve_workspace_bytes(int dtype, int64_t n);
int hsl2_spd_solve(int dtype, const void* A, const void* b, void* x, int64_t n, void* workspace,
                   size_t workspace_bytes, int* ok, void* stream);

/* ---- pgba: Sim3 pose-graph bundle adjustment ----------------------------------
 * Replaces droid_backends.pgba (src/droid.cpp:171-195 -> pgba_cuda,
 * src/droid_kernels.cu:1940-2044, pixel terms projective_transform_sim3_kernel 608-858).
 * One Gauss-Newton step over the Sim3 poses of [t0, t1) (7 parameters: translation, rotation,
 * scale) and the disparities of the sorted unique frames kx of [t0, t1) ++ ii, with the
 * caller's pose-graph blocks added to the pose system.  The plan (host) holds the index
 * structures: ii, jj [E] and iip, jjp [Np] are host int64 arrays with every index in [0, t1)
 * (HSL2_ERR_INDEX otherwise; the reference leaves such indices undefined), 0 <= t0 < t1 <=
 * num_poses. */
typedef struct hsl2_pgba_plan hsl2_pgba_plan;
int hsl2_pgba_plan_create(const int64_t* ii, const int64_t* jj, int64_t num_edges, const int64_t* iip,
                          const int64_t* jjp, int64_t num_pose_graph, int64_t num_poses, int64_t ht, int64_t wd,
                          int t0, int t1, hsl2_pgba_plan** plan);
void hsl2_pgba_plan_destroy(hsl2_pgba_plan* plan);
/* |kx| and kx itself (host int64 [F]) */
int64_t hsl2_pgba_plan_num_frames(const hsl2_pgba_plan* plan);
int hsl2_pgba_plan_frames(const hsl2_pgba_plan* plan, int64_t* kx);
size_t hsl2_pgba_workspace_bytes(const hsl2_pgba_plan* plan);
/* Device pointers: poses [N, 8] (t, q xyzw, s), disps [N, ht, wd] (updated in place: disps[kx]
 * += dz, no clamp), intrinsics [4], targets / weights [E, 2, ht, wd], eta [F, ht, wd] (or one
 * [ht, wd] row), Hsp [4, Np, 7, 7] (blocks (iip, iip), (iip, jjp), (jjp, iip), (jjp, jjp)), vsp
 * [2, Np, 7]; outputs dx [t1 - t0, 7] and dz [F, ht * wd].  The lower triangle of each 7x7
 * block enters the FP64 factorisation of the damped reduced system (diag += ep + lm * diag);
 * when it fails dx = 0, dz is still applied and *solve_ok = 0.  The back-substitution leaves
 * out pose slot 0, as the reference's EvT6x1_kernel does (src/droid_kernels.cu:1572).  The call
 * synchronizes the stream before it returns.  Workspace: hsl2_pgba_workspace_bytes, 256-byte
 * aligned. */
int hsl2_pgba_run(const hsl2_pgba_plan* plan, const float* poses, float* disps, const float* intrinsics,
                  const float* targets, const float* weights, const float* eta, int64_t eta_rows, const float* Hsp,
                  const float* vsp, float lm, float ep, float* dx, float* dz, void* workspace,
                  size_t workspace_bytes, int* solve_ok, void* stream);

/* Debug/test view of the last iteration's intermediates inside a workspace after
 * hsl2_ba_run: per-edge Hessian blocks Hs [4, E, 6, 6] and gradients vs [2, E, 6], the
 * buffers of src/droid_kernels.cu:1856-1857 but accumulated in f64, and the damped reduced
 * system (f64 column-major [6P, 6P] and rhs [6P] in the split mode, f32 in the plain mode;
 * the rhs is overwritten by the solution).  Device pointers into workspace. */
int hsl2_ba_debug_views(const hsl2_ba_plan* plan, void* workspace, double** Hs, double** vs,
                        void** reduced_matrix, void** reduced_rhs);

#ifdef __cplusplus
}
#endif

#endif /* HSL2_H_ */
