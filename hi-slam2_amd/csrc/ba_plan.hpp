// Host-side plan of one dense-BA call: the index structures of ba_cuda
// (src/droid_kernels.cu:1837-1849), the per-frame segment lists that replace
// accum_cuda's host CSR (src/droid_kernels.cu:1413-1463), and the Schur pair list
// and block-assembly lists that replace schur_block / SparseBlock
// (src/droid_kernels.cu:1584-1779).  Built once per call; the iterations then run
// entirely on the device.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

struct hsl2_ba_plan {
  int64_t E = 0;          // real edges
  int64_t num_poses = 0;  // rows of the poses / disps buffers
  int64_t ht = 0, wd = 0, HW = 0;
  int t0 = 0, t1 = 0, P = 0;
  int motion_only = 0;
  int precision = 0;
  int semantics = 0;      // HSL2_SEM_CUDA or HSL2_SEM_TORCH
  int64_t F = 0;          // unique depth frames |kx|
  int64_t NS = 0;         // Schur slots (expanded edges with a pose slot in [0, P))
  int64_t NPAIR = 0;      // slot pairs (s <= s') inside a frame

  std::vector<int64_t> kx;       // [F]
  std::vector<int64_t> kk_exp;   // [P + E]

  // ---- int32 tables uploaded to the workspace (all in one blob) ----
  std::vector<int32_t> blob;
  // a device copy of blob registered by hsl2_ba_plan_upload_tables and the device it lives on;
  // runs on that device read the tables there instead of uploading blob into the workspace
  const int32_t* dev_blob = nullptr;
  int dev_blob_device = -1;
  // offsets (in int32 elements) of each table inside blob
  int64_t o_edge_ii = 0, o_edge_jj = 0;        // [E]
  int64_t o_edge_frame = 0;                    // [E] depth-frame index of ii[e]
  int64_t o_frame_kx = 0, o_frame_pose = 0;    // [F]
  int64_t o_frame_eptr = 0, o_frame_edges = 0; // [F+1], [E]
  int64_t o_frame_order = 0;                   // [F] frames by descending edge count (stable)
  int64_t o_slot_ptr = 0;                      // [F+1]
  int64_t o_slot_row = 0, o_slot_pose = 0;     // [NS]
  int64_t o_task = 0;                          // gram tasks [NT][4] = (kind, a, b, frame)
  int64_t NT = 0;                              // gram tasks = NPAIR + NS
  int64_t o_blk_ptr = 0, o_blk_src = 0;        // [P*P+1], [*]
  int64_t o_rhs_ptr = 0, o_rhs_src = 0;        // [P+1], [*]
  int64_t o_frame_pair0 = 0;                   // [F] first Schur pair of each frame
  int64_t o_gram_group_item = 0;               // [NGGI][3] Gram group items (frame, row group, column group)
  int64_t NGGI = 0;
  int gram_slices = 1;                         // pixel slices per Gram item (partial sums when > 1)
  int lin_slices = 1;                          // pixel slices per edge in the linearisation

  // ---- workspace layout (byte offsets) ----
  size_t w_blob = 0;
  size_t w_Eall = 0;     // f32 [P+E][6][HW]
  size_t w_Eii = 0;      // f32 [E][6][HW] (torch semantics only)
  size_t w_Cii = 0;      // f32 [E][HW] (torch semantics only)
  size_t w_bz = 0;       // f32 [E][HW] (torch semantics only)
  size_t w_Q = 0;        // f32 [F][HW]
  size_t w_w = 0;        // f32 [F][HW]
  size_t w_Hs = 0;       // f64 [4][E][6][6]
  size_t w_vs = 0;       // f64 [2][E][6]
  size_t w_S = 0;        // f64 [NPAIR][6][6]
  size_t w_Ev = 0;       // f64 [NS][6]
  size_t w_Hpart = 0;    // f64 or f32 [lin_slices][E][90]: partial edge sums (lin_slices > 1)
  size_t w_Spart = 0;    // f64 or f32 [gram_slices][NPAIR 36 + NS 6]: partial Gram sums (gram_slices > 1)
  int64_t n_pad = 0;     // 6P rounded up to a multiple of 64 (identity padding)
  size_t w_A = 0;        // f64 or f32 [n_pad][n_pad], column-major
  size_t w_b = 0;        // f64 or f32 [n_pad]: rhs, then the forward-substitution result
  size_t w_x = 0;        // f64 or f32 [n_pad]: solution
  size_t w_Ld = 0;       // f64 or f32 [n_pad / 64][64][64]: diagonal Cholesky factors
  size_t w_info = 0;     // int32 [4 + n_pad / 64]: solver info, all-ok flag; back-substitution counters
  // torch-semantics step only
  size_t w_lz = 0;       // f32 [F][HW]: depth part of the adjoint solution
  size_t w_dpart = 0;    // f32 [E][HW]: per-edge disparity-state gradients of the adjoint
  size_t w_W = 0;        // f64 [n_pad][n_pad]: inverse of the Cholesky factor (dzcov)
  size_t w_Sinv = 0;     // f64 [n_pad][n_pad]: inverse of the reduced system (dzcov)
  size_t total = 0;
};

namespace hsl2 {
// kinds of block contributions (low 3 bits of a blk_src code; index in the high bits)
enum { kBlkHii = 0, kBlkHij = 1, kBlkHji = 2, kBlkHjj = 3, kBlkS = 4, kBlkSt = 5 };
// kinds of rhs contributions (low 2 bits)
enum { kRhsVi = 0, kRhsVj = 1, kRhsEv = 2 };
// gram task kinds
enum { kTaskPair = 0, kTaskEv = 1 };
// Gram items: a frame's 6m x 6m Gram matrix is cut into 16 x 16 tiles, and the row tiles into
// groups of kGramGroup; one item (one block) covers the tile pairs of one or two groups
constexpr int kGramGroup = 4;
// below this many Gram items, items are cut into pixel slices to reach about this many blocks
constexpr int kGramTargetBlocks = 1024;
// below this many edges, each edge's pixels are cut into slices (of at least kLinMinSlice
// pixels) so that the linearisation launch has about kLinTargetBlocks blocks
constexpr int kLinTargetBlocks = 512;
constexpr int kLinMinSlice = 768;
}  // namespace hsl2
