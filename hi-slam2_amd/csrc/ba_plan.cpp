// Host-side planning of a dense-BA call (see ba_plan.hpp).
#include "ba_plan.hpp"

#include <algorithm>
#include <new>
#include <string>

#include "hsl2.h"

namespace hsl2 {
void set_error(const std::string& msg);
}

namespace {

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// (key, code) pushes, bucketed stably by key into a CSR table (ptr [nkeys + 1], src).
struct Pushes {
  std::vector<int32_t> key, code;
  void reserve(int64_t n) {
    key.reserve(static_cast<size_t>(n));
    code.reserve(static_cast<size_t>(n));
  }
  size_t size() const { return key.size(); }
  void push(int64_t k, int32_t c) {
    key.push_back(static_cast<int32_t>(k));
    code.push_back(c);
  }
  void to_csr(int64_t nkeys, std::vector<int32_t>& ptr, std::vector<int32_t>& src) const {
    ptr.assign(static_cast<size_t>(nkeys + 1), 0);
    for (int32_t k : key) ptr[k + 1]++;
    for (int64_t k = 0; k < nkeys; ++k) ptr[k + 1] += ptr[k];
    src.resize(key.size());
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (size_t i = 0; i < key.size(); ++i) src[fill[key[i]]++] = code[i];
  }
};

}  // namespace

extern "C" int hsl2_ba_plan_create(const int64_t* ii, const int64_t* jj, int64_t num_edges,
                                   int64_t num_poses, int64_t ht, int64_t wd, int t0, int t1,
                                   int motion_only, int precision, hsl2_ba_plan** out) {
  return hsl2_ba_plan_create_ex(ii, jj, num_edges, num_poses, ht, wd, t0, t1, motion_only,
                                precision, HSL2_SEM_CUDA, out);
}

extern "C" int hsl2_ba_plan_create_ex(const int64_t* ii, const int64_t* jj, int64_t num_edges,
                                      int64_t num_poses, int64_t ht, int64_t wd, int t0, int t1,
                                      int motion_only, int precision, int semantics,
                                      hsl2_ba_plan** out) {
  using hsl2::set_error;
  if (!out) { set_error("ba plan: null output pointer"); return HSL2_ERR_ARGUMENT; }
  *out = nullptr;
  if (num_edges < 0 || (num_edges > 0 && (!ii || !jj))) {
    set_error("ba plan: null edge arrays");
    return HSL2_ERR_ARGUMENT;
  }
  if (ht <= 0 || wd <= 0 || ht * wd >= (1ll << 31)) {
    set_error("ba plan: bad disparity map size");
    return HSL2_ERR_SHAPE;
  }
  if (!(0 <= t0 && t0 < t1 && t1 <= num_poses)) {
    set_error("ba plan: need 0 <= t0 < t1 <= number of poses");
    return HSL2_ERR_ARGUMENT;
  }
  if (precision != HSL2_BA_SPLIT && precision != HSL2_BA_FP32) {
    set_error("ba plan: unknown precision");
    return HSL2_ERR_ARGUMENT;
  }
  if (semantics != HSL2_SEM_CUDA && semantics != HSL2_SEM_TORCH) {
    set_error("ba plan: unknown semantics");
    return HSL2_ERR_ARGUMENT;
  }
  if (semantics == HSL2_SEM_TORCH && precision != HSL2_BA_SPLIT) {
    set_error("ba plan: the torch-semantics step solves in FP64 only");
    return HSL2_ERR_UNSUPPORTED;
  }
  if (num_edges >= (1ll << 27)) { set_error("ba plan: too many edges"); return HSL2_ERR_SHAPE; }
  for (int64_t e = 0; e < num_edges; ++e) {
    if (ii[e] < 0 || jj[e] < 0 || ii[e] >= t1 || jj[e] >= t1) {
      set_error("ba plan: edge " + std::to_string(e) + " (" + std::to_string(ii[e]) + ", " +
                std::to_string(jj[e]) + ") indexes a frame outside [0, t1)");
      return HSL2_ERR_INDEX;
    }
  }

  hsl2_ba_plan* p = new (std::nothrow) hsl2_ba_plan();
  if (!p) { set_error("ba plan: out of host memory"); return HSL2_ERR_ARGUMENT; }
  p->E = num_edges;
  p->num_poses = num_poses;
  p->ht = ht;
  p->wd = wd;
  p->HW = ht * wd;
  p->t0 = t0;
  p->t1 = t1;
  p->P = t1 - t0;
  p->motion_only = motion_only ? 1 : 0;
  p->precision = precision;
  p->semantics = semantics;
  const int64_t E = num_edges, P = p->P, HW = p->HW;

  // ---- kx = sorted unique([t0, t1) ++ ii), kk_exp (src/droid_kernels.cu:1841-1849) ----
  // The torch formulation's depth frames are unique(ii) only (hislam2/geom/ba.py:70); a
  // window frame without edges then has no depth block and its pseudo-edge is unused
  // (kk_exp = -1).
  std::vector<int64_t> ii_exp(P + E), jj_exp(P + E);
  for (int64_t k = 0; k < P; ++k) ii_exp[k] = jj_exp[k] = t0 + k;
  for (int64_t e = 0; e < E; ++e) {
    ii_exp[P + e] = ii[e];
    jj_exp[P + e] = jj[e];
  }
  std::vector<int32_t> frame_of(static_cast<size_t>(t1), -1);
  for (int64_t n = (semantics == HSL2_SEM_TORCH ? P : 0); n < P + E; ++n) frame_of[ii_exp[n]] = 0;
  for (int64_t k = 0; k < t1; ++k)
    if (frame_of[k] == 0) {
      frame_of[k] = static_cast<int32_t>(p->kx.size());
      p->kx.push_back(k);
    }
  p->F = static_cast<int64_t>(p->kx.size());
  const int64_t F = p->F;
  p->kk_exp.resize(P + E);
  for (int64_t n = 0; n < P + E; ++n) p->kk_exp[n] = frame_of[ii_exp[n]];

  // ---- per-frame lists ----
  std::vector<int32_t> frame_eptr(F + 1, 0), frame_edges(E);
  for (int64_t e = 0; e < E; ++e) frame_eptr[p->kk_exp[P + e] + 1]++;
  for (int64_t f = 0; f < F; ++f) frame_eptr[f + 1] += frame_eptr[f];
  {
    std::vector<int32_t> fill(frame_eptr.begin(), frame_eptr.end() - 1);
    for (int64_t e = 0; e < E; ++e) frame_edges[fill[p->kk_exp[P + e]]++] = static_cast<int32_t>(e);
  }
  std::vector<int32_t> frame_pose(F);
  for (int64_t f = 0; f < F; ++f) {
    const int64_t k = p->kx[f];
    frame_pose[f] = (k >= t0 && k < t1) ? static_cast<int32_t>(k - t0) : -1;
  }

  // Schur slots: expanded edges n (ascending) with pose slot jj_exp - t0 in [0, P),
  // grouped by depth frame (the graph of src/droid_kernels.cu:1706-1716).
  // Block (P x P) and rhs (P) contribution lists are collected as (key, code) pushes and
  // bucketed stably into CSR tables at the end, so each list keeps its push order.
  std::vector<int32_t> slot_ptr(F + 1, 0), slot_row, slot_pose;
  std::vector<int32_t> tasks;
  Pushes blk, rhs;
  blk.reserve(4 * E);
  rhs.reserve(2 * E);
  // H-block and v contributions (src/droid_kernels.cu:1882-1889)
  for (int64_t e = 0; e < E; ++e) {
    const int64_t a = ii[e] - t0, b = jj[e] - t0;
    const int32_t ie = static_cast<int32_t>(e) << 3;
    if (a >= 0) blk.push(a * P + a, ie | hsl2::kBlkHii);
    if (a >= 0 && b >= 0) {
      blk.push(a * P + b, ie | hsl2::kBlkHij);
      blk.push(b * P + a, ie | hsl2::kBlkHji);
    }
    if (b >= 0) blk.push(b * P + b, ie | hsl2::kBlkHjj);
    if (a >= 0) rhs.push(a, (static_cast<int32_t>(e) << 2) | hsl2::kRhsVi);
    if (b >= 0) rhs.push(b, (static_cast<int32_t>(e) << 2) | hsl2::kRhsVj);
  }
  if (!p->motion_only) {
    // slots bucketed by depth frame, ascending n inside a frame
    for (int64_t n = 0; n < P + E; ++n) {
      const int64_t a = jj_exp[n] - t0;
      if (a >= 0 && a < P && p->kk_exp[n] >= 0) slot_ptr[p->kk_exp[n] + 1]++;
    }
    for (int64_t f = 0; f < F; ++f) slot_ptr[f + 1] += slot_ptr[f];
    p->NS = slot_ptr[F];
    slot_row.resize(p->NS);
    slot_pose.resize(p->NS);
    {
      std::vector<int32_t> fill(slot_ptr.begin(), slot_ptr.end() - 1);
      for (int64_t n = 0; n < P + E; ++n) {
        const int64_t a = jj_exp[n] - t0;
        if (a >= 0 && a < P && p->kk_exp[n] >= 0) {
          const int32_t s = fill[p->kk_exp[n]]++;
          slot_row[s] = static_cast<int32_t>(n);
          slot_pose[s] = static_cast<int32_t>(a);
        }
      }
    }
    int64_t npair = 0;
    for (int64_t f = 0; f < F; ++f) {
      const int64_t m = slot_ptr[f + 1] - slot_ptr[f];
      npair += m * (m + 1) / 2;
    }
    if (npair >= (1ll << 27)) {
      delete p;
      set_error("ba plan: too many Schur pairs");
      return HSL2_ERR_SHAPE;
    }
    p->NPAIR = npair;
    tasks.reserve(4 * (npair + p->NS));
    blk.reserve(blk.size() + 2 * npair);
    npair = 0;
    for (int64_t f = 0; f < F; ++f) {
      for (int32_t s = slot_ptr[f]; s < slot_ptr[f + 1]; ++s) {
        for (int32_t s2 = s; s2 < slot_ptr[f + 1]; ++s2) {
          tasks.insert(tasks.end(), {hsl2::kTaskPair, s, s2, static_cast<int32_t>(f)});
          const int64_t a = slot_pose[s], b = slot_pose[s2];
          blk.push(a * P + b, static_cast<int32_t>(npair << 3) | hsl2::kBlkS);
          if (s2 != s) blk.push(b * P + a, static_cast<int32_t>(npair << 3) | hsl2::kBlkSt);
          ++npair;
        }
      }
    }
    for (int64_t f = 0; f < F; ++f)
      for (int32_t s = slot_ptr[f]; s < slot_ptr[f + 1]; ++s) {
        tasks.insert(tasks.end(), {hsl2::kTaskEv, s, s, static_cast<int32_t>(f)});
        rhs.push(slot_pose[s], (s << 2) | hsl2::kRhsEv);
      }
  }
  p->NT = static_cast<int64_t>(tasks.size() / 4);

  // ---- flatten into the int32 blob ----
  auto& B = p->blob;
  auto put = [&B](const std::vector<int32_t>& v) {
    const int64_t off = static_cast<int64_t>(B.size());
    B.insert(B.end(), v.begin(), v.end());
    while (B.size() % 64) B.push_back(0);
    return off;
  };
  std::vector<int32_t> eii(E), ejj(E), efr(E), fkx(F);
  for (int64_t e = 0; e < E; ++e) {
    eii[e] = static_cast<int32_t>(ii[e]);
    ejj[e] = static_cast<int32_t>(jj[e]);
    efr[e] = static_cast<int32_t>(p->kk_exp[P + e]);
  }
  for (int64_t f = 0; f < F; ++f) fkx[f] = static_cast<int32_t>(p->kx[f]);
  p->o_edge_ii = put(eii);
  p->o_edge_jj = put(ejj);
  p->o_edge_frame = put(efr);
  p->o_frame_kx = put(fkx);
  p->o_frame_pose = put(frame_pose);
  p->o_frame_eptr = put(frame_eptr);
  p->o_frame_edges = put(frame_edges);
  {
    // per-frame work grows with the frame's edge count: the heaviest frames are dispatched first
    std::vector<int32_t> order(F);
    for (int64_t f = 0; f < F; ++f) order[f] = static_cast<int32_t>(f);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
      return frame_eptr[a + 1] - frame_eptr[a] > frame_eptr[b + 1] - frame_eptr[b];
    });
    p->o_frame_order = put(order);
  }
  p->o_slot_ptr = put(slot_ptr);
  p->o_slot_row = put(slot_row);
  p->o_slot_pose = put(slot_pose);
  p->o_task = put(tasks);
  std::vector<int32_t> bptr, bsrc;
  blk.to_csr(P * P, bptr, bsrc);
  p->o_blk_ptr = put(bptr);
  p->o_blk_src = put(bsrc);
  std::vector<int32_t> rptr, rsrc;
  rhs.to_csr(P, rptr, rsrc);
  p->o_rhs_ptr = put(rptr);
  p->o_rhs_src = put(rsrc);
  // Gram group items in frame order, and each frame's first pair
  {
    std::vector<int32_t> pair0(F, 0), items;
    int64_t acc = 0;
    for (int64_t f = 0; f < F; ++f) {
      pair0[f] = static_cast<int32_t>(acc);
      const int64_t m = slot_ptr[f + 1] - slot_ptr[f];
      acc += m * (m + 1) / 2;
    }
    if (!p->motion_only) {
      for (int64_t f = 0; f < F; ++f) {
        const int64_t m = slot_ptr[f + 1] - slot_ptr[f];
        const int32_t nt = static_cast<int32_t>((6 * m + 15) / 16);
        const int32_t ng = (nt + hsl2::kGramGroup - 1) / hsl2::kGramGroup;
        for (int32_t ga = 0; ga < ng; ++ga)
          for (int32_t gb = ga; gb < ng; ++gb) items.insert(items.end(), {static_cast<int32_t>(f), ga, gb});
      }
    }
    p->NGGI = static_cast<int64_t>(items.size() / 3);
    // few items (small problems): cut each item's pixels into slices of whole 256-pixel folds
    // so that the launch has about kGramTargetBlocks blocks
    if (p->NGGI > 0) {
      const int64_t nfold = (HW + 255) / 256;
      const int64_t want = (hsl2::kGramTargetBlocks + p->NGGI - 1) / p->NGGI;
      p->gram_slices = static_cast<int>(std::max<int64_t>(1, std::min(want, nfold)));
    }
    p->o_frame_pair0 = put(pair0);
    p->o_gram_group_item = put(items);
  }

  // ---- workspace layout ----
  size_t off = 0;
  auto take = [&off](size_t bytes) {
    const size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  const size_t f4 = sizeof(float);
  const bool schur = !p->motion_only;
  const size_t real = (p->precision == HSL2_BA_SPLIT) ? sizeof(double) : sizeof(float);
  p->w_blob = take(B.size() * sizeof(int32_t));
  p->w_Eall = take(schur ? (P + E) * 6 * HW * f4 : 0);
  // the per-edge planes of ba_frame_reduce: the torch-semantics step only (the ba_cuda loop
  // forms the frame sums from the pixels)
  const bool planes = schur && semantics == HSL2_SEM_TORCH;
  p->w_Eii = take(planes ? E * 6 * HW * f4 : 0);
  p->w_Cii = take(planes ? E * HW * f4 : 0);
  p->w_bz = take(planes ? E * HW * f4 : 0);
  p->w_Q = take(schur ? F * HW * f4 : 0);
  p->w_w = take(schur ? F * HW * f4 : 0);
  const size_t f8 = sizeof(double);
  p->w_Hs = take(4 * E * 36 * f8);
  p->w_vs = take(2 * E * 6 * f8);
  p->w_S = take(p->NPAIR * 36 * f8);
  p->w_Ev = take(p->NS * 6 * f8);
  if (E > 0) {
    const int64_t want = (hsl2::kLinTargetBlocks + E - 1) / E;
    p->lin_slices = static_cast<int>(std::max<int64_t>(1, std::min(want, HW / hsl2::kLinMinSlice)));
  }
  p->w_Hpart = take(p->lin_slices > 1 ? p->lin_slices * E * 90 * real : 0);
  p->w_Spart = take(p->gram_slices > 1 ? p->gram_slices * (p->NPAIR * 36 + p->NS * 6) * real : 0);
  p->n_pad = ((6 * P + 63) / 64) * 64;
  p->w_A = take(static_cast<size_t>(p->n_pad) * p->n_pad * real);
  p->w_b = take(static_cast<size_t>(p->n_pad) * real);
  p->w_x = take(static_cast<size_t>(p->n_pad) * real);
  p->w_Ld = take(static_cast<size_t>(p->n_pad) * 64 * real);
  p->w_info = take((4 + p->n_pad / 64) * sizeof(int32_t));
  if (semantics == HSL2_SEM_TORCH) {
    p->w_lz = take(schur ? F * HW * f4 : 0);
    p->w_dpart = take(E * HW * f4);
    p->w_W = take(static_cast<size_t>(p->n_pad) * p->n_pad * f8);
    p->w_Sinv = take(static_cast<size_t>(p->n_pad) * p->n_pad * f8);
  }
  p->total = off;
  *out = p;
  return HSL2_OK;
}

extern "C" void hsl2_ba_plan_destroy(hsl2_ba_plan* plan) { delete plan; }

extern "C" int64_t hsl2_ba_plan_num_frames(const hsl2_ba_plan* plan) { return plan ? plan->F : -1; }

extern "C" int hsl2_ba_plan_frames(const hsl2_ba_plan* plan, int64_t* kx) {
  if (!plan || !kx) { hsl2::set_error("ba plan: null pointer"); return HSL2_ERR_ARGUMENT; }
  std::copy(plan->kx.begin(), plan->kx.end(), kx);
  return HSL2_OK;
}

extern "C" int hsl2_ba_plan_inverse(const hsl2_ba_plan* plan, int64_t* kk_exp) {
  if (!plan || !kk_exp) { hsl2::set_error("ba plan: null pointer"); return HSL2_ERR_ARGUMENT; }
  std::copy(plan->kk_exp.begin(), plan->kk_exp.end(), kk_exp);
  return HSL2_OK;
}

extern "C" size_t hsl2_ba_plan_workspace_bytes(const hsl2_ba_plan* plan) {
  return plan ? plan->total : 0;
}
