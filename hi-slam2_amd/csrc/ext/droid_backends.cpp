// droid_backends: the torch extension module the reference's Python code imports
// (hislam2/geom/ba.py:4, hislam2/depth_video.py:3, hislam2/modules/corr.py:4),
// rebuilt over the C ABI of libhsl2 (hsl2.h).  Same function names, positional
// argument order and return conventions as src/droid.cpp:141-168, 254-305, 325-341.
//
// This layer owns validation (device, dtype, shape, contiguity), allocation through
// the torch caching allocator, and the current HIP stream; the kernels live behind
// the C ABI.  A non-contiguous input raises RuntimeError exactly like the
// reference's CHECK_INPUT; inconsistent shapes and out-of-range edge indices, which
// are undefined behaviour in the reference, raise ValueError.
#include <torch/extension.h>

#include <ATen/hip/impl/HIPGuardImplMasqueradingAsCUDA.h>
#include <ATen/hip/impl/HIPStreamMasqueradingAsCUDA.h>

#include <array>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hsl2.h"

namespace py = pybind11;

namespace {

#define CHECK_CONTIGUOUS(x) TORCH_CHECK((x).is_contiguous(), #x " must be contiguous")

void*current_stream() {
  return reinterpret_cast<void*>(c10::hip::getCurrentHIPStreamMasqueradingAsCUDA().stream());
}

void raise_status(int st, const char* what) {
  if (st == HSL2_OK) return;
  std::string msg = std::string(what) + ": " + hsl2_status_string(st) + ": " + hsl2_last_error();
  if (st == HSL2_ERR_ARGUMENT || st == HSL2_ERR_SHAPE || st == HSL2_ERR_INDEX) throw py::value_error(msg);
  if (st == HSL2_ERR_UNSUPPORTED) throw py::type_error(msg);
  throw std::runtime_error(msg);
}

void require(bool cond, const std::string& msg) {
  if (!cond) throw py::value_error(msg);
}

void require_device(const torch::Tensor& t, const char* name) {
  require(t.is_cuda(), std::string(name) + " must be a GPU tensor");
}

int dtype_code(const torch::Tensor& t) {
  switch (t.scalar_type()) {
    case torch::kFloat32: return HSL2_F32;
    case torch::kFloat16: return HSL2_F16;
    case torch::kBFloat16: return HSL2_BF16;
    case torch::kFloat64: return HSL2_F64;
    default: throw py::type_error("unsupported dtype");
  }
}

// ---- ba (src/droid.cpp:141-168) ------------------------------------------------
// The last few ba plans, keyed by the complete edge list and every planning argument (device
// included), so that repeated calls on one factor graph (each call of a global BA over the same
// graph) skip the host planning and the upload of the plan's tables, which stay resident on the
// device.  Lookup first by identity: the same ii / jj storage (held by the entry, so its address
// cannot be taken by another tensor) at the same version counter, which torch advances on every
// in-place write, holds the same edges, and the call needs no host copy of the list and no
// stream synchronisation.  Otherwise the list is copied to the host and compared in full.
// Entries are shared, so an evicted plan (and its device tables) stays valid for a call still
// using it.
class PlanCache {
 public:
  using Args = std::array<int64_t, 9>;   // num_poses, ht, wd, t0, t1, motion_only, precision, semantics, device
  using Ptr = std::shared_ptr<const hsl2_ba_plan>;

  Ptr get(const torch::Tensor& ii, const torch::Tensor& jj, const Args& a, const torch::TensorOptions& dev) {
    const int64_t E = ii.size(0);
    {
      std::lock_guard<std::mutex> lock(mu_);
      for (auto it = entries_.begin(); it != entries_.end(); ++it) {
        if (it->args == a && same_edges(it->ii_ref, it->ii_version, ii) && same_edges(it->jj_ref, it->jj_version, jj)) {
          entries_.splice(entries_.begin(), entries_, it);
          return entries_.front().plan;
        }
      }
    }
    torch::Tensor ii_h = ii.to(torch::kCPU), jj_h = jj.to(torch::kCPU);
    const int64_t* iip = ii_h.data_ptr<int64_t>();
    const int64_t* jjp = jj_h.data_ptr<int64_t>();
    {
      std::lock_guard<std::mutex> lock(mu_);
      for (auto it = entries_.begin(); it != entries_.end(); ++it) {
        if (it->args == a && static_cast<int64_t>(it->ii.size()) == E &&
            std::memcmp(it->ii.data(), iip, E * sizeof(int64_t)) == 0 &&
            std::memcmp(it->jj.data(), jjp, E * sizeof(int64_t)) == 0) {
          it->remember(ii, jj);
          entries_.splice(entries_.begin(), entries_, it);
          return entries_.front().plan;
        }
      }
    }
    hsl2_ba_plan* raw = nullptr;
    raise_status(hsl2_ba_plan_create_ex(iip, jjp, E, a[0], a[1], a[2], static_cast<int>(a[3]),
                                        static_cast<int>(a[4]), static_cast<int>(a[5]),
                                        static_cast<int>(a[6]), static_cast<int>(a[7]), &raw),
                 "ba");
    std::unique_ptr<hsl2_ba_plan, void (*)(hsl2_ba_plan*)> owned(raw, hsl2_ba_plan_destroy);
    // the plan's tables on the device, owned by the plan's deleter
    torch::Tensor tables = torch::empty({static_cast<int64_t>(hsl2_ba_plan_table_bytes(raw)) + 256},
                                        dev.dtype(torch::kUInt8));
    const uintptr_t tb = reinterpret_cast<uintptr_t>(tables.data_ptr());
    raise_status(hsl2_ba_plan_upload_tables(raw, reinterpret_cast<void*>((tb + 255) & ~uintptr_t(255)),
                                            current_stream()),
                 "ba");
    Ptr plan(owned.release(), [tables](hsl2_ba_plan* q) mutable {
      hsl2_ba_plan_destroy(q);
      tables = torch::Tensor();
    });
    std::lock_guard<std::mutex> lock(mu_);
    Entry e{a, std::vector<int64_t>(iip, iip + E), std::vector<int64_t>(jjp, jjp + E), {}, 0, {}, 0, plan};
    e.remember(ii, jj);
    entries_.push_front(std::move(e));
    if (entries_.size() > kCapacity) entries_.pop_back();
    return plan;
  }

 private:
  struct Entry {
    Args args;
    std::vector<int64_t> ii, jj;
    torch::Tensor ii_ref;     // the caller's tensors of the last hit, and their versions then
    int64_t ii_version;
    torch::Tensor jj_ref;
    int64_t jj_version;
    Ptr plan;
    void remember(const torch::Tensor& i, const torch::Tensor& j) {
      ii_ref = i;
      ii_version = i._version();
      jj_ref = j;
      jj_version = j._version();
    }
  };
  static bool same_edges(const torch::Tensor& held, int64_t version, const torch::Tensor& t) {
    return held.defined() && held.data_ptr() == t.data_ptr() && held.numel() == t.numel() &&
           held.device() == t.device() && t._version() == version;
  }
  static constexpr size_t kCapacity = 4;
  std::mutex mu_;
  std::list<Entry> entries_;
};

PlanCache& plan_cache() {
  static PlanCache cache;
  return cache;
}

std::vector<torch::Tensor> ba_impl(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics,
                                   torch::Tensor targets, torch::Tensor weights, torch::Tensor eta,
                                   torch::Tensor ii, torch::Tensor jj, int t0, int t1, int iterations,
                                   double lm, double ep, bool motion_only, int precision,
                                   bool want_status, const torch::Tensor* disps_sens = nullptr) {
  CHECK_CONTIGUOUS(targets);
  CHECK_CONTIGUOUS(weights);
  CHECK_CONTIGUOUS(poses);
  CHECK_CONTIGUOUS(disps);
  CHECK_CONTIGUOUS(intrinsics);
  CHECK_CONTIGUOUS(ii);
  CHECK_CONTIGUOUS(jj);
  for (auto* t : {&poses, &disps, &intrinsics, &targets, &weights}) {
    require_device(*t, "ba tensors");
    require(t->scalar_type() == torch::kFloat32, "ba: poses, disps, intrinsics, targets, weights must be float32");
  }
  require(poses.dim() == 2 && poses.size(1) == 7, "ba: poses must be [N, 7]");
  require(disps.dim() == 3 && disps.size(0) == poses.size(0), "ba: disps must be [N, ht, wd]");
  require(intrinsics.numel() >= 4, "ba: intrinsics must hold fx, fy, cx, cy");
  require(ii.dim() == 1 && jj.dim() == 1 && ii.size(0) == jj.size(0), "ba: ii, jj must be [E]");
  require(ii.scalar_type() == torch::kInt64 && jj.scalar_type() == torch::kInt64, "ba: ii, jj must be int64");
  const int64_t E = ii.size(0), ht = disps.size(1), wd = disps.size(2);
  require(targets.numel() == E * 2 * ht * wd && weights.numel() == E * 2 * ht * wd,
          "ba: targets and weights must be [E, 2, ht, wd]");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());

  // at most one host copy of the edge list per call, none when the same ii / jj tensors repeat
  // (the reference copies the list 16+ times per iteration)
  const PlanCache::Ptr plan_ref = plan_cache().get(
      ii, jj, {poses.size(0), ht, wd, t0, t1, motion_only ? 1 : 0, precision, HSL2_SEM_CUDA,
               static_cast<int64_t>(poses.device().index())},
      poses.options());
  const hsl2_ba_plan* plan = plan_ref.get();
  const int64_t F = hsl2_ba_plan_num_frames(plan);
  int64_t eta_rows = 1;
  const float* eta_ptr = nullptr;
  torch::Tensor eta_c;
  if (!motion_only) {
    require_device(eta, "eta");
    require(eta.scalar_type() == torch::kFloat32, "ba: eta must be float32");
    eta_c = eta.contiguous();
    require(eta_c.numel() % (ht * wd) == 0, "ba: eta must hold whole [ht, wd] rows");
    eta_rows = eta_c.numel() / (ht * wd);
    require(eta_rows == F || eta_rows == 1,
            "ba: eta must have one [ht, wd] row per unique frame of [t0, t1) ++ ii (" + std::to_string(F) +
                ") or a single row");
    eta_ptr = eta_c.data_ptr<float>();
  }
  auto opts = poses.options();
  torch::Tensor ws = torch::empty({static_cast<int64_t>(hsl2_ba_plan_workspace_bytes(plan)) + 256},
                                  opts.dtype(torch::kUInt8));
  uintptr_t base = reinterpret_cast<uintptr_t>(ws.data_ptr());
  void* ws_ptr = reinterpret_cast<void*>((base + 255) & ~uintptr_t(255));
  torch::Tensor dx = torch::zeros({t1 - t0, 6}, opts);
  torch::Tensor dz = motion_only ? torch::Tensor() : torch::zeros({F, ht * wd}, opts);
  const float* sens_ptr = nullptr;
  if (disps_sens) {
    CHECK_CONTIGUOUS((*disps_sens));
    require_device(*disps_sens, "disps_sens");
    require(disps_sens->scalar_type() == torch::kFloat32 && disps_sens->sizes() == disps.sizes(),
            "ba: disps_sens must be float32 shaped like disps [N, ht, wd]");
    sens_ptr = disps_sens->data_ptr<float>();
  }
  int ok = 1;
  raise_status(hsl2_ba_run_sens(plan, poses.data_ptr<float>(), disps.data_ptr<float>(),
                                intrinsics.data_ptr<float>(), sens_ptr, targets.data_ptr<float>(),
                                weights.data_ptr<float>(), eta_ptr, eta_rows, iterations,
                                static_cast<float>(lm), static_cast<float>(ep), dx.data_ptr<float>(),
                                motion_only ? nullptr : dz.data_ptr<float>(),
                                ws_ptr, static_cast<size_t>(ws.numel() - 256), want_status ? &ok : nullptr,
                                current_stream()),
               "ba");
  std::vector<torch::Tensor> out = {iterations > 0 ? dx : torch::Tensor(),
                                    iterations > 0 ? dz : torch::Tensor(), torch::Tensor()};
  if (want_status) out.push_back(torch::tensor({ok}, torch::kInt32));
  return out;
}

std::vector<torch::Tensor> ba(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics,
                              torch::Tensor targets, torch::Tensor weights, torch::Tensor eta,
                              torch::Tensor ii, torch::Tensor jj, int t0, int t1, int iterations,
                              double lm, double ep, bool motion_only) {
  return ba_impl(poses, disps, intrinsics, targets, weights, eta, ii, jj, t0, t1, iterations, lm, ep,
                 motion_only, HSL2_BA_SPLIT, false);
}

// The 15-argument form with the sensor disparities after the intrinsics (upstream DROID-SLAM's
// binding; see hsl2_ba_run_sens).  An all-zero disps_sens gives the 14-argument results exactly.
std::vector<torch::Tensor> ba_sens(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics,
                                   torch::Tensor disps_sens, torch::Tensor targets, torch::Tensor weights,
                                   torch::Tensor eta, torch::Tensor ii, torch::Tensor jj, int t0, int t1,
                                   int iterations, double lm, double ep, bool motion_only) {
  return ba_impl(poses, disps, intrinsics, targets, weights, eta, ii, jj, t0, t1, iterations, lm, ep,
                 motion_only, HSL2_BA_SPLIT, false, &disps_sens);
}

std::vector<torch::Tensor> ba_ex(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics,
                                 torch::Tensor targets, torch::Tensor weights, torch::Tensor eta,
                                 torch::Tensor ii, torch::Tensor jj, int t0, int t1, int iterations,
                                 double lm, double ep, bool motion_only, int precision) {
  return ba_impl(poses, disps, intrinsics, targets, weights, eta, ii, jj, t0, t1, iterations, lm, ep,
                 motion_only, precision, true);
}

// kx and kk_exp of a ba call, computed by the library's planner (host only)
std::vector<torch::Tensor> ba_index(torch::Tensor ii, torch::Tensor jj, int t0, int t1, int64_t num_poses) {
  torch::Tensor ii_h = ii.to(torch::kCPU).to(torch::kInt64).contiguous();
  torch::Tensor jj_h = jj.to(torch::kCPU).to(torch::kInt64).contiguous();
  require(ii_h.numel() == jj_h.numel(), "ba_index: ii and jj differ in length");
  hsl2_ba_plan* plan = nullptr;
  raise_status(hsl2_ba_plan_create(ii_h.data_ptr<int64_t>(), jj_h.data_ptr<int64_t>(), ii_h.numel(),
                                   num_poses, 1, 1, t0, t1, 0, HSL2_BA_SPLIT, &plan),
               "ba_index");
  std::unique_ptr<hsl2_ba_plan, void (*)(hsl2_ba_plan*)> guard_plan(plan, hsl2_ba_plan_destroy);
  torch::Tensor kx = torch::empty({hsl2_ba_plan_num_frames(plan)}, torch::kInt64);
  torch::Tensor kk = torch::empty({(t1 - t0) + ii_h.numel()}, torch::kInt64);
  raise_status(hsl2_ba_plan_frames(plan, kx.data_ptr<int64_t>()), "ba_index");
  raise_status(hsl2_ba_plan_inverse(plan, kk.data_ptr<int64_t>()), "ba_index");
  return {kx, kk};
}

// Dense SPD solve with the library's tiled Cholesky (the reduced-system solver of ba).
std::vector<torch::Tensor> spd_solve(torch::Tensor A, torch::Tensor b) {
  require_device(A, "A");
  require(A.dim() == 2 && A.size(0) == A.size(1) && b.numel() == A.size(0), "spd_solve: A [n, n], b [n]");
  require(A.scalar_type() == b.scalar_type() &&
              (A.scalar_type() == torch::kFloat64 || A.scalar_type() == torch::kFloat32),
          "spd_solve: float64 or float32");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(A.device());
  const int64_t n = A.size(0);
  torch::Tensor Acm = A.t().contiguous();            // column-major
  torch::Tensor bc = b.contiguous();
  const int dt = dtype_code(A);
  torch::Tensor ws = torch::empty({static_cast<int64_t>(hsl2_spd_solve_workspace_bytes(dt, n)) + 256},
                                  A.options().dtype(torch::kUInt8));
  void* wp = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(ws.data_ptr()) + 255) & ~uintptr_t(255));
  torch::Tensor x = torch::empty({n}, A.options());
  int ok = 0;
  raise_status(hsl2_spd_solve(dt, Acm.data_ptr(), bc.data_ptr(), x.data_ptr(), n, wp,
                              static_cast<size_t>(ws.numel() - 256), &ok, current_stream()),
               "spd_solve");
  return {x, torch::tensor({ok}, torch::kInt32)};
}

// ---- corr_index (src/droid.cpp:254-275) ----------------------------------------------
std::vector<torch::Tensor> corr_index_forward(torch::Tensor volume, torch::Tensor coords, int radius) {
  CHECK_CONTIGUOUS(volume);
  CHECK_CONTIGUOUS(coords);
  require_device(volume, "volume");
  require_device(coords, "coords");
  require(volume.dim() == 5, "corr_index_forward: volume must be [B, h1, w1, h2, w2]");
  require(coords.scalar_type() == torch::kFloat32, "corr_index_forward: coords must be float32");
  const int64_t B = volume.size(0), h1 = volume.size(1), w1 = volume.size(2);
  require(coords.dim() == 4 && coords.size(0) == B && coords.size(1) == 2 && coords.size(2) == h1 &&
              coords.size(3) == w1,
          "corr_index_forward: coords must be [B, 2, h1, w1]");
  require(radius >= 0, "corr_index_forward: negative radius");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(volume.device());
  torch::Tensor corr = torch::empty({B, 2 * radius + 1, 2 * radius + 1, h1, w1}, volume.options());
  raise_status(hsl2_corr_index_forward(volume.data_ptr(), coords.data_ptr<float>(), corr.data_ptr(),
                                       dtype_code(volume), B, h1, w1, volume.size(3), volume.size(4),
                                       radius, current_stream()),
               "corr_index_forward");
  return {corr};
}

std::vector<torch::Tensor> corr_index_backward(torch::Tensor volume, torch::Tensor coords,
                                               torch::Tensor corr_grad, int radius) {
  CHECK_CONTIGUOUS(volume);
  CHECK_CONTIGUOUS(coords);
  CHECK_CONTIGUOUS(corr_grad);
  require_device(volume, "volume");
  require(volume.dim() == 5, "corr_index_backward: volume must be [B, h1, w1, h2, w2]");
  const int64_t B = volume.size(0), h1 = volume.size(1), w1 = volume.size(2);
  require(coords.scalar_type() == torch::kFloat32 && coords.numel() == B * 2 * h1 * w1,
          "corr_index_backward: coords must be float32 [B, 2, h1, w1]");
  require(corr_grad.scalar_type() == volume.scalar_type() &&
              corr_grad.numel() == B * (2 * radius + 1) * (2 * radius + 1) * h1 * w1,
          "corr_index_backward: corr_grad must match the forward output");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(volume.device());
  torch::Tensor vgrad = torch::empty_like(volume);
  raise_status(hsl2_corr_index_backward(coords.data_ptr<float>(), corr_grad.data_ptr(), vgrad.data_ptr(),
                                        dtype_code(volume), B, h1, w1, volume.size(3), volume.size(4),
                                        radius, current_stream()),
               "corr_index_backward");
  return {vgrad};
}

// Fused CorrBlock.__call__ (hislam2/modules/corr.py:40-50): all levels, one launch.
torch::Tensor corr_pyramid_lookup(std::vector<torch::Tensor> levels, torch::Tensor coords, int radius,
                                  bool tiled) {
  CHECK_CONTIGUOUS(coords);
  require(!levels.empty() && levels.size() <= 4, "corr_pyramid_lookup: 1..4 levels");
  require(coords.dim() == 5 && coords.size(4) == 2 && coords.scalar_type() == torch::kFloat32,
          "corr_pyramid_lookup: coords must be float32 [B, N, h, w, 2]");
  const int64_t B = coords.size(0), N = coords.size(1), h = coords.size(2), w = coords.size(3);
  std::vector<const void*> ptrs;
  for (size_t i = 0; i < levels.size(); ++i) {
    CHECK_CONTIGUOUS(levels[i]);
    require_device(levels[i], "levels");
    require(levels[i].scalar_type() == levels[0].scalar_type(), "corr_pyramid_lookup: mixed level dtypes");
    require(levels[i].dim() == 5 && levels[i].size(0) == B * N && levels[i].size(1) == h &&
                levels[i].size(2) == w && levels[i].size(3) == (h >> i) && levels[i].size(4) == (w >> i),
            "corr_pyramid_lookup: level " + std::to_string(i) + " must be [B*N, h, w, h>>i, w>>i]");
    ptrs.push_back(levels[i].data_ptr());
  }
  c10::hip::HIPGuardMasqueradingAsCUDA guard(coords.device());
  const int64_t R = 2 * radius + 1;
  torch::Tensor out = torch::empty({B, N, static_cast<int64_t>(levels.size()) * R * R, h, w},
                                   levels[0].options());
  raise_status(hsl2_corr_pyramid_lookup(ptrs.data(), static_cast<int>(levels.size()), coords.data_ptr<float>(),
                                        out.data_ptr(), dtype_code(levels[0]), B * N, h, w, radius,
                                        tiled ? 1 : 0, current_stream()),
               "corr_pyramid_lookup");
  return out;
}

// The CorrBlock lookup fused with a 1x1 convolution (196 -> C_out) + optional ReLU: bf16
// pyramid of 4 levels, radius 3; weight [C_out, 196] (or [C_out, 196, 1, 1]) bf16, bias [C_out]
// float32 -> [B, N, C_out, h, w] bf16.
torch::Tensor corr_lookup_conv(std::vector<torch::Tensor> levels, torch::Tensor coords, int radius, bool tiled,
                               torch::Tensor weight, torch::Tensor bias, bool relu) {
  CHECK_CONTIGUOUS(coords);
  CHECK_CONTIGUOUS(weight);
  CHECK_CONTIGUOUS(bias);
  require(levels.size() == 4 && radius == 3, "corr_lookup_conv: 4 levels of radius 3");
  require(coords.dim() == 5 && coords.size(4) == 2 && coords.scalar_type() == torch::kFloat32,
          "corr_lookup_conv: coords must be float32 [B, N, h, w, 2]");
  const int64_t B = coords.size(0), N = coords.size(1), h = coords.size(2), w = coords.size(3);
  std::vector<const void*> ptrs;
  for (size_t i = 0; i < levels.size(); ++i) {
    CHECK_CONTIGUOUS(levels[i]);
    require_device(levels[i], "levels");
    require(levels[i].scalar_type() == torch::kBFloat16, "corr_lookup_conv: bf16 pyramid levels");
    require(levels[i].dim() == 5 && levels[i].size(0) == B * N && levels[i].size(1) == h &&
                levels[i].size(2) == w && levels[i].size(3) == (h >> i) && levels[i].size(4) == (w >> i),
            "corr_lookup_conv: level " + std::to_string(i) + " must be [B*N, h, w, h>>i, w>>i]");
    ptrs.push_back(levels[i].data_ptr());
  }
  require_device(weight, "weight");
  require_device(bias, "bias");
  require(weight.scalar_type() == torch::kBFloat16 && weight.size(0) > 0 && weight.numel() == weight.size(0) * 196,
          "corr_lookup_conv: weight must be bf16 [C_out, 196]");
  const int64_t cout = weight.size(0);
  require(bias.scalar_type() == torch::kFloat32 && bias.numel() == cout, "corr_lookup_conv: bias must be float32 [C_out]");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(coords.device());
  torch::Tensor out = torch::empty({B, N, cout, h, w}, levels[0].options());
  raise_status(hsl2_corr_lookup_conv(ptrs.data(), 4, coords.data_ptr<float>(), weight.data_ptr(),
                                     bias.data_ptr<float>(), relu ? 1 : 0, out.data_ptr(), B * N, h, w, radius,
                                     tiled ? 1 : 0, cout, current_stream()),
               "corr_lookup_conv");
  return out;
}

// ---- altcorr (src/droid.cpp:277-305) ---------------------------------------------------
struct AltDims {
  int64_t B, N1, N2, C, H, W, H2, W2, M;
};

AltDims alt_dims(const torch::Tensor& fmap1, const torch::Tensor& fmap2, const torch::Tensor& coords,
                 const torch::Tensor& ii, const torch::Tensor& jj) {
  require(fmap1.dim() == 5 && fmap2.dim() == 5, "altcorr: fmaps must be [B, N, C, H, W]");
  require(coords.dim() == 5 && coords.size(2) == 2, "altcorr: coords must be [B, M, 2, H, W]");
  require(fmap1.scalar_type() == fmap2.scalar_type(), "altcorr: fmap dtypes differ");
  require(coords.scalar_type() == torch::kFloat32, "altcorr: coords must be float32");
  AltDims d{fmap1.size(0), fmap1.size(1), fmap2.size(1), fmap1.size(2), fmap1.size(3), fmap1.size(4),
            fmap2.size(3), fmap2.size(4), coords.size(1)};
  require(fmap2.size(0) == d.B && fmap2.size(2) == d.C && coords.size(0) == d.B && coords.size(3) == d.H &&
              coords.size(4) == d.W,
          "altcorr: inconsistent shapes");
  require(ii.numel() == d.M && jj.numel() == d.M && ii.scalar_type() == torch::kInt64 &&
              jj.scalar_type() == torch::kInt64 && ii.is_cuda() && jj.is_cuda(),
          "altcorr: ii, jj must be int64 GPU tensors with one entry per edge");
  return d;
}

torch::Tensor to_channels_last(const torch::Tensor& f, const AltDims&) {
  torch::Tensor out = torch::empty_like(f);
  raise_status(hsl2_channels_last(f.data_ptr(), out.data_ptr(), dtype_code(f), f.size(0) * f.size(1), f.size(2),
                                  f.size(3) * f.size(4), current_stream()),
               "altcorr");
  return out;
}

std::vector<torch::Tensor> altcorr_forward(torch::Tensor fmap1, torch::Tensor fmap2, torch::Tensor coords,
                                           torch::Tensor ii, torch::Tensor jj, int radius) {
  CHECK_CONTIGUOUS(fmap1);
  CHECK_CONTIGUOUS(fmap2);
  CHECK_CONTIGUOUS(coords);
  require_device(fmap1, "fmap1");
  const AltDims d = alt_dims(fmap1, fmap2, coords, ii, jj);
  c10::hip::HIPGuardMasqueradingAsCUDA guard(fmap1.device());
  torch::Tensor ii_c = ii.contiguous(), jj_c = jj.contiguous();
  torch::Tensor f1 = to_channels_last(fmap1, d), f2 = to_channels_last(fmap2, d);
  torch::Tensor corr = torch::empty({d.B, d.M, 2 * radius + 1, 2 * radius + 1, d.H, d.W}, fmap1.options());
  raise_status(hsl2_altcorr_forward(f1.data_ptr(), f2.data_ptr(), coords.data_ptr<float>(),
                                    ii_c.data_ptr<int64_t>(), jj_c.data_ptr<int64_t>(), corr.data_ptr(),
                                    dtype_code(fmap1), d.B, d.N1, d.N2, d.C, d.H, d.W, d.H2, d.W2, d.M,
                                    radius, current_stream()),
               "altcorr_forward");
  return {corr};
}

std::vector<torch::Tensor> altcorr_backward(torch::Tensor fmap1, torch::Tensor fmap2, torch::Tensor coords,
                                            torch::Tensor corr_grad, torch::Tensor ii, torch::Tensor jj,
                                            int radius) {
  CHECK_CONTIGUOUS(fmap1);
  CHECK_CONTIGUOUS(fmap2);
  CHECK_CONTIGUOUS(coords);
  CHECK_CONTIGUOUS(corr_grad);
  require_device(fmap1, "fmap1");
  const AltDims d = alt_dims(fmap1, fmap2, coords, ii, jj);
  require(corr_grad.numel() == d.B * d.M * (2 * radius + 1) * (2 * radius + 1) * d.H * d.W,
          "altcorr_backward: corr_grad must match the forward output");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(fmap1.device());
  torch::Tensor g = corr_grad.to(torch::kFloat32).contiguous();
  torch::Tensor ii_c = ii.contiguous(), jj_c = jj.contiguous();
  torch::Tensor f1 = to_channels_last(fmap1, d), f2 = to_channels_last(fmap2, d);
  auto fopts = fmap1.options().dtype(torch::kFloat32);
  torch::Tensor g1 = torch::empty({d.B, d.N1, d.H, d.W, d.C}, fopts);
  torch::Tensor g2 = torch::empty({d.B, d.N2, d.H2, d.W2, d.C}, fopts);
  raise_status(hsl2_altcorr_backward(f1.data_ptr(), f2.data_ptr(), coords.data_ptr<float>(),
                                     g.data_ptr<float>(), ii_c.data_ptr<int64_t>(), jj_c.data_ptr<int64_t>(),
                                     g1.data_ptr<float>(), g2.data_ptr<float>(), dtype_code(fmap1), d.B, d.N1,
                                     d.N2, d.C, d.H, d.W, d.H2, d.W2, d.M, radius, current_stream()),
               "altcorr_backward");
  torch::Tensor o1 = torch::empty({d.B, d.N1, d.C, d.H, d.W}, fopts);
  torch::Tensor o2 = torch::empty({d.B, d.N2, d.C, d.H2, d.W2}, fopts);
  raise_status(hsl2_channels_first_f32(g1.data_ptr<float>(), o1.data_ptr<float>(), d.B * d.N1, d.C, d.H * d.W,
                                       current_stream()),
               "altcorr_backward");
  raise_status(hsl2_channels_first_f32(g2.data_ptr<float>(), o2.data_ptr<float>(), d.B * d.N2, d.C,
                                       d.H2 * d.W2, current_stream()),
               "altcorr_backward");
  return {o1.to(fmap1.scalar_type()), o2.to(fmap2.scalar_type())};
}

// ---- fp8 feature maps for the on-the-fly correlation (additions of this build) -------------
// fmap [B, N, C, H, W] (f32 or bf16) -> [q uint8 [B, N, H, W, C] (e4m3fn bits, channel-last),
// scales f32 [C]]
std::vector<torch::Tensor> fp8_quantize(torch::Tensor fmap) {
  CHECK_CONTIGUOUS(fmap);
  require_device(fmap, "fmap");
  require(fmap.dim() == 5, "fp8_quantize: fmap must be [B, N, C, H, W]");
  require(fmap.scalar_type() == torch::kFloat32 || fmap.scalar_type() == torch::kBFloat16,
          "fp8_quantize: fmap must be float32 or bfloat16");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(fmap.device());
  const int64_t B = fmap.size(0), N = fmap.size(1), C = fmap.size(2), H = fmap.size(3), W = fmap.size(4);
  torch::Tensor cl = torch::empty_like(fmap);
  raise_status(hsl2_channels_last(fmap.data_ptr(), cl.data_ptr(), dtype_code(fmap), B * N, C, H * W, current_stream()),
               "fp8_quantize");
  torch::Tensor q = torch::empty({B, N, H, W, C}, fmap.options().dtype(torch::kUInt8));
  torch::Tensor sc = torch::empty({C}, fmap.options().dtype(torch::kFloat32));
  raise_status(hsl2_fp8_quantize(cl.data_ptr(), dtype_code(fmap), B * N * H * W, C, q.data_ptr<uint8_t>(),
                                 sc.data_ptr<float>(), current_stream()),
               "fp8_quantize");
  return {q, sc};
}

// q1 [B, N1, H, W, C], q2 [B, N2, H2, W2, C] uint8 (fp8_quantize output), scales [C] f32,
// coords [B, M, 2, H, W] f32 -> corr [B, M, 2r+1, 2r+1, H, W] (bf16, or f32 when out_float)
std::vector<torch::Tensor> altcorr_forward_fp8(torch::Tensor q1, torch::Tensor s1, torch::Tensor q2,
                                               torch::Tensor s2, torch::Tensor coords, torch::Tensor ii,
                                               torch::Tensor jj, int radius, bool out_float) {
  for (auto* t : {&q1, &s1, &q2, &s2, &coords, &ii, &jj}) {
    CHECK_CONTIGUOUS(*t);
    require_device(*t, "altcorr_forward_fp8 tensors");
  }
  require(q1.dim() == 5 && q2.dim() == 5 && q1.scalar_type() == torch::kUInt8 && q2.scalar_type() == torch::kUInt8,
          "altcorr_forward_fp8: q1, q2 must be uint8 [B, N, H, W, C]");
  const int64_t B = q1.size(0), N1 = q1.size(1), H = q1.size(2), W = q1.size(3), C = q1.size(4);
  const int64_t N2 = q2.size(1), H2 = q2.size(2), W2 = q2.size(3);
  require(q2.size(0) == B && q2.size(4) == C, "altcorr_forward_fp8: inconsistent fmaps");
  require(s1.numel() == C && s2.numel() == C && s1.scalar_type() == torch::kFloat32 &&
              s2.scalar_type() == torch::kFloat32,
          "altcorr_forward_fp8: scales must be float32 [C]");
  require(coords.dim() == 5 && coords.size(0) == B && coords.size(2) == 2 && coords.size(3) == H &&
              coords.size(4) == W && coords.scalar_type() == torch::kFloat32,
          "altcorr_forward_fp8: coords must be float32 [B, M, 2, H, W]");
  const int64_t M = coords.size(1);
  require(ii.numel() == M && jj.numel() == M && ii.scalar_type() == torch::kInt64 && jj.scalar_type() == torch::kInt64,
          "altcorr_forward_fp8: ii, jj must be int64 with one entry per edge");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(q1.device());
  auto odt = out_float ? torch::kFloat32 : torch::kBFloat16;
  torch::Tensor corr = torch::empty({B, M, 2 * radius + 1, 2 * radius + 1, H, W}, coords.options().dtype(odt));
  raise_status(hsl2_altcorr_forward_fp8(q1.data_ptr<uint8_t>(), s1.data_ptr<float>(), q2.data_ptr<uint8_t>(),
                                        s2.data_ptr<float>(), coords.data_ptr<float>(), ii.data_ptr<int64_t>(),
                                        jj.data_ptr<int64_t>(), corr.data_ptr(), out_float ? HSL2_F32 : HSL2_BF16,
                                        B, N1, N2, C, H, W, H2, W2, M, radius, current_stream()),
               "altcorr_forward_fp8");
  return {corr};
}

// ---- correlation build (hislam2/modules/corr.py:24-38, 63-71) ---------------------------
std::vector<torch::Tensor> corr_build(torch::Tensor fmap1, torch::Tensor fmap2, int num_levels,
                                      py::object out_dtype, bool tiled) {
  CHECK_CONTIGUOUS(fmap1);
  CHECK_CONTIGUOUS(fmap2);
  require_device(fmap1, "fmap1");
  require(fmap1.dim() == 5 && fmap1.sizes() == fmap2.sizes() && fmap1.scalar_type() == fmap2.scalar_type(),
          "corr_build: fmaps must both be [B, N, C, h, w] of one dtype");
  torch::ScalarType od = out_dtype.is_none() ? fmap1.scalar_type()
                                             : torch::python::detail::py_object_to_dtype(out_dtype);
  const int64_t B = fmap1.size(0), N = fmap1.size(1), C = fmap1.size(2), h = fmap1.size(3), w = fmap1.size(4);
  c10::hip::HIPGuardMasqueradingAsCUDA guard(fmap1.device());
  std::vector<torch::Tensor> levels;
  std::vector<void*> ptrs;
  for (int i = 0; i < num_levels; ++i) {
    levels.push_back(torch::empty({B * N, h, w, h >> i, w >> i}, fmap1.options().dtype(od)));
    ptrs.push_back(levels.back().data_ptr());
  }
  int odc = od == torch::kBFloat16 ? HSL2_BF16 : od == torch::kFloat32 ? HSL2_F32 : od == torch::kFloat16 ? HSL2_F16 : -1;
  // the kernel reads the maps in MFMA fragment tiles: [B*N, C, h*w] -> [B*N, h*w/32, C/G, 32, G]
  torch::Tensor c1 = torch::empty_like(fmap1), c2 = torch::empty_like(fmap2);
  raise_status(hsl2_fmap_tiles(fmap1.data_ptr(), c1.data_ptr(), dtype_code(fmap1), B * N, C, h * w,
                               current_stream()),
               "corr_build");
  raise_status(hsl2_fmap_tiles(fmap2.data_ptr(), c2.data_ptr(), dtype_code(fmap2), B * N, C, h * w,
                               current_stream()),
               "corr_build");
  raise_status(hsl2_corr_build(c1.data_ptr(), c2.data_ptr(), dtype_code(fmap1), ptrs.data(), odc, B * N, C,
                               h, w, num_levels, tiled ? 1 : 0, current_stream()),
               "corr_build");
  return levels;
}

// ---- on-the-fly correlation pyramid (hislam2/modules/corr.py:89-117) ------------------------
// The feature pyramid of AltCorrBlock.__init__, channel-last: level l [B, N, H >> l, W >> l, C] in
// `dtype` (default: the fmaps' dtype), built once per block.
std::vector<torch::Tensor> altcorr_pyramid(torch::Tensor fmaps, int num_levels, py::object dtype) {
  require_device(fmaps, "fmaps");
  require(fmaps.dim() == 5, "altcorr_pyramid: fmaps must be [B, N, C, H, W]");
  require(num_levels >= 1 && num_levels <= 4, "altcorr_pyramid: 1 to 4 levels");
  const torch::ScalarType dt = dtype.is_none() ? fmaps.scalar_type() : torch::python::detail::py_object_to_dtype(dtype);
  require(dt == torch::kFloat32 || dt == torch::kBFloat16 || dt == torch::kFloat16,
          "altcorr_pyramid: dtype must be float32, float16 or bfloat16");
  const int64_t B = fmaps.size(0), N = fmaps.size(1), C = fmaps.size(2), H = fmaps.size(3), W = fmaps.size(4);
  require((H >> (num_levels - 1)) > 0 && (W >> (num_levels - 1)) > 0 && C % 8 == 0,
          "altcorr_pyramid: maps too small for the levels, or channels not a multiple of 8");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(fmaps.device());
  torch::Tensor f = fmaps.to(dt).contiguous();
  std::vector<torch::Tensor> levels;
  levels.push_back(torch::empty({B, N, H, W, C}, f.options()));
  raise_status(hsl2_channels_last(f.data_ptr(), levels[0].data_ptr(), dtype_code(f), B * N, C, H * W,
                                  current_stream()),
               "altcorr_pyramid");
  for (int l = 1; l < num_levels; ++l) {
    levels.push_back(torch::empty({B, N, H >> l, W >> l, C}, f.options()));
    raise_status(hsl2_pool_channels_last(levels[0].data_ptr(), levels[l].data_ptr(), dtype_code(f), B * N, H, W, C,
                                         l, current_stream()),
                 "altcorr_pyramid");
  }
  return levels;
}

// AltCorrBlock.__call__ in one launch: coords [B, M, H, W, 2] f32 -> [B, M, L (2r+1)^2, H, W]
torch::Tensor altcorr_pyramid_forward(std::vector<torch::Tensor> levels, torch::Tensor coords, torch::Tensor ii,
                                      torch::Tensor jj, int radius) {
  require(!levels.empty() && levels.size() <= 4, "altcorr_pyramid_forward: 1 to 4 levels");
  const torch::Tensor& l0 = levels[0];
  require_device(l0, "levels");
  require(l0.dim() == 5, "altcorr_pyramid_forward: levels must be [B, N, H >> l, W >> l, C]");
  const int64_t B = l0.size(0), N = l0.size(1), H = l0.size(2), W = l0.size(3), C = l0.size(4);
  std::vector<const void*> ptrs;
  for (size_t l = 0; l < levels.size(); ++l) {
    const torch::Tensor& t = levels[l];
    CHECK_CONTIGUOUS(t);
    require(t.scalar_type() == l0.scalar_type() && t.dim() == 5 && t.size(0) == B && t.size(1) == N &&
                t.size(2) == (H >> l) && t.size(3) == (W >> l) && t.size(4) == C,
            "altcorr_pyramid_forward: level " + std::to_string(l) + " must be [B, N, H >> l, W >> l, C]");
    ptrs.push_back(t.data_ptr());
  }
  CHECK_CONTIGUOUS(coords);
  require(coords.scalar_type() == torch::kFloat32 && coords.dim() == 5 && coords.size(0) == B &&
              coords.size(2) == H && coords.size(3) == W && coords.size(4) == 2,
          "altcorr_pyramid_forward: coords must be float32 [B, M, H, W, 2]");
  const int64_t M = coords.size(1);
  require(ii.numel() == M && jj.numel() == M && ii.scalar_type() == torch::kInt64 &&
              jj.scalar_type() == torch::kInt64 && ii.is_cuda() && jj.is_cuda(),
          "altcorr_pyramid_forward: ii, jj must be int64 GPU tensors with one entry per edge");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(l0.device());
  torch::Tensor ii_c = ii.contiguous(), jj_c = jj.contiguous();
  const int64_t R = 2 * radius + 1;
  torch::Tensor corr = torch::empty({B, M, static_cast<int64_t>(levels.size()) * R * R, H, W}, l0.options());
  raise_status(hsl2_altcorr_pyramid_forward(ptrs.data(), static_cast<int>(levels.size()), coords.data_ptr<float>(),
                                            ii_c.data_ptr<int64_t>(), jj_c.data_ptr<int64_t>(), corr.data_ptr(),
                                            dtype_code(l0), B, N, C, H, W, M, radius, current_stream()),
               "altcorr_pyramid_forward");
  return corr;
}

}  // namespace

// ---- edge-selection and point-cloud geometry (src/droid.cpp:170-241) ----------------------
namespace {
void check_geometry(torch::Tensor& poses, torch::Tensor& disps, torch::Tensor& intrinsics, const char* what) {
  for (auto* t : {&poses, &disps, &intrinsics}) {
    CHECK_CONTIGUOUS(*t);
    require_device(*t, what);
    require(t->scalar_type() == torch::kFloat32, std::string(what) + ": poses, disps, intrinsics must be float32");
  }
  require(poses.dim() == 2 && poses.size(1) == 7 && disps.dim() == 3 && disps.size(0) <= poses.size(0),
          std::string(what) + ": poses [N, 7], disps [N, h, w]");
  require(intrinsics.numel() >= 4, std::string(what) + ": intrinsics needs fx, fy, cx, cy");
}

void check_index(torch::Tensor& idx, int64_t lo_ok, int64_t hi_ok, int64_t lo_shift, int64_t hi_shift,
                 const char* what) {
  CHECK_CONTIGUOUS(idx);
  require_device(idx, what);
  require(idx.scalar_type() == torch::kInt64 && idx.dim() == 1, std::string(what) + ": indices must be int64 [E]");
  if (idx.numel() > 0) {
    const int64_t lo = idx.min().item<int64_t>() + lo_shift, hi = idx.max().item<int64_t>() + hi_shift;
    if (lo < lo_ok || hi > hi_ok) throw py::value_error(std::string(what) + ": an index refers to a frame outside the range");
  }
}
}  // namespace

torch::Tensor frame_distance(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics, torch::Tensor ii,
                             torch::Tensor jj, double beta) {
  check_geometry(poses, disps, intrinsics, "frame_distance");
  const int64_t N = disps.size(0);
  check_index(ii, 0, N - 1, 0, 0, "frame_distance");
  check_index(jj, 0, N - 1, 0, 0, "frame_distance");
  require(ii.numel() == jj.numel(), "frame_distance: ii and jj differ in length");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());
  torch::Tensor dist = torch::empty({ii.numel()}, poses.options());
  raise_status(hsl2_frame_distance(poses.data_ptr<float>(), disps.data_ptr<float>(), intrinsics.data_ptr<float>(),
                                   ii.data_ptr<int64_t>(), jj.data_ptr<int64_t>(), ii.numel(), disps.size(1),
                                   disps.size(2), static_cast<float>(beta), dist.data_ptr<float>(), current_stream()),
               "frame_distance");
  return dist;
}

torch::Tensor covis_distance(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics, torch::Tensor ii) {
  check_geometry(poses, disps, intrinsics, "covis_distance");
  const int64_t N = disps.size(0);
  check_index(ii, 0, N - 1, -1, 1, "covis_distance");   // ii - 1 and ii + 1 must be frames
  c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());
  torch::Tensor dist = torch::empty({ii.numel()}, poses.options());
  raise_status(hsl2_covis_distance(poses.data_ptr<float>(), disps.data_ptr<float>(), intrinsics.data_ptr<float>(),
                                   ii.data_ptr<int64_t>(), ii.numel(), disps.size(1), disps.size(2),
                                   dist.data_ptr<float>(), current_stream()),
               "covis_distance");
  return dist;
}

torch::Tensor depth_filter(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics, torch::Tensor ix,
                           torch::Tensor thresh) {
  check_geometry(poses, disps, intrinsics, "depth_filter");
  const int64_t N = disps.size(0);
  check_index(ix, 0, N - 1, 0, 0, "depth_filter");
  CHECK_CONTIGUOUS(thresh);
  require_device(thresh, "depth_filter");
  require(thresh.scalar_type() == torch::kFloat32 && thresh.numel() == ix.numel(),
          "depth_filter: thresh must be float32 with one value per index");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());
  torch::Tensor counter = torch::empty({ix.numel(), disps.size(1), disps.size(2)}, disps.options());
  raise_status(hsl2_depth_filter(poses.data_ptr<float>(), disps.data_ptr<float>(), intrinsics.data_ptr<float>(),
                                 ix.data_ptr<int64_t>(), thresh.data_ptr<float>(), ix.numel(), N, disps.size(1),
                                 disps.size(2), counter.data_ptr<float>(), current_stream()),
               "depth_filter");
  return counter;
}

torch::Tensor iproj(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics) {
  check_geometry(poses, disps, intrinsics, "iproj");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());
  torch::Tensor points = torch::empty({disps.size(0), disps.size(1), disps.size(2), 3}, disps.options());
  raise_status(hsl2_iproj(poses.data_ptr<float>(), disps.data_ptr<float>(), intrinsics.data_ptr<float>(),
                          disps.size(0), disps.size(1), disps.size(2), points.data_ptr<float>(), current_stream()),
               "iproj");
  return points;
}

// ---- mono-prior step pieces (src/droid.cpp:121-138, 244-250) ----------------------------
// proj_trans(poses [N, 7], disps [N, h, w], intrinsics (first 4 values used), targets / weights
// [E, 2, h, w], ii, jj) -> [C, w], each [M, h * w] over the sorted unique ii frames.
std::vector<torch::Tensor> proj_trans(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics,
                                      torch::Tensor targets, torch::Tensor weights, torch::Tensor ii,
                                      torch::Tensor jj) {
  for (auto* t : {&poses, &disps, &intrinsics, &targets, &weights, &ii, &jj}) {
    CHECK_CONTIGUOUS(*t);
    require_device(*t, "proj_trans tensors");
  }
  for (auto* t : {&poses, &disps, &intrinsics, &targets, &weights})
    require(t->scalar_type() == torch::kFloat32, "proj_trans: poses, disps, intrinsics, targets, weights must be float32");
  require(ii.scalar_type() == torch::kInt64 && jj.scalar_type() == torch::kInt64 && ii.dim() == 1 &&
              jj.dim() == 1 && ii.size(0) == jj.size(0),
          "proj_trans: ii, jj must be int64 [E]");
  const int64_t N = poses.size(0), E = ii.size(0);
  require(poses.dim() == 2 && poses.size(1) == 7, "proj_trans: poses must be [N, 7]");
  require(disps.dim() == 3 && disps.size(0) == N, "proj_trans: disps must be [N, h, w]");
  require(intrinsics.numel() >= 4, "proj_trans: intrinsics needs fx, fy, cx, cy");
  const int64_t h = disps.size(1), w = disps.size(2);
  require(targets.numel() == E * 2 * h * w && weights.numel() == E * 2 * h * w,
          "proj_trans: targets, weights must be [E, 2, h, w]");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());
  torch::Tensor ii_h = ii.to(torch::kCPU).contiguous(), jj_h = jj.to(torch::kCPU).contiguous();
  if (E > 0) {
    auto lo = std::min(ii_h.min().item<int64_t>(), jj_h.min().item<int64_t>());
    auto hi = std::max(ii_h.max().item<int64_t>(), jj_h.max().item<int64_t>());
    require(lo >= 0 && hi < N, "proj_trans: an edge indexes a frame outside [0, N)");
  }
  torch::Tensor kx = torch::empty({std::max<int64_t>(E, 1)}, torch::kInt64);
  torch::Tensor fptr = torch::empty({E + 1}, torch::kInt32), fedge = torch::empty({std::max<int64_t>(E, 1)}, torch::kInt32);
  int64_t M = 0;
  raise_status(hsl2_edge_frames(ii_h.data_ptr<int64_t>(), E, kx.data_ptr<int64_t>(), fptr.data_ptr<int32_t>(),
                                fedge.data_ptr<int32_t>(), &M),
               "proj_trans");
  auto opts = poses.options();
  torch::Tensor C = torch::empty({M, h * w}, opts), wv = torch::empty({M, h * w}, opts);
  if (M == 0) return {C, wv};
  torch::Tensor fptr_d = fptr.slice(0, 0, M + 1).to(poses.device()), fedge_d = fedge.slice(0, 0, E).to(poses.device());
  raise_status(hsl2_proj_trans(poses.data_ptr<float>(), disps.data_ptr<float>(), intrinsics.data_ptr<float>(),
                               targets.data_ptr<float>(), weights.data_ptr<float>(), ii.data_ptr<int64_t>(),
                               jj.data_ptr<int64_t>(), fptr_d.data_ptr<int32_t>(), fedge_d.data_ptr<int32_t>(), M, h,
                               w, C.data_ptr<float>(), wv.data_ptr<float>(), current_stream()),
               "proj_trans");
  return {C, wv};
}

// bi_inter(scales [M, hs, ws], grids [M, h, w, 2]) -> [outputs [M, h, w], Jacobis [M, h, w, hs, ws]]
std::vector<torch::Tensor> bi_inter(torch::Tensor scales, torch::Tensor grids) {
  for (auto* t : {&scales, &grids}) {
    CHECK_CONTIGUOUS(*t);
    require_device(*t, "bi_inter tensors");
    require(t->scalar_type() == torch::kFloat32, "bi_inter: float32 tensors");
  }
  require(scales.dim() == 3 && grids.dim() == 4 && grids.size(3) == 2 && grids.size(0) == scales.size(0),
          "bi_inter: scales [M, hs, ws], grids [M, h, w, 2]");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(scales.device());
  const int64_t M = grids.size(0), h = grids.size(1), w = grids.size(2), hs = scales.size(1), ws = scales.size(2);
  auto opts = grids.options();
  torch::Tensor out = torch::empty({M, h, w}, opts), J = torch::empty({M, h, w, hs, ws}, opts);
  raise_status(hsl2_bi_inter(scales.data_ptr<float>(), grids.data_ptr<float>(), M, h, w, hs, ws,
                             out.data_ptr<float>(), J.data_ptr<float>(), current_stream()),
               "bi_inter");
  return {out, J};
}

// jdsa_solve(C, w, disps [N, h, w], prior, dscales [N, hs, ws], eta [M, h, w], kx [M] int64,
// slot_src [M] int32, gx [w], gy [h], alpha, ep, lm) -> [disps_out, dzcov]; dscales updated in place.
std::vector<torch::Tensor> jdsa_solve(torch::Tensor C, torch::Tensor w, torch::Tensor disps, torch::Tensor prior,
                                      torch::Tensor dscales, torch::Tensor eta, torch::Tensor kx,
                                      torch::Tensor slot_src, torch::Tensor gx, torch::Tensor gy, double alpha,
                                      double ep, double lm) {
  for (auto* t : {&C, &w, &disps, &prior, &dscales, &eta, &kx, &slot_src, &gx, &gy}) {
    CHECK_CONTIGUOUS(*t);
    require_device(*t, "jdsa tensors");
  }
  for (auto* t : {&C, &w, &disps, &prior, &dscales, &eta, &gx, &gy})
    require(t->scalar_type() == torch::kFloat32, "jdsa: float32 tensors");
  require(kx.scalar_type() == torch::kInt64 && slot_src.scalar_type() == torch::kInt32,
          "jdsa: kx int64, slot_src int32");
  const int64_t M = C.size(0), h = disps.size(1), wd = disps.size(2), hs = dscales.size(1), ws = dscales.size(2);
  require(disps.dim() == 3 && dscales.dim() == 3 && dscales.size(0) == disps.size(0), "jdsa: disps [N, h, w], dscales [N, hs, ws]");
  require(C.dim() == 2 && C.size(1) == h * wd && w.sizes() == C.sizes() && eta.numel() == M * h * wd &&
              kx.numel() == M && slot_src.numel() == M && gx.numel() == wd && gy.numel() == h &&
              prior.numel() >= disps.numel() && prior.size(-1) == wd,
          "jdsa: inconsistent shapes");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(C.device());
  torch::Tensor disps_out = disps.clone(), dscales_in = dscales.clone();
  torch::Tensor dzcov = torch::empty({M, h * wd}, C.options());
  torch::Tensor ws_t = torch::empty({static_cast<int64_t>(hsl2_jdsa_workspace_bytes(M)) + 256},
                                    C.options().dtype(torch::kUInt8));
  void* wp = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(ws_t.data_ptr()) + 255) & ~uintptr_t(255));
  raise_status(hsl2_jdsa_solve(C.data_ptr<float>(), w.data_ptr<float>(), disps.data_ptr<float>(),
                               prior.data_ptr<float>(), dscales_in.data_ptr<float>(), eta.data_ptr<float>(),
                               kx.data_ptr<int64_t>(), slot_src.data_ptr<int32_t>(), gx.data_ptr<float>(),
                               gy.data_ptr<float>(), M, h, wd, hs, ws, static_cast<float>(alpha), static_cast<float>(ep),
                               static_cast<float>(lm), wp, static_cast<size_t>(ws_t.numel() - 256),
                               disps_out.data_ptr<float>(), dscales.data_ptr<float>(), dzcov.data_ptr<float>(),
                               current_stream()),
               "jdsa_solve");
  return {disps_out, dzcov};
}

// ---- projective transform of the torch formulation (hislam2/geom/projective_ops.py:44-74) ----
void check_projective(torch::Tensor& poses, torch::Tensor& disps, torch::Tensor& intrinsics, torch::Tensor& ii,
                      torch::Tensor& jj) {
  for (auto* t : {&poses, &disps, &intrinsics, &ii, &jj}) {
    CHECK_CONTIGUOUS(*t);
    require_device(*t, "projective_transform tensors");
  }
  for (auto* t : {&poses, &disps, &intrinsics})
    require(t->scalar_type() == torch::kFloat32, "projective_transform: poses, disps, intrinsics must be float32");
  require(ii.scalar_type() == torch::kInt64 && jj.scalar_type() == torch::kInt64 && ii.dim() == 1 &&
              jj.dim() == 1 && ii.size(0) == jj.size(0),
          "projective_transform: ii, jj must be int64 [E]");
  const int64_t N = poses.size(0);
  require(poses.dim() == 2 && poses.size(1) == 7, "projective_transform: poses must be [N, 7]");
  require(disps.dim() == 3 && disps.size(0) == N, "projective_transform: disps must be [N, h, w]");
  require(intrinsics.numel() == N * 4, "projective_transform: intrinsics must be [N, 4]");
  if (ii.size(0) > 0) {
    auto lo = torch::stack({ii.min(), jj.min()}).min().item<int64_t>();
    auto hi = torch::stack({ii.max(), jj.max()}).max().item<int64_t>();
    require(lo >= 0 && hi < N, "projective_transform: an edge indexes a frame outside [0, N)");
  }
}

std::vector<torch::Tensor> projective_transform(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics,
                                                torch::Tensor ii, torch::Tensor jj, bool jacobian,
                                                bool return_depth) {
  check_projective(poses, disps, intrinsics, ii, jj);
  c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());
  const int64_t E = ii.size(0), h = disps.size(1), w = disps.size(2);
  auto opts = poses.options();
  torch::Tensor coords = torch::empty({E, h, w, return_depth ? 3 : 2}, opts);
  torch::Tensor valid = torch::empty({E, h, w, 1}, opts);
  torch::Tensor Ji, Jj, Jz;
  if (jacobian) {
    Ji = torch::empty({E, h, w, 2, 6}, opts);
    Jj = torch::empty({E, h, w, 2, 6}, opts);
    Jz = torch::empty({E, h, w, 2, 1}, opts);
  }
  raise_status(hsl2_projective_transform(poses.data_ptr<float>(), disps.data_ptr<float>(),
                                         intrinsics.data_ptr<float>(), ii.data_ptr<int64_t>(),
                                         jj.data_ptr<int64_t>(), E, h, w, return_depth ? 1 : 0,
                                         coords.data_ptr<float>(), valid.data_ptr<float>(),
                                         jacobian ? Ji.data_ptr<float>() : nullptr,
                                         jacobian ? Jj.data_ptr<float>() : nullptr,
                                         jacobian ? Jz.data_ptr<float>() : nullptr, current_stream()),
               "projective_transform");
  return {coords, valid, Ji, Jj, Jz};
}

// returns [grad_poses (left tangent, [N, 6]), grad_disps [N, h, w]]
std::vector<torch::Tensor> projective_transform_backward(torch::Tensor poses, torch::Tensor disps,
                                                         torch::Tensor intrinsics, torch::Tensor ii,
                                                         torch::Tensor jj, torch::Tensor grad_coords,
                                                         bool return_depth) {
  check_projective(poses, disps, intrinsics, ii, jj);
  CHECK_CONTIGUOUS(grad_coords);
  const int64_t N = poses.size(0), E = ii.size(0), h = disps.size(1), w = disps.size(2);
  require(grad_coords.is_cuda() && grad_coords.scalar_type() == torch::kFloat32 &&
              grad_coords.numel() == E * h * w * (return_depth ? 3 : 2),
          "projective_transform backward: grad_coords must be float32 [E, h, w, 2 or 3]");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());
  auto opts = poses.options();
  torch::Tensor gp = torch::empty({N, 6}, opts), gd = torch::empty({N, h, w}, opts);
  const size_t bytes = hsl2_projective_transform_workspace_bytes(E, h, w);
  torch::Tensor ws = torch::empty({static_cast<int64_t>(bytes) + 256}, opts.dtype(torch::kUInt8));
  uintptr_t base = reinterpret_cast<uintptr_t>(ws.data_ptr());
  void* ws_ptr = reinterpret_cast<void*>((base + 255) & ~uintptr_t(255));
  raise_status(hsl2_projective_transform_backward(poses.data_ptr<float>(), disps.data_ptr<float>(),
                                                  intrinsics.data_ptr<float>(), ii.data_ptr<int64_t>(),
                                                  jj.data_ptr<int64_t>(), E, N, h, w, return_depth ? 1 : 0,
                                                  grad_coords.data_ptr<float>(), gp.data_ptr<float>(),
                                                  gd.data_ptr<float>(), ws_ptr, bytes, current_stream()),
               "projective_transform backward");
  return {gp, gd};
}

// ---- one GN step of the torch formulation (hislam2/geom/ba.py:32-158) -------------
// Holds the plan of one (ii, jj, fixedp) graph and the workspace of its last forward step,
// which the backward of the same step reads.
class GNStep {
 public:
  GNStep(torch::Tensor ii, torch::Tensor jj, int64_t num_poses, int64_t ht, int64_t wd, int fixedp,
         bool motion_only)
      : N_(num_poses), ht_(ht), wd_(wd), motion_only_(motion_only) {
    require(ii.dim() == 1 && jj.dim() == 1 && ii.size(0) == jj.size(0), "gn_step: ii, jj must be [E]");
    require(ii.scalar_type() == torch::kInt64 && jj.scalar_type() == torch::kInt64,
            "gn_step: ii, jj must be int64");
    torch::Tensor ii_h = ii.to(torch::kCPU).contiguous(), jj_h = jj.to(torch::kCPU).contiguous();
    E_ = ii.size(0);
    raise_status(hsl2_ba_plan_create_ex(ii_h.data_ptr<int64_t>(), jj_h.data_ptr<int64_t>(), E_, num_poses,
                                        ht, wd, fixedp, static_cast<int>(num_poses), motion_only ? 1 : 0,
                                        HSL2_BA_SPLIT, HSL2_SEM_TORCH, &plan_),
                 "gn_step");
    F_ = hsl2_ba_plan_num_frames(plan_);
    P_ = num_poses - fixedp;
    kx_ = torch::empty({F_}, torch::kInt64);
    raise_status(hsl2_ba_plan_frames(plan_, kx_.data_ptr<int64_t>()), "gn_step");
  }
  ~GNStep() { hsl2_ba_plan_destroy(plan_); }
  GNStep(const GNStep&) = delete;
  GNStep& operator=(const GNStep&) = delete;

  torch::Tensor frames() const { return kx_; }

  // poses [N, 7] and disps [N, ht, wd] are updated in place; returns [dx, dz, dzcov, ok]
  std::vector<torch::Tensor> forward(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics,
                                     torch::Tensor targets, torch::Tensor weights, torch::Tensor eta,
                                     double lm, double ep, bool want_cov) {
    check_state(poses, disps, intrinsics, targets, weights);
    const int64_t HW = ht_ * wd_;
    const float* eta_ptr = nullptr;
    if (!motion_only_) {
      CHECK_CONTIGUOUS(eta);
      require_device(eta, "eta");
      require(eta.scalar_type() == torch::kFloat32 && eta.numel() == F_ * HW,
              "gn_step: eta must be float32 with |unique(ii)| x ht x wd elements");
      eta_ptr = eta.data_ptr<float>();
    }
    c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());
    auto opts = poses.options();
    ws_ = torch::empty({static_cast<int64_t>(hsl2_ba_plan_workspace_bytes(plan_)) + 256}, opts.dtype(torch::kUInt8));
    torch::Tensor dx = torch::zeros({P_, 6}, opts);
    torch::Tensor dz = motion_only_ ? torch::Tensor() : torch::zeros({F_, HW}, opts);
    torch::Tensor cov = (!motion_only_ && want_cov) ? torch::zeros({F_, HW}, opts) : torch::Tensor();
    int ok = 1;
    raise_status(hsl2_gn_step(plan_, poses.data_ptr<float>(), disps.data_ptr<float>(), intrinsics.data_ptr<float>(),
                              E_ ? targets.data_ptr<float>() : nullptr, E_ ? weights.data_ptr<float>() : nullptr,
                              eta_ptr, static_cast<float>(lm), static_cast<float>(ep), dx.data_ptr<float>(),
                              dz.defined() ? dz.data_ptr<float>() : nullptr,
                              cov.defined() ? cov.data_ptr<float>() : nullptr, ws_ptr(), ws_bytes(), &ok,
                              current_stream()),
                 "gn_step");
    return {dx, dz, cov, torch::tensor({ok}, torch::kInt32)};
  }

  // gradients of one step given the step's input state and outputs; returns
  // [grad_targets, grad_weights, grad_eta, grad_poses (tangent, [N, 6]), grad_disps]
  std::vector<torch::Tensor> backward(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics,
                                      torch::Tensor targets, torch::Tensor weights, double lm, double ep,
                                      torch::Tensor dx, torch::Tensor dz, torch::Tensor grad_dx,
                                      torch::Tensor grad_dz) {
    require(ws_.defined(), "gn_step: backward before forward");
    check_state(poses, disps, intrinsics, targets, weights);
    const int64_t HW = ht_ * wd_;
    for (auto* t : {&dx, &grad_dx}) {
      CHECK_CONTIGUOUS(*t);
      require(t->is_cuda() && t->scalar_type() == torch::kFloat32 && t->numel() == P_ * 6,
              "gn_step: dx and its gradient must be float32 [N - fixedp, 6]");
    }
    if (!motion_only_)
      for (auto* t : {&dz, &grad_dz}) {
        CHECK_CONTIGUOUS(*t);
        require(t->is_cuda() && t->scalar_type() == torch::kFloat32 && t->numel() == F_ * HW,
                "gn_step: dz and its gradient must be float32 [|unique(ii)|, ht * wd]");
      }
    c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());
    auto opts = poses.options();
    torch::Tensor gt = torch::empty({E_, 2, ht_, wd_}, opts), gw = torch::empty({E_, 2, ht_, wd_}, opts);
    torch::Tensor geta = motion_only_ ? torch::Tensor() : torch::empty({F_, ht_, wd_}, opts);
    torch::Tensor gp = torch::empty({N_, 6}, opts), gd = torch::empty({N_, ht_, wd_}, opts);
    raise_status(hsl2_gn_step_backward(
                     plan_, poses.data_ptr<float>(), disps.data_ptr<float>(), intrinsics.data_ptr<float>(),
                     E_ ? targets.data_ptr<float>() : nullptr, E_ ? weights.data_ptr<float>() : nullptr,
                     static_cast<float>(lm), static_cast<float>(ep), dx.data_ptr<float>(),
                     motion_only_ ? nullptr : dz.data_ptr<float>(), grad_dx.data_ptr<float>(),
                     motion_only_ ? nullptr : grad_dz.data_ptr<float>(), E_ ? gt.data_ptr<float>() : nullptr,
                     E_ ? gw.data_ptr<float>() : nullptr, motion_only_ ? nullptr : geta.data_ptr<float>(),
                     gp.data_ptr<float>(), gd.data_ptr<float>(), ws_ptr(), ws_bytes(), current_stream()),
                 "gn_step backward");
    return {gt, gw, geta, gp, gd};
  }

 private:
  void check_state(torch::Tensor& poses, torch::Tensor& disps, torch::Tensor& intrinsics, torch::Tensor& targets,
                   torch::Tensor& weights) {
    for (auto* t : {&poses, &disps, &intrinsics, &targets, &weights}) {
      CHECK_CONTIGUOUS(*t);
      require_device(*t, "gn_step tensors");
      require(t->scalar_type() == torch::kFloat32, "gn_step: tensors must be float32");
    }
    require(poses.dim() == 2 && poses.size(0) == N_ && poses.size(1) == 7, "gn_step: poses must be [N, 7]");
    require(disps.dim() == 3 && disps.size(0) == N_ && disps.size(1) == ht_ && disps.size(2) == wd_,
            "gn_step: disps must be [N, ht, wd]");
    require(intrinsics.numel() == N_ * 4, "gn_step: intrinsics must be [N, 4]");
    require(targets.numel() == E_ * 2 * ht_ * wd_ && weights.numel() == E_ * 2 * ht_ * wd_,
            "gn_step: targets and weights must be [E, 2, ht, wd]");
  }
  void* ws_ptr() const {
    uintptr_t base = reinterpret_cast<uintptr_t>(ws_.data_ptr());
    return reinterpret_cast<void*>((base + 255) & ~uintptr_t(255));
  }
  size_t ws_bytes() const { return static_cast<size_t>(ws_.numel() - 256); }

  hsl2_ba_plan* plan_ = nullptr;
  int64_t N_, ht_, wd_, E_ = 0, F_ = 0, P_ = 0;
  bool motion_only_;
  torch::Tensor kx_, ws_;
};

// ---- pgba (src/droid.cpp:171-195) ----------------------------------------------
// The last few pgba plans, keyed by the concatenated index lists ii ++ jj ++ iip ++ jjp and the
// sizes; a hit compares the whole key.
class PgbaPlanCache {
 public:
  using Args = std::array<int64_t, 5>;   // num_poses, ht, wd, t0, t1
  using Ptr = std::shared_ptr<const hsl2_pgba_plan>;

  Ptr get(const int64_t* idx, int64_t E, int64_t Np, const Args& a) {
    const size_t n = static_cast<size_t>(2 * E + 2 * Np);
    {
      std::lock_guard<std::mutex> lock(mu_);
      for (auto it = entries_.begin(); it != entries_.end(); ++it) {
        if (it->args == a && it->E == E && it->idx.size() == n &&
            std::memcmp(it->idx.data(), idx, n * sizeof(int64_t)) == 0) {
          entries_.splice(entries_.begin(), entries_, it);
          return entries_.front().plan;
        }
      }
    }
    hsl2_pgba_plan* raw = nullptr;
    raise_status(hsl2_pgba_plan_create(idx, idx + E, E, idx + 2 * E, idx + 2 * E + Np, Np, a[0], a[1], a[2],
                                       static_cast<int>(a[3]), static_cast<int>(a[4]), &raw),
                 "pgba");
    Ptr plan(raw, hsl2_pgba_plan_destroy);
    std::lock_guard<std::mutex> lock(mu_);
    entries_.push_front(Entry{a, E, std::vector<int64_t>(idx, idx + n), plan});
    if (entries_.size() > kCapacity) entries_.pop_back();
    return plan;
  }

 private:
  struct Entry {
    Args args;
    int64_t E;
    std::vector<int64_t> idx;
    Ptr plan;
  };
  static constexpr size_t kCapacity = 4;
  std::mutex mu_;
  std::list<Entry> entries_;
};

PgbaPlanCache& pgba_plan_cache() {
  static PgbaPlanCache cache;
  return cache;
}

// Sim3 pose-graph BA step: [dx [t1 - t0, 7], dz [F, ht * wd]], disps updated in place.
std::vector<torch::Tensor> pgba(torch::Tensor poses, torch::Tensor disps, torch::Tensor intrinsics,
                                torch::Tensor targets, torch::Tensor weights, torch::Tensor eta, torch::Tensor Hsp,
                                torch::Tensor vsp, torch::Tensor ii, torch::Tensor jj, torch::Tensor iip,
                                torch::Tensor jjp, int t0, int t1, double lm, double ep) {
  for (auto* t : {&poses, &disps, &intrinsics, &targets, &weights, &eta, &Hsp, &vsp}) {
    CHECK_CONTIGUOUS(*t);
    require_device(*t, "pgba tensors");
    require(t->scalar_type() == torch::kFloat32, "pgba: poses, disps, intrinsics, targets, weights, eta, Hsp, vsp "
                                                 "must be float32");
  }
  for (auto* t : {&ii, &jj, &iip, &jjp}) {
    CHECK_CONTIGUOUS(*t);
    require(t->dim() == 1 && t->scalar_type() == torch::kInt64, "pgba: ii, jj, iip, jjp must be int64 [*]");
  }
  require(poses.dim() == 2 && poses.size(1) == 8, "pgba: poses must be [N, 8] (t, q, s)");
  require(disps.dim() == 3 && disps.size(0) >= poses.size(0), "pgba: disps must be [N, ht, wd]");
  require(intrinsics.numel() >= 4, "pgba: intrinsics must hold fx, fy, cx, cy");
  require(ii.size(0) == jj.size(0) && iip.size(0) == jjp.size(0), "pgba: ii / jj and iip / jjp differ in length");
  const int64_t E = ii.size(0), Np = iip.size(0), ht = disps.size(1), wd = disps.size(2);
  require(targets.numel() == E * 2 * ht * wd && weights.numel() == E * 2 * ht * wd,
          "pgba: targets and weights must be [E, 2, ht, wd]");
  require(Hsp.numel() == Np * 4 * 49 && vsp.numel() == Np * 2 * 7, "pgba: Hsp must be [4, Np, 7, 7], vsp [2, Np, 7]");
  c10::hip::HIPGuardMasqueradingAsCUDA guard(poses.device());
  // one host copy of the four index lists (the caller's loop repeats the same graph: update_pgba
  // makes 6 x 2 calls per loop closure), then the cached plan when the lists and sizes repeat
  torch::Tensor idx_h = torch::cat({ii, jj, iip, jjp}).to(torch::kCPU);
  const int64_t* ih = idx_h.data_ptr<int64_t>();
  const std::shared_ptr<const hsl2_pgba_plan> plan =
      pgba_plan_cache().get(ih, E, Np, {poses.size(0), ht, wd, t0, t1});
  const int64_t F = hsl2_pgba_plan_num_frames(plan.get());
  require(eta.numel() % (ht * wd) == 0, "pgba: eta must hold whole [ht, wd] rows");
  const int64_t eta_rows = eta.numel() / (ht * wd);
  require(eta_rows == F || eta_rows == 1, "pgba: eta must have one [ht, wd] row per unique frame of [t0, t1) ++ ii (" +
                                              std::to_string(F) + ") or a single row");
  auto opts = poses.options();
  torch::Tensor ws = torch::empty({static_cast<int64_t>(hsl2_pgba_workspace_bytes(plan.get())) + 256},
                                  opts.dtype(torch::kUInt8));
  uintptr_t base = reinterpret_cast<uintptr_t>(ws.data_ptr());
  void* ws_ptr = reinterpret_cast<void*>((base + 255) & ~uintptr_t(255));
  torch::Tensor dx = torch::empty({t1 - t0, 7}, opts);
  torch::Tensor dz = torch::empty({F, ht * wd}, opts);
  raise_status(hsl2_pgba_run(plan.get(), poses.data_ptr<float>(), disps.data_ptr<float>(), intrinsics.data_ptr<float>(),
                             E ? targets.data_ptr<float>() : nullptr, E ? weights.data_ptr<float>() : nullptr,
                             eta.data_ptr<float>(), eta_rows, Np ? Hsp.data_ptr<float>() : nullptr,
                             Np ? vsp.data_ptr<float>() : nullptr, static_cast<float>(lm), static_cast<float>(ep),
                             dx.data_ptr<float>(), dz.data_ptr<float>(), ws_ptr,
                             static_cast<size_t>(ws.numel() - 256), nullptr, current_stream()),
               "pgba");
  return {dx, dz};
}

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "MI355X (gfx950) dense-tracking kernels behind the droid_backends interface";
  // The calls run without the GIL (they touch no Python object; the plan caches are guarded by
  // their own mutexes and the library's error message is per thread), so that a tracking thread
  // in ba does not stall the other Python threads.  corr_build and altcorr_pyramid read a
  // Python dtype argument and keep the GIL.
  const auto kNoGil = py::call_guard<py::gil_scoped_release>();
  // bundle adjustment (src/droid.cpp:327)
  m.def("ba", &ba, "dense bundle adjustment (split precision: FP32 linearization, FP64 reduced solve)", kNoGil);
  m.def("ba", &ba_sens, "dense bundle adjustment with a sensor-disparity prior (15-argument form)", kNoGil);
  m.def("pgba", &pgba, "Sim3 pose-graph bundle adjustment step: [dx, dz], disps updated in place", kNoGil);
  // correlation volume kernels (src/droid.cpp:337-340)
  m.def("altcorr_forward", &altcorr_forward, "on-the-fly correlation lookup", kNoGil);
  m.def("altcorr_backward", &altcorr_backward, "on-the-fly correlation lookup adjoint", kNoGil);
  m.def("corr_index_forward", &corr_index_forward, "radius-r lookup into a correlation volume", kNoGil);
  m.def("corr_index_backward", &corr_index_backward, "adjoint of corr_index_forward", kNoGil);
  // additions of this build
  m.def("ba_ex", &ba_ex, "ba with an explicit precision mode (0 split, 1 all-FP32); returns [dx, dz, None, ok]", kNoGil);
  m.def("ba_index", &ba_index, "kx and kk_exp of a ba call (host-side planner)", kNoGil);
  m.def("spd_solve", &spd_solve, "dense SPD solve with the tiled GPU Cholesky; returns [x, ok]", kNoGil);
  m.def("corr_pyramid_lookup", &corr_pyramid_lookup, "CorrBlock lookup over all levels in one launch", kNoGil,
        py::arg("levels"), py::arg("coords"), py::arg("radius"), py::arg("tiled") = false);
  m.def("corr_lookup_conv", &corr_lookup_conv, "CorrBlock lookup fused with a 1x1 convolution (+ ReLU)", kNoGil,
        py::arg("levels"), py::arg("coords"), py::arg("radius"), py::arg("tiled"), py::arg("weight"),
        py::arg("bias"), py::arg("relu") = true);
  m.def("corr_build", &corr_build, "all-pairs correlation + average-pool pyramid", py::arg("fmap1"),
        py::arg("fmap2"), py::arg("num_levels") = 4, py::arg("out_dtype") = py::none(),
        py::arg("tiled") = false);
  m.def("altcorr_pyramid", &altcorr_pyramid, "channel-last feature pyramid of AltCorrBlock", py::arg("fmaps"),
        py::arg("num_levels") = 4, py::arg("dtype") = py::none());
  m.def("altcorr_pyramid_forward", &altcorr_pyramid_forward,
        "AltCorrBlock lookup over all levels in one launch on the matrix cores", kNoGil, py::arg("levels"), py::arg("coords"),
        py::arg("ii"), py::arg("jj"), py::arg("radius") = 3);
  m.def("fp8_quantize", &fp8_quantize, "per-channel-scaled e4m3fn quantisation: [q (channel-last uint8), scales]", kNoGil);
  m.def("altcorr_forward_fp8", &altcorr_forward_fp8, "altcorr lookup on fp8 feature maps", kNoGil, py::arg("q1"),
        py::arg("s1"), py::arg("q2"), py::arg("s2"), py::arg("coords"), py::arg("ii"), py::arg("jj"),
        py::arg("radius"), py::arg("out_float") = false);
  m.def("proj_trans", &proj_trans, "depth-only terms summed per ii frame: [C, w]", kNoGil, py::arg("poses"),
        py::arg("disps"), py::arg("intrinsics"), py::arg("targets"), py::arg("weights"), py::arg("ii"),
        py::arg("jj"));
  m.def("frame_distance", &frame_distance, "beta-weighted mean flow of each edge", kNoGil, py::arg("poses"),
        py::arg("disps"), py::arg("intrinsics"), py::arg("ii"), py::arg("jj"), py::arg("beta"));
  m.def("covis_distance", &covis_distance, "fraction of pixels leaving both neighbours", kNoGil, py::arg("poses"),
        py::arg("disps"), py::arg("intrinsics"), py::arg("ii"));
  m.def("depth_filter", &depth_filter, "multi-view depth consistency counts", kNoGil, py::arg("poses"), py::arg("disps"),
        py::arg("intrinsics"), py::arg("ix"), py::arg("thresh"));
  m.def("iproj", &iproj, "pixels lifted to 3-D points", kNoGil, py::arg("poses"), py::arg("disps"), py::arg("intrinsics"));
  m.def("jdsa_solve", &jdsa_solve, "JDSA reduced system per frame slot and update: [disps, dzcov]", kNoGil,
        py::arg("C"), py::arg("w"), py::arg("disps"), py::arg("prior"), py::arg("dscales"), py::arg("eta"),
        py::arg("kx"), py::arg("slot_src"), py::arg("gx"), py::arg("gy"), py::arg("alpha"), py::arg("ep"),
        py::arg("lm"));
  m.def("bi_inter", &bi_inter, "bilinear interpolation of scale grids: [outputs, Jacobis]", kNoGil, py::arg("scales"),
        py::arg("grids"));
  m.def("projective_transform", &projective_transform,
        "torch-formulation projective transform: [coords, valid, Ji, Jj, Jz]", kNoGil, py::arg("poses"),
        py::arg("disps"), py::arg("intrinsics"), py::arg("ii"), py::arg("jj"), py::arg("jacobian") = false,
        py::arg("return_depth") = false);
  m.def("projective_transform_backward", &projective_transform_backward,
        "adjoint of projective_transform: [grad_poses (tangent), grad_disps]", kNoGil);
  py::class_<GNStep>(m, "GNStep", "one Gauss-Newton step of the torch formulation (hislam2/geom/ba.py) "
                                  "with its hand-written adjoint")
      .def(py::init<torch::Tensor, torch::Tensor, int64_t, int64_t, int64_t, int, bool>(), py::arg("ii"),
           py::arg("jj"), py::arg("num_poses"), py::arg("ht"), py::arg("wd"), py::arg("fixedp"),
           py::arg("motion_only") = false)
      .def("frames", &GNStep::frames)
      .def("forward", &GNStep::forward, kNoGil)
      .def("backward", &GNStep::backward, kNoGil);
  m.attr("PRECISION_SPLIT") = py::int_(static_cast<int>(HSL2_BA_SPLIT));
  m.attr("PRECISION_FP32") = py::int_(static_cast<int>(HSL2_BA_FP32));
}
