// C++ host mirror of the reference's global-placement API over the C ABI.
//
// EPlacer_2D keeps the method names, call order and stage semantics of the
// reference's EPlacer_2D (EPlace/eplace.h:95-200); FirstOrderOptimizer and
// EplaceNesterovOpt keep opt() (Optimization/opt.hpp:18-36,
// nesterov.hpp:59-415).  The optimiser loop itself runs device-resident
// through ep_nesterov_run; positions and gradients are std::vector<VECTOR_3D>
// in the reference's stage order with z = 0.  Errors throw
// std::runtime_error carrying ep_last_error(); nothing calls exit().
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/eplace_amd.h"

namespace eplace_amd {

struct VECTOR_3D {
  float x = 0, y = 0, z = 0;
};

enum PLACEMENT_STAGE { mGP = EP_STAGE_MGP, FILLERONLY = EP_STAGE_FILLERONLY, cGP = EP_STAGE_CGP };

struct PlacerOptions {
  int bins_x = 0, bins_y = 0;  // 0: the reference's bin rule
  bool lse = false;            // -LSE
  bool true_exp = false;       // exp instead of (1+a/1024)^1024
  uint32_t filler_seed = 1;
  int poisson_mode = 0;        // ep_params.poisson_mode (0 FP32 MFMA, 2 split-FP16, ...)
  int device = 0;
};

class EPlacer_2D {
 public:
  EPlacer_2D(const ep_design &design, float targetDensity, const PlacerOptions &o = PlacerOptions()) {
    ep_params p{};
    p.target_density = targetDensity;
    p.bins_x = o.bins_x;
    p.bins_y = o.bins_y;
    p.wirelength_model = o.lse ? 1 : 0;
    p.exp_mode = o.true_exp ? 1 : 0;
    p.filler_seed = o.filler_seed;
    p.filler_xy = nullptr;
    p.filler_count = -1;
    p.poisson_mode = o.poisson_mode;
    p.device = o.device;
    int rc = ep_create(&design, &p, &ctx_);
    if (rc != EP_OK) {
      std::string msg = ctx_ ? ep_last_error(ctx_) : "ep_create failed";
      if (ctx_) ep_destroy(ctx_);
      ctx_ = nullptr;
      throw std::runtime_error("ep_create: " + msg);
    }
    targetDensity_ = targetDensity;
  }
  ~EPlacer_2D() {
    if (ctx_) ep_destroy(ctx_);
  }
  EPlacer_2D(const EPlacer_2D &) = delete;
  EPlacer_2D &operator=(const EPlacer_2D &) = delete;

  // eplace.cpp:39-50 (fillers and bins are built by the constructor)
  void initialization() {
    totalGradientUpdate();
    penaltyFactorInitilization();
  }
  void totalGradientUpdate() { check(ep_total_gradient_update(ctx_)); }
  std::vector<VECTOR_3D> getGradient() { return pairs(&ep_get_gradient); }
  std::vector<VECTOR_3D> getPosition() { return pairs(&ep_get_position); }
  void setPosition(const std::vector<VECTOR_3D> &v) {
    std::vector<float> xy(2 * v.size());
    for (size_t i = 0; i < v.size(); ++i) {
      xy[2 * i] = v[i].x;
      xy[2 * i + 1] = v[i].y;
    }
    check(ep_set_position(ctx_, xy.data(), (int64_t)v.size()));
  }
  void penaltyFactorInitilization() { check(ep_penalty_factor_init(ctx_)); }
  void updatePenaltyFactor() { check(ep_update_penalty_factor(ctx_)); }
  void switch2FillerOnly() {
    check(ep_switch_to_filler_only(ctx_));
    placementStage = FILLERONLY;
  }
  void switch2cGP() {
    check(ep_switch_to_cgp(ctx_, mGPIterationCount));
    placementStage = cGP;
  }
  ep_stats stats() {
    ep_stats s{};
    check(ep_get_stats(ctx_, &s));
    return s;
  }
  double calcHPWL() {
    double v = 0;
    check(ep_hpwl(ctx_, &v));
    return v;
  }
  std::string showInfo() {
    ep_stats s = stats();
    return "Overflow: " + std::to_string(s.overflow) + "\npenalty factor: " + std::to_string(s.lambda) +
           "\nHPWL: " + std::to_string(s.last_hpwl) + "\n";
  }
  ep_info info() {
    ep_info i{};
    check(ep_get_info(ctx_, &i));
    return i;
  }
  ep_ctx *context() { return ctx_; }
  void check(int rc) {
    if (rc != EP_OK) throw std::runtime_error(std::string("eplace_amd: ") + ep_last_error(ctx_));
  }

  int placementStage = mGP;
  int mGPIterationCount = 0;

 private:
  std::vector<VECTOR_3D> pairs(int (*fn)(ep_ctx *, float *, int64_t)) {
    int64_t n = ep_stage_size(ctx_);
    std::vector<float> xy(2 * (size_t)n);
    check(fn(ctx_, xy.data(), n));
    std::vector<VECTOR_3D> v((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      v[i].x = xy[2 * i];
      v[i].y = xy[2 * i + 1];
    }
    return v;
  }
  ep_ctx *ctx_ = nullptr;
  float targetDensity_ = 1.f;
};

// QPPlacer (QPlace/qplace.h:24-44).  quadraticPlacement() runs the
// reference's quadratic placement on the device and moves the movable nodes
// (node_xy: [N][2] centres) to its result, as the reference moves the
// modules of its database.  log receives one row per outer iteration: HPWL,
// X error, Y error, BiCGSTAB iterations (the line qplace.cpp:89 prints).
class QPPlacer {
 public:
  explicit QPPlacer(const ep_design &design, int device = 0) : design_(design), device_(device) {}
  ep_qp_stats quadraticPlacement(std::vector<float> &node_xy, int maxIterations = 20,
                                 std::vector<double> *log = nullptr) {
    if (node_xy.size() != 2 * (size_t)design_.num_nodes) throw std::runtime_error("QPPlacer: node_xy size");
    ep_qp_params p{};
    p.max_iterations = maxIterations;   // -IPiteCount (qplace.cpp:35-42)
    p.solver_iterations = 100;          // qplace.cpp:68
    p.target_error = 0.000001f;         // qplace.cpp:45
    p.solver_tolerance = 1.1920928955078125e-07f;  // Eigen's default for float
    p.start_at_center = 1;              // moveNodesCenterToCenter (qplace.cpp:23)
    p.device = device_;
    std::vector<double> rows(4 * (size_t)(maxIterations + 1));
    ep_qp_stats st{};
    if (ep_qplace(&design_, &p, node_xy.data(), rows.data(), &st) != EP_OK)
      throw std::runtime_error(std::string("ep_qplace: ") + ep_qp_last_error());
    if (log) log->assign(rows.begin(), rows.begin() + 4 * (size_t)st.iterations);
    return st;
  }

 private:
  ep_design design_;
  int device_;
};

// SAMacroLegalizer (Legalization/legalizer.h:304-420) on a placer, as
// ePlace_main.cpp:174-186 runs it: positions and the rand() stream are the
// placer's.
class SAMacroLegalizer {
 public:
  SAMacroLegalizer(EPlacer_2D *placer, const ep_design &design) : placer_(placer), design_(design) {}
  void setTargetDensity(float td) { targetDensity_ = td; }
  ep_mlg_stats legalization() {
    ep_mlg_params p{};
    p.target_density = targetDensity_;
    p.rand_seed = 1;
    ep_mlg_stats st{};
    placer_->check(ep_macro_legalize(placer_->context(), &design_, &p, &st));
    return st;
  }

 private:
  EPlacer_2D *placer_;
  ep_design design_;
  float targetDensity_ = 1.0f;
};

template <typename T>
class FirstOrderOptimizer {
 public:
  virtual ~FirstOrderOptimizer() {}
  virtual void opt() = 0;
};

// Step-length variants of EplaceNesterovOpt::opt_step (nesterov.hpp:152-165):
// vanilla (default), backtracking (-bktrk), Barzilai-Borwein (-bb, which
// the reference checks first), and backtracking led by the learned
// controller (no counterpart in the reference).  The values are
// ep_opt_params.method.
enum StepMethod { STEP_VANILLA = 0, STEP_BKTRK = 1, STEP_LEARNED = 2, STEP_BB = 3 };

// The float32 values of a little-endian, C-order .npy file (the format the
// controller's parameters are saved in).
inline std::vector<float> read_npy_f32(const std::string &path) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  std::unique_ptr<FILE, int (*)(FILE *)> guard(f, fclose);
  unsigned char pre[10];
  if (fread(pre, 1, 8, f) != 8 || memcmp(pre, "\x93NUMPY", 6) != 0) throw std::runtime_error(path + ": not a .npy file");
  size_t hlen = 0;
  if (pre[6] == 1) {
    if (fread(pre + 8, 1, 2, f) != 2) throw std::runtime_error(path + ": truncated header");
    hlen = pre[8] | (size_t)pre[9] << 8;
  } else {
    unsigned char l[4];
    if (fread(l, 1, 4, f) != 4) throw std::runtime_error(path + ": truncated header");
    hlen = l[0] | (size_t)l[1] << 8 | (size_t)l[2] << 16 | (size_t)l[3] << 24;
  }
  std::string header(hlen, ' ');
  if (fread(&header[0], 1, hlen, f) != hlen) throw std::runtime_error(path + ": truncated header");
  if (header.find("'descr': '<f4'") == std::string::npos || header.find("'fortran_order': False") == std::string::npos)
    throw std::runtime_error(path + ": expected little-endian float32 in C order");
  std::vector<float> v;
  float buf[1024];
  size_t n;
  while ((n = fread(buf, sizeof(float), 1024, f)) > 0) v.insert(v.end(), buf, buf + n);
  return v;
}

// The learned step / penalty controller on one device, from saved
// parameters (EP_CTRL_PARAMS floats).
class Controller {
 public:
  // The parameters of a saved controller, checked (no device work).
  static std::vector<float> load(const std::string &npy_path) {
    std::vector<float> p = read_npy_f32(npy_path);
    if (p.size() != (size_t)EP_CTRL_PARAMS)
      throw std::runtime_error(npy_path + ": expected " + std::to_string((int)EP_CTRL_PARAMS) + " parameters, found " +
                               std::to_string(p.size()));
    return p;
  }
  explicit Controller(const std::vector<float> &p, int device = 0) {
    if (p.size() != (size_t)EP_CTRL_PARAMS) throw std::runtime_error("controller: wrong number of parameters");
    if (ep_ctrl_create(device, p.data(), 1, &h_) != EP_OK) {
      std::string msg = h_ ? ep_ctrl_last_error(h_) : "ep_ctrl_create failed";
      if (h_) ep_ctrl_destroy(h_);
      throw std::runtime_error(msg);
    }
  }
  ~Controller() {
    if (h_) ep_ctrl_destroy(h_);
  }
  Controller(const Controller &) = delete;
  Controller &operator=(const Controller &) = delete;
  const ep_ctrl *get() const { return h_; }

 private:
  ep_ctrl *h_ = nullptr;
};

// Nesterov with the chosen step length.
template <typename T>
class EplaceNesterovOpt : public FirstOrderOptimizer<T> {
 public:
  explicit EplaceNesterovOpt(EPlacer_2D *placer, int method = STEP_VANILLA, float targetOverflow = 0.10f)
      : placer_(placer), method_(method), targetOverflow_(targetOverflow) {}
  void opt() override {
    ep_opt_params p{};
    p.method = method_;
    p.target_overflow = targetOverflow_;
    p.max_iterations = 1000;  // MAX_ITERATION (nesterov.hpp:36)
    p.max_backtracks = 30;
    p.sync_every = 16;
    p.log_capacity = 0;
    p.log = nullptr;
    p.phase_seconds = nullptr;
    p.controller = controller_;
    if (keepLog_) {
      log_.assign((size_t)(p.max_iterations + 2) * EP_LOG_COLS, 0.0);
      p.log_capacity = p.max_iterations + 2;
      p.log = log_.data();
    }
    placer_->check(ep_nesterov_run(placer_->context(), &p, &last_));
    if (keepLog_) log_.resize((size_t)std::min(last_.iterations, p.log_capacity) * EP_LOG_COLS);
    if (placer_->placementStage == mGP) placer_->mGPIterationCount = last_.iterations;
  }
  const ep_run_stats &last() const { return last_; }
  // The per-iteration log of the next opt() (EP_LOG_COLS doubles per
  // iteration, columns EP_LOG_*): what the reference prints per iteration
  // (nesterov.hpp:154, 333, 395) plus evaluations, backtracks and the
  // controller's inputs.
  void keepLog(bool on) { keepLog_ = on; }
  // The controller STEP_LEARNED runs with (on the placer's device).
  void setController(const Controller *c) { controller_ = c ? c->get() : nullptr; }
  const std::vector<double> &log() const { return log_; }

 private:
  EPlacer_2D *placer_;
  int method_;
  float targetOverflow_;
  ep_run_stats last_{};
  bool keepLog_ = false;
  std::vector<double> log_;
  const ep_ctrl *controller_ = nullptr;
};

}  // namespace eplace_amd
