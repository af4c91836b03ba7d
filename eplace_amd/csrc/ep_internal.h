// Internal declarations shared by the engine's HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/eplace_amd.h"

namespace ep {

// glibc rand() (TYPE_3 additive feedback after srand(seed)).  The
// reference never seeds it: fillers (placedb.cpp:353-379) and the macro
// legaliser's annealing (legalizer.cpp:1414-1517) draw from one stream.
struct Rand {
  std::vector<uint32_t> r;
  int64_t draws = 0;
  explicit Rand(uint32_t seed) {
    if (seed == 0) seed = 1;
    r.resize(344);
    int32_t w = (int32_t)seed;
    r[0] = (uint32_t)w;
    for (int i = 1; i < 31; ++i) {
      int32_t hi = w / 127773, lo = w % 127773;
      w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      r[i] = (uint32_t)w;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 34; i < 344; ++i) r[i] = r[i - 31] + r[i - 3];
  }
  int32_t next() {
    size_t i = r.size();
    uint32_t v = r[i - 31] + r[i - 3];
    r.push_back(v);
    if (r.size() > (1u << 20)) r.erase(r.begin(), r.end() - 64);
    ++draws;
    return (int32_t)(v >> 1);
  }
};

enum ObjKind : uint8_t { KIND_CELL = 0, KIND_MACRO = 1, KIND_FILLER = 2 };

constexpr double kEps = 1.0e-15;  // global.h:39
constexpr int kBlock = 256;
constexpr int kFinish = 1024;  // single-block kernels that sum per-block partials
constexpr int kNetLaneMax = 32;  // nets above this get a wave each; the rest a power-of-two lane group each
// Objects that may cover more bins get a block each; smaller ones go through
// the tile-sorted scatter and gather, one lane per object (a lane's serial
// loop over up to 64 bins costs less than a block per object, and it sums
// in the reference's order).
constexpr int kBigObjectBins = 64;
constexpr int kXcds = 8;  // MI355X: workgroups go to the 8 XCDs round-robin
// phase events recorded per evaluation: 9 between the phases, a start and an
// end event around each of the 4 transform GEMM launches, and one recorded
// right after the last GEMM's end event: the two bracket nothing but the
// processing of one event on the stream, which every event pair around a
// kernel also holds and which is subtracted from the GEMM times
constexpr int kEvPerEval = 18;
// transform GEMM launches per Poisson solve
constexpr int kGemmStamped = 4;

// Wall-clock stamps of the transform GEMM launches.  Every block of launch L
// of an evaluation stores its start and its end (plain stores, one pair per
// block) in the stamp set of the evaluation count's parity,
// t[2 ((parity kGemmStamped + L) stride + block)].  The WL reduction of the
// next evaluation, on the second stream and off the critical path, reduces
// the other set to (first start, last end) per launch while this
// evaluation's GEMMs write theirs, and marks a launch folded by clearing its
// block 0 end, which the next launch into that set rewrites.
struct GemmStamps {
  unsigned long long *t;     // [2][kGemmStamped][stride][2]
  const int *evals;          // the evaluation count (DevScalars::evals)
  int stride;                // blocks reserved per launch
  int nblk[kGemmStamped];    // blocks of each launch of an evaluation's solve
};
constexpr int kEvGemm = 9;
constexpr int kEvMarker = 17;
// timed phases of an evaluation: scatter, density finalise, wiring kernels,
// node sums of the pin gradients, Poisson solve, field gather, combine
constexpr int kPhases = 7;
// ... and the transform GEMM launches alone (inside the Poisson phase)
constexpr int kPhasesTimed = kPhases + 1;

// Renumbers blocks so that the blocks of each XCD cover one contiguous
// 1/8 of the grid (block b runs on XCD b % 8).  Work ranges that are
// contiguous in memory then stay inside one XCD's L2.  Grids that are not a
// multiple of 8 keep their numbering.
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, n = gridDim.x;
  if (n % kXcds) return b;
  return (b % kXcds) * (n / kXcds) + b / kXcds;
}

// The core region, for the position clamp.
struct Core {
  float llx, lly, urx, ury;
};

// PlaceDB::setModuleCenter_2D clamp (placedb.cpp:253-288): compares in double,
// +/- EPS, result stored as float.
__device__ __forceinline__ float2 clamp_center(float x, float y, float w, float h, const Core &k) {
  double hw = 0.5 * (double)w, hh = 0.5 * (double)h;
  if ((double)x - hw < (double)k.llx) x = (float)((double)k.llx + hw + kEps);
  if ((double)x + hw > (double)k.urx) x = (float)((double)k.urx - hw - kEps);
  if ((double)y - hh < (double)k.lly) y = (float)((double)k.lly + hh + kEps);
  if ((double)y + hh > (double)k.ury) y = (float)((double)k.ury - hh - kEps);
  return make_float2(x, y);
}

// One Nesterov move (nesterov.hpp:400-411 + setPosition):
// x_{k+1} = y_k + step g_k;  y_{k+1} = x_{k+1} + coef (x_{k+1} - x_k);
// positions := clamp(y_{k+1}).  ref == nullptr reads y_k from the
// positions; save_ref/save_grad receive y_k and g_k for the next Lipschitz
// estimate.
struct MoveArgs {
  const uint8_t *in_stage;
  const float2 *ref, *grad, *main_in;
  float2 *main_out, *ref_unclamped_out, *save_ref, *save_grad;
  Core k;
};
__device__ __forceinline__ float2 move_object(const MoveArgs &a, int o, float2 *pos, float w, float h, float step,
                                              float coef) {
  const float2 r = a.ref ? a.ref[o] : pos[o];
  const float2 g = a.grad[o];
  const float2 m = a.main_in[o];
  if (a.save_ref) a.save_ref[o] = r;
  if (a.save_grad) a.save_grad[o] = g;
  const float npx = r.x + g.x * step, npy = r.y + g.y * step;
  const float nrx = npx + (npx - m.x) * coef, nry = npy + (npy - m.y) * coef;
  a.main_out[o] = make_float2(npx, npy);
  if (a.ref_unclamped_out) a.ref_unclamped_out[o] = make_float2(nrx, nry);
  const float2 p = clamp_center(nrx, nry, w, h, a.k);
  pos[o] = p;
  return p;
}

// Blocks of kBlock threads covering n items, rounded up to a multiple of
// kXcds so that xcd_block() applies (the extra blocks find no work).
inline int object_blocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return (int)((b + kXcds - 1) / kXcds * kXcds);
}

// Bin grid, passed by value to the density kernels.  Edges are the
// reference's float bin coordinates (eplace.cpp:292-314).
struct Grid {
  int nx, ny;
  float bsx, bsy, core_llx, core_lly;
  float bin_area, inv_bin_area, td;
};

// Device-resident scalars of the evaluator and optimiser.
struct DevScalars {
  double overflow;     // tau of the last scatter
  double inv_gamma_x, inv_gamma_y;
  double lambda;
  double hpwl;         // HPWL at the positions of the last wirelength pass
  double last_hpwl;
  double wl_value;
  double step, a_k, a_next, coef, new_step;
  double lip_num, lip_den;
  double init_lambda_num, init_lambda_den;
  double rho_mean;     // mean charge density; removed before the transform (DC is discarded)
  double rho_max;      // largest charge density (sets the operand scale of the FP16 transforms)
  int iter, done, accepted, stage;
  int evals, backtracks, max_iter, bad_numeric;
  float target_overflow, pad0;
  // controller features and actions (backtracking / learned runs)
  double gwl_norm2, gd_norm2;         // sum |grad WL|^2, sum |grad D|^2 of the last combine
  double lambda0, hpwl0, step0;       // at the start of the run / first accepted step
  double step_hist[3];                // accepted 1/L' of the last three iterations
  double tau_prev, hpwl_prev;         // at y_{k-1}
  double step_prop, step_used, lambda_before, lam_floor;
  // running mean of log(reference multiplier), weight 1/16 per update: how
  // fast the reference rule itself would grow lambda (a controller feature)
  double ref_mult_ema;
  // updatePenaltyFactor staged by a backtracking trial's evaluation: lambda
  // and the reference multiplier it would take if the trial is the
  // iteration's last; the stop check applies them
  double lambda_next, ref_mult_next;
  float feat[EP_CTRL_FEATURES];
  int backtracks_mark, trial, fallbacks, ctrl_active;
  int bb_short;                       // iterations that took the short Barzilai-Borwein step
  int pad1;
  // transform GEMM launches folded by the density finisher: summed wall-clock
  // ticks from each launch's first block start to its last block's end
  unsigned long long gemm_ticks;
  int gemm_launches, pad2;
};

// Slots of the partial-sum buffer (doubles).
constexpr int kPartialTau = 0;          // up to kMaxPartials
constexpr int kMaxPartials = 32768;
constexpr int kPartialWL = kMaxPartials;        // hpwl, wl pairs
constexpr int kPartialLip = 3 * kMaxPartials;   // num, den pairs
constexpr int kPartialLambda = 5 * kMaxPartials;  // |gwl|, |gd| pairs
constexpr int kPartialRho = 7 * kMaxPartials;      // sum of rho per block
constexpr int kPartialNorm = 8 * kMaxPartials;     // |gwl|^2, |gd|^2 pairs
constexpr int kPartialRhoMax = 10 * kMaxPartials;  // max of rho per block
constexpr int kPartialSY = 11 * kMaxPartials;      // (y - y_prev).(g - g_prev), stride 2
constexpr int kPartialLipNext = 13 * kMaxPartials; // num, den pairs at the staged lambda
constexpr int kPartialTotal = 15 * kMaxPartials;

struct Poisson;  // poisson.hip

// Counting sort of the objects by (kind, 32x32-bin tile of the footprint's
// lower-left bin), shared by the scatter and the gather of one evaluation
// (density.hip).  key = kind * ntiles + tile; key 3 * ntiles marks objects
// handled by the block-per-object kernels.  Order inside a key is not
// fixed; every consumer's result is independent of it.
struct TileSort {
  int tyc = 0, ntiles = 0, nkeys = 0, nblocks = 0, per = 0, grid = 0;
  uint32_t *okey = nullptr;    // [O] key of each object
  uint32_t *bcount = nullptr;  // [nblocks][nkeys] offset of each block's run inside the key
  uint32_t *ktotal = nullptr;  // [nkeys] running key totals, zero between sorts
  uint32_t *sobj = nullptr;    // [O] objects in key order
  float4 *svals = nullptr;     // [O] their (x, y, w, h)
  int4 *chunks = nullptr;      // (key, begin, end) runs of <= kTileChunk objects of one key
  int *nchunks = nullptr;
};

struct Ctx {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream: the wiring kernels of an evaluation run on it, beside the
  // Poisson solve and the field gather (fork after the density finalise,
  // join before the combine)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // Pinned coherent host copy of the scalars and its sequence number: a
  // read queues a kernel that stores the scalars there and then the number,
  // and the host polls the number (no copy engine, no stream wait).
  struct HostScalars {
    DevScalars s;
    unsigned int seq;
  };
  HostScalars *h_sc = nullptr;
  unsigned int pub_seq = 0;
  // per-evaluation phase events: 0 start, 1 scatter end, 2 finalise end,
  // 3 Poisson end, 4 field gather end, 5 combine end (stream), 6 wiring
  // start, 8 wiring kernels end, 7 node sums end (stream2)
  hipEvent_t ev[kEvPerEval] = {};
  double last_timings[kPhasesTimed] = {};
  // event pool for per-phase timing inside the optimiser loop (kEvPerEval per evaluation)
  std::vector<hipEvent_t> prof_events;
  // Captured evaluations (HIP graphs), one per distinct set of launch
  // arguments; an evaluation without phase events replays its graph with
  // one launch instead of ~20 kernel launches and event operations.
  struct EvalGraph {
    bool lambda_update, lipschitz;
    const float2 *ref_now, *ref_prev, *grad_prev;
    int next_lambda;
    float2 *pos_out, *gtot_out;
    bool counted;  // the keys were counted by the move before it
    hipGraphExec_t exec;
  };
  std::vector<EvalGraph> eval_graphs;
  bool graphs_ok = true;  // cleared if capture or instantiation fails
  size_t prof_used = 0;
  unsigned long long prof_hold_ticks = 0;
  bool profiling = false;

  // sizes
  int N = 0, F = 0, O = 0, NT = 0, E = 0;
  int64_t P = 0;
  int nx = 0, ny = 0;
  int shift = 0;
  int stage = EP_STAGE_MGP;
  bool eval_valid = false;  // gd/gwl/tau describe the current positions
  int wl_model = 0, exp_mode = 0, poisson_mode = 0;
  float td = 1.f;

  // host copies needed for stage handling and filler re-randomisation
  std::vector<float> h_w, h_h;
  std::vector<uint8_t> h_kind;
  std::vector<int32_t> h_stage_idx;  // objects of the current stage, in stage order
  std::vector<int32_t> h_grad_lane;  // pin_grad entry of each design pin (-1: terminal pin)
  float core_llx = 0, core_lly = 0, core_urx = 0, core_ury = 0, total_row_area = 0;
  float bsx = 0, bsy = 0;
  float std_area = 0, macro_area = 0, filler_w = 0, filler_h = 0;
  struct Rand *rng = nullptr;
  double lambda_host_init = 0;

  // device: centres of the objects [O] followed by the terminals [NT]
  float2 *pos = nullptr;
  // device: objects [O]
  float *w = nullptr, *h = nullptr, *area = nullptr;
  uint8_t *kind = nullptr;
  uint8_t *in_stage = nullptr;
  int32_t *npins = nullptr;
  int32_t *big = nullptr;   // objects handled by the block-per-object kernels
  int num_big = 0;
  uint8_t *is_big = nullptr;
  // nets / pins, nets ordered by their lowest movable node
  int32_t *net_start = nullptr;  // [E+1]
  int32_t *pin_obj = nullptr;    // index into pos: objects, then O + terminal
  float *pin_dx = nullptr, *pin_dy = nullptr;
  int32_t *large_nets = nullptr;
  int num_large = 0;
  // nets of 2..kNetLaneMax pins by degree class c (D = 2 << c lanes per net):
  // first stored net, net count, first 64-lane group (wl_gstart[5] = all)
  int wl_first[5] = {0, 0, 0, 0, 0}, wl_count[5] = {0, 0, 0, 0, 0}, wl_gstart[6] = {0, 0, 0, 0, 0, 0};
  int wl_groups = 0;
  // their pins in lane order: object (-1 = padding), offset
  int32_t *lane_obj = nullptr;
  float2 *lane_off = nullptr;
  int32_t *pin_net = nullptr;  // net of each pin
  int32_t *node_pin_start = nullptr;  // [N+1] runs of node_pin_lane per node
  int32_t *node_pin_lane = nullptr;   // [node pins] pin_grad entry of each node pin, modulePins order
  float2 *pin_grad = nullptr;         // WA gradient of every pin, in the wiring kernels' lane order
  int64_t grad_lanes = 0;             // written entries of pin_grad: lane groups, then the pins of large nets;
                                      // entry grad_lanes stays zero (pins of one-pin nets)
  int32_t large_lane_off = 0;         // pin_grad entry of large-net pin p = p + large_lane_off
  // bins
  unsigned long long *node_q = nullptr, *filler_q = nullptr;
  float *term_map = nullptr, *base_map = nullptr;
  float *rho = nullptr, *phi = nullptr, *ex = nullptr, *ey = nullptr;
  // gradients [O]
  float2 *gd = nullptr, *gwl = nullptr, *gtot = nullptr;
  // optimiser buffers [O] (float2)
  float2 *cur_main = nullptr, *cur_ref = nullptr, *last_ref = nullptr, *last_grad = nullptr;
  float2 *new_main = nullptr, *new_ref = nullptr, *cur_grad = nullptr;
  // scalars
  DevScalars *sc = nullptr;
  double *partials = nullptr;
  unsigned int *tickets = nullptr;
  GemmStamps gemm_stamps = {};  // set up with the transforms (poisson.hip)
  int wall_khz = 100000;        // wall clock rate
  Poisson *poisson = nullptr;
  TileSort tiles;

  Grid grid() const {
    Grid g;
    g.nx = nx;
    g.ny = ny;
    g.bsx = bsx;
    g.bsy = bsy;
    g.core_llx = core_llx;
    g.core_lly = core_lly;
    g.bin_area = bsx * bsy;
    g.inv_bin_area = (float)(1.0 / (double)(bsx * bsy));
    g.td = td;
    return g;
  }
};

// ---- launchers (density.hip) ----
// counted: the keys were counted and the maps cleared by launch_move_hist
void launch_scatter(Ctx &c, bool counted = false);
// The Nesterov move of the stage's objects fused with the scatter's key count
// (the evaluation that follows calls launch_scatter(c, true)).
void launch_move_hist(Ctx &c, const MoveArgs &a);
// Zeroes the sort's key totals that a move counted for no evaluation.
void launch_clear_key_totals(Ctx &c);
void launch_finalize_density(Ctx &c);
// pos_out / gtot_out (optional): copies of the positions and of the total
// gradient written by the same pass
void launch_gather_combine(Ctx &c, bool recompute_gather, bool lipschitz, const float2 *ref_now,
                           const float2 *ref_prev, const float2 *grad_prev, float2 *pos_out = nullptr,
                           float2 *gtot_out = nullptr);
void launch_field_gather(Ctx &c);
void launch_combine(Ctx &c, bool recompute, bool lipschitz, const float2 *ref_now, const float2 *ref_prev,
                    const float2 *grad_prev, float2 *pos_out = nullptr, float2 *gtot_out = nullptr,
                    bool at_next_lambda = false);
bool tile_sort_create(Ctx &c, std::string &err);
void tile_sort_destroy(Ctx &c);
// ---- wirelength.hip ----
// st: the stream to launch on (nullptr: the context's stream)
void launch_wirelength(Ctx &c, bool gradient, hipStream_t st = nullptr);
// with_value: the wiring pass computed WA values (its gradient variant).
// lambda_update: 0 none, 1 updatePenaltyFactor, 2 the same with the
// controller's floor; 3 and 4 stage 1 and 2 in lambda_next / ref_mult_next
// without applying them.
void launch_reduce_wirelength(Ctx &c, int lambda_update, hipStream_t st = nullptr, bool with_value = true);
// the node sums of the pin gradients with the same reduction (with_value)
// in block 0 of the launch
void launch_node_wl_reduce(Ctx &c, int lambda_update, hipStream_t st);
// ---- poisson.hip ----
bool poisson_create(Ctx &c, std::string &err);
void poisson_destroy(Ctx &c);
// gemm_ev (optional): 8 events, recorded before and after each GEMM launch
void launch_poisson(Ctx &c, bool with_phi, hipEvent_t *gemm_ev = nullptr);
// ---- nesterov.hip ----
void launch_positions_to_ref(Ctx &c, float2 *dst);
// bb: the Barzilai-Borwein step (nesterov.hpp:231-302) instead of 1/L
void launch_step_scalar(Ctx &c, int bb, int lip_slot = kPartialLip);
void launch_accept(Ctx &c);
void launch_lambda_init(Ctx &c);
void launch_set_positions(Ctx &c, const float2 *xy_stage, const int32_t *stage_idx, int64_t n);
void launch_copy_float2(Ctx &c, float2 *dst, const float2 *src, int64_t n);
void launch_publish_scalars(Ctx &c, unsigned int seq);
void launch_hold(Ctx &c, unsigned long long ticks);
void launch_done_check_ex(Ctx &c, int next_iter, double *log, int log_cap, int set_iter, int apply_next = 0);
// launch_done_check_ex(next_iter, log, log_cap, 1, 1) and then launch_step_scalar(0, lip_slot)
// in one launch
void launch_done_step(Ctx &c, int next_iter, double *log, int log_cap, int lip_slot);
void launch_nesterov_move_ex(Ctx &c, const float2 *ref, const float2 *grad, const float2 *main_in,
                             float2 *main_out, float2 *ref_unclamped_out, float2 *save_ref, float2 *save_grad);
// the same move with the next evaluation's key count (launch_move_hist)
void launch_nesterov_move_hist(Ctx &c, const float2 *ref, const float2 *grad, const float2 *main_in,
                             float2 *main_out, float2 *ref_unclamped_out, float2 *save_ref, float2 *save_grad);
int poisson_ld(const Ctx &c);
// ---- controller.hip ----
struct CtrlDev;
void launch_ctrl_step(Ctx &c, const CtrlDev *ctrl, int active);
const float *ctrl_params(const CtrlDev *ctrl);
const CtrlDev *ctrl_of(const ep_ctrl *x);
int ctrl_device(const CtrlDev *ctrl);

#define EP_HIP(call)                                                           \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      c.err = std::string(#call) + ": " + hipGetErrorString(e_);              \
      return EP_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

// Annealing macro legaliser (macro_legalize.cpp); xy: [N][2] node centres,
// updated in place; draws from rng.
int macro_legalize_sa(const ep_design &d, std::vector<float> &xy, const ep_mlg_params &p, Rand &rng,
                      ep_mlg_stats &st);

}  // namespace ep

struct ep_ctx {
  ep::Ctx c;
};
