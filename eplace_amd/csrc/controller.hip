// Learned step / lambda controller: a 12 -> 64 -> 64 -> 2 MLP (tanh hidden
// layers, FP32) with hand-written inference, backward and Adam kernels.
// No counterpart in the reference; it drives method 2 of the optimiser (the
// reference's backtracking Nesterov, nesterov.hpp:310-369, with the first
// trial step proposed by the network and the penalty multiplier of
// updatePenaltyFactor, eplace.cpp:846-883, held at or above the network's
// floor).
//
// Features of iteration k, computed on the device at y_k (after the
// combine at the current lambda and the plain step proposal 1/L):
//   0 tau_k                      1 tau_k - tau_{k-1}
//   2 log(lambda_k / lambda_0)   3 log(HPWL_k / HPWL_0)
//   4 100 (HPWL_k - HPWL_{k-1}) / HPWL_k
//   5 log(|grad WL| / (lambda_k |grad D|))
//   6 log(s_{k-1} / s_0)         7 log(s_{k-1} / s_{k-2})   8 log(s_{k-2} / s_{k-3})
//   9 k / 1000 (MAX_ITERATION)   10 log(HPWL_k / 350000)
//   11 100 m_k, m the running mean (weight 1/16, from log 1.05) of the
//      log of the reference's own multiplier 1.05^(1 - dHPWL/350000)
// with s_j the accepted 1/L' of iteration j; terms that need missing history
// are 0.  Features 10 and 11 carry the design's scale against the
// reference's absolute dHPWL constant (DELTA_HPWL_REF, eplace.h:11), which
// on large designs keeps the reference's multiplier near 1 for hundreds of
// iterations.
// Outputs: o0 = log(step / proposal), clamped to [log 1/4, 0] (the
// controller only shortens the plain proposal), and o1, clamped to
// [0, log(1.05/0.95)]: the iteration's multiplier is max(reference
// multiplier, 0.95 e^o1), inside the reference's range [0.95, 1.05].
// o0 = o1 = 0 reproduces the reference.
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ep_internal.h"
#include "reduce.h"

namespace ep {

constexpr int F = EP_CTRL_FEATURES, H = EP_CTRL_HIDDEN, NO = EP_CTRL_OUTPUTS;
constexpr int oW1 = 0, ob1 = oW1 + H * F, oW2 = ob1 + H, ob2 = oW2 + H * H, oW3 = ob2 + H, ob3 = oW3 + NO * H;
constexpr int NW = ob3 + NO;  // trained weights
constexpr int oXmu = NW, oXsd = oXmu + F, oYmu = oXsd + F, oYsd = oYmu + NO;
constexpr int NP = oYsd + NO;
static_assert(NW == EP_CTRL_WEIGHTS && NP == EP_CTRL_PARAMS, "controller parameter layout");
constexpr int kSamplesPerBlock = 32;

struct CtrlDev {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  float *P = nullptr, *m = nullptr, *v = nullptr;  // parameters, Adam moments
  int adam_t = 0;
};

const float *ctrl_params(const CtrlDev *c) { return c->P; }

// One sample through the network with the weights staged in LDS: x (F
// normalised inputs) -> h1, h2 -> out (normalised outputs).  64 threads.
__device__ __forceinline__ void mlp_forward(const float *W, const float *x, float *h1, float *h2, float *out) {
  const int t = threadIdx.x;
  float a = W[ob1 + t];
#pragma unroll
  for (int i = 0; i < F; ++i) a += W[oW1 + t * F + i] * x[i];
  h1[t] = tanhf(a);
  __syncthreads();
  float b = W[ob2 + t];
  for (int j = 0; j < H; ++j) b += W[oW2 + t * H + j] * h1[j];
  h2[t] = tanhf(b);
  __syncthreads();
  if (t < NO) {
    float o = W[ob3 + t];
    for (int j = 0; j < H; ++j) o += W[oW3 + t * H + j] * h2[j];
    out[t] = o;
  }
  __syncthreads();
}

__device__ __forceinline__ void stage_weights(const float *__restrict__ P, float *W) {
  for (int i = threadIdx.x; i < NP; i += blockDim.x) W[i] = P[i];
  __syncthreads();
}

__device__ __forceinline__ float flog(double v) { return isfinite(v) ? (float)v : 0.f; }

// The rows of the network that lane t needs (row t of W1 and W2, row t of W3
// for t < NO, and their biases), loaded into registers at the start of the
// step kernel so that their latency overlaps the feature computation.
struct LaneWeights {
  float4 w1[F / 4], w2[H / 4], w3[H / 4];
  float b1, b2, b3;
};

__device__ __forceinline__ void load_lane_weights(const float *__restrict__ P, LaneWeights &r) {
  static_assert(F % 4 == 0 && H % 4 == 0 && oW2 % 4 == 0 && oW3 % 4 == 0, "float4 rows");
  const int t = threadIdx.x, t3 = t < NO ? t : 0;
  const float4 *w1 = reinterpret_cast<const float4 *>(P + oW1 + t * F);
  const float4 *w2 = reinterpret_cast<const float4 *>(P + oW2 + t * H);
  const float4 *w3 = reinterpret_cast<const float4 *>(P + oW3 + t3 * H);
#pragma unroll
  for (int q = 0; q < F / 4; ++q) r.w1[q] = w1[q];
#pragma unroll
  for (int q = 0; q < H / 4; ++q) r.w2[q] = w2[q];
#pragma unroll
  for (int q = 0; q < H / 4; ++q) r.w3[q] = w3[q];
  r.b1 = P[ob1 + t];
  r.b2 = P[ob2 + t];
  r.b3 = P[ob3 + t3];
}

// a += w . v over 4*n terms, in index order (the order of mlp_forward).
template <int n>
__device__ __forceinline__ float dot_rows(float a, const float4 (&w)[n], const float *v) {
#pragma unroll
  for (int q = 0; q < n; ++q) {
    a += w[q].x * v[4 * q];
    a += w[q].y * v[4 * q + 1];
    a += w[q].z * v[4 * q + 2];
    a += w[q].w * v[4 * q + 3];
  }
  return a;
}

// mlp_forward with the weights in registers: the same sums in the same order.
__device__ __forceinline__ void mlp_forward_regs(const LaneWeights &r, const float *x, float *h1, float *h2,
                                                 float *out) {
  const int t = threadIdx.x;
  h1[t] = tanhf(dot_rows(r.b1, r.w1, x));
  __syncthreads();
  h2[t] = tanhf(dot_rows(r.b2, r.w2, h1));
  __syncthreads();
  if (t < NO) out[t] = dot_rows(r.b3, r.w3, h2);
  __syncthreads();
}

// Features of iteration k and, when active, the controller's step and lambda
// multiplier.  One block of H threads; lane i < F computes feature i (the
// logarithms of the ratios in parallel).
__global__ void __launch_bounds__(H) ctrl_step_kernel(DevScalars *sc, const float *__restrict__ P, int active) {
  LaneWeights r;
  if (active) load_lane_weights(P, r);
  if (sc->done) return;
  __shared__ float x[F], h1[H], h2[H], out[NO];
  if (threadIdx.x < F) {
    const int i = threadIdx.x, k = sc->iter;
    const double tau = sc->overflow, lam = sc->lambda, hp = sc->hpwl;
    double num = 1.0, den = 1.0;  // feature = log(num / den) when `ratio`
    bool ratio = true, have = true;
    float direct = 0.f;
    switch (i) {
      case 0: ratio = false; direct = (float)tau; break;
      case 1: ratio = false; direct = k > 0 ? (float)(tau - sc->tau_prev) : 0.f; break;
      case 2: num = lam; den = sc->lambda0; break;
      case 3: num = hp; den = sc->hpwl0; break;
      case 4: ratio = false; direct = k > 0 ? flog(100.0 * (hp - sc->hpwl_prev) / hp) : 0.f; break;
      case 5: num = sqrt(sc->gwl_norm2); den = lam * sqrt(sc->gd_norm2); break;
      case 6: have = k >= 1; num = sc->step_hist[0]; den = sc->step0; break;
      case 7: have = k >= 2; num = sc->step_hist[0]; den = sc->step_hist[1]; break;
      case 8: have = k >= 3; num = sc->step_hist[1]; den = sc->step_hist[2]; break;
      case 9: ratio = false; direct = (float)k / 1000.f; break;
      case 10: num = hp; den = 350000.0; break;  // HPWL in units of DELTA_HPWL_REF (eplace.h:11)
      default: ratio = false; direct = (float)(100.0 * sc->ref_mult_ema); break;
    }
    const float f = ratio ? (have ? flog(log(num / den)) : 0.f) : direct;
    sc->feat[i] = f;
    x[i] = active ? (f - P[oXmu + i]) / P[oXsd + i] : f;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double lam = sc->lambda, hp = sc->hpwl, prop = sc->step;
    sc->step_prop = prop;
    sc->lambda_before = lam;
    sc->tau_prev = sc->overflow;
    sc->hpwl_prev = hp;
    sc->backtracks_mark = sc->backtracks;
    sc->trial = 0;
    sc->ctrl_active = active;
    sc->lam_floor = 0.95;
    sc->step_used = prop;
  }
  if (!active) return;
  mlp_forward_regs(r, x, h1, h2, out);
  if (threadIdx.x == 0) {
    const float o0 = out[0] * P[oYsd] + P[oYmu], o1 = out[1] * P[oYsd + 1] + P[oYmu + 1];
    const double s = fmin(fmax((double)o0, -1.3862943611198906), 0.0);
    // floor under the reference's multiplier: 0.95 e^l, l in [0, log(1.05/0.95)]
    const double l = fmin(fmax((double)o1, 0.0), 0.10008345855698253);
    if (isfinite(s) && isfinite(l)) {
      sc->step = sc->step_prop * exp(s);
      sc->lam_floor = 0.95 * exp(l);
    }
    sc->step_used = sc->step;
  }
}

void launch_ctrl_step(Ctx &c, const CtrlDev *ctrl, int active) {
  hipLaunchKernelGGL(ctrl_step_kernel, dim3(1), dim3(H), 0, c.stream, c.sc, ctrl ? ctrl->P : nullptr,
                     ctrl ? active : 0);
}

// ---------------------------------------------------------------------------
// Training: mean over samples of sum over outputs of (out - (y - mu)/sd)^2.
// Each block takes kSamplesPerBlock samples, thread t owns hidden unit t
// (row t of W1 and W2, column t of W3) and accumulates its gradient in
// registers; per-block partial gradients are summed in a fixed order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(H) mlp_grad_kernel(const float *__restrict__ P, const float *__restrict__ X,
                                                    const float *__restrict__ Y, int n, float *gpart,
                                                    float *lpart) {
  __shared__ float W[NP];
  __shared__ float x[F], h1[H], h2[H], out[NO], dout[NO], dz2[H];
  stage_weights(P, W);
  const int t = threadIdx.x;
  float gw1[F], gw2[H], gb1 = 0.f, gb2 = 0.f, gw3[NO] = {0.f, 0.f}, gb3 = 0.f, loss = 0.f;
#pragma unroll
  for (int i = 0; i < F; ++i) gw1[i] = 0.f;
#pragma unroll
  for (int j = 0; j < H; ++j) gw2[j] = 0.f;
  const int s0 = blockIdx.x * kSamplesPerBlock, s1 = min(n, s0 + kSamplesPerBlock);
  const float inv_n = 1.f / (float)n;
  for (int s = s0; s < s1; ++s) {
    if (t < F) x[t] = (X[(size_t)s * F + t] - W[oXmu + t]) / W[oXsd + t];
    __syncthreads();
    mlp_forward(W, x, h1, h2, out);
    if (t < NO) {
      const float d = out[t] - (Y[(size_t)s * NO + t] - W[oYmu + t]) / W[oYsd + t];
      dout[t] = 2.f * d * inv_n;
      loss += d * d;
    }
    __syncthreads();
#pragma unroll
    for (int o = 0; o < NO; ++o) gw3[o] += dout[o] * h2[t];
    if (t < NO) gb3 += dout[t];
    float dh2 = 0.f;
#pragma unroll
    for (int o = 0; o < NO; ++o) dh2 += dout[o] * W[oW3 + o * H + t];
    const float z2 = dh2 * (1.f - h2[t] * h2[t]);
    dz2[t] = z2;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < H; ++j) gw2[j] += z2 * h1[j];
    gb2 += z2;
    float dh1 = 0.f;
    for (int j = 0; j < H; ++j) dh1 += dz2[j] * W[oW2 + j * H + t];
    const float z1 = dh1 * (1.f - h1[t] * h1[t]);
#pragma unroll
    for (int i = 0; i < F; ++i) gw1[i] += z1 * x[i];
    gb1 += z1;
    __syncthreads();
  }
  float *g = gpart + (size_t)blockIdx.x * NW;
#pragma unroll
  for (int i = 0; i < F; ++i) g[oW1 + t * F + i] = gw1[i];
  g[ob1 + t] = gb1;
#pragma unroll
  for (int j = 0; j < H; ++j) g[oW2 + t * H + j] = gw2[j];
  g[ob2 + t] = gb2;
#pragma unroll
  for (int o = 0; o < NO; ++o) g[oW3 + o * H + t] = gw3[o];
  if (t < NO) g[ob3 + t] = gb3;
  // block loss: lanes 0..NO-1 hold their output's sum
  float l = loss;
  for (int o = 1; o < NO; ++o) l += __shfl(loss, o, 64);
  if (t == 0) lpart[blockIdx.x] = l;
}

// g = sum over blocks of the partial gradients (fixed order); update = 1
// applies one Adam step to the weights, else g is stored in gout.
__global__ void adam_kernel(float *P, float *m, float *v, const float *__restrict__ gpart, int nblocks, float lr,
                            float bc1, float bc2, int update, float *gout) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NW) return;
  float g = 0.f;
  for (int b = 0; b < nblocks; ++b) g += gpart[(size_t)b * NW + p];
  if (!update) {
    gout[p] = g;
    return;
  }
  const float mm = 0.9f * m[p] + 0.1f * g;
  const float vv = 0.999f * v[p] + 0.001f * g * g;
  m[p] = mm;
  v[p] = vv;
  P[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
}

__global__ void loss_sum_kernel(const float *__restrict__ lpart, int nblocks, int n, float *dst) {
  double s = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += lpart[b];
  s = block_sum(s);
  if (threadIdx.x == 0) *dst = (float)(s / n);
}

__global__ void __launch_bounds__(H) mlp_predict_kernel(const float *__restrict__ P, const float *__restrict__ X,
                                                       int n, float *Y) {
  __shared__ float W[NP];
  __shared__ float x[F], h1[H], h2[H], out[NO];
  stage_weights(P, W);
  const int t = threadIdx.x;
  const int s0 = blockIdx.x * kSamplesPerBlock, s1 = min(n, s0 + kSamplesPerBlock);
  for (int s = s0; s < s1; ++s) {
    if (t < F) x[t] = (X[(size_t)s * F + t] - W[oXmu + t]) / W[oXsd + t];
    __syncthreads();
    mlp_forward(W, x, h1, h2, out);
    if (t < NO) Y[(size_t)s * NO + t] = out[t] * W[oYsd + t] + W[oYmu + t];
    __syncthreads();
  }
}

}  // namespace ep

struct ep_ctrl {
  ep::CtrlDev d;
};

namespace ep {
const CtrlDev *ctrl_of(const ep_ctrl *x) { return x ? &x->d : nullptr; }
int ctrl_device(const CtrlDev *c) { return c->device; }
}  // namespace ep

namespace {

using ep::CtrlDev;

#define CT_HIP(call)                                                           \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      c.err = std::string(#call) + ": " + hipGetErrorString(e_);              \
      return EP_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

void init_params(std::vector<float> &p, uint32_t seed) {
  std::mt19937 rng(seed ? seed : 1);
  auto fill = [&](int off, int rows, int cols) {
    std::normal_distribution<float> nd(0.f, std::sqrt(1.f / (float)cols));
    for (int i = 0; i < rows * cols; ++i) p[off + i] = nd(rng);
  };
  std::fill(p.begin(), p.end(), 0.f);
  fill(ep::oW1, ep::H, ep::F);
  fill(ep::oW2, ep::H, ep::H);
  fill(ep::oW3, ep::NO, ep::H);
  for (int i = 0; i < ep::F; ++i) p[ep::oXsd + i] = 1.f;
  for (int o = 0; o < ep::NO; ++o) p[ep::oYsd + o] = 1.f;
}

int check_params(CtrlDev &c, const float *p) {
  for (int i = 0; i < ep::NP; ++i)
    if (!std::isfinite(p[i])) return c.err = "non-finite controller parameter", EP_ERR_ARG;
  for (int i = 0; i < ep::F; ++i)
    if (!(p[ep::oXsd + i] > 0.f)) return c.err = "input std must be > 0", EP_ERR_ARG;
  for (int o = 0; o < ep::NO; ++o)
    if (!(p[ep::oYsd + o] > 0.f)) return c.err = "output std must be > 0", EP_ERR_ARG;
  return EP_OK;
}

int upload_params(CtrlDev &c, const float *p) {
  int rc = check_params(c, p);
  if (rc) return rc;
  CT_HIP(hipMemcpy(c.P, p, ep::NP * sizeof(float), hipMemcpyHostToDevice));
  CT_HIP(hipMemset(c.m, 0, ep::NW * sizeof(float)));
  CT_HIP(hipMemset(c.v, 0, ep::NW * sizeof(float)));
  c.adam_t = 0;
  return EP_OK;
}

struct Batch {
  float *X = nullptr, *Y = nullptr, *gpart = nullptr, *lpart = nullptr, *scalar = nullptr;
  int nblocks = 0;
  ~Batch() {
    void *p[] = {X, Y, gpart, lpart, scalar};
    for (void *q : p)
      if (q) (void)hipFree(q);
  }
};

int make_batch(CtrlDev &c, Batch &b, const float *X, const float *Y, int64_t n) {
  if (n <= 0 || n > (1 << 24)) return c.err = "sample count out of range [1, 2^24]", EP_ERR_ARG;
  b.nblocks = (int)((n + ep::kSamplesPerBlock - 1) / ep::kSamplesPerBlock);
  CT_HIP(hipMalloc(&b.X, (size_t)n * ep::F * sizeof(float)));
  CT_HIP(hipMemcpy(b.X, X, (size_t)n * ep::F * sizeof(float), hipMemcpyHostToDevice));
  if (Y) {
    CT_HIP(hipMalloc(&b.Y, (size_t)n * ep::NO * sizeof(float)));
    CT_HIP(hipMemcpy(b.Y, Y, (size_t)n * ep::NO * sizeof(float), hipMemcpyHostToDevice));
    CT_HIP(hipMalloc(&b.gpart, (size_t)b.nblocks * ep::NW * sizeof(float)));
    CT_HIP(hipMalloc(&b.lpart, (size_t)b.nblocks * sizeof(float)));
  }
  CT_HIP(hipMalloc(&b.scalar, ep::NW * sizeof(float)));
  return EP_OK;
}

}  // namespace

extern "C" {

int ep_ctrl_create(int32_t device, const float *params, uint32_t seed, ep_ctrl **out) {
  if (!out) return EP_ERR_ARG;
  *out = nullptr;
  ep_ctrl *x = new ep_ctrl();
  *out = x;
  CtrlDev &c = x->d;
  c.device = device;
  CT_HIP(hipSetDevice(device));
  CT_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  CT_HIP(hipMalloc(&c.P, ep::NP * sizeof(float)));
  CT_HIP(hipMalloc(&c.m, ep::NW * sizeof(float)));
  CT_HIP(hipMalloc(&c.v, ep::NW * sizeof(float)));
  std::vector<float> p(ep::NP);
  if (params) std::memcpy(p.data(), params, ep::NP * sizeof(float));
  else init_params(p, seed);
  return upload_params(c, p.data());
}

int ep_ctrl_destroy(ep_ctrl *x) {
  if (!x) return EP_ERR_ARG;
  CtrlDev &c = x->d;
  (void)hipSetDevice(c.device);
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  void *p[] = {c.P, c.m, c.v};
  for (void *q : p)
    if (q) (void)hipFree(q);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete x;
  return EP_OK;
}

const char *ep_ctrl_last_error(const ep_ctrl *x) { return x ? x->d.err.c_str() : "null controller"; }

int ep_ctrl_get_params(ep_ctrl *x, float *params) {
  if (!x || !params) return EP_ERR_ARG;
  CtrlDev &c = x->d;
  CT_HIP(hipSetDevice(c.device));
  CT_HIP(hipStreamSynchronize(c.stream));
  CT_HIP(hipMemcpy(params, c.P, ep::NP * sizeof(float), hipMemcpyDeviceToHost));
  return EP_OK;
}

int ep_ctrl_set_params(ep_ctrl *x, const float *params) {
  if (!x || !params) return EP_ERR_ARG;
  CtrlDev &c = x->d;
  CT_HIP(hipSetDevice(c.device));
  CT_HIP(hipStreamSynchronize(c.stream));
  return upload_params(c, params);
}

int ep_ctrl_train(ep_ctrl *x, const float *X, const float *Y, int64_t n, int32_t steps, float lr,
                  float *loss_out) {
  if (!x || !X || !Y || steps < 0 || !(lr > 0.f)) return EP_ERR_ARG;
  CtrlDev &c = x->d;
  CT_HIP(hipSetDevice(c.device));
  Batch b;
  int rc = make_batch(c, b, X, Y, n);
  if (rc) return rc;
  float *dloss = nullptr;
  if (loss_out && steps > 0) CT_HIP(hipMalloc(&dloss, (size_t)steps * sizeof(float)));
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(ep::mlp_grad_kernel, dim3(b.nblocks), dim3(ep::H), 0, c.stream, c.P, b.X, b.Y, (int)n,
                       b.gpart, b.lpart);
    if (dloss)
      hipLaunchKernelGGL(ep::loss_sum_kernel, dim3(1), dim3(256), 0, c.stream, b.lpart, b.nblocks, (int)n,
                         dloss + s);
    ++c.adam_t;
    const float bc1 = 1.f - std::pow(0.9f, (float)c.adam_t), bc2 = 1.f - std::pow(0.999f, (float)c.adam_t);
    hipLaunchKernelGGL(ep::adam_kernel, dim3((ep::NW + 255) / 256), dim3(256), 0, c.stream, c.P, c.m, c.v, b.gpart,
                       b.nblocks, lr, bc1, bc2, 1, nullptr);
  }
  hipError_t e = hipStreamSynchronize(c.stream);
  if (e == hipSuccess && dloss) e = hipMemcpy(loss_out, dloss, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost);
  if (dloss) (void)hipFree(dloss);
  if (e != hipSuccess) return c.err = std::string("training: ") + hipGetErrorString(e), EP_ERR_HIP;
  return EP_OK;
}

int ep_ctrl_predict(ep_ctrl *x, const float *X, int64_t n, float *Y) {
  if (!x || !X || !Y) return EP_ERR_ARG;
  CtrlDev &c = x->d;
  CT_HIP(hipSetDevice(c.device));
  Batch b;
  int rc = make_batch(c, b, X, nullptr, n);
  if (rc) return rc;
  float *dy = nullptr;
  CT_HIP(hipMalloc(&dy, (size_t)n * ep::NO * sizeof(float)));
  hipLaunchKernelGGL(ep::mlp_predict_kernel, dim3(b.nblocks), dim3(ep::H), 0, c.stream, c.P, b.X, (int)n, dy);
  hipError_t e = hipStreamSynchronize(c.stream);
  if (e == hipSuccess) e = hipMemcpy(Y, dy, (size_t)n * ep::NO * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(dy);
  if (e != hipSuccess) return c.err = std::string("predict: ") + hipGetErrorString(e), EP_ERR_HIP;
  return EP_OK;
}

int ep_ctrl_loss_grad(ep_ctrl *x, const float *X, const float *Y, int64_t n, float *grad, float *loss) {
  if (!x || !X || !Y || !grad) return EP_ERR_ARG;
  CtrlDev &c = x->d;
  CT_HIP(hipSetDevice(c.device));
  Batch b;
  int rc = make_batch(c, b, X, Y, n);
  if (rc) return rc;
  hipLaunchKernelGGL(ep::mlp_grad_kernel, dim3(b.nblocks), dim3(ep::H), 0, c.stream, c.P, b.X, b.Y, (int)n,
                     b.gpart, b.lpart);
  hipLaunchKernelGGL(ep::adam_kernel, dim3((ep::NW + 255) / 256), dim3(256), 0, c.stream, c.P, c.m, c.v, b.gpart,
                     b.nblocks, 0.f, 1.f, 1.f, 0, b.scalar);
  float *dl = nullptr;
  CT_HIP(hipMalloc(&dl, sizeof(float)));
  hipLaunchKernelGGL(ep::loss_sum_kernel, dim3(1), dim3(256), 0, c.stream, b.lpart, b.nblocks, (int)n, dl);
  hipError_t e = hipStreamSynchronize(c.stream);
  if (e == hipSuccess) e = hipMemcpy(grad, b.scalar, ep::NW * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && loss) e = hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(dl);
  if (e != hipSuccess) return c.err = std::string("loss_grad: ") + hipGetErrorString(e), EP_ERR_HIP;
  return EP_OK;
}

}  // extern "C"
