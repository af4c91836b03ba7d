// Per-net bound pins, HPWL and the weighted-average (or log-sum-exp)
// wirelength gradient, pins sorted by net.
//
// Reference: Net::calcBoundPin (objects.cpp:75-144),
// Net::calcWirelengthWA_2D / getWirelengthGradientWA_2D (objects.cpp:248-415),
// LSE (objects.cpp:161-241, 423-437), PlaceDB::calcHPWL (placedb.cpp:424-432).
//
// The WA gradient is evaluated in the max/min-shifted form
//   g+ = e+_i/D+ (1 + ig (u_i - N'+/D+)),   u_i = x_i - x_max,  N'+ = sum u e+
//   g- = e-_i/D- (1 - ig (v_i - N'-/D-)),   v_i = x_i - x_min,  N'- = sum v e-
// which equals the reference's ((e + x ig e) D - ig e N)/D^2 form but does not
// cancel catastrophically in float when |x| >> 1/ig.  Exponents are the
// reference's (x - x_max) ig and (x_min - x) ig with its -300 cut-off
// (global.h:46); exp_mode 0 evaluates (1 + a/1024)^1024 (global.h:457-471) in
// double, exp_mode 1 uses expf.
#include "ep_internal.h"
#include "reduce.h"

namespace ep {

// Pin positions: centre of the pin's object or terminal (one packed 8-byte
// read) plus the pin offset (objects.cpp:454-459).
struct PinSrc {
  const int32_t *net_start, *pin_obj, *slot;
  const float *pdx, *pdy;
  const float2 *pos;
  __device__ __forceinline__ float2 at(int p) const {
    const float2 c = pos[pin_obj[p]];
    return make_float2(c.x + pdx[p], c.y + pdy[p]);
  }
};

template <int EXPM>
__device__ __forceinline__ float wl_exp(float a) {
  if (a <= -300.f) return 0.f;
  if (EXPM == 0) {
    double b = 1.0 + (double)a * (1.0 / 1024.0);
#pragma unroll
    for (int i = 0; i < 10; ++i) b *= b;
    return (float)b;
  }
  return expf(a);
}

struct NetAcc {
  float dpx, npx, dmx, nmx, dpy, npy, dmy, nmy;
};

template <int MODEL, int EXPM>
__device__ __forceinline__ void pin_terms(float x, float xmax, float xmin, float igx, float &ep, float &em) {
  ep = wl_exp<EXPM>((x - xmax) * igx);
  em = wl_exp<EXPM>((xmin - x) * igx);
}

// Pin gradient along one axis from the net sums, evaluated in FP64: the
// factor 1 - ig (v - N'-/D-) cancels when ig * (pin spread) ~ 1, which in
// float alone costs ~1e-5 of the node gradient on small 2-pin designs.
template <int MODEL>
__device__ __forceinline__ float pin_grad_axis(float x, float xmax, float xmin, float ig, float ep, float em,
                                               float dp, float np, float dm, float nm) {
  if (MODEL == 1) return (float)((double)ep / dp - (double)em / dm);  // LSE
  double gp = 0.0, gm = 0.0;
  const double g = ig;
  if (ep != 0.f) gp = (double)ep / dp * (1.0 + g * ((double)(x - xmax) - (double)np / dp));
  if (em != 0.f) gm = (double)em / dm * (1.0 - g * ((double)(x - xmin) - (double)nm / dm));
  return (float)(gp - gm);
}

template <int MODEL>
__device__ __forceinline__ double net_value(float xmax, float xmin, float ig, float dp, float np, float dm,
                                            float nm) {
  if (MODEL == 1)
    return (double)(xmax - xmin) + (double)(logf(dp) / ig) + (double)(logf(dm) / ig);
  return ((double)xmax + (double)(np / dp)) - ((double)xmin + (double)(nm / dm));
}

__device__ __forceinline__ float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_fsum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One wave per net of degree > 64, grid-stride over nets.
template <int MODEL, int EXPM, int GRAD>
__global__ void __launch_bounds__(kBlock) wl_large_kernel(int num, const int32_t *__restrict__ nets,
                                                           PinSrc src, const DevScalars *sc,
                                                           float2 *__restrict__ pin_grad, double *partials,
                                                           int slot0) {
  if (sc->done) return;
  float igx = (float)sc->inv_gamma_x, igy = (float)sc->inv_gamma_y;
  int lane = threadIdx.x & 63;
  int wpb = blockDim.x >> 6;
  double hp = 0, wv = 0;
  for (int t = blockIdx.x * wpb + (threadIdx.x >> 6); t < num; t += gridDim.x * wpb) {
    int e = nets[t];
    int s = src.net_start[e], n = src.net_start[e + 1];
    float xmax = -INFINITY, xmin = INFINITY, ymax = -INFINITY, ymin = INFINITY;
    for (int p = s + lane; p < n; p += 64) {
      const float2 q = src.at(p);
      const float px = q.x, py = q.y;
      xmax = fmaxf(xmax, px);
      xmin = fminf(xmin, px);
      ymax = fmaxf(ymax, py);
      ymin = fminf(ymin, py);
    }
    xmax = wave_max(xmax);
    xmin = wave_min(xmin);
    ymax = wave_max(ymax);
    ymin = wave_min(ymin);
    if (lane == 0) hp += ((double)xmax - (double)xmin) + ((double)ymax - (double)ymin);
    if (!GRAD) continue;
    NetAcc a = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = s + lane; p < n; p += 64) {
      const float2 q = src.at(p);
      const float px = q.x, py = q.y;
      float ep, em, fp, fm;
      pin_terms<MODEL, EXPM>(px, xmax, xmin, igx, ep, em);
      pin_terms<MODEL, EXPM>(py, ymax, ymin, igy, fp, fm);
      a.dpx += ep;
      a.npx += (px - xmax) * ep;
      a.dmx += em;
      a.nmx += (px - xmin) * em;
      a.dpy += fp;
      a.npy += (py - ymax) * fp;
      a.dmy += fm;
      a.nmy += (py - ymin) * fm;
    }
    a.dpx = wave_fsum(a.dpx);
    a.npx = wave_fsum(a.npx);
    a.dmx = wave_fsum(a.dmx);
    a.nmx = wave_fsum(a.nmx);
    a.dpy = wave_fsum(a.dpy);
    a.npy = wave_fsum(a.npy);
    a.dmy = wave_fsum(a.dmy);
    a.nmy = wave_fsum(a.nmy);
    if (lane == 0)
      wv += net_value<MODEL>(xmax, xmin, igx, a.dpx, a.npx, a.dmx, a.nmx) +
            net_value<MODEL>(ymax, ymin, igy, a.dpy, a.npy, a.dmy, a.nmy);
    for (int p = s + lane; p < n; p += 64) {
      const float2 q = src.at(p);
      const float px = q.x, py = q.y;
      float ep, em, fp, fm;
      pin_terms<MODEL, EXPM>(px, xmax, xmin, igx, ep, em);
      pin_terms<MODEL, EXPM>(py, ymax, ymin, igy, fp, fm);
      float gx = pin_grad_axis<MODEL>(px, xmax, xmin, igx, ep, em, a.dpx, a.npx, a.dmx, a.nmx);
      float gy = pin_grad_axis<MODEL>(py, ymax, ymin, igy, fp, fm, a.dpy, a.npy, a.dmy, a.nmy);
      const int sl = src.slot[p];
      if (sl >= 0) pin_grad[sl] = make_float2(gx, gy);
    }
  }
  hp = block_sum(hp);
  wv = block_sum(wv);
  if (threadIdx.x == 0) {
    partials[kPartialWL + 2 * (slot0 + blockIdx.x)] = hp;
    partials[kPartialWL + 2 * (slot0 + blockIdx.x) + 1] = wv;
  }
}

// Segmented reductions inside a wave over runs of lanes holding the same
// net.  same[k] says lane+2^k holds the same net as this lane.
template <class Op>
__device__ __forceinline__ float seg_reduce(float v, const bool *same, Op op) {
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    float o = __shfl_down(v, 1 << k, 64);
    if (same[k]) v = op(v, o);
  }
  return v;
}

// One wave per group of consecutive nets of total degree <= 64, one pin per
// lane (coalesced pin loads, no divergence on net degree).  Bound pins,
// HPWL, the WA/LSE sums and the per-pin gradient come from segmented
// shuffle reductions over each net's lanes; the net's head lane holds the
// result and broadcasts it.  Each wave takes a contiguous run of groups and
// the blocks are numbered XCD-contiguously, so with nets stored by their
// lowest node one XCD reads the centres of a narrow band of objects.
template <int MODEL, int EXPM, int GRAD>
__global__ void __launch_bounds__(kBlock) wl_group_kernel(int num_groups, const int32_t *__restrict__ gbeg,
                                                           const int32_t *__restrict__ gend,
                                                           const int32_t *__restrict__ pin_net, PinSrc src,
                                                           const DevScalars *sc, float2 *__restrict__ pin_grad,
                                                           double *partials) {
  if (sc->done) return;
  const float igx = (float)sc->inv_gamma_x, igy = (float)sc->inv_gamma_y;
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int waves = gridDim.x * wpb, wid = xcd_block() * wpb + (threadIdx.x >> 6);
  const int per = (num_groups + waves - 1) / waves;
  const int gi0 = wid * per, gi1 = min(num_groups, gi0 + per);
  double hp = 0, wv = 0;
  for (int gi = gi0; gi < gi1; ++gi) {
    const int p0 = gbeg[gi], p1 = gend[gi];
    const int p = p0 + lane;
    const bool act = p < p1;
    int net = act ? pin_net[p] : -1 - lane;  // inactive lanes: unique segments
    float px = 0.f, py = 0.f;
    if (act) {
      const float2 q = src.at(p);
      px = q.x;
      py = q.y;
    }
    bool same[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      int on = __shfl_down(net, 1 << k, 64);
      same[k] = (lane + (1 << k) < 64) && on == net;
    }
    int head = act ? src.net_start[net] - p0 : lane;
    auto fmax_ = [](float a, float b) { return fmaxf(a, b); };
    auto fmin_ = [](float a, float b) { return fminf(a, b); };
    auto fadd_ = [](float a, float b) { return a + b; };
    float xmax = __shfl(seg_reduce(px, same, fmax_), head, 64);
    float xmin = __shfl(seg_reduce(px, same, fmin_), head, 64);
    float ymax = __shfl(seg_reduce(py, same, fmax_), head, 64);
    float ymin = __shfl(seg_reduce(py, same, fmin_), head, 64);
    const bool is_head = act && head == lane;
    if (is_head) hp += ((double)xmax - (double)xmin) + ((double)ymax - (double)ymin);
    if (!GRAD) continue;
    float ep = 0.f, em = 0.f, fp = 0.f, fm = 0.f;
    if (act) {
      pin_terms<MODEL, EXPM>(px, xmax, xmin, igx, ep, em);
      pin_terms<MODEL, EXPM>(py, ymax, ymin, igy, fp, fm);
    }
    float dpx = __shfl(seg_reduce(ep, same, fadd_), head, 64);
    float npx = __shfl(seg_reduce((px - xmax) * ep, same, fadd_), head, 64);
    float dmx = __shfl(seg_reduce(em, same, fadd_), head, 64);
    float nmx = __shfl(seg_reduce((px - xmin) * em, same, fadd_), head, 64);
    float dpy = __shfl(seg_reduce(fp, same, fadd_), head, 64);
    float npy = __shfl(seg_reduce((py - ymax) * fp, same, fadd_), head, 64);
    float dmy = __shfl(seg_reduce(fm, same, fadd_), head, 64);
    float nmy = __shfl(seg_reduce((py - ymin) * fm, same, fadd_), head, 64);
    if (is_head)
      wv += net_value<MODEL>(xmax, xmin, igx, dpx, npx, dmx, nmx) + net_value<MODEL>(ymax, ymin, igy, dpy, npy, dmy, nmy);
    if (act) {
      float gx = pin_grad_axis<MODEL>(px, xmax, xmin, igx, ep, em, dpx, npx, dmx, nmx);
      float gy = pin_grad_axis<MODEL>(py, ymax, ymin, igy, fp, fm, dpy, npy, dmy, nmy);
      const int sl = src.slot[p];
      if (sl >= 0) pin_grad[sl] = make_float2(gx, gy);
    }
  }
  hp = block_sum(hp);
  wv = block_sum(wv);
  if (threadIdx.x == 0) {
    partials[kPartialWL + 2 * blockIdx.x] = hp;
    partials[kPartialWL + 2 * blockIdx.x + 1] = wv;
  }
}

// Sums the per-block HPWL/WL partials in a fixed order; optionally applies
// updatePenaltyFactor (eplace.cpp:846-883) with FP64 HPWL and lambda
// (lambda_update 1), times the controller's factor (lambda_update 2).
__global__ void __launch_bounds__(kFinish) reduce_wl_kernel(int nslots, double *partials, DevScalars *sc,
                                                            int lambda_update) {
  if (sc->done) return;
  double hp = block_sum_partials_strided(partials + kPartialWL, nslots, 2);
  double wv = block_sum_partials_strided(partials + kPartialWL + 1, nslots, 2);
  if (threadIdx.x == 0) {
    sc->hpwl = hp;
    sc->wl_value = wv;
    if (lambda_update) {
      double delta = hp - sc->last_hpwl;
      double mult = delta < 0.0 ? 1.05 : pow(1.05, -delta / 350000.0 + 1.0);
      mult = fmin(1.05, fmax(0.95, mult));
      // the controller's factor of this iteration, kept inside the
      // reference's own multiplier range
      if (lambda_update == 2) mult = fmin(1.05, fmax(0.95, mult * sc->lam_mult));
      sc->lambda *= mult;
      sc->last_hpwl = hp;
    }
  }
}

static int g_small_blocks(const Ctx &c) {
  int wpb = kBlock / 64;
  int b = (c.num_groups + wpb - 1) / wpb;
  b = b < 1 ? 1 : (b > 8192 ? 8192 : b);
  return (b + kXcds - 1) / kXcds * kXcds;
}
static int g_large_blocks(const Ctx &c) {
  int wpb = kBlock / 64;
  int b = (c.num_large + wpb - 1) / wpb;
  return b < 1 ? 1 : (b > 4096 ? 4096 : b);
}

template <int MODEL, int EXPM, int GRAD>
static void launch_wl_t(Ctx &c) {
  PinSrc src{c.net_start, c.pin_obj, c.pin_slot, c.pin_dx, c.pin_dy, c.pos};
  int bs = g_small_blocks(c);
  hipLaunchKernelGGL((wl_group_kernel<MODEL, EXPM, GRAD>), dim3(bs), dim3(kBlock), 0, c.stream, c.num_groups,
                     c.group_begin, c.group_end, c.pin_net, src, c.sc, c.pin_grad, c.partials);
  if (c.num_large) {
    hipLaunchKernelGGL((wl_large_kernel<MODEL, EXPM, GRAD>), dim3(g_large_blocks(c)), dim3(kBlock), 0,
                       c.stream, c.num_large, c.large_nets, src, c.sc, c.pin_grad, c.partials, bs);
  }
}

void launch_wirelength(Ctx &c, bool gradient) {
  if (!gradient) {
    launch_wl_t<0, 0, 0>(c);
    return;
  }
  if (c.wl_model == 0) {
    if (c.exp_mode == 0) launch_wl_t<0, 0, 1>(c);
    else launch_wl_t<0, 1, 1>(c);
  } else {
    if (c.exp_mode == 0) launch_wl_t<1, 0, 1>(c);
    else launch_wl_t<1, 1, 1>(c);
  }
}

void launch_reduce_wirelength(Ctx &c, int lambda_update) {
  int n = g_small_blocks(c) + (c.num_large ? g_large_blocks(c) : 0);
  hipLaunchKernelGGL(reduce_wl_kernel, dim3(1), dim3(kFinish), 0, c.stream, n, c.partials, c.sc, lambda_update);
}

}  // namespace ep
