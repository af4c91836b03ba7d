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
  const int32_t *net_start, *pin_obj;
  int32_t lane_off;  // pin_grad entry of pin p is p + lane_off
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
  // one reciprocal per sum (the division's rounding differs from a/b by an
  // ulp of double, far below the float result)
  double gp = 0.0, gm = 0.0;
  const double g = ig;
  if (ep != 0.f) {
    const double r = 1.0 / (double)dp;
    gp = (double)ep * r * (1.0 + g * ((double)(x - xmax) - (double)np * r));
  }
  if (em != 0.f) {
    const double r = 1.0 / (double)dm;
    gm = (double)em * r * (1.0 - g * ((double)(x - xmin) - (double)nm * r));
  }
  return (float)(gp - gm);
}

template <int MODEL>
__device__ __forceinline__ double net_value(float xmax, float xmin, float ig, float dp, float np, float dm,
                                            float nm) {
  if (MODEL == 1)
    return (double)(xmax - xmin) + (double)(logf(dp) / ig) + (double)(logf(dm) / ig);
  return ((double)xmax + (double)(np / dp)) - ((double)xmin + (double)(nm / dm));
}

// DPP lane moves (no LDS): row_shr:k reads lane i-k of the same 16-lane
// row, row_bcast:15 / row_bcast:31 pass lane 15 / 31 on to the following
// rows.  Lanes whose source falls outside get `old`.
constexpr int kDppRowShr = 0x110, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ float dppf(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, ROWMASK, 0xF, false));
}

// Whole-wave reduction: scan inside each row, then the row totals through
// lanes 15 and 31; lane 63 holds the total, read as a wave-uniform value.
template <class Op>
__device__ __forceinline__ float wave_reduce_dpp(float v, float id, Op op) {
  v = op(v, dppf<kDppRowShr + 1>(id, v));
  v = op(v, dppf<kDppRowShr + 2>(id, v));
  v = op(v, dppf<kDppRowShr + 4>(id, v));
  v = op(v, dppf<kDppRowShr + 8>(id, v));
  v = op(v, dppf<kDppRowBcast15, 0xA>(id, v));
  v = op(v, dppf<kDppRowBcast31, 0xC>(id, v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ float wave_max(float v) {
  return wave_reduce_dpp(v, -INFINITY, [](float a, float b) { return fmaxf(a, b); });
}
__device__ __forceinline__ float wave_min(float v) {
  return wave_reduce_dpp(v, INFINITY, [](float a, float b) { return fminf(a, b); });
}
__device__ __forceinline__ float wave_fsum(float v) {
  return wave_reduce_dpp(v, 0.f, [](float a, float b) { return a + b; });
}

// One wave per net of more than kNetLaneMax pins, grid-stride over nets.
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
      pin_grad[p + src.lane_off] = make_float2(gx, gy);
    }
  }
  hp = block_sum(hp);
  wv = block_sum(wv);
  if (threadIdx.x == 0) {
    partials[kPartialWL + 2 * (slot0 + blockIdx.x)] = hp;
    partials[kPartialWL + 2 * (slot0 + blockIdx.x) + 1] = wv;
  }
}

// Butterfly all-reduce inside aligned groups of D lanes (D = 2..32): lane
// xor 1 and xor 2 by quad permutes, xor 4 by the half-row mirror (the other
// quad of an 8-lane half holds that quad's result), xor 8 by the row mirror,
// xor 16 by a lane permute.  Every lane of a group ends with the same value.
constexpr int kDppQuadXor1 = 0xB1, kDppQuadXor2 = 0x4E, kDppRowMirror = 0x140, kDppRowHalfMirror = 0x141;

template <int D, class Op>
__device__ __forceinline__ float group_allreduce(float v, Op op) {
  if (D >= 2) v = op(v, dppf<kDppQuadXor1>(v, v));
  if (D >= 4) v = op(v, dppf<kDppQuadXor2>(v, v));
  if (D >= 8) v = op(v, dppf<kDppRowHalfMirror>(v, v));
  if (D >= 16) v = op(v, dppf<kDppRowMirror>(v, v));
  if (D >= 32) v = op(v, __shfl_xor(v, 16, 64));
  return v;
}

// Nets of 2..kNetLaneMax pins, grouped by degree class: a net of degree d in
// (D/2, D] takes D aligned lanes (pins in the first d, padding after), 64/D
// nets per 64-lane group.  The pins are stored in that lane order (object
// index, offset, gradient slot; object -1 for padding), so a group reads its
// pins with coalesced loads and one gather of centres.  Net-wide bound pins
// and the WA / LSE sums are butterfly all-reduces inside the D lanes (no
// segment flags, no broadcast); the net's first lane adds the net's HPWL and
// value.  Gradients are written in lane order (coalesced); the node sums
// read them back through node_pin_lane.
struct WlClasses {
  int group_start[6];  // first 64-lane group of class c (D = 2 << c); group_start[5] = all groups
};

struct LanePins {
  const int32_t *obj;
  const float2 *off;
  const float2 *pos;
};

// One 64-lane group from its loaded pin positions (act: the lane holds a
// pin, idx: the lane's slot).
template <int D, int MODEL, int EXPM, int GRAD>
__device__ __forceinline__ void wl_group_d(int idx, bool act, float px, float py, float igx, float igy,
                                           float2 *__restrict__ pin_grad, double &hp, double &wv) {
  const int lane = threadIdx.x & 63;
  auto fmax_ = [](float a, float b) { return fmaxf(a, b); };
  auto fmin_ = [](float a, float b) { return fminf(a, b); };
  auto fadd_ = [](float a, float b) { return a + b; };
  const float xmax = group_allreduce<D>(act ? px : -INFINITY, fmax_);
  const float xmin = group_allreduce<D>(act ? px : INFINITY, fmin_);
  const float ymax = group_allreduce<D>(act ? py : -INFINITY, fmax_);
  const float ymin = group_allreduce<D>(act ? py : INFINITY, fmin_);
  const bool head = act && (lane % D) == 0;
  if (head) hp += ((double)xmax - (double)xmin) + ((double)ymax - (double)ymin);
  if (!GRAD) return;
  float ep = 0.f, em = 0.f, fp = 0.f, fm = 0.f;
  if (act) {
    pin_terms<MODEL, EXPM>(px, xmax, xmin, igx, ep, em);
    pin_terms<MODEL, EXPM>(py, ymax, ymin, igy, fp, fm);
  }
  const float dpx = group_allreduce<D>(ep, fadd_);
  const float npx = group_allreduce<D>((px - xmax) * ep, fadd_);
  const float dmx = group_allreduce<D>(em, fadd_);
  const float nmx = group_allreduce<D>((px - xmin) * em, fadd_);
  const float dpy = group_allreduce<D>(fp, fadd_);
  const float npy = group_allreduce<D>((py - ymax) * fp, fadd_);
  const float dmy = group_allreduce<D>(fm, fadd_);
  const float nmy = group_allreduce<D>((py - ymin) * fm, fadd_);
  if (head)
    wv += net_value<MODEL>(xmax, xmin, igx, dpx, npx, dmx, nmx) + net_value<MODEL>(ymax, ymin, igy, dpy, npy, dmy, nmy);
  if (act) {
    const float gx = pin_grad_axis<MODEL>(px, xmax, xmin, igx, ep, em, dpx, npx, dmx, nmx);
    const float gy = pin_grad_axis<MODEL>(py, ymax, ymin, igy, fp, fm, dpy, npy, dmy, nmy);
    pin_grad[idx] = make_float2(gx, gy);
  }
}

// Each wave takes a contiguous run of 64-lane groups (XCD-contiguous block
// numbering); the class of a group is wave-uniform.
template <int MODEL, int EXPM, int GRAD>
__global__ void __launch_bounds__(kBlock) wl_group_kernel(WlClasses cls, LanePins lp, const DevScalars *sc,
                                                           float2 *__restrict__ pin_grad, double *partials) {
  if (sc->done) return;
  const float igx = (float)sc->inv_gamma_x, igy = (float)sc->inv_gamma_y;
  const int wpb = blockDim.x >> 6;
  const int waves = gridDim.x * wpb, wid = xcd_block() * wpb + (threadIdx.x >> 6);
  const int total = cls.group_start[5];
  const int per = (total + waves - 1) / waves;
  const int g0 = wid * per, g1 = min(total, g0 + per);
  double hp = 0, wv = 0;
  // Software pipeline over the wave's groups: the pin slots of group g + 2
  // and the centre gather of group g + 1 are in flight while group g is
  // reduced, so the dependent load chain (slot -> object -> centre) is paid
  // once per wave instead of once per group.
  const int lane = threadIdx.x & 63;
  int obj1 = -1, obj2 = -1;
  float2 off1 = make_float2(0.f, 0.f), off2 = off1, c1 = off1;
  if (g0 < g1) {
    obj1 = lp.obj[g0 * 64 + lane];
    off1 = lp.off[g0 * 64 + lane];
  }
  if (g0 + 1 < g1) {
    obj2 = lp.obj[(g0 + 1) * 64 + lane];
    off2 = lp.off[(g0 + 1) * 64 + lane];
  }
  if (obj1 >= 0) c1 = lp.pos[obj1];
  for (int g = g0; g < g1; ++g) {
    const int obj = obj1;
    const float2 off = off1, c0 = c1;
    obj1 = obj2;
    off1 = off2;
    obj2 = -1;
    if (g + 2 < g1) {
      obj2 = lp.obj[(g + 2) * 64 + lane];
      off2 = lp.off[(g + 2) * 64 + lane];
    }
    if (obj1 >= 0) c1 = lp.pos[obj1];
    const bool act = obj >= 0;
    const float px = act ? c0.x + off.x : 0.f, py = act ? c0.y + off.y : 0.f;
    const int idx = g * 64 + lane;
    int c = 0;
    while (c < 4 && g >= cls.group_start[c + 1]) ++c;
    switch (c) {
      case 0: wl_group_d<2, MODEL, EXPM, GRAD>(idx, act, px, py, igx, igy, pin_grad, hp, wv); break;
      case 1: wl_group_d<4, MODEL, EXPM, GRAD>(idx, act, px, py, igx, igy, pin_grad, hp, wv); break;
      case 2: wl_group_d<8, MODEL, EXPM, GRAD>(idx, act, px, py, igx, igy, pin_grad, hp, wv); break;
      case 3: wl_group_d<16, MODEL, EXPM, GRAD>(idx, act, px, py, igx, igy, pin_grad, hp, wv); break;
      default: wl_group_d<32, MODEL, EXPM, GRAD>(idx, act, px, py, igx, igy, pin_grad, hp, wv); break;
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
// (lambda_update 1), or no lower than the controller's floor (lambda_update 2).  The
// WA value is stored only when the wiring pass computed it (with_value).
// Spans of the previous evaluation's transform GEMM launches, from their
// per-block wall-clock stamps (the stamp set of the other parity): each
// launch's first block start to its last block end, added to the run's sum.
__device__ void fold_gemm_stamps(const GemmStamps &st, DevScalars *sc) {
  __shared__ unsigned long long fmin_[kGemmStamped][kFinish / 64], fmax_[kGemmStamped][kFinish / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long *set = st.t + (size_t)((sc->evals - 1) & 1) * 2 * kGemmStamped * st.stride;
  unsigned long long mark[kGemmStamped], mn[kGemmStamped], mx[kGemmStamped];
#pragma unroll
  for (int L = 0; L < kGemmStamped; ++L) mark[L] = set[(size_t)2 * L * st.stride + 1];
#pragma unroll
  for (int L = 0; L < kGemmStamped; ++L) {
    const unsigned long long *t = set + (size_t)2 * L * st.stride;
    mn[L] = ~0ull;
    mx[L] = 0;
    if (mark[L] != 0)
      for (int b = threadIdx.x; b < st.nblk[L]; b += blockDim.x) {
        mn[L] = min(mn[L], t[2 * b]);
        mx[L] = max(mx[L], t[2 * b + 1]);
      }
  }
#pragma unroll
  for (int L = 0; L < kGemmStamped; ++L) {
    for (int off = 32; off > 0; off >>= 1) {
      mn[L] = min(mn[L], __shfl_xor(mn[L], off, 64));
      mx[L] = max(mx[L], __shfl_xor(mx[L], off, 64));
    }
    if (lane == 0) {
      fmin_[L][wave] = mn[L];
      fmax_[L][wave] = mx[L];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int L = 0; L < kGemmStamped; ++L) {
      unsigned long long a = ~0ull, z = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        a = min(a, fmin_[L][w]);
        z = max(z, fmax_[L][w]);
      }
      if (z > a) {
        sc->gemm_ticks += z - a;
        sc->gemm_launches += 1;
        set[(size_t)2 * L * st.stride + 1] = 0;
      }
    }
  }
}

__device__ void reduce_wl_block(int nslots, double *partials, DevScalars *sc, int lambda_update, int with_value,
                                const GemmStamps &st) {
  if (st.t) fold_gemm_stamps(st, sc);
  if (sc->done) return;
  const double *const src[2] = {partials + kPartialWL, partials + kPartialWL + 1};
  double sum[2];
  block_sum_partials_strided_n(src, nslots, 2, sum);
  const double hp = sum[0], wv = sum[1];
  if (threadIdx.x == 0) {
    sc->hpwl = hp;
    if (with_value) sc->wl_value = wv;
    if (lambda_update) {
      double delta = hp - sc->last_hpwl;
      double mult = delta < 0.0 ? 1.05 : pow(1.05, -delta / 350000.0 + 1.0);
      mult = fmin(1.05, fmax(0.95, mult));
      const double ref_mult = mult;
      // the controller's floor of this iteration (0.95: none), kept inside
      // the reference's own multiplier range
      if (lambda_update == 2 || lambda_update == 4) mult = fmin(1.05, fmax(mult, sc->lam_floor));
      if (lambda_update <= 2) {
        sc->ref_mult_ema += (log(ref_mult) - sc->ref_mult_ema) * (1.0 / 16.0);
        sc->lambda *= mult;
        sc->last_hpwl = hp;
      } else {
        sc->lambda_next = sc->lambda * mult;
        sc->ref_mult_next = ref_mult;
      }
    }
  }
}

__global__ void __launch_bounds__(kBlock) reduce_wl_kernel(int nslots, double *partials, DevScalars *sc,
                                                            int lambda_update, int with_value, GemmStamps st) {
  reduce_wl_block(nslots, partials, sc, lambda_update, with_value, st);
}

// Node wirelength gradients: each node's pin gradients, read through
// node_pin_lane in modulePins order (objects.cpp:380-415 summed per module,
// eplace.cpp:694-737).  Block 0 is the WL reduction above (the same block
// size, so the same summation order): it needs the wiring kernels' partials
// as the node sums need their pin gradients, and as a block of this launch
// it runs beside the node sums instead of before them.  The other blocks
// take 256 objects each in XCD-contiguous order.  Blocks of 256 threads
// find room beside the transform GEMM and the field gather on the other
// stream sooner than blocks of 1024.
__global__ void __launch_bounds__(kBlock) node_wl_kernel(int N, const int32_t *__restrict__ node_pin_start,
                                                           const int32_t *__restrict__ node_pin_lane,
                                                           const float2 *__restrict__ pin_grad, float2 *gwl,
                                                           DevScalars *sc, int nslots, double *partials,
                                                           int lambda_update, GemmStamps st) {
  if (blockIdx.x == 0) {
    reduce_wl_block(nslots, partials, sc, lambda_update, 1, st);
    return;
  }
  if (sc->done) return;
  const int nb = gridDim.x - 1, b = blockIdx.x - 1;
  const int blk = nb % kXcds ? b : (b % kXcds) * (nb / kXcds) + b / kXcds;
  const int o = blk * blockDim.x + threadIdx.x;
  if (o >= N) return;
  float2 wl = make_float2(0.f, 0.f);
  int q = node_pin_start[o];
  const int e = node_pin_start[o + 1];
  // four gathers in flight, summed in pin order
  for (; q + 4 <= e; q += 4) {
    const int l0 = node_pin_lane[q], l1 = node_pin_lane[q + 1], l2 = node_pin_lane[q + 2],
              l3 = node_pin_lane[q + 3];
    const float2 p0 = pin_grad[l0], p1 = pin_grad[l1], p2 = pin_grad[l2], p3 = pin_grad[l3];
    wl.x += p0.x;
    wl.y += p0.y;
    wl.x += p1.x;
    wl.y += p1.y;
    wl.x += p2.x;
    wl.y += p2.y;
    wl.x += p3.x;
    wl.y += p3.y;
  }
  for (; q < e; ++q) {
    const float2 pg = pin_grad[node_pin_lane[q]];
    wl.x += pg.x;
    wl.y += pg.y;
  }
  gwl[o] = wl;
}

static int g_small_blocks(const Ctx &c) {
  const int wpb = kBlock / 64;
  const int groups = c.wl_groups;
  int b = (groups + wpb - 1) / wpb;
  // three blocks per CU leave room for the transform GEMM blocks running
  // beside this kernel on the other stream (measured at 1M cells: with four
  // the GEMMs are slowed more than the wiring kernel gains, with two the
  // field gather and node sums lose more than the GEMMs gain)
  b = b < 1 ? 1 : (b > 768 ? 768 : b);
  return (b + kXcds - 1) / kXcds * kXcds;
}
static int g_large_blocks(const Ctx &c) {
  int wpb = kBlock / 64;
  int b = (c.num_large + wpb - 1) / wpb;
  return b < 1 ? 1 : (b > 4096 ? 4096 : b);
}

template <int MODEL, int EXPM, int GRAD>
static void launch_wl_t(Ctx &c, hipStream_t st) {
  PinSrc src{c.net_start, c.pin_obj, c.large_lane_off, c.pin_dx, c.pin_dy, c.pos};
  int bs = g_small_blocks(c);
  WlClasses cls;
  for (int k = 0; k < 6; ++k) cls.group_start[k] = c.wl_gstart[k];
  LanePins lp{c.lane_obj, c.lane_off, c.pos};
  hipLaunchKernelGGL((wl_group_kernel<MODEL, EXPM, GRAD>), dim3(bs), dim3(kBlock), 0, st, cls, lp, c.sc,
                     c.pin_grad, c.partials);
  if (c.num_large) {
    hipLaunchKernelGGL((wl_large_kernel<MODEL, EXPM, GRAD>), dim3(g_large_blocks(c)), dim3(kBlock), 0,
                       st, c.num_large, c.large_nets, src, c.sc, c.pin_grad, c.partials, bs);
  }
}

void launch_wirelength(Ctx &c, bool gradient, hipStream_t st) {
  if (!st) st = c.stream;
  if (!gradient) {
    launch_wl_t<0, 0, 0>(c, st);
    return;
  }
  if (c.wl_model == 0) {
    if (c.exp_mode == 0) launch_wl_t<0, 0, 1>(c, st);
    else launch_wl_t<0, 1, 1>(c, st);
  } else {
    if (c.exp_mode == 0) launch_wl_t<1, 0, 1>(c, st);
    else launch_wl_t<1, 1, 1>(c, st);
  }
}

void launch_reduce_wirelength(Ctx &c, int lambda_update, hipStream_t st, bool with_value) {
  if (!st) st = c.stream;
  int n = g_small_blocks(c) + (c.num_large ? g_large_blocks(c) : 0);
  hipLaunchKernelGGL(reduce_wl_kernel, dim3(1), dim3(kBlock), 0, st, n, c.partials, c.sc, lambda_update,
                     with_value ? 1 : 0, c.gemm_stamps);
}

void launch_node_wl_reduce(Ctx &c, int lambda_update, hipStream_t st) {
  const int n = g_small_blocks(c) + (c.num_large ? g_large_blocks(c) : 0);
  int b = (c.N + kBlock - 1) / kBlock;
  b = (b + kXcds - 1) / kXcds * kXcds;
  hipLaunchKernelGGL(node_wl_kernel, dim3(b + 1), dim3(kBlock), 0, st, c.N, c.node_pin_start, c.node_pin_lane,
                     c.pin_grad, c.gwl, c.sc, n, c.partials, lambda_update, c.gemm_stamps);
}

}  // namespace ep
