// Host side of the engine: design import, one-time initialisation (fillers,
// bin grid, terminal/base density), device buffers and the C ABI.
//
// The initialisation restates EPlacer_2D::fillerInitialization and
// binInitialization (EPlace/eplace.cpp:61-412) in the reference's float
// arithmetic, so the filler set, bin grid and fixed density maps equal the
// reference's for the same design.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ep_internal.h"

namespace ep {

namespace {

inline bool f_less(float a, float b) { return (double)(a - b) < -kEps; }
inline bool f_greater(float a, float b) { return (double)(a - b) > kEps; }
inline bool f_equal(float a, float b) { return std::fabs((double)(a - b)) < kEps; }

inline double olen(double x1, double x2, double x3, double x4) {
  double s = std::max(x1, x3), e = std::min(x2, x4);
  return s >= e ? 0.0 : s - e;
}
inline double oarea(double l1, double b1, double r1, double t1, double l2, double b2, double r2, double t2) {
  double h = olen(l1, r1, l2, r2);
  if (h == 0) return 0;
  double v = olen(b1, t1, b2, t2);
  if (v == 0) return 0;
  return h * v;
}

template <class T>
bool dev_alloc(Ctx &c, T *&p, size_t n) {
  if (n == 0) n = 1;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) {
    c.err = "hipMalloc failed";
    return false;
  }
  return true;
}
template <class T>
bool dev_upload(Ctx &c, T *&p, const T *h, size_t n) {
  if (!dev_alloc(c, p, n)) return false;
  if (n && hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
    c.err = "hipMemcpy failed";
    return false;
  }
  return true;
}
template <class T>
bool dev_zero(Ctx &c, T *&p, size_t n) {
  if (!dev_alloc(c, p, n)) return false;
  if (hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    c.err = "hipMemset failed";
    return false;
  }
  return true;
}

// Host-side state built once (kept small; the device owns the bulk).
struct Setup {
  std::vector<float> x, y, w, h, area;
  std::vector<uint8_t> kind;
  std::vector<int32_t> npins;
  std::vector<float> bllx, burx, blly, bury, term_map, base_map;
};

// PlaceDB::setModuleLocation_2D (placedb.cpp:200-240) for a movable object,
// returning the centre Module::setLocation_2D derives (objects.cpp:518-527).
void location_to_center(const Ctx &c, float w, float h, float x, float y, float &cx, float &cy) {
  if (x < c.core_llx) x = c.core_llx;
  if (x + w > c.core_urx) x = (float)((double)(c.core_urx - w) - kEps);
  if (y < c.core_lly) y = c.core_lly;
  if (y + h > c.core_ury) y = (float)((double)(c.core_ury - h) - kEps);
  cx = x + (float)0.5 * w;
  cy = y + (float)0.5 * h;
}

// PlaceDB::setModuleCenter_2D clamp (placedb.cpp:253-288).
void clamp_center(const Ctx &c, float w, float h, float &x, float &y) {
  double hw = 0.5 * (double)w, hh = 0.5 * (double)h;
  if ((double)x - hw < (double)c.core_llx) x = (float)((double)c.core_llx + hw + kEps);
  if ((double)x + hw > (double)c.core_urx) x = (float)((double)c.core_urx - hw - kEps);
  if ((double)y - hh < (double)c.core_lly) y = (float)((double)c.core_lly + hh + kEps);
  if ((double)y + hh > (double)c.core_ury) y = (float)((double)c.core_ury - hh - kEps);
}

// Random filler position (setModuleLocation_2D_random, placedb.cpp:353-379).
void random_filler(Ctx &c, float w, float h, float &cx, float &cy) {
  float x = (float)c.rng->next();
  float y = (float)c.rng->next();
  float pw = c.core_urx - c.core_llx, ph = c.core_ury - c.core_lly;
  ph -= h;
  pw -= w;
  float inv = (float)1.0 / (float)2147483647;
  x = x * inv * pw;
  y = y * inv * ph;
  location_to_center(c, w, h, x, y, cx, cy);
}

int validate(Ctx &c, const ep_design *d, const ep_params *p) {
  if (!d || !p) return c.err = "null design or params", EP_ERR_ARG;
  if (d->num_nodes <= 0 || d->num_rows <= 0 || d->num_nets < 0 || d->num_terminals < 0 || d->num_pins < 0)
    return c.err = "bad design sizes", EP_ERR_ARG;
  if (!d->node_w || !d->node_h || !d->node_x || !d->node_y || !d->row_bottom || !d->row_height || !d->row_x0 ||
      !d->row_x1 || !d->net_start)
    return c.err = "null design array", EP_ERR_ARG;
  if (d->num_pins > 0 && (!d->pin_obj || !d->pin_dx || !d->pin_dy)) return c.err = "null pin array", EP_ERR_ARG;
  if (d->num_terminals > 0 && (!d->term_w || !d->term_h || !d->term_x || !d->term_y || !d->term_ni))
    return c.err = "null terminal array", EP_ERR_ARG;
  if (d->num_pins >= (int64_t)INT32_MAX) return c.err = "too many pins", EP_ERR_ARG;
  if (d->net_start[0] != 0 || d->net_start[d->num_nets] != d->num_pins)
    return c.err = "net_start must run from 0 to num_pins", EP_ERR_ARG;
  for (int e = 0; e < d->num_nets; ++e)
    if (d->net_start[e + 1] < d->net_start[e]) return c.err = "net_start not monotone", EP_ERR_ARG;
  int64_t no = (int64_t)d->num_nodes + d->num_terminals;
  for (int64_t q = 0; q < d->num_pins; ++q)
    if (d->pin_obj[q] < 0 || d->pin_obj[q] >= no) return c.err = "pin_obj out of range", EP_ERR_ARG;
  if (!(p->target_density > 0.f)) return c.err = "target_density must be > 0", EP_ERR_ARG;
  // partial-sum slots bound the grid: nx*ny/256 and objects/256 <= kMaxPartials
  if (p->bins_x < 0 || p->bins_y < 0 || p->bins_x > 2048 || p->bins_y > 2048)
    return c.err = "bins out of range [0, 2048]", EP_ERR_ARG;
  if (p->poisson_mode < 0 || p->poisson_mode > 3)
    return c.err = "poisson_mode must be 0 (FP32), 1 (split-BF16), 2 (split-FP16) or 3 (3-plane split-BF16)",
           EP_ERR_ARG;
  if ((p->bins_x > 0 && p->bins_x % 2) || (p->bins_y > 0 && p->bins_y % 2))
    return c.err = "bins per side must be even", EP_ERR_ARG;
  if (p->wirelength_model < 0 || p->wirelength_model > 1 || p->exp_mode < 0 || p->exp_mode > 1)
    return c.err = "bad wirelength_model/exp_mode", EP_ERR_ARG;
  return EP_OK;
}

void rebuild_stage(Ctx &c) {
  c.h_stage_idx.clear();
  std::vector<uint8_t> in(c.O, 0);
  for (int o = 0; o < c.O; ++o) {
    bool s = c.stage == EP_STAGE_MGP || (c.stage == EP_STAGE_FILLERONLY && c.h_kind[o] == KIND_FILLER) ||
             (c.stage == EP_STAGE_CGP && c.h_kind[o] != KIND_MACRO);
    if (s) {
      c.h_stage_idx.push_back(o);
      in[o] = 1;
    }
  }
  (void)hipMemcpy(c.in_stage, in.data(), c.O, hipMemcpyHostToDevice);
}

int create(Ctx &c, const ep_design *d, const ep_params *p) {
  int rc = validate(c, d, p);
  if (rc) return rc;
  c.device = p->device;
  EP_HIP(hipSetDevice(c.device));
  EP_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  EP_HIP(hipStreamCreateWithFlags(&c.stream2, hipStreamNonBlocking));
  EP_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
  EP_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
  // timing events release at device scope only: a system-scope release writes
  // back the L2 at every event and stretches the kernels between them
  for (auto &e : c.ev) EP_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  EP_HIP(hipHostMalloc((void **)&c.h_sc, sizeof(Ctx::HostScalars), hipHostMallocCoherent));
  c.h_sc->seq = 0;
  c.td = p->target_density;
  c.wl_model = p->wirelength_model;
  c.exp_mode = p->exp_mode;
  c.poisson_mode = p->poisson_mode;
  c.N = d->num_nodes;
  c.NT = d->num_terminals;
  c.E = d->num_nets;
  c.P = d->num_pins;
  Setup s;
  // ---- nodes (PlaceDB::addNode, placedb.cpp:102-111; Module ctor area) ----
  float rowh = (float)d->common_row_height;
  for (int i = 0; i < c.N; ++i) {
    float w = d->node_w[i], h = d->node_h[i];
    s.w.push_back(w);
    s.h.push_back(h);
    s.area.push_back(w * h);
    s.kind.push_back(f_greater(h, rowh) ? KIND_MACRO : KIND_CELL);
    s.x.push_back(d->node_x[i]);
    s.y.push_back(d->node_y[i]);
  }
  s.npins.assign(c.N, 0);
  for (int64_t q = 0; q < c.P; ++q)
    if (d->pin_obj[q] < c.N) s.npins[d->pin_obj[q]]++;
  // ---- core region (PlaceDB::setCoreRegion, placedb.cpp:52-77) ----
  int R = d->num_rows;
  float bottom = (float)d->row_bottom[0];
  float top = (float)(d->row_bottom[R - 1] + d->row_height[R - 1]);
  float left = (float)d->row_x0[0], right = (float)d->row_x1[0];
  float tra = 0;
  for (int r = 0; r < R; ++r) {
    left = std::min(left, (float)d->row_x0[r]);
    right = std::max(right, (float)d->row_x1[r]);
    tra += (float)((double)((float)d->row_x1[r] - (float)d->row_x0[r]) * d->row_height[r]);
  }
  c.core_llx = left;
  c.core_lly = bottom;
  c.core_urx = right;
  c.core_ury = top;
  c.total_row_area = tra;
  auto row_lly = [&](int r) { return (float)d->row_bottom[r]; };
  auto row_ury = [&](int r) { return (float)((double)(float)d->row_bottom[r] + d->row_height[r]); };
  std::vector<float> tllx(c.NT), tlly(c.NT);
  for (int t = 0; t < c.NT; ++t) {
    tllx[t] = d->term_x[t] - (float)0.5 * d->term_w[t];
    tlly[t] = d->term_y[t] - (float)0.5 * d->term_h[t];
  }
  // ---- fillers (eplace.cpp:61-208) ----
  float overlap = 0;
  for (int t = 0; t < c.NT; ++t) {
    if (d->term_ni[t]) continue;
    for (int r = 0; r < R; ++r)
      overlap = (float)((double)overlap + oarea(tllx[t], tlly[t], tllx[t] + d->term_w[t], tlly[t] + d->term_h[t],
                                                 (float)d->row_x0[r], row_lly(r), (float)d->row_x1[r], row_ury(r)));
  }
  float ws = c.total_row_area - overlap;
  float sa = 0, ma = 0;
  for (int i = 0; i < c.N; ++i) {
    if (s.kind[i] == KIND_MACRO) ma += s.area[i];
    else sa += s.area[i];
  }
  c.std_area = sa;
  c.macro_area = ma;
  float total_filler = ws * c.td - (sa + ma * c.td);
  std::vector<float> sorted(s.area.begin(), s.area.end());
  std::sort(sorted.begin(), sorted.end());
  int lo = (int)(0.05 * (float)c.N), hi = (int)(0.95 * (float)c.N);
  float tot = 0;
  for (int i = lo; i < hi; ++i) tot += sorted[i];
  float favg = hi > lo ? tot / (float)(hi - lo) : sorted[0];
  c.filler_h = (float)d->common_row_height;
  c.filler_w = favg / c.filler_h;
  int count = (int)(total_filler / favg + 0.5);
  if (p->filler_count >= 0) count = p->filler_count;
  if (count < 0) count = 0;
  if ((int64_t)c.N + count > (int64_t)kMaxPartials * kBlock)
    return c.err = "too many objects (nodes + fillers > 8388608)", EP_ERR_ARG;
  c.F = count;
  c.O = c.N + c.F;
  c.rng = new Rand(p->filler_seed);
  for (int f = 0; f < c.F; ++f) {
    float cx, cy;
    if (p->filler_xy) {
      cx = p->filler_xy[2 * f];
      cy = p->filler_xy[2 * f + 1];
      clamp_center(c, c.filler_w, c.filler_h, cx, cy);
    } else {
      random_filler(c, c.filler_w, c.filler_h, cx, cy);
    }
    s.x.push_back(cx);
    s.y.push_back(cy);
    s.w.push_back(c.filler_w);
    s.h.push_back(c.filler_h);
    s.area.push_back(c.filler_w * c.filler_h);
    s.kind.push_back(KIND_FILLER);
    s.npins.push_back(0);
  }
  // ---- bin grid (eplace.cpp:229-314) ----
  float cw = c.core_urx - c.core_llx, ch = c.core_ury - c.core_lly;
  float core_area = cw * ch;
  float avg = (float)(1.0 * (sa + ma) / c.N);
  float ideal_area = avg / c.td;
  int ideal = (int)(1.0 * (core_area / ideal_area) + 0.5f);
  int dim = 1024;
  for (int i = 1; i < 10; ++i)
    if ((2 << i) * (2 << i) <= ideal && (2 << (i + 1)) * (2 << (i + 1)) > ideal) {
      dim = 2 << i;
      break;
    }
  c.nx = p->bins_x > 0 ? p->bins_x : dim;
  c.ny = p->bins_y > 0 ? p->bins_y : dim;
  c.bsx = cw / (float)c.nx;
  c.bsy = ch / (float)c.ny;
  s.bllx.resize(c.nx);
  s.burx.resize(c.nx);
  s.blly.resize(c.ny);
  s.bury.resize(c.ny);
  for (int i = 0; i < c.nx; ++i) {
    s.bllx[i] = (float)i * c.bsx + c.core_llx;
    s.burx[i] = s.bllx[i] + c.bsx;
  }
  for (int j = 0; j < c.ny; ++j) {
    s.blly[j] = (float)j * c.bsy + c.core_lly;
    s.bury[j] = s.blly[j] + c.bsy;
  }
  size_t nb = (size_t)c.nx * c.ny;
  s.term_map.assign(nb, 0.f);
  s.base_map.assign(nb, 0.f);
  // terminal density (eplace.cpp:327-374)
  for (int t = 0; t < c.NT; ++t) {
    float tl = tllx[t], tb = tlly[t], tr = tllx[t] + d->term_w[t], tt = tlly[t] + d->term_h[t];
    if (tl < c.core_llx || tl + d->term_w[t] > c.core_urx) continue;
    if (tb < c.core_lly || tb + d->term_h[t] > c.core_ury) continue;
    int i0 = (int)((tl - c.core_llx) / c.bsx), i1 = (int)((tr - c.core_llx) / c.bsx);
    int j0 = (int)((tb - c.core_lly) / c.bsy), j1 = (int)((tt - c.core_lly) / c.bsy);
    i1 = std::min(i1, c.nx - 1);
    j1 = std::min(j1, c.ny - 1);
    for (int i = std::max(i0, 0); i <= i1; ++i)
      for (int j = std::max(j0, 0); j <= j1; ++j) {
        float &v = s.term_map[(size_t)i * c.ny + j];
        v = (float)((double)v + (double)c.td * oarea(s.bllx[i], s.blly[j], s.burx[i], s.bury[j], tl, tb, tr, tt));
      }
  }
  // base density (eplace.cpp:385-409); rows that miss a bin add exactly 0
  std::vector<std::vector<int>> rows_of(c.ny);
  for (int r = 0; r < R; ++r) {
    float rl = row_lly(r), ru = row_ury(r);
    int j0 = std::max(0, (int)std::floor((rl - c.core_lly) / c.bsy) - 1);
    int j1 = std::min(c.ny - 1, (int)std::floor((ru - c.core_lly) / c.bsy) + 1);
    for (int j = j0; j <= j1; ++j)
      if (olen(s.blly[j], s.bury[j], rl, ru) != 0) rows_of[j].push_back(r);
  }
  float ba = c.bsx * c.bsy;
  for (int i = 0; i < c.nx; ++i)
    for (int j = 0; j < c.ny; ++j) {
      float avail = 0;
      for (int r : rows_of[j]) {
        std::vector<int> &v = rows_of[j];
        (void)v;
        avail = (float)((double)avail + oarea(s.bllx[i], s.blly[j], s.burx[i], s.bury[j], (float)d->row_x0[r],
                                              row_lly(r), (float)d->row_x1[r], row_ury(r)));
      }
      s.base_map[(size_t)i * c.ny + j] = f_equal(ba, avail) ? 0.f : c.td * (ba - avail);
    }
  // ---- fixed point: the largest bin sum is below the total charge ----
  double total = 0;
  for (int o = 0; o < c.O; ++o) total += (double)s.area[o] * (s.kind[o] == KIND_MACRO ? c.td : 1.0);
  c.shift = std::max(8, std::min(60, (int)std::floor(61.0 - std::log2(total + 1.0))));
  // ---- big objects (may cover more than kBigObjectBins bins) ----
  std::vector<int32_t> big;
  std::vector<uint8_t> is_big(c.O, 0);
  for (int o = 0; o < c.O; ++o) {
    double sx = std::floor((double)s.w[o] / c.bsx) + 2, sy = std::floor((double)s.h[o] / c.bsy) + 2;
    if (sx * sy > kBigObjectBins) {
      big.push_back(o);
      is_big[o] = 1;
    }
  }
  c.num_big = (int)big.size();
  // ---- nets, stored by their lowest movable node (stable), so that runs of
  // consecutive nets read the centres of a narrow band of objects; pins keep
  // their order inside a net.  Objects are numbered nodes, fillers, then
  // terminals (pin_obj of terminal t becomes O + t). ----
  std::vector<int32_t> order(c.E);
  {
    std::vector<int32_t> key(c.E), cnt(c.N + 2, 0);
    for (int e = 0; e < c.E; ++e) {
      int32_t k = c.N;
      for (int64_t q = d->net_start[e]; q < d->net_start[e + 1]; ++q)
        if (d->pin_obj[q] < c.N) k = std::min(k, d->pin_obj[q]);
      key[e] = k;
      cnt[k + 1]++;
    }
    for (int k = 0; k <= c.N; ++k) cnt[k + 1] += cnt[k];
    std::vector<int32_t> by_node(c.E);
    for (int e = 0; e < c.E; ++e) by_node[cnt[key[e]]++] = e;
    // then stable by degree class: nets of 0-1 pins (no HPWL, no gradient),
    // then 2, 3-4, 5-8, 9-16, 17-32 pins, then the larger ones
    std::vector<int32_t> dcnt(9, 0);
    auto dkey = [&](int e) {
      const int64_t dg = d->net_start[e + 1] - d->net_start[e];
      if (dg <= 1) return 0;
      if (dg > kNetLaneMax) return 6;
      int k = 1;
      while ((2 << (k - 1)) < dg) ++k;  // 2 -> 1, 3-4 -> 2, 5-8 -> 3, 9-16 -> 4, 17-32 -> 5
      return k;
    };
    for (int e = 0; e < c.E; ++e) dcnt[dkey(e) + 1]++;
    for (int k = 0; k < 8; ++k) dcnt[k + 1] += dcnt[k];
    for (int k = 0; k < 5; ++k) c.wl_first[k] = dcnt[k + 1], c.wl_count[k] = dcnt[k + 2] - dcnt[k + 1];
    for (int e : by_node) order[dcnt[dkey(e)]++] = e;
  }
  std::vector<int32_t> ns32(c.E + 1), pobj(std::max<int64_t>(c.P, 1)), pin_net(std::max<int64_t>(c.P, 1));
  std::vector<float> pdx(std::max<int64_t>(c.P, 1)), pdy(std::max<int64_t>(c.P, 1));
  std::vector<int32_t> pin_of(std::max<int64_t>(c.P, 1));  // device pin of each design pin
  {
    int32_t q2 = 0;
    for (int e2 = 0; e2 < c.E; ++e2) {
      int e = order[e2];
      ns32[e2] = q2;
      for (int64_t q = d->net_start[e]; q < d->net_start[e + 1]; ++q, ++q2) {
        int32_t o = d->pin_obj[q];
        pobj[q2] = o < c.N ? o : c.O + (o - c.N);
        pdx[q2] = d->pin_dx[q];
        pdy[q2] = d->pin_dy[q];
        pin_net[q2] = e2;
        pin_of[q] = q2;
      }
    }
    ns32[c.E] = q2;
  }
  // degree > kNetLaneMax (stored last) get a wave each; the others fill
  // 64-lane groups, D = 2 << c lanes per net of class c
  std::vector<int32_t> large;
  {
    const int n_lane = c.wl_first[4] + c.wl_count[4];
    for (int e = n_lane; e < c.E; ++e) large.push_back(e);
    int g = 0;
    for (int k = 0; k < 5; ++k) {
      c.wl_gstart[k] = g;
      g += (c.wl_count[k] * (2 << k) + 63) / 64;
    }
    c.wl_gstart[5] = g;
    c.wl_groups = g;
  }
  c.num_large = (int)large.size();
  // ---- pin gradient entries: the lane groups' pins in lane order, then the
  // pins of the large nets (stored last, contiguous); node_pin_lane lists,
  // node by node in modulePins (file) order, the entry of each node pin ----
  const size_t nlanes = std::max<size_t>((size_t)c.wl_groups * 64, 1);
  std::vector<int32_t> lane_obj(nlanes, -1), lane_of_pin(std::max<int64_t>(c.P, 1), -1);
  std::vector<float2> lane_off(nlanes, make_float2(0.f, 0.f));
  for (int k = 0; k < 5; ++k) {
    const int D = 2 << k;
    for (int j = 0; j < c.wl_count[k]; ++j) {
      const int e = c.wl_first[k] + j;
      const size_t base = (size_t)c.wl_gstart[k] * 64 + (size_t)j * D;
      for (int32_t p = ns32[e]; p < ns32[e + 1]; ++p) {
        const size_t l = base + (p - ns32[e]);
        lane_obj[l] = pobj[p];
        lane_off[l] = make_float2(pdx[p], pdy[p]);
        lane_of_pin[p] = (int32_t)l;
      }
    }
  }
  {
    const int32_t p_large0 = ns32[c.wl_first[4] + c.wl_count[4]];
    c.large_lane_off = (int32_t)((int64_t)c.wl_groups * 64 - p_large0);
    for (int32_t p = p_large0; p < ns32[c.E]; ++p) lane_of_pin[p] = p + c.large_lane_off;
    c.grad_lanes = (int64_t)c.wl_groups * 64 + (ns32[c.E] - p_large0);
    // pins of one-pin nets (no gradient) read the extra entry, kept at zero
    for (int64_t p = 0; p < c.P; ++p)
      if (lane_of_pin[p] < 0) lane_of_pin[p] = (int32_t)c.grad_lanes;
  }
  std::vector<int32_t> nps(c.N + 1, 0), npl(std::max<int64_t>(c.P, 1), -1);
  for (int i = 0; i < c.N; ++i) nps[i + 1] = nps[i] + s.npins[i];
  c.h_grad_lane.assign(c.P, -1);
  {
    std::vector<int32_t> fill(nps.begin(), nps.end() - 1);
    for (int64_t q = 0; q < c.P; ++q)
      if (d->pin_obj[q] < c.N) {
        const int32_t ln = lane_of_pin[pin_of[q]];
        npl[fill[d->pin_obj[q]]++] = ln;
        c.h_grad_lane[q] = ln;
      }
  }
  c.h_w = s.w;
  c.h_h = s.h;
  c.h_kind = s.kind;
  // ---- device buffers ----
  std::vector<float2> pos((size_t)c.O + c.NT);
  for (int o = 0; o < c.O; ++o) pos[o] = make_float2(s.x[o], s.y[o]);
  for (int t = 0; t < c.NT; ++t) pos[(size_t)c.O + t] = make_float2(d->term_x[t], d->term_y[t]);
  bool ok = dev_upload(c, c.pos, pos.data(), pos.size()) && dev_upload(c, c.w, s.w.data(), c.O) && dev_upload(c, c.h, s.h.data(), c.O) &&
            dev_upload(c, c.area, s.area.data(), c.O) && dev_upload(c, c.kind, s.kind.data(), c.O) &&
            dev_zero(c, c.in_stage, c.O) && dev_upload(c, c.npins, s.npins.data(), c.O) &&
            dev_upload(c, c.big, big.data(), big.size()) && dev_upload(c, c.is_big, is_big.data(), c.O) &&
            dev_upload(c, c.net_start, ns32.data(), c.E + 1) && dev_upload(c, c.pin_obj, pobj.data(), c.P) &&
            dev_upload(c, c.pin_dx, pdx.data(), c.P) && dev_upload(c, c.pin_dy, pdy.data(), c.P) &&

            dev_upload(c, c.pin_net, pin_net.data(), pin_net.size()) &&
            dev_upload(c, c.lane_obj, lane_obj.data(), nlanes) && dev_upload(c, c.lane_off, lane_off.data(), nlanes) &&
            dev_upload(c, c.large_nets, large.data(), large.size()) &&
            dev_upload(c, c.node_pin_start, nps.data(), c.N + 1) &&
            dev_upload(c, c.node_pin_lane, npl.data(), npl.size()) &&
            dev_zero(c, c.pin_grad, (size_t)c.grad_lanes + 1) &&
            dev_zero(c, c.node_q, nb) && dev_zero(c, c.filler_q, nb) &&
            dev_upload(c, c.term_map, s.term_map.data(), nb) && dev_upload(c, c.base_map, s.base_map.data(), nb) &&
            dev_zero(c, c.gd, c.O) && dev_zero(c, c.gwl, c.O) && dev_zero(c, c.gtot, c.O) &&
            dev_zero(c, c.cur_main, c.O) && dev_zero(c, c.cur_ref, c.O) && dev_zero(c, c.last_ref, c.O) &&
            dev_zero(c, c.last_grad, c.O) && dev_zero(c, c.new_main, c.O) && dev_zero(c, c.new_ref, c.O) &&
            dev_zero(c, c.cur_grad, c.O) && dev_zero(c, c.sc, 1) && dev_zero(c, c.partials, kPartialTotal) &&
            dev_zero(c, c.tickets, 8);
  if (!ok) return EP_ERR_NOMEM;
  if (hipDeviceGetAttribute(&c.wall_khz, hipDeviceAttributeWallClockRate, c.device) != hipSuccess || c.wall_khz <= 0)
    c.wall_khz = 100000;
  std::string err;
  if (!poisson_create(c, err)) return c.err = err, EP_ERR_NOMEM;
  if (!tile_sort_create(c, err)) return c.err = err, EP_ERR_NOMEM;
  rebuild_stage(c);
  DevScalars h = {};
  h.a_k = 1;
  h.step = 1;
  h.max_iter = 1000;
  h.target_overflow = 0.1f;
  h.stage = c.stage;
  EP_HIP(hipMemcpy(c.sc, &h, sizeof(h), hipMemcpyHostToDevice));
  EP_HIP(hipDeviceSynchronize());
  return EP_OK;
}

void destroy(Ctx &c) {
  (void)hipSetDevice(c.device);
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  poisson_destroy(c);
  tile_sort_destroy(c);
  void *ptrs[] = {c.pos, c.w, c.h, c.area, c.kind, c.in_stage, c.npins, c.big, c.is_big,
                  c.net_start, c.pin_obj, c.pin_dx, c.pin_dy, c.pin_net, c.lane_obj, c.lane_off, c.large_nets, c.node_pin_start,
                  c.node_pin_lane, c.pin_grad, c.node_q, c.filler_q, c.term_map,
                  c.base_map, c.rho, c.phi, c.ex, c.ey, c.gd, c.gwl, c.gtot, c.cur_main, c.cur_ref, c.last_ref,
                  c.last_grad, c.new_main, c.new_ref, c.cur_grad, c.sc, c.partials, c.tickets, c.gemm_stamps.t};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (auto &e : c.ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c.prof_events) (void)hipEventDestroy(e);
  for (auto &g : c.eval_graphs) (void)hipGraphExecDestroy(g.exec);
  c.eval_graphs.clear();
  if (c.h_sc) (void)hipHostFree(c.h_sc);
  if (c.ev_fork) (void)hipEventDestroy(c.ev_fork);
  if (c.ev_join) (void)hipEventDestroy(c.ev_join);
  if (c.stream2) {
    (void)hipStreamSynchronize(c.stream2);
    (void)hipStreamDestroy(c.stream2);
  }
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete c.rng;
}

// The kernels of one evaluation of totalGradientUpdate.  The wiring kernels
// and the node sums of their pin gradients need only gamma (set by the
// density finalise), so they run on the second stream beside the Poisson
// solve and the field gather; the combine waits for them.
// A backtracking trial (next_lambda 1, or 2 with the controller's floor)
// also stages its updatePenaltyFactor and writes the gradient at that lambda
// to gtot_out with the positions to pos_out (see gather_combine_kernel).
static void enqueue_evaluation(Ctx &c, bool lambda_update, bool lipschitz, const float2 *ref_now,
                               const float2 *ref_prev, const float2 *grad_prev, int next_lambda, float2 *pos_out,
                               float2 *gtot_out, bool counted, hipEvent_t *ev) {
  if (ev) (void)hipEventRecord(ev[0], c.stream);
  launch_scatter(c, counted);
  if (ev) (void)hipEventRecord(ev[1], c.stream);
  launch_finalize_density(c);
  if (ev) (void)hipEventRecord(ev[2], c.stream);
  (void)hipEventRecord(c.ev_fork, c.stream);
  // The transform GEMMs go first: their blocks (8 waves, 64 KB LDS) take one
  // CU each before the wiring kernel's blocks fill the rest; launched after
  // it, they waited for wiring blocks to retire.
  launch_poisson(c, false, ev ? ev + kEvGemm : nullptr);
  if (ev) (void)hipEventRecord(ev[kEvMarker], c.stream);
  if (ev) (void)hipEventRecord(ev[3], c.stream);
  (void)hipStreamWaitEvent(c.stream2, c.ev_fork, 0);
  if (ev) (void)hipEventRecord(ev[6], c.stream2);
  launch_wirelength(c, true, c.stream2);
  if (ev) (void)hipEventRecord(ev[8], c.stream2);
  launch_node_wl_reduce(c, next_lambda ? 2 + next_lambda : lambda_update ? 1 : 0, c.stream2);
  if (ev) (void)hipEventRecord(ev[7], c.stream2);
  (void)hipEventRecord(c.ev_join, c.stream2);
  launch_field_gather(c);
  if (ev) (void)hipEventRecord(ev[4], c.stream);
  (void)hipStreamWaitEvent(c.stream, c.ev_join, 0);
  launch_combine(c, false, lipschitz, ref_now, ref_prev, grad_prev, pos_out, gtot_out, next_lambda != 0);
  if (ev) (void)hipEventRecord(ev[5], c.stream);
}

// The evaluation's graph for these arguments, captured on first use (the
// second stream joins the capture through the fork event and leaves it at
// the join).  nullptr when capture is not available.
static hipGraphExec_t eval_graph(Ctx &c, bool lambda_update, bool lipschitz, const float2 *ref_now,
                                 const float2 *ref_prev, const float2 *grad_prev, int next_lambda, float2 *pos_out,
                                 float2 *gtot_out, bool counted) {
  for (const Ctx::EvalGraph &g : c.eval_graphs)
    if (g.lambda_update == lambda_update && g.lipschitz == lipschitz && g.ref_now == ref_now &&
        g.ref_prev == ref_prev && g.grad_prev == grad_prev && g.next_lambda == next_lambda &&
        g.pos_out == pos_out && g.gtot_out == gtot_out && g.counted == counted)
      return g.exec;
  if (!c.graphs_ok) return nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    c.graphs_ok = false;
    return nullptr;
  }
  enqueue_evaluation(c, lambda_update, lipschitz, ref_now, ref_prev, grad_prev, next_lambda, pos_out, gtot_out,
                     counted, nullptr);
  hipError_t e = hipStreamEndCapture(c.stream, &graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    c.graphs_ok = false;
    (void)hipGetLastError();
    return nullptr;
  }
  c.eval_graphs.push_back({lambda_update, lipschitz, ref_now, ref_prev, grad_prev, next_lambda, pos_out, gtot_out,
                           counted, exec});
  return exec;
}

// One full evaluation of totalGradientUpdate at the current positions:
// replayed from its graph, or launched kernel by kernel when its phases are
// timed.
void evaluate(Ctx &c, bool lambda_update, bool lipschitz, const float2 *ref_now, const float2 *ref_prev,
              const float2 *grad_prev, bool timed, int next_lambda = 0, float2 *pos_out = nullptr,
              float2 *gtot_out = nullptr, bool counted = false) {
  hipEvent_t *ev = nullptr;
  if (timed) {
    ev = c.ev;
  } else if (c.profiling && c.prof_used + kEvPerEval <= c.prof_events.size()) {
    ev = &c.prof_events[c.prof_used];
    c.prof_used += kEvPerEval;
  }
  if (!ev) {
    hipGraphExec_t g = eval_graph(c, lambda_update, lipschitz, ref_now, ref_prev, grad_prev, next_lambda, pos_out,
                                  gtot_out, counted);
    if (g && hipGraphLaunch(g, c.stream) == hipSuccess) return;
  }
  // profiled: the device waits until the whole evaluation is enqueued, so the
  // phases run back to back as in a graph replay and the events time kernels,
  // not the host's submissions
  if (ev && !timed) launch_hold(c, c.prof_hold_ticks);
  enqueue_evaluation(c, lambda_update, lipschitz, ref_now, ref_prev, grad_prev, next_lambda, pos_out, gtot_out,
                     counted, ev);
}

// Phase times of one evaluation from its events: scatter, finalise, wiring
// kernels, node sums, Poisson, field gather, combine (the combine includes
// any wait for the second stream).
void phase_times(const hipEvent_t *ev, double *out) {
  static const int pairs[kPhases][2] = {{0, 1}, {1, 2}, {6, 8}, {8, 7}, {2, 3}, {3, 4}, {4, 5}};
  for (int k = 0; k < kPhases; ++k) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[pairs[k][0]], ev[pairs[k][1]]);
    out[k] = ms * 1e-3;
  }
  // the GEMM launches: each start/end pair less the cost of one event, timed
  // by the pair that follows the last GEMM's end event
  float marker = 0;
  (void)hipEventElapsedTime(&marker, ev[kEvGemm + 7], ev[kEvMarker]);
  out[kPhases] = 0;
  for (int g = 0; g < 4; ++g) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[kEvGemm + 2 * g], ev[kEvGemm + 2 * g + 1]);
    out[kPhases] += (ms - marker) * 1e-3;
  }
}

int sync(Ctx &c) {
  hipError_t e = hipStreamSynchronize(c.stream);
  if (e != hipSuccess) {
    c.err = std::string("kernel failure: ") + hipGetErrorString(e);
    return EP_ERR_HIP;
  }
  return EP_OK;
}

// The scalars once the stream reaches this point: the host polls the
// published sequence number; after 20 ms of polling it waits on the stream
// instead, which also reports a failed kernel.
int read_scalars(Ctx &c, DevScalars &h) {
  const unsigned int seq = ++c.pub_seq;
  launch_publish_scalars(c, seq);
  EP_HIP(hipGetLastError());
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned int n = 1; __atomic_load_n(&c.h_sc->seq, __ATOMIC_ACQUIRE) != seq; ++n) {
    if ((n & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
      int rc = sync(c);
      if (rc) return rc;
      if (__atomic_load_n(&c.h_sc->seq, __ATOMIC_ACQUIRE) != seq)
        return c.err = "scalars were not published", EP_ERR_HIP;
      break;
    }
  }
  h = c.h_sc->s;
  return EP_OK;
}

template <class F>
int update_scalars(Ctx &c, F f) {
  DevScalars h;
  int rc = read_scalars(c, h);
  if (rc) return rc;
  f(h);
  EP_HIP(hipMemcpy(c.sc, &h, sizeof(h), hipMemcpyHostToDevice));
  return EP_OK;
}

// HPWL (and the optional lambda update) at the current positions; the
// wiring kernels run without gradients, so the stored WA value is kept.
int hpwl_pass(Ctx &c, int lambda_update) {
  launch_wirelength(c, false);
  launch_reduce_wirelength(c, lambda_update, nullptr, false);
  return sync(c);
}

// Stage-ordered float pairs from an [O] float2 device array.
int gather_stage(Ctx &c, const float2 *d2, float *out, int64_t n) {
  if (n != (int64_t)c.h_stage_idx.size()) return c.err = "size does not match the stage", EP_ERR_ARG;
  int rc = sync(c);
  if (rc) return rc;
  std::vector<float> a((size_t)c.O * 2);
  EP_HIP(hipMemcpy(a.data(), d2, (size_t)c.O * sizeof(float2), hipMemcpyDeviceToHost));
  for (int64_t k = 0; k < n; ++k) {
    int o = c.h_stage_idx[k];
    out[2 * k] = a[2 * o];
    out[2 * k + 1] = a[2 * o + 1];
  }
  return EP_OK;
}

int set_stage_positions(Ctx &c, const float *xy, int64_t n) {
  if (n != (int64_t)c.h_stage_idx.size()) return c.err = "size does not match the stage", EP_ERR_ARG;
  if (n == 0) return EP_OK;
  float2 *dxy = nullptr;
  int32_t *didx = nullptr;
  EP_HIP(hipMalloc(&dxy, n * sizeof(float2)));
  EP_HIP(hipMalloc(&didx, n * sizeof(int32_t)));
  EP_HIP(hipMemcpy(dxy, xy, n * sizeof(float2), hipMemcpyHostToDevice));
  EP_HIP(hipMemcpy(didx, c.h_stage_idx.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
  launch_set_positions(c, dxy, didx, n);
  c.eval_valid = false;
  int rc = sync(c);
  (void)hipFree(dxy);
  (void)hipFree(didx);
  return rc;
}

// Device-resident EplaceNesterovOpt::opt (nesterov.hpp:207-223 + opt.hpp:29-36);
// method 2 is the backtracking method with the controller's first trial step
// and lambda multiplier.
int nesterov_run(Ctx &c, const ep_opt_params *op, ep_run_stats *out) {
  ep_opt_params o = {};
  o.method = 0;
  o.target_overflow = 0.10f;
  o.max_iterations = 1000;
  o.max_backtracks = 30;
  o.sync_every = 16;
  if (op) o = *op;
  if (o.sync_every <= 0) o.sync_every = 16;
  if (o.max_backtracks <= 0) o.max_backtracks = 30;
  if (o.log_capacity < 0 || (o.log_capacity > 0 && !o.log)) return c.err = "bad log buffer", EP_ERR_ARG;
  if (o.method < 0 || o.method > 3)
    return c.err = "method must be 0 (vanilla), 1 (backtracking), 2 (learned) or 3 (Barzilai-Borwein)", EP_ERR_ARG;
  const CtrlDev *ctrl = ctrl_of(o.controller);
  if (o.method == 2 && !ctrl) return c.err = "method 2 needs a controller", EP_ERR_ARG;
  if (ctrl && ctrl_device(ctrl) != c.device) return c.err = "controller is on another device", EP_ERR_ARG;
  double *dlog = nullptr;
  // frees the device log and ends phase profiling on every exit path
  struct RunScope {
    Ctx &c;
    double *&dlog;
    ~RunScope() {
      if (dlog) (void)hipFree(dlog);
      c.profiling = false;
    }
  } scope{c, dlog};
  int cap = o.log_capacity;
  if (cap > 0) {
    EP_HIP(hipMalloc(&dlog, (size_t)cap * EP_LOG_COLS * sizeof(double)));
    EP_HIP(hipMemset(dlog, 0, (size_t)cap * EP_LOG_COLS * sizeof(double)));
  }
  if (o.phase_seconds) {
    // events for every evaluation, created before the timed loop
    size_t need = (size_t)kEvPerEval * ((size_t)o.max_iterations + 2) * (o.method == 1 ? 3 : 1);
    need = std::min<size_t>(need, (size_t)kEvPerEval * 20000);
    while (c.prof_events.size() < need) {
      hipEvent_t e;
      EP_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
      c.prof_events.push_back(e);
    }
    c.prof_used = 0;
    c.profiling = true;
    // 0.4 ms of the wall clock (its rate in kHz) ahead of each timed evaluation
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c.device) != hipSuccess || khz <= 0)
      khz = 100000;
    c.prof_hold_ticks = (unsigned long long)khz * 4 / 10;
  }
  auto t0 = std::chrono::steady_clock::now();
  // init (nesterov.hpp:207-214): a_0 = 1, main = current positions
  int rc = update_scalars(c, [&](DevScalars &h) {
    h.iter = 0;
    h.a_k = 1;
    h.step = 1;
    h.done = 0;
    h.evals = 0;
    h.backtracks = 0;
    h.bad_numeric = 0;
    h.stage = c.stage;
    h.max_iter = o.max_iterations;
    h.target_overflow = o.target_overflow;
    h.lambda0 = h.lambda;
    h.hpwl0 = h.hpwl;
    h.step0 = 1;
    h.step_hist[0] = h.step_hist[1] = h.step_hist[2] = 1;
    h.tau_prev = h.overflow;
    h.hpwl_prev = h.hpwl;
    h.step_prop = h.step_used = 1;
    h.lambda_before = h.lambda;
    h.lam_floor = 0.95;
    h.ref_mult_ema = 0.04879016416943205;  // log 1.05: no stall seen yet
    h.backtracks_mark = h.trial = h.fallbacks = h.ctrl_active = 0;
    h.bb_short = 0;
    h.gemm_ticks = 0;
    h.gemm_launches = 0;
  });
  if (rc) return rc;
  // GEMM launches before the run are not the run's: all stamps zero reads as folded
  EP_HIP(hipMemsetAsync(c.gemm_stamps.t, 0, (size_t)4 * kGemmStamped * c.gemm_stamps.stride * sizeof(unsigned long long),
                        c.stream));
  launch_positions_to_ref(c, c.cur_main);
  // stop check before the first step uses tau of the last evaluation
  launch_done_check_ex(c, 0, nullptr, 0, 0);
  int k = 0;
  const bool bk = o.method == 1 || o.method == 2, learned = o.method == 2, bb = o.method == 3;
  const bool features = learned || dlog;  // controller inputs are logged for training
  while (true) {
    if (!bk) {
      int chunk = o.sync_every;
      for (int j = 0; j < chunk; ++j, ++k) {
        // evaluation at y_k: iteration 0 reuses the caller's evaluation at
        // these positions (only lambda changed since), later ones recompute
        // (from the second iteration on, the previous move counted the keys)
        if (k == 0 && c.eval_valid) launch_gather_combine(c, false, true, nullptr, c.last_ref, c.last_grad);
        else evaluate(c, k > 0, true, nullptr, c.last_ref, c.last_grad, false, 0, nullptr, nullptr, k > 0);
        launch_step_scalar(c, bb ? 1 : 0);
        launch_nesterov_move_hist(c, nullptr, c.gtot, c.cur_main, c.cur_main, nullptr, c.last_ref, c.last_grad);
        launch_done_check_ex(c, k + 1, dlog, cap, 1);
      }
    } else {
      // combine-only re-evaluation at y_k with the updated lambda (the
      // positions were evaluated by the accepted trial or the caller)
      // (y_k and g_k are also copied for the trials: cur_ref, cur_grad)
      // From the second iteration on, the accepted trial's combine already
      // computed g_k at the updated lambda, y_k, and the Lipschitz sums of
      // that pass (kPartialLipNext).
      if (k == 0 && !c.eval_valid) {
        evaluate(c, false, true, nullptr, c.last_ref, c.last_grad, false);
        launch_positions_to_ref(c, c.cur_ref);
        launch_copy_float2(c, c.cur_grad, c.gtot, c.O);
        launch_step_scalar(c, 0);
      } else if (k == 0) {
        launch_gather_combine(c, false, true, nullptr, c.last_ref, c.last_grad, c.cur_ref, c.cur_grad);
        launch_step_scalar(c, 0);
      }
      // from the second iteration on, the step came with the stop check
      // at the end of the previous iteration (launch_done_step)
      if (features) launch_ctrl_step(c, ctrl, learned ? 1 : 0);
      int tries = 0;
      bool stopped = false;
      while (true) {
        // every trial's move is followed by its evaluation, so the move also
        // counts the keys of the scatter (the sort's totals are zero again
        // only after that evaluation's scatter)
        launch_nesterov_move_hist(c, c.cur_ref, c.cur_grad, c.cur_main, c.new_main, c.new_ref, nullptr, nullptr);
        // y_k, g_k of the previous iteration (last_ref, last_grad) are free:
        // the trial's positions and its gradient at the staged lambda go there
        evaluate(c, false, true, c.new_ref, c.cur_ref, c.cur_grad, false, learned ? 2 : 1, c.last_ref, c.last_grad,
                 true);
        launch_accept(c);
        DevScalars h;
        if ((rc = read_scalars(c, h))) return rc;
        ++tries;
        // Only the stop check at the end of an iteration sets done, and the
        // iteration is enqueued without waiting for it: done seen here means
        // the previous iteration ended the run and this one's kernels
        // returned without work.
        if (h.done) {
          stopped = true;
          // that evaluation's phase events time kernels that did nothing
          if (c.profiling && c.prof_used >= (size_t)kEvPerEval) c.prof_used -= kEvPerEval;
          break;
        }
        if (h.accepted || tries >= o.max_backtracks) break;
      }
      if (stopped) break;
      // cur_ref, cur_grad := the last trial's y and g at the staged lambda;
      // last_ref, last_grad := y_k, g_k
      std::swap(c.last_ref, c.cur_ref);
      std::swap(c.last_grad, c.cur_grad);
      std::swap(c.cur_main, c.new_main);
      // updatePenaltyFactor with the HPWL of the last trial (or the
      // controller's floor), as staged by that trial, then the stop check
      // and the next iteration's step from that trial's Lipschitz sums
      launch_done_step(c, k + 1, dlog, cap, kPartialLipNext);
      ++k;
      continue;
    }
    DevScalars h;
    if ((rc = read_scalars(c, h))) {
      if (!bk) launch_clear_key_totals(c);
      return rc;
    }
    if (h.done) break;
  }
  // the vanilla run's last move counted keys for an evaluation that does not
  // follow: the sort's totals go back to zero
  if (!bk && k > 0) launch_clear_key_totals(c);
  // lambda update pending after the last move (vanilla evaluates it lazily)
  DevScalars h;
  if ((rc = read_scalars(c, h))) return rc;
  int iters = h.iter;
  h.done = 0;
  EP_HIP(hipMemcpy(c.sc, &h, sizeof(h), hipMemcpyHostToDevice));
  if (!bk && iters > 0)
    if ((rc = hpwl_pass(c, 1))) return rc;
  double final_tau = h.overflow;
  // wrap_up: positions := main solution (clamped by setPosition)
  {
    std::vector<float> m((size_t)c.O * 2), sxy;
    EP_HIP(hipMemcpy(m.data(), c.cur_main, (size_t)c.O * 8, hipMemcpyDeviceToHost));
    for (int32_t idx : c.h_stage_idx) {
      sxy.push_back(m[2 * idx]);
      sxy.push_back(m[2 * idx + 1]);
    }
    if ((rc = set_stage_positions(c, sxy.data(), (int64_t)c.h_stage_idx.size()))) return rc;
  }
  // final HPWL at the written-back positions (does not touch lambda)
  {
    if ((rc = hpwl_pass(c, 0))) return rc;
    DevScalars h3;
    if ((rc = read_scalars(c, h3))) return rc;
    if (out) out->final_hpwl = h3.hpwl;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (c.profiling) {
    c.profiling = false;
    double acc[kPhasesTimed] = {};
    size_t n = c.prof_used / kEvPerEval;
    for (size_t e = 0; e < n; ++e) {
      double t[kPhasesTimed];
      ep::phase_times(&c.prof_events[kEvPerEval * e], t);
      for (int k = 0; k < kPhasesTimed; ++k) acc[k] += t[k];
    }
    for (int k = 0; k < kPhasesTimed; ++k) o.phase_seconds[k] = acc[k];
    o.phase_seconds[kPhasesTimed] = (double)n;
  }
  if (cap > 0) EP_HIP(hipMemcpy(o.log, dlog, (size_t)cap * EP_LOG_COLS * sizeof(double), hipMemcpyDeviceToHost));
  if (out) {
    out->iterations = iters;
    out->evaluations = h.evals;
    out->backtracks = h.backtracks;
    out->stopped = 1;
    out->final_overflow = final_tau;
    out->seconds = std::chrono::duration<double>(t1 - t0).count();
    out->controller_fallbacks = h.fallbacks;
    out->bb_short_steps = h.bb_short;
    out->gemm_seconds = (double)h.gemm_ticks / (1e3 * c.wall_khz);
    out->gemm_launches = h.gemm_launches;
    out->numeric_fallbacks = h.bad_numeric;
  }
  return EP_OK;
}

}  // namespace
}  // namespace ep

using ep::Ctx;

extern "C" {

int ep_create(const ep_design *design, const ep_params *params, ep_ctx **out) {
  if (!out) return EP_ERR_ARG;
  *out = nullptr;
  ep_ctx *x = new ep_ctx();
  int rc = ep::create(x->c, design, params);
  *out = x;
  return rc;
}

int ep_destroy(ep_ctx *x) {
  if (!x) return EP_ERR_ARG;
  ep::destroy(x->c);
  delete x;
  return EP_OK;
}

const char *ep_last_error(const ep_ctx *x) { return x ? x->c.err.c_str() : "null context"; }

int ep_get_info(ep_ctx *x, ep_info *i) {
  if (!x || !i) return EP_ERR_ARG;
  Ctx &c = x->c;
  *i = ep_info{};
  i->num_nodes = c.N;
  i->num_fillers = c.F;
  i->num_objects = c.O;
  i->num_terminals = c.NT;
  i->num_nets = c.E;
  i->num_pins = c.P;
  i->bins_x = c.nx;
  i->bins_y = c.ny;
  i->fixed_shift = c.shift;
  i->num_big_objects = c.num_big;
  i->num_large_nets = c.num_large;
  i->bin_step_x = c.bsx;
  i->bin_step_y = c.bsy;
  i->core_llx = c.core_llx;
  i->core_lly = c.core_lly;
  i->core_urx = c.core_urx;
  i->core_ury = c.core_ury;
  i->filler_w = c.filler_w;
  i->filler_h = c.filler_h;
  i->std_cell_area = c.std_area;
  i->macro_area = c.macro_area;
  i->total_row_area = c.total_row_area;
  return EP_OK;
}

int ep_total_gradient_update(ep_ctx *x) {
  if (!x) return EP_ERR_ARG;
  Ctx &c = x->c;
  ep::evaluate(c, false, false, nullptr, nullptr, nullptr, true);
  int rc = ep::sync(c);
  if (rc) return rc;
  c.eval_valid = true;
  ep::phase_times(c.ev, c.last_timings);
  return EP_OK;
}

int64_t ep_stage_size(ep_ctx *x) { return x ? (int64_t)x->c.h_stage_idx.size() : -1; }

int ep_get_position(ep_ctx *x, float *xy, int64_t n) {
  if (!x || (!xy && n)) return EP_ERR_ARG;
  return ep::gather_stage(x->c, x->c.pos, xy, n);
}

int ep_set_position(ep_ctx *x, const float *xy, int64_t n) {
  if (!x || (!xy && n)) return EP_ERR_ARG;
  return ep::set_stage_positions(x->c, xy, n);
}

int ep_get_gradient(ep_ctx *x, float *xy, int64_t n) {
  if (!x || (!xy && n)) return EP_ERR_ARG;
  return ep::gather_stage(x->c, x->c.gtot, xy, n);
}

int ep_penalty_factor_init(ep_ctx *x) {
  if (!x) return EP_ERR_ARG;
  Ctx &c = x->c;
  int rc = ep::hpwl_pass(c, 0);
  if (rc) return rc;
  ep::launch_lambda_init(c);
  return ep::sync(c);
}

int ep_update_penalty_factor(ep_ctx *x) {
  if (!x) return EP_ERR_ARG;
  return ep::hpwl_pass(x->c, 1);
}

int ep_switch_to_filler_only(ep_ctx *x) {
  if (!x) return EP_ERR_ARG;
  Ctx &c = x->c;
  int rc = ep::sync(c);
  if (rc) return rc;
  std::vector<float2> p(c.O);
  EP_HIP(hipMemcpy(p.data(), c.pos, (size_t)c.O * sizeof(float2), hipMemcpyDeviceToHost));
  for (int o = c.N; o < c.O; ++o) ep::random_filler(c, c.h_w[o], c.h_h[o], p[o].x, p[o].y);
  EP_HIP(hipMemcpy(c.pos, p.data(), (size_t)c.O * sizeof(float2), hipMemcpyHostToDevice));
  c.eval_valid = false;
  c.stage = EP_STAGE_FILLERONLY;
  ep::rebuild_stage(c);
  return ep::update_scalars(c, [&](ep::DevScalars &h) { h.stage = c.stage; });
}

int ep_macro_legalize(ep_ctx *x, const ep_design *d, const ep_mlg_params *p, ep_mlg_stats *st) {
  if (!x || !d || !p) return EP_ERR_ARG;
  Ctx &c = x->c;
  if (d->num_nodes != c.N) return c.err = "design does not match the placer", EP_ERR_ARG;
  if (!(p->target_density > 0)) return c.err = "target_density must be > 0", EP_ERR_ARG;
  int rc = ep::sync(c);
  if (rc) return rc;
  std::vector<float2> pos(c.N);
  EP_HIP(hipMemcpy(pos.data(), c.pos, (size_t)c.N * sizeof(float2), hipMemcpyDeviceToHost));
  std::vector<float> xy(2 * (size_t)c.N);
  for (int i = 0; i < c.N; ++i) xy[2 * i] = pos[i].x, xy[2 * i + 1] = pos[i].y;
  ep_mlg_stats s{};
  rc = ep::macro_legalize_sa(*d, xy, *p, *c.rng, s);
  if (rc) return c.err = "macro legalisation failed", rc;
  for (int i = 0; i < c.N; ++i) pos[i] = make_float2(xy[2 * i], xy[2 * i + 1]);
  EP_HIP(hipMemcpy(c.pos, pos.data(), (size_t)c.N * sizeof(float2), hipMemcpyHostToDevice));
  c.eval_valid = false;
  if (st) *st = s;
  return EP_OK;
}

int ep_switch_to_cgp(ep_ctx *x, int32_t mgp_iterations) {
  if (!x) return EP_ERR_ARG;
  Ctx &c = x->c;
  c.stage = EP_STAGE_CGP;
  ep::rebuild_stage(c);
  return ep::update_scalars(c, [&](ep::DevScalars &h) {
    h.lambda = h.lambda / std::pow(1.1, mgp_iterations * 0.1);
    h.stage = c.stage;
  });
}

int ep_get_stats(ep_ctx *x, ep_stats *s) {
  if (!x || !s) return EP_ERR_ARG;
  ep::DevScalars h;
  int rc = ep::read_scalars(x->c, h);
  if (rc) return rc;
  s->overflow = h.overflow;
  s->lambda = h.lambda;
  s->last_hpwl = h.last_hpwl;
  s->hpwl = h.hpwl;
  s->inv_gamma_x = h.inv_gamma_x;
  s->inv_gamma_y = h.inv_gamma_y;
  s->wirelength = h.wl_value;
  s->stage = x->c.stage;
  s->mgp_iterations = h.iter;
  return EP_OK;
}

int ep_hpwl(ep_ctx *x, double *out) {
  if (!x || !out) return EP_ERR_ARG;
  Ctx &c = x->c;
  int rc = ep::hpwl_pass(c, 0);
  if (rc) return rc;
  ep::DevScalars h;
  if ((rc = ep::read_scalars(c, h))) return rc;
  *out = h.hpwl;
  return EP_OK;
}

int ep_nesterov_run(ep_ctx *x, const ep_opt_params *opt, ep_run_stats *out) {
  if (!x) return EP_ERR_ARG;
  return ep::nesterov_run(x->c, opt, out);
}

int ep_poisson(ep_ctx *x, const float *rho, float *phi, float *ex, float *ey) {
  if (!x || !rho || !ex || !ey) return EP_ERR_ARG;
  Ctx &c = x->c;
  int ld = ep::poisson_ld(c);
  std::vector<float> buf((size_t)c.nx * ld, 0.f);
  for (int i = 0; i < c.nx; ++i)
    std::memcpy(&buf[(size_t)i * ld], rho + (size_t)i * c.ny, (size_t)c.ny * 4);
  int rc = ep::sync(c);
  if (rc) return rc;
  double mean = 0, mx = 0;
  for (size_t k = 0; k < (size_t)c.nx * c.ny; ++k) {
    mean += rho[k];
    mx = std::max(mx, std::fabs((double)rho[k]));
  }
  mean /= (double)c.nx * c.ny;
  if ((rc = ep::update_scalars(c, [&](ep::DevScalars &h) {
         h.rho_mean = mean;
         h.rho_max = mx;
       })))
    return rc;
  EP_HIP(hipMemcpy(c.rho, buf.data(), buf.size() * 4, hipMemcpyHostToDevice));
  ep::launch_poisson(c, phi != nullptr);
  if ((rc = ep::sync(c))) return rc;
  // this solve's launches belong to no run
  EP_HIP(hipMemset(c.gemm_stamps.t, 0,
                   (size_t)4 * ep::kGemmStamped * c.gemm_stamps.stride * sizeof(unsigned long long)));
  auto down = [&](float *dst, const float *src) -> int {
    EP_HIP(hipMemcpy(buf.data(), src, buf.size() * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < c.nx; ++i) std::memcpy(dst + (size_t)i * c.ny, &buf[(size_t)i * ld], (size_t)c.ny * 4);
    return EP_OK;
  };
  if ((rc = down(ex, c.ex))) return rc;
  if ((rc = down(ey, c.ey))) return rc;
  if (phi && (rc = down(phi, c.phi))) return rc;
  return EP_OK;
}

int ep_read(ep_ctx *x, int32_t which, void *host, int64_t count) {
  if (!x || !host || count < 0) return EP_ERR_ARG;
  Ctx &c = x->c;
  int rc = ep::sync(c);
  if (rc) return rc;
  size_t nb = (size_t)c.nx * c.ny;
  int ld = ep::poisson_ld(c);
  auto map_down = [&](const float *src) -> int {
    if (count != (int64_t)nb) return c.err = "count mismatch", EP_ERR_ARG;
    std::vector<float> buf((size_t)c.nx * ld);
    EP_HIP(hipMemcpy(buf.data(), src, buf.size() * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < c.nx; ++i) std::memcpy((float *)host + (size_t)i * c.ny, &buf[(size_t)i * ld], c.ny * 4);
    return EP_OK;
  };
  auto plain = [&](const void *src, int64_t n, size_t es) -> int {
    if (count != n) return c.err = "count mismatch", EP_ERR_ARG;
    EP_HIP(hipMemcpy(host, src, (size_t)n * es, hipMemcpyDeviceToHost));
    return EP_OK;
  };
  switch (which) {
    case EP_ARR_POSITIONS: return plain(c.pos, 2 * (int64_t)c.O, 4);
    case EP_ARR_NODE_Q: return plain(c.node_q, nb, 8);
    case EP_ARR_FILLER_Q: return plain(c.filler_q, nb, 8);
    case EP_ARR_RHO: return map_down(c.rho);
    case EP_ARR_PHI: return map_down(c.phi);
    case EP_ARR_EX: return map_down(c.ex);
    case EP_ARR_EY: return map_down(c.ey);
    case EP_ARR_GRAD_DENSITY: return plain(c.gd, 2 * (int64_t)c.O, 4);
    case EP_ARR_GRAD_WL: {
      if (count != 2 * (int64_t)c.N) return c.err = "count mismatch", EP_ERR_ARG;
      EP_HIP(hipMemcpy(host, c.gwl, (size_t)c.N * 8, hipMemcpyDeviceToHost));
      return EP_OK;
    }
    case EP_ARR_GRAD_TOTAL: return plain(c.gtot, 2 * (int64_t)c.O, 4);
    case EP_ARR_TERM_MAP: return plain(c.term_map, nb, 4);
    case EP_ARR_BASE_MAP: return plain(c.base_map, nb, 4);
    case EP_ARR_PIN_GRAD: {  // back to design pin order; terminal pins read 0
      if (count != 2 * c.P) return c.err = "count mismatch", EP_ERR_ARG;
      std::vector<float2> g((size_t)c.grad_lanes + 1);
      EP_HIP(hipMemcpy(g.data(), c.pin_grad, g.size() * sizeof(float2), hipMemcpyDeviceToHost));
      for (int64_t q = 0; q < c.P; ++q)
        ((float2 *)host)[q] = c.h_grad_lane[q] >= 0 ? g[c.h_grad_lane[q]] : make_float2(0.f, 0.f);
      return EP_OK;
    }
    case EP_ARR_MAIN: return plain(c.cur_main, 2 * (int64_t)c.O, 4);
  }
  c.err = "unknown array";
  return EP_ERR_ARG;
}

int ep_write(ep_ctx *x, int32_t which, const void *host, int64_t count) {
  if (!x || !host || count < 0) return EP_ERR_ARG;
  Ctx &c = x->c;
  int rc = ep::sync(c);
  if (rc) return rc;
  size_t nb = (size_t)c.nx * c.ny;
  int ld = ep::poisson_ld(c);
  auto map_up = [&](float *dst) -> int {
    if (count != (int64_t)nb) return c.err = "count mismatch", EP_ERR_ARG;
    std::vector<float> buf((size_t)c.nx * ld, 0.f);
    for (int i = 0; i < c.nx; ++i) std::memcpy(&buf[(size_t)i * ld], (const float *)host + (size_t)i * c.ny, c.ny * 4);
    EP_HIP(hipMemcpy(dst, buf.data(), buf.size() * 4, hipMemcpyHostToDevice));
    return EP_OK;
  };
  switch (which) {
    case EP_ARR_POSITIONS: {
      if (count != 2 * (int64_t)c.O) return c.err = "count mismatch", EP_ERR_ARG;
      EP_HIP(hipMemcpy(c.pos, host, (size_t)c.O * sizeof(float2), hipMemcpyHostToDevice));
      c.eval_valid = false;
      return EP_OK;
    }
    case EP_ARR_EX: return map_up(c.ex);
    case EP_ARR_EY: return map_up(c.ey);
    case EP_ARR_RHO: return map_up(c.rho);
  }
  c.err = "array not writable";
  return EP_ERR_ARG;
}

int ep_last_timings(ep_ctx *x, double *s) {
  if (!x || !s) return EP_ERR_ARG;
  for (int k = 0; k < ep::kPhases; ++k) s[k] = x->c.last_timings[k];
  return EP_OK;
}

}  // extern "C"
