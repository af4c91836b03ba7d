// Macro legalisation: the simulated annealing of SAMacroLegalizer
// (Legalization/legalizer.cpp:701-1530, ePlace-MS eq. 30) on the host.
//
// The annealing is one sequential Markov chain over a few macros: each step
// draws from the process-wide rand() stream, moves one macro and accepts or
// rolls back.  It is restated here with the reference's float / double / int
// conversions, its rand() draw order and its quirks (integer truncation of
// the centre before a move, int overlap sums, the union area returned as
// int), so that from the same positions and rand() state the result is the
// reference's, bit for bit.  Departures, both on inputs where the reference
// reads out of bounds: a macro index of exactly dbMacros.size() (rand() ==
// RAND_MAX) is clamped to the last macro, and bin ranges of objects outside
// the core are clamped to the grid.
//
// Cost pieces per macro:
//   HPWL         PlaceDB::calcModuleHPWLfast (placedb.cpp:537-582)
//   cell cover   bins overlapped x frozen non-macro density (legalizer.cpp:1296-1345)
//   overlap      pairwise overlap with the other macros, summed into an int
//                (legalizer.cpp:1355-1367)
// and the global overlap check is total macro area - union area of the
// integer-truncated macro rectangles (legalizer.cpp:1243-1260).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ep_internal.h"

namespace ep {
namespace {

constexpr double kMlgEps = 1.0e-15;  // global.h:39
constexpr int kRandMax = 2147483647;
// RAND_MAX in the reference's float expressions converts to 2^31.
constexpr float kRandMaxF = 2147483648.0f;

// getOverlap / getOverlapArea (global.h:477-518): double, the two negative
// lengths multiplied.
inline double ov_len(double x1, double x2, double x3, double x4) {
  double s = std::max(x1, x3), e = std::min(x2, x4);
  return s >= e ? 0.0 : s - e;
}
inline double ov_area(float l1, float b1, float r1, float t1, float l2, float b2, float r2, float t2) {
  double h = ov_len(l1, r1, l2, r2);
  if (h == 0) return 0;
  double v = ov_len(b1, t1, b2, t2);
  if (v == 0) return 0;
  return h * v;
}

// Union area of integer rectangles (RectangleAreaSolution::rectangleArea,
// legalizer.h:441-500): exact, returned through int as the reference does.
int union_area(const std::vector<int> &rl, const std::vector<int> &rb, const std::vector<int> &rr,
               const std::vector<int> &rt) {
  const size_t n = rl.size();
  std::vector<int> ys;
  ys.reserve(2 * n);
  for (size_t i = 0; i < n; ++i) ys.push_back(rb[i]), ys.push_back(rt[i]);
  std::sort(ys.begin(), ys.end());
  ys.erase(std::unique(ys.begin(), ys.end()), ys.end());
  if (ys.size() < 2) return 0;
  struct Ev {
    int x, idx, diff;
  };
  std::vector<Ev> ev;
  for (size_t i = 0; i < n; ++i) ev.push_back({rl[i], (int)i, 1}), ev.push_back({rr[i], (int)i, -1});
  std::sort(ev.begin(), ev.end(), [](const Ev &a, const Ev &b) {
    if (a.x != b.x) return a.x < b.x;
    if (a.idx != b.idx) return a.idx < b.idx;
    return a.diff < b.diff;
  });
  std::vector<int> cover(ys.size() - 1, 0);
  long long ans = 0;
  for (size_t i = 0; i < ev.size();) {
    size_t j = i;
    while (j + 1 < ev.size() && ev[j + 1].x == ev[i].x) ++j;
    if (j + 1 == ev.size()) break;
    for (size_t k = i; k <= j; ++k) {
      const int idx = ev[k].idx;
      const size_t lo = std::lower_bound(ys.begin(), ys.end(), rb[idx]) - ys.begin();
      const size_t hi = std::lower_bound(ys.begin(), ys.end(), rt[idx]) - ys.begin();
      for (size_t s = lo; s < hi; ++s) cover[s] += ev[k].diff;
    }
    long long len = 0;
    for (size_t s = 0; s + 1 < ys.size(); ++s)
      if (cover[s] > 0) len += ys[s + 1] - ys[s];
    ans += len * (long long)(ev[j + 1].x - ev[j].x);
    i = j + 1;
  }
  return (int)ans;
}

struct MacroSA {
  const ep_design &d;
  const ep_mlg_params &p;
  Rand &rng;
  int N = 0, T = 0;
  float cllx = 0, clly = 0, curx = 0, cury = 0;  // coreRegion (placedb.cpp:52-77)
  double rowh = 1;                              // PlaceDB::commonRowHeight
  std::vector<float> cx, cy, llx, lly;          // Module::center / coor of the nodes
  std::vector<float> tllx, tlly;                // terminals' coor
  std::vector<float> px, py;                    // Pin::absolutePos
  std::vector<int32_t> npin_start, npin;        // Module::modulePins, in pin order
  std::vector<int32_t> npin_net;                // net of each of those pins
  std::vector<int> macros;
  double totalMacroArea = 0, totalCellArea = 0;
  int bdx = 0, bdy = 0;
  float bsx = 0, bsy = 0, barea = 0;
  std::vector<float> bllx, blly, burx, bury, dens;  // dens[i * bdy + j] = nonMacroDensity
  double totalHPWL = 0, totalCellAreaCovered = 0;
  int totalMacroOverlap = 0;
  bool overlapFree = false;
  float miuD = 0, miuO = 0, beta = 0, temp = 0, tempCoef = 0;
  float rx = 0, ry = 0, ux = 0, uy = 0, stepx = 0, stepy = 0;
  int jLimit = 1000, kLimit = 1000;
  int64_t perturbations = 0, accepted = 0;

  MacroSA(const ep_design &d_, const ep_mlg_params &p_, Rand &r) : d(d_), p(p_), rng(r) {}

  float term_cx(int t) const { return d.term_x[t]; }
  float term_cy(int t) const { return d.term_y[t]; }

  // Module::setCenter_2D + Pin::calculateAbsolutePos (objects.cpp:454-539).
  void set_center_raw(int m, float x, float y) {
    cx[m] = x;
    cy[m] = y;
    llx[m] = x - (float)0.5 * d.node_w[m];
    lly[m] = y - (float)0.5 * d.node_h[m];
    for (int k = npin_start[m]; k < npin_start[m + 1]; ++k) {
      const int q = npin[k];
      px[q] = cx[m] + d.pin_dx[q];
      py[q] = cy[m] + d.pin_dy[q];
    }
  }
  // PlaceDB::setModuleCenter_2D (placedb.cpp:253-290), movable modules.
  void set_center(int m, float x, float y) {
    const float w = d.node_w[m], h = d.node_h[m];
    if (x - 0.5 * w < cllx) x = (float)(cllx + 0.5 * w + kMlgEps);
    if (x + 0.5 * w > curx) x = (float)(curx - 0.5 * w - kMlgEps);
    if (y - 0.5 * h < clly) y = (float)(clly + 0.5 * h + kMlgEps);
    if (y + 0.5 * h > cury) y = (float)(cury - 0.5 * h - kMlgEps);
    set_center_raw(m, x, y);
  }
  // PlaceDB::setModuleLocation_2D (placedb.cpp:203-235) + Module::setLocation_2D.
  void set_location(int m, float x, float y) {
    const float w = d.node_w[m], h = d.node_h[m];
    if (x < cllx) x = cllx;
    if (x + w > curx) x = (float)(curx - w - kMlgEps);
    if (y < clly) y = clly;
    if (y + h > cury) y = (float)(cury - h - kMlgEps);
    llx[m] = x;
    lly[m] = y;
    cx[m] = llx[m] + (float)0.5 * w;
    cy[m] = lly[m] + (float)0.5 * h;
    for (int k = npin_start[m]; k < npin_start[m + 1]; ++k) {
      const int q = npin[k];
      px[q] = cx[m] + d.pin_dx[q];
      py[q] = cy[m] + d.pin_dy[q];
    }
  }

  void setup(const float *xy) {
    N = d.num_nodes;
    T = d.num_terminals;
    const int R = d.num_rows;
    cllx = (float)d.row_x0[0];
    curx = (float)d.row_x1[0];
    for (int r = 0; r < R; ++r) cllx = std::min(cllx, (float)d.row_x0[r]), curx = std::max(curx, (float)d.row_x1[r]);
    clly = (float)d.row_bottom[0];
    cury = (float)(d.row_bottom[R - 1] + d.row_height[R - 1]);
    rowh = d.common_row_height;
    const int64_t P = d.num_pins;
    npin_start.assign(N + 1, 0);
    for (int64_t q = 0; q < P; ++q)
      if (d.pin_obj[q] < N) ++npin_start[d.pin_obj[q] + 1];
    for (int i = 0; i < N; ++i) npin_start[i + 1] += npin_start[i];
    npin.resize(npin_start[N]);
    std::vector<int32_t> cur(npin_start.begin(), npin_start.end() - 1);
    npin_net.resize(npin_start[N]);
    for (int e = 0; e < d.num_nets; ++e)
      for (int64_t q = d.net_start[e]; q < d.net_start[e + 1]; ++q)
        if (d.pin_obj[q] < N) {
          const int k = cur[d.pin_obj[q]]++;
          npin[k] = (int32_t)q;
          npin_net[k] = e;
        }
    cx.resize(N), cy.resize(N), llx.resize(N), lly.resize(N);
    px.resize(P), py.resize(P);
    for (int64_t q = 0; q < P; ++q) {
      const int o = d.pin_obj[q];
      if (o >= N) px[q] = term_cx(o - N) + d.pin_dx[q], py[q] = term_cy(o - N) + d.pin_dy[q];
    }
    for (int i = 0; i < N; ++i) set_center_raw(i, xy[2 * i], xy[2 * i + 1]);
    tllx.resize(T), tlly.resize(T);
    for (int t = 0; t < T; ++t) {
      tllx[t] = d.term_x[t] - (float)0.5 * d.term_w[t];
      tlly[t] = d.term_y[t] - (float)0.5 * d.term_h[t];
    }
  }

  bool is_macro(int i) const {  // PlaceDB::addNode (placedb.cpp:102-111)
    return (double)(d.node_h[i] - (float)d.common_row_height) > kMlgEps;
  }

  // initializeMacros (legalizer.cpp:840-870)
  void init_macros() {
    for (int i = 0; i < N; ++i) {
      const float area = d.node_w[i] * d.node_h[i];  // float_mul (objects.h:168)
      if (is_macro(i)) {
        macros.push_back(i);
        totalMacroArea += area;
        float lx = (float)(int)(llx[i] + 0.5);
        float ly = (float)(((int)((lly[i] + 0.5 * rowh - clly) / (double)rowh)) * rowh + (int)(clly + 0.5));
        set_location(i, lx, ly);
      } else {
        totalCellArea += area;
      }
    }
  }

  void bin_range(float l, float b, float r, float t, int &i0, int &i1, int &j0, int &j1) const {
    i0 = (int)((l - cllx) / bsx);
    i1 = (int)((r - cllx) / bsx);
    j0 = (int)((b - clly) / bsy);
    j1 = (int)((t - clly) / bsy);
    i0 = std::max(i0, 0), j0 = std::max(j0, 0);
    if (i1 >= bdx) i1 = bdx - 1;
    if (j1 >= bdy) j1 = bdy - 1;
  }

  // initializeBins (legalizer.cpp:906-1186)
  void init_bins() {
    const int nodeCount = N;
    const float nodeArea = (float)(totalCellArea + totalMacroArea);
    const float cw = curx - cllx, ch = cury - clly;
    const float coreArea = cw * ch;
    const float averageNodeArea = (float)(1.0 * nodeArea / nodeCount);
    const float idealBinArea = averageNodeArea / p.target_density;
    const int idealBinCount = (int)(1.0 * (coreArea / idealBinArea) + 0.5f);
    bool upd = false;
    for (int i = 1; i < 10; i++) {
      if ((2 << i) * (2 << i) <= idealBinCount && (2 << (i + 1)) * (2 << (i + 1)) > idealBinCount) {
        bdx = bdy = 2 << i;
        upd = true;
        break;
      }
    }
    if (!upd) bdx = bdy = 1024;
    bsx = cw / (float)bdx;
    bsy = ch / (float)bdy;
    bllx.resize(bdx), burx.resize(bdx), blly.resize(bdy), bury.resize(bdy);
    for (int i = 0; i < bdx; ++i) bllx[i] = i * bsx + cllx, burx[i] = bllx[i] + bsx;
    for (int j = 0; j < bdy; ++j) blly[j] = j * bsy + clly, bury[j] = blly[j] + bsy;
    barea = bsx * bsy;
    const size_t nb = (size_t)bdx * bdy;
    std::vector<float> cell(nb, 0.f), term(nb, 0.f), base(nb, 0.f);
    const float td = p.target_density;
    // terminals entirely inside the core
    for (int t = 0; t < T; ++t) {
      const float l = tllx[t], b = tlly[t], r = l + d.term_w[t], tp = b + d.term_h[t];
      if (l < cllx || l + d.term_w[t] > curx) continue;
      if (b < clly || b + d.term_h[t] > cury) continue;
      int i0, i1, j0, j1;
      bin_range(l, b, r, tp, i0, i1, j0, j1);
      for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) {
          float &a = term[(size_t)i * bdy + j];
          a = (float)(a + td * ov_area(bllx[i], blly[j], burx[i], bury[j], l, b, r, tp));
        }
    }
    // base: bin area outside the rows.  Rows without overlap add +0 to the
    // float sum, so only the rows meeting the bin's y range are visited.
    const int R = d.num_rows;
    std::vector<float> rl(R), rb(R), rr(R), rt(R);
    for (int r = 0; r < R; ++r) {
      rl[r] = (float)d.row_x0[r];
      rb[r] = (float)d.row_bottom[r];
      rr[r] = (float)d.row_x1[r];
      rt[r] = (float)((double)rb[r] + d.row_height[r]);
    }
    for (int j = 0; j < bdy; ++j) {
      std::vector<int> rows;
      for (int r = 0; r < R; ++r)
        if (ov_len(blly[j], bury[j], rb[r], rt[r]) != 0) rows.push_back(r);
      for (int i = 0; i < bdx; ++i) {
        float avail = 0;
        for (int r : rows) avail = (float)(avail + ov_area(bllx[i], blly[j], burx[i], bury[j], rl[r], rb[r], rr[r], rt[r]));
        base[(size_t)i * bdy + j] = std::fabs((double)(barea - avail)) < kMlgEps ? 0.f : td * (barea - avail);
      }
    }
    // standard cells (macros only feed the unused macroArea)
    for (int m = 0; m < N; ++m) {
      if (is_macro(m)) continue;
      const float l = llx[m], b = lly[m], r = l + d.node_w[m], t = b + d.node_h[m];
      int i0, i1, j0, j1;
      bin_range(l, b, r, t, i0, i1, j0, j1);
      for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) {
          const float a = (float)ov_area(bllx[i], blly[j], burx[i], bury[j], l, b, r, t);
          cell[(size_t)i * bdy + j] += a;
        }
    }
    dens.resize(nb);
    for (size_t k = 0; k < nb; ++k) dens[k] = (cell[k] + term[k] + base[k]) / barea;
  }

  // Net::calcNetHPWL summed over nets (PlaceDB::calcHPWL, placedb.cpp:424-435)
  double calc_hpwl() const {
    double h = 0;
    for (int e = 0; e < d.num_nets; ++e) {
      double mnx = 1.7976931348623157e308, mxx = -mnx, mny = mnx, mxy = -mnx;
      for (int64_t q = d.net_start[e]; q < d.net_start[e + 1]; ++q) {
        mnx = std::min(mnx, (double)px[q]), mxx = std::max(mxx, (double)px[q]);
        mny = std::min(mny, (double)py[q]), mxy = std::max(mxy, (double)py[q]);
      }
      if (d.net_start[e + 1] > d.net_start[e]) h += (mxx - mnx) + (mxy - mny);
    }
    return h;
  }

  // PlaceDB::calcModuleHPWLfast (placedb.cpp:537-582)
  double module_hpwl(int m) const {
    double h = 0;
    for (int k = npin_start[m]; k < npin_start[m + 1]; ++k) {
      const int64_t q0 = npin_net[k];
      float maxX = -3.40282347e+38f, minX = 3.40282347e+38f, maxY = -3.40282347e+38f, minY = 3.40282347e+38f;
      float maxZ = -3.40282347e+38f, minZ = 3.40282347e+38f;
      for (int64_t q = d.net_start[q0]; q < d.net_start[q0 + 1]; ++q) {
        minX = std::min(minX, px[q]);
        maxX = std::max(maxX, px[q]);
        minY = std::min(minY, py[q]);
        maxY = std::max(maxY, py[q]);
        minZ = std::min(minZ, 0.0f);
        maxZ = std::max(maxZ, 0.0f);
      }
      h += ((maxX - minX) + (maxY - minY) + (maxZ - minZ));
    }
    return h;
  }

  float cell_cover(int m) const {
    float cost = 0;
    const float l = llx[m], b = lly[m], r = l + d.node_w[m], t = b + d.node_h[m];
    int i0, i1, j0, j1;
    bin_range(l, b, r, t, i0, i1, j0, j1);
    for (int i = i0; i <= i1; ++i)
      for (int j = j0; j <= j1; ++j)
        cost = (float)(cost + ov_area(bllx[i], blly[j], burx[i], bury[j], l, b, r, t) * dens[(size_t)i * bdy + j]);
    return cost;
  }

  int macro_overlap(int m) const {
    int tot = 0;
    const float l = llx[m], b = lly[m], r = l + d.node_w[m], t = b + d.node_h[m];
    for (int o : macros) {
      if (o == m) continue;
      tot = (int)(tot + ov_area(l, b, r, t, llx[o], lly[o], llx[o] + d.node_w[o], lly[o] + d.node_h[o]));
    }
    return tot;
  }

  int covered_by_macros() const {
    std::vector<int> rl, rb, rr, rt;
    for (int m : macros) {
      rl.push_back((int)llx[m]);
      rb.push_back((int)lly[m]);
      rr.push_back((int)(llx[m] + d.node_w[m]));
      rt.push_back((int)(lly[m] + d.node_h[m]));
    }
    return union_area(rl, rb, rr, rt);
  }

  float macro_cost(int m, float &hp, float &cc, float &mo) const {
    hp = (float)module_hpwl(m);
    cc = cell_cover(m);
    mo = (float)macro_overlap(m);
    return hp + cc * miuD + mo * miuO;
  }

  bool accept(float oldCost, float newCost) {
    const float costDelta = (newCost - oldCost) / oldCost;
    const int rnd = rng.next();
    const float drnd = (float)rnd / kRandMaxF;
    const float expValue = (float)std::exp(-1.0 * costDelta / temp);
    return drnd < expValue;
  }

  // SAperturb (legalizer.cpp:1405-1500)
  void perturb() {
    const int rnd_idx = rng.next();
    const double drnd_idx = (double)rnd_idx / kRandMax;
    int idx = (int)(drnd_idx * (double)macros.size());
    if (idx >= (int)macros.size()) idx = (int)macros.size() - 1;
    const int m = macros[idx];
    const float ccx = cx[m], ccy = cy[m];
    const int rnx = rng.next(), rny = rng.next();
    const float inv = (float)1.0 / kRandMaxF;
    const float drx = (float)((float)rnx * inv - 0.5);
    const float dry = (float)((float)rny * inv - 0.5);
    const int dx = (int)(std::round(drx * rx) * ux);
    const int dy = (int)(std::round(dry * ry / rowh) * rowh * uy);
    float ch, cc, co, nh, nc, no;
    const float curCost = macro_cost(m, ch, cc, co);
    // moveModule_2D(Module *, VECTOR_2D_INT): the centre truncated to int first
    const int ix = (int)cx[m], iy = (int)cy[m];
    set_center(m, (float)(ix + dx), (float)(iy + dy));
    const float newCost = macro_cost(m, nh, nc, no);
    ++perturbations;
    if (accept(curCost, newCost)) {
      ++accepted;
      totalHPWL += nh - ch;
      totalCellAreaCovered += nc - cc;
      totalMacroOverlap = (int)((float)totalMacroOverlap + (no - co));
      if (totalMacroOverlap <= 0 && no <= 0 && co > 0) {
        totalMacroOverlap = (int)(totalMacroArea - covered_by_macros());
        if (totalMacroOverlap <= 0) overlapFree = true;
      }
    } else {
      set_center(m, ccx, ccy);
    }
  }

  int outer = 0;

  void run() {
    jLimit = p.max_outer > 0 ? p.max_outer : 1000;
    kLimit = p.max_inner > 0 ? p.max_inner : 1000;
    init_macros();
    if (macros.empty()) {
      overlapFree = true;
      return;
    }
    init_bins();
    // initializeCost (legalizer.cpp:1191-1205)
    totalHPWL = calc_hpwl();
    {
      float c = 0;
      for (int m : macros) c += cell_cover(m);
      totalCellAreaCovered = c;
    }
    totalMacroOverlap = (int)(totalMacroArea - covered_by_macros());
    if (totalMacroOverlap <= 0) overlapFree = true;
    // initializeSAparams (legalizer.cpp:1221-1234)
    miuD = (float)(totalHPWL / totalCellAreaCovered);
    miuO = (float)((totalHPWL + totalCellAreaCovered * miuD) / (float)totalMacroOverlap);
    beta = 1.5;
    // legalization (legalizer.cpp:701-790)
    for (int j = 0; j < jLimit && !overlapFree; j++) {
      outer = j + 1;
      const float init_rate = (float)(0.03 * std::pow(beta, (float)j));
      const float last_rate = (float)(0.0001 * std::pow(beta, (float)j));
      const float init_t = (float)(init_rate / std::log(2.0));
      temp = init_t;
      tempCoef = (float)std::pow(last_rate / init_rate, 1.0 / (float)kLimit);
      const float sa_n = std::sqrt(float(macros.size()));
      const float ncof = (float)(0.05 * std::pow(beta, (float)j));
      const float maxrx = (curx - cllx) / sa_n * ncof, maxry = (cury - clly) / sa_n * ncof;
      rx = maxrx;
      ry = maxry;
      ux = uy = 1.0;
      stepx = (float)((maxrx - 1.0) / (float)kLimit);
      stepy = (float)((maxry - 1.0) / (float)kLimit);
      const int inner = (int)macros.size();
      for (int k = 0; k < kLimit && !overlapFree; k++) {
        for (int i = 0; i < inner && !overlapFree; i++) perturb();
        temp *= tempCoef;
        rx -= stepx;
        ry -= stepy;
      }
      miuO *= beta;
      totalMacroOverlap = (int)(totalMacroArea - covered_by_macros());
      if (totalMacroOverlap <= 0) overlapFree = true;
    }
    totalMacroOverlap = (int)(totalMacroArea - covered_by_macros());
  }
};

}  // namespace

int macro_legalize_sa(const ep_design &d, std::vector<float> &xy, const ep_mlg_params &p, Rand &rng,
                      ep_mlg_stats &st) {
  auto t0 = std::chrono::steady_clock::now();
  MacroSA sa(d, p, rng);
  const int64_t draws0 = rng.draws;
  sa.setup(xy.data());
  st = ep_mlg_stats{};
  st.initial_hpwl = sa.calc_hpwl();
  sa.run();
  for (int i = 0; i < sa.N; ++i) xy[2 * i] = sa.cx[i], xy[2 * i + 1] = sa.cy[i];
  st.macros = (int32_t)sa.macros.size();
  st.bins_x = sa.bdx;
  st.bins_y = sa.bdy;
  st.outer_iterations = sa.outer;
  st.perturbations = sa.perturbations;
  st.accepted = sa.accepted;
  st.overlap_free = sa.overlapFree ? 1 : 0;
  st.final_overlap = sa.totalMacroOverlap;
  st.final_hpwl = sa.calc_hpwl();
  st.cell_area_covered = sa.totalCellAreaCovered;
  st.rand_draws = rng.draws - draws0;
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return EP_OK;
}

}  // namespace ep

using namespace ep;

namespace {
thread_local std::string g_mlg_err;
}

extern "C" {

const char *ep_mlg_last_error(void) { return g_mlg_err.c_str(); }

int ep_legalize_macros(const ep_design *design, const float *node_xy, const ep_mlg_params *params, float *out_xy,
                       ep_mlg_stats *stats) {
  g_mlg_err.clear();
  if (!design || !params || !out_xy || design->num_nodes <= 0 || design->num_rows <= 0 || !design->net_start ||
      !design->pin_obj || !design->node_w || !design->node_h || !design->row_bottom || !design->row_height ||
      !design->row_x0 || !design->row_x1 || (design->num_terminals > 0 && (!design->term_x || !design->term_y ||
                                                                           !design->term_w || !design->term_h))) {
    g_mlg_err = "bad arguments";
    return EP_ERR_ARG;
  }
  if (!(params->target_density > 0) || params->rand_skip < 0) {
    g_mlg_err = "target_density must be > 0 and rand_skip >= 0";
    return EP_ERR_ARG;
  }
  const int N = design->num_nodes;
  std::vector<float> xy(2 * (size_t)N);
  for (int i = 0; i < N; ++i) {
    xy[2 * i] = node_xy ? node_xy[2 * i] : design->node_x[i];
    xy[2 * i + 1] = node_xy ? node_xy[2 * i + 1] : design->node_y[i];
  }
  Rand rng(params->rand_seed);
  for (int64_t k = 0; k < params->rand_skip; ++k) rng.next();
  ep_mlg_stats st{};
  macro_legalize_sa(*design, xy, *params, rng, st);
  std::memcpy(out_xy, xy.data(), xy.size() * sizeof(float));
  if (stats) *stats = st;
  return EP_OK;
}

}  // extern "C"
