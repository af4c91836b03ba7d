// Row geometry shared by the host-side legaliser and detailed placer: node
// rectangles from centres, the reference's float comparisons, and the
// site-row intervals left after removing blocked sites
// (PlaceDB::removeBlockedSite, placedb.cpp:629-849).
#pragma once

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/eplace_amd.h"

namespace ep_rows {

// The message ep_host_last_error() returns (this thread's).
std::string &host_error();

constexpr double kEps = 1.0e-15;  // global.h:39
// global.h:345-367 take float arguments.
inline bool f_greater(float a, float b) { return (double)(a - b) > kEps; }
inline bool f_less(float a, float b) { return (double)(a - b) < -kEps; }
inline bool f_equal(float a, float b) { return std::fabs((double)(a - b)) < kEps; }
inline bool f_lessorequal(float a, float b) { return f_less(a, b) || f_equal(a, b); }
inline bool f_greaterorequal(float a, float b) { return f_greater(a, b) || f_equal(a, b); }

struct Interval {
  float start, end;
};

struct Geometry {
  int N = 0;
  std::vector<float> llx, lly, w, h;
  std::vector<uint8_t> macro;
  float core_llx = 0, core_lly = 0, core_urx = 0, core_ury = 0;
};

inline Geometry geometry(const ep_design *d, const float *xy) {
  Geometry g;
  g.N = d->num_nodes;
  g.llx.resize(g.N);
  g.lly.resize(g.N);
  g.w.assign(d->node_w, d->node_w + g.N);
  g.h.assign(d->node_h, d->node_h + g.N);
  g.macro.resize(g.N);
  const float rowh = (float)d->common_row_height;
  for (int i = 0; i < g.N; ++i) {
    const float cx = xy ? xy[2 * i] : d->node_x[i], cy = xy ? xy[2 * i + 1] : d->node_y[i];
    g.llx[i] = cx - (float)0.5 * g.w[i];  // Module::setCenter_2D (objects.cpp:530-539)
    g.lly[i] = cy - (float)0.5 * g.h[i];
    g.macro[i] = f_greater(g.h[i], rowh);  // PlaceDB::addNode (placedb.cpp:102-111)
  }
  const int R = d->num_rows;
  g.core_lly = (float)d->row_bottom[0];
  g.core_ury = (float)(d->row_bottom[R - 1] + d->row_height[R - 1]);
  g.core_llx = (float)d->row_x0[0];
  g.core_urx = (float)d->row_x1[0];
  for (int r = 0; r < R; ++r) {
    g.core_llx = std::min(g.core_llx, (float)d->row_x0[r]);
    g.core_urx = std::max(g.core_urx, (float)d->row_x1[r]);
  }
  return g;
}

struct Rect {
  float llx, lly, urx, ury;
};

// Obstacles of removeBlockedSite (placedb.cpp:629-690): terminals that meet
// the core, and every macro.
inline std::vector<Rect> obstacles(const ep_design *d, const Geometry &g) {
  std::vector<Rect> ob;
  for (int t = 0; t < d->num_terminals; ++t) {
    Rect r{d->term_x[t] - (float)0.5 * d->term_w[t], d->term_y[t] - (float)0.5 * d->term_h[t], 0, 0};
    r.urx = r.llx + d->term_w[t];
    r.ury = r.lly + d->term_h[t];
    if (f_greater(g.core_ury, r.lly) && f_less(g.core_lly, r.ury) && f_greater(g.core_urx, r.llx) &&
        f_less(g.core_llx, r.urx))
      ob.push_back(r);
  }
  for (int i = 0; i < g.N; ++i)
    if (g.macro[i]) ob.push_back(Rect{g.llx[i], g.lly[i], g.llx[i] + g.w[i], g.lly[i] + g.h[i]});
  return ob;
}

// Row intervals after removing blocked sites (placedb.cpp:629-849), from the
// full rows.  The reference recomputes them in place from the previous
// intervals each time it is called; applications > 1 repeats the removal
// (the Abacus legaliser calls it once, detailed placement once more).  On
// an error err is set and false returned.
inline bool row_intervals(const ep_design *d, const Geometry &g, double step, std::vector<std::vector<Interval>> &iv,
                          std::string &err, int applications = 1) {
  const int R = d->num_rows;
  iv.assign(R, {});
  for (int r = 0; r < R; ++r) iv[r].push_back(Interval{(float)d->row_x0[r], (float)d->row_x1[r]});
  std::vector<Rect> ob = obstacles(d, g);
  bool overlap = false;
  std::sort(ob.begin(), ob.end(), [&](const Rect &a, const Rect &b) {
    if (!f_equal(a.llx, b.llx)) return f_less(a.llx, b.llx);
    if (!f_equal(a.lly, b.lly)) return f_less(a.lly, b.lly);
    overlap = true;
    return false;
  });
  if (overlap) {
    err = "terminals or macros share a lower-left corner (they overlap)";
    return false;
  }
  for (int pass = 0; pass < applications; ++pass) {
    for (const Rect &o : ob) {
      int rb = 0;
      while (rb < R && !(d->row_bottom[rb] + d->row_height[rb] > o.lly)) ++rb;
      if (rb == R) continue;  // above every row
      int re = rb;
      while (re < R && !(d->row_bottom[re] + d->row_height[re] >= o.ury)) ++re;
      if (re == R) --re;
      for (int r = rb; r <= re; ++r) {
        std::vector<Interval> &v = iv[r];
        for (int i = 0; i < (int)v.size(); ++i) {
          const Interval t = v[i];
          if (t.start >= o.urx || t.end <= o.llx) continue;
          if (t.start >= o.llx && t.end <= o.urx) {
            v.erase(v.begin() + i);  // the next interval shifts into i and is skipped, as in the reference
          } else if (t.end > o.urx && t.start >= o.llx) {
            v[i].start = o.urx;
          } else if (t.start < o.llx && t.end <= o.urx) {
            v[i].end = o.llx;
          } else if (t.start < o.llx && t.end > o.urx) {
            v[i].end = o.llx;
            v.insert(v.begin() + i + 1, Interval{o.urx, t.end});
          }
        }
      }
    }
    for (auto &v : iv) {
      for (auto it = v.begin(); it != v.end();) {
        const double width = (double)(it->end - it->start);
        if (f_less((float)width, (float)step)) {
          it = v.erase(it);
          continue;
        }
        const double left = std::ceil((it->start - g.core_llx) / step) * step + g.core_llx;
        const double right = std::floor((it->end - g.core_llx) / step) * step + g.core_llx;
        if (f_greater((float)(right - left), 0.f)) {
          it->start = (float)left;
          it->end = (float)right;
          ++it;
        } else {
          it = v.erase(it);
        }
      }
    }
  }
  return true;
}

}  // namespace ep_rows
