// Abacus standard-cell legalisation and a legality checker (host C++).
//
// Reference: AbacusLegalizer (Legalization/legalizer.cpp:58-680) on the
// site-row intervals of PlaceDB::removeBlockedSite (placedb.cpp:629-849).
// Mode 0 keeps the reference's arithmetic and its quirks so that results
// match it exactly:
//   * cells are visited in ascending lower-left x (std::sort, float_less);
//   * the nearest row is the first whose bottom is within half a row height
//     of the cell's lower-left y; rows above are searched first, then rows
//     below, stopping when |dy| alone exceeds the best cost; the downward
//     search never reaches subrow 0 (legalizer.cpp:128-131);
//   * a cluster sits at (int)(q / e / step) * step clamped into the subrow
//     (a new cluster's int(x + 0.5) start, legalizer.cpp:533, is always
//     overwritten by this collapse);
//   * cluster fields x, e, w, q are float, the row's used width is double;
//   * the written x is an int that accumulates int(x + width) per cell
//     (legalizer.cpp:163-173), then PlaceDB::setModuleLocation_2D clamps it
//     into the core.
// Mode 1 changes exactly these points: row 0 is searched, cluster positions
// round to the site grid of the core origin (floor((x - llx) / step + 1/2)),
// positions stay float, and a cell whose nearest row is past the last one
// starts from the last row.  In both modes a cell that fits no subrow is an
// error (the reference asserts).
//
// Trial placements do not copy the subrow (the reference copies its whole
// cluster list per trial): a trial replays the merges of the row's last
// clusters on a small local stack, which gives the same numbers.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/eplace_amd.h"
#include "host_rows.h"

namespace {

using namespace ep_rows;

thread_local std::string g_err;

struct Cluster {
  float x = 0, e = 0, w = 0, q = 0;
  std::vector<int> cells;
};

struct Subrow {
  float start = 0, end = 0;
  double bottom = 0, height = 0, step = 1, width = 0;
  std::vector<Cluster> clusters;
};

struct Abacus {
  const Geometry &g;
  int mode;
  std::vector<Subrow> rows;

  float collapse_x(const Subrow &r, const Cluster &c) const {
    float x;
    if (mode == 0) {
      x = (float)((int)(1.0 * c.q / c.e / r.step) * r.step);
    } else {
      x = (float)(std::floor((1.0 * c.q / c.e - g.core_llx) / r.step + 0.5) * r.step + g.core_llx);
    }
    if (f_less(x, r.start)) x = r.start;
    if (f_greater(x + c.w, r.end)) x = r.end - c.w;
    return x;
  }
  static void add_cell(Cluster &c, float llx, float w) {  // AbacusRow::addCell
    c.e += 1;
    c.q += llx - c.w;
    c.w += w;
  }
  static void merge(Cluster &prev, const Cluster &c) {  // AbacusRow::addCluster
    prev.e += c.e;
    prev.q += c.q - c.e * prev.w;
    prev.w += c.w;
  }

  // placeRow (legalizer.cpp:495-590): the cell's cost on subrow r; with
  // commit the row is updated, otherwise the last clusters are replayed on a
  // local stack.
  double place(int r, int cell, bool commit) {
    Subrow &row = rows[r];
    std::vector<Cluster> &cl = row.clusters;
    const float lx = g.llx[cell], w = g.w[cell];
    // stack of (x, e, w, q) for the clusters touched, deepest last
    struct Top {
      float x, e, w, q;
    };
    std::vector<Top> st;
    int keep = (int)cl.size();  // clusters below the stack are untouched
    Top cur;
    if (cl.empty() || f_lessorequal(cl.back().x + cl.back().w, lx)) {
      // a new cluster; its start position (legalizer.cpp:533-544) is
      // replaced by collapse before it is used
      cur = Top{0, 0, 0, 0};
    } else {
      const Cluster &b = cl.back();
      cur = Top{b.x, b.e, b.w, b.q};
      keep = (int)cl.size() - 1;
    }
    Cluster tmp;
    tmp.e = cur.e;
    tmp.w = cur.w;
    tmp.q = cur.q;
    add_cell(tmp, lx, w);
    cur.e = tmp.e;
    cur.w = tmp.w;
    cur.q = tmp.q;
    // collapse, merging backwards while the predecessor overlaps
    int pred = keep - 1;
    std::vector<int> merged;  // indices of existing clusters merged into cur, in order
    while (true) {
      Cluster c;
      c.e = cur.e;
      c.w = cur.w;
      c.q = cur.q;
      cur.x = collapse_x(row, c);
      if (pred < 0) break;
      const Cluster &p = cl[pred];
      if (!(p.x + p.w > cur.x)) break;
      Cluster m;
      m.e = p.e;
      m.w = p.w;
      m.q = p.q;
      merge(m, c);
      cur = Top{p.x, m.e, m.w, m.q};
      merged.push_back(pred);
      --pred;
    }
    // cost (legalizer.cpp:573-590): the cell is the last one of the resulting
    // cluster; x accumulates the widths of the cells before it, in order
    double cost;
    {
      float x = cur.x;
      for (auto it = merged.rbegin(); it != merged.rend(); ++it)
        for (int o : cl[*it].cells) x += g.w[o];
      if (keep < (int)cl.size())
        for (int o : cl[keep].cells) x += g.w[o];
      const float dx = lx - x;
      const double dy = (double)g.lly[cell] - row.bottom;
      cost = std::sqrt((double)(dx * dx) + dy * dy);
    }
    if (commit) {
      row.width += w;
      const int first = merged.empty() ? (keep < (int)cl.size() ? keep : (int)cl.size()) : merged.back();
      Cluster out;
      out.x = cur.x;
      out.e = cur.e;
      out.w = cur.w;
      out.q = cur.q;
      for (int i = first; i < (int)cl.size(); ++i)
        out.cells.insert(out.cells.end(), cl[i].cells.begin(), cl[i].cells.end());
      out.cells.push_back(cell);
      cl.resize(first);
      cl.push_back(std::move(out));
    }
    return cost;
  }
};

}  // namespace

std::string &ep_rows::host_error() { return g_err; }

extern "C" {

const char *ep_host_last_error(void) { return g_err.c_str(); }

int ep_legalize_abacus(const ep_design *d, const float *node_xy, const ep_legal_params *p, float *out_xy,
                       ep_legal_stats *stats) {
  g_err.clear();
  if (!d || !p || !out_xy || d->num_nodes <= 0 || d->num_rows <= 0) return g_err = "bad arguments", EP_ERR_ARG;
  if (!(p->site_step > 0) || (p->mode != 0 && p->mode != 1)) return g_err = "bad site_step or mode", EP_ERR_ARG;
  Geometry g = geometry(d, node_xy);
  std::vector<std::vector<Interval>> iv;
  if (!row_intervals(d, g, p->site_step, iv, g_err)) return EP_ERR_ARG;
  Abacus a{g, p->mode, {}};
  for (int r = 0; r < d->num_rows; ++r)
    for (const Interval &t : iv[r]) {
      Subrow s;
      s.start = t.start;
      s.end = t.end;
      s.bottom = d->row_bottom[r];
      s.height = d->row_height[r];
      s.step = p->site_step;
      a.rows.push_back(s);
    }
  bool dup = false;
  std::sort(a.rows.begin(), a.rows.end(), [&](const Subrow &x, const Subrow &y) {
    if (!f_equal((float)x.bottom, (float)y.bottom)) return f_less((float)x.bottom, (float)y.bottom);
    if (!f_equal(x.start, y.start)) return f_less(x.start, y.start);
    dup = true;
    return false;
  });
  if (dup) return g_err = "subrows overlap", EP_ERR_ARG;
  std::vector<int> cells;
  for (int i = 0; i < g.N; ++i)
    if (!g.macro[i]) cells.push_back(i);
  std::sort(cells.begin(), cells.end(), [&](int x, int y) { return f_less(g.llx[x], g.llx[y]); });
  const int S = (int)a.rows.size();
  for (int cell : cells) {
    double cost = 1.79769313486231570815e+308;
    int best = -1, closest = 0;
    for (const Subrow &r : a.rows) {
      if (f_lessorequal((float)std::fabs(r.bottom - g.lly[cell]), (float)(0.5 * r.height))) break;
      ++closest;
    }
    if (p->mode == 1 && closest >= S) closest = S - 1;
    auto consider = [&](int i) -> bool {  // false: stop this direction
      const Subrow &r = a.rows[i];
      if (f_greater((float)(r.width + g.w[cell]), r.end - r.start)) return true;
      const double lb = std::fabs(g.lly[cell] - r.bottom);
      if (f_greaterorequal((float)lb, (float)cost)) return false;
      const double c = a.place(i, cell, false);
      if (f_lessorequal((float)c, (float)cost)) {
        best = i;
        cost = c;
      }
      return true;
    };
    for (int i = closest; i < S; ++i)
      if (!consider(i)) break;
    for (int i = closest - 1; i >= 0; --i) {
      if (p->mode == 0 && i == 0) break;
      if (!consider(i)) break;
    }
    if (best < 0) {
      g_err = "cell " + std::to_string(cell) + " fits no subrow";
      return EP_ERR_STATE;
    }
    a.place(best, cell, true);
  }
  // write back (legalizer.cpp:160-175 + setModuleLocation_2D, placedb.cpp:203-235)
  std::vector<float> llx = g.llx, lly = g.lly;
  for (const Subrow &r : a.rows)
    for (const Cluster &c : r.clusters) {
      if (p->mode == 0) {
        int x = (int)c.x;
        for (int o : c.cells) {
          llx[o] = (float)x;
          lly[o] = (float)r.bottom;
          x = (int)((float)x + g.w[o]);
        }
      } else {
        float x = c.x;
        for (int o : c.cells) {
          llx[o] = x;
          lly[o] = (float)r.bottom;
          x += g.w[o];
        }
      }
    }
  double total = 0, worst = 0;
  for (int o : cells) {
    float x = llx[o], y = lly[o];
    if (x < g.core_llx) x = g.core_llx;
    if (x + g.w[o] > g.core_urx) x = (float)((double)(g.core_urx - g.w[o]) - kEps);
    if (y < g.core_lly) y = g.core_lly;
    if (y + g.h[o] > g.core_ury) y = (float)((double)(g.core_ury - g.h[o]) - kEps);
    const double dx = (double)x - g.llx[o], dy = (double)y - g.lly[o];
    const double m = std::sqrt(dx * dx + dy * dy);
    total += m;
    worst = std::max(worst, m);
    llx[o] = x;
    lly[o] = y;
  }
  for (int i = 0; i < g.N; ++i) {
    if (g.macro[i]) {
      out_xy[2 * i] = node_xy ? node_xy[2 * i] : d->node_x[i];
      out_xy[2 * i + 1] = node_xy ? node_xy[2 * i + 1] : d->node_y[i];
    } else {  // Module::setLocation_2D (objects.cpp:518-527)
      out_xy[2 * i] = llx[i] + (float)0.5 * g.w[i];
      out_xy[2 * i + 1] = lly[i] + (float)0.5 * g.h[i];
    }
  }
  if (stats) {
    stats->cells = (int64_t)cells.size();
    stats->subrows = S;
    stats->total_displacement = total;
    stats->max_displacement = worst;
  }
  return EP_OK;
}

int ep_check_legal(const ep_design *d, const float *node_xy, double site_step, ep_legal_check *out) {
  g_err.clear();
  if (!d || !out || d->num_nodes <= 0 || d->num_rows <= 0 || !(site_step > 0)) return g_err = "bad arguments", EP_ERR_ARG;
  *out = ep_legal_check{};
  Geometry g = geometry(d, node_xy);
  const double tol = 1e-4;
  const int R = d->num_rows;
  std::vector<Rect> ob = obstacles(d, g);
  // macros against macros
  std::vector<int> macros;
  for (int i = 0; i < g.N; ++i)
    if (g.macro[i]) macros.push_back(i);
  for (size_t a = 0; a < macros.size(); ++a)
    for (size_t b = a + 1; b < macros.size(); ++b) {
      const int i = macros[a], j = macros[b];
      const double ox = std::min(g.llx[i] + g.w[i], g.llx[j] + g.w[j]) - std::max(g.llx[i], g.llx[j]);
      const double oy = std::min(g.lly[i] + g.h[i], g.lly[j] + g.h[j]) - std::max(g.lly[i], g.lly[j]);
      if (ox > tol && oy > tol) ++out->macro_overlaps;
    }
  // cells: core, row, site, blockage; then overlaps row by row
  std::vector<std::vector<int>> by_row(R);
  for (int i = 0; i < g.N; ++i) {
    if (g.macro[i]) continue;
    ++out->cells;
    const double lx = g.llx[i], ly = g.lly[i], ux = lx + g.w[i], uy = ly + g.h[i];
    if (lx < g.core_llx - tol || ux > g.core_urx + tol || ly < g.core_lly - tol || uy > g.core_ury + tol)
      ++out->out_of_core;
    int row = -1;
    for (int r = 0; r < R; ++r)
      if (std::fabs(ly - d->row_bottom[r]) <= tol) {
        row = r;
        break;
      }
    if (row < 0 || lx < d->row_x0[row] - tol || ux > d->row_x1[row] + tol) {
      ++out->off_row;
    } else {
      const double s = (lx - d->row_x0[row]) / site_step;
      if (std::fabs(s - std::round(s)) > tol) ++out->off_site;
      by_row[row].push_back(i);
    }
    for (const Rect &o : ob) {
      const double ox = std::min<double>(ux, o.urx) - std::max<double>(lx, o.llx);
      const double oy = std::min<double>(uy, o.ury) - std::max<double>(ly, o.lly);
      if (ox > tol && oy > tol) {
        ++out->in_blockage;
        break;
      }
    }
  }
  for (auto &v : by_row) {
    std::sort(v.begin(), v.end(), [&](int a, int b) { return g.llx[a] < g.llx[b]; });
    double reach = -1e300;
    for (int i : v) {
      const double lx = g.llx[i];
      if (lx < reach - tol) {
        ++out->overlaps;
        out->overlap_area += (std::min<double>(reach, lx + g.w[i]) - lx) * g.h[i];
      }
      reach = std::max<double>(reach, lx + g.w[i]);
    }
  }
  return EP_OK;
}

}  // extern "C"
