// Detailed placement (host C++): the local-search rounds of
// DetailedPlacer::detailedPlacement (DetailedPlacement/detailed.cpp:74-85)
// on a legal standard-cell placement.  Each round runs three moves, each of
// which only ever produces legal positions:
//   * local reordering (runLocalReordering, detailed.cpp:241-262,
//     1422-2050): every three neighbouring cells of a row segment, with the
//     segment's free space inside their span as one more block, are placed
//     in the order of least wirelength (depth-first search over the orders,
//     a net's HPWL counted once all its window cells are placed, pruned at
//     the best complete order so far); twice per round;
//   * independent-set matching (runISM, detailed.cpp:102-216, 264-750):
//     square windows of 20 - k row heights sweep the core; in each window
//     the cells (widest first) that can take a slot of the widest cell's
//     width are matched to those slots and to free slots by a linear
//     assignment of their wirelength (Jonker-Volgenant, lap2::lapSolve,
//     detailed.cpp:1117-1420); sweeps repeat while they gain 0.2%, then
//     with independent cells and a second window, until a sweep gains less
//     than 0.6%;
//   * global swap (runGlobalSwap, detailed.cpp:218-239, 2053-2371): every
//     cell outside its optimal region (the median box of its nets' other
//     pins, placedb.cpp:1480-1533) moves to the first free slot there that
//     shortens its nets, else swaps with the first cell of equal width there
//     for which the pair's wirelength drops; twice per round.
// Positions, widths and pin offsets are float, wirelength sums double, the
// row books std::map keyed by rounded x, as in the reference, and the
// moves visit cells, windows, slots and candidates in the reference's
// orders, so that from the same legal placement the result follows the
// reference's.  Departures, all on inputs where the reference reads out of
// bounds or aborts: a cell without nets has no optimal region and is left
// alone by the global swap (the reference indexes an empty vector), and
// module row books skip rows past the last one.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <list>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/eplace_amd.h"
#include "host_rows.h"

namespace {

using namespace ep_rows;

constexpr double kDoubleMax = 1.7976931348623157e308;  // DOUBLE_MAX (global.h:40)
constexpr float kFloatMax = 3.40282347e38f;            // FLOAT_MAX (global.h:42)

// The placement the moves work on: nodes' lower-left corners (Module::coor),
// sizes, the nets and pins of the design, and the core and rows.
struct Board {
  const ep_design *d = nullptr;
  int N = 0, T = 0, E = 0, R = 0;
  std::vector<float> x, y, w, h;  // lower-left corners and sizes, float
  std::vector<uint8_t> macro;
  std::vector<float> tcx, tcy;    // terminal centres
  std::vector<int> npin_start, npin;  // each node's pins in module order (net order, then pin order)
  std::vector<int> pin_net;
  float core_llx = 0, core_lly = 0, core_urx = 0, core_ury = 0;
  double rowh = 1, site_step = 1;
  std::vector<double> row_bottom;
  std::vector<std::vector<Interval>> iv;  // placeable intervals per row

  // Pin position: the owner's centre (coor + w/2 in float) plus the offset
  // (Pin::calculateAbsolutePos, objects.cpp:454-459).
  float pin_x(int64_t q) const {
    const int o = d->pin_obj[q];
    return (o < N ? x[o] + (float)0.5 * w[o] : tcx[o - N]) + d->pin_dx[q];
  }
  float pin_y(int64_t q) const {
    const int o = d->pin_obj[q];
    return (o < N ? y[o] + (float)0.5 * h[o] : tcy[o - N]) + d->pin_dy[q];
  }

  // Net::calcNetHPWL (objects.cpp:35-75): double extremes of float positions.
  double net_hpwl(int e) const {
    double mnx = kDoubleMax, mxx = -kDoubleMax, mny = kDoubleMax, mxy = -kDoubleMax;
    for (int64_t q = d->net_start[e]; q < d->net_start[e + 1]; ++q) {
      const double px = pin_x(q), py = pin_y(q);
      mnx = std::min(mnx, px);
      mxx = std::max(mxx, px);
      mny = std::min(mny, py);
      mxy = std::max(mxy, py);
    }
    return (mxx - mnx) + (mxy - mny) + 0.0;
  }

  // PlaceDB::calcHPWL (placedb.cpp:424-432).
  double hpwl() const {
    double s = 0;
    for (int e = 0; e < E; ++e)
      if (d->net_start[e + 1] > d->net_start[e]) s += net_hpwl(e);
    return s;
  }

  // PlaceDB::calcModuleHPWLfast (placedb.cpp:537-582): one float box per
  // pin of the node (a net with two of its pins counts twice).
  double module_hpwl(int o) const {
    double s = 0;
    for (int k = npin_start[o]; k < npin_start[o + 1]; ++k) {
      const int e = pin_net[npin[k]];
      float mnx = kFloatMax, mxx = -kFloatMax, mny = kFloatMax, mxy = -kFloatMax;
      for (int64_t q = d->net_start[e]; q < d->net_start[e + 1]; ++q) {
        const float px = pin_x(q), py = pin_y(q);
        mnx = std::min(mnx, px);
        mxx = std::max(mxx, px);
        mny = std::min(mny, py);
        mxy = std::max(mxy, py);
      }
      s += ((mxx - mnx) + (mxy - mny)) + 0.0f;
    }
    return s;
  }

  // The boxes of node o's nets over the pins of other objects, one per pin
  // of o in module order.  With them, module_hpwl at any position of o is
  // moved_hpwl (min and max are exact, so adding o's own pins to these boxes
  // gives the same floats as module_hpwl's walk over whole nets); the moves
  // that try one node at many slots compute the boxes once.
  struct Box {
    float mnx, mxx, mny, mxy;
  };
  void other_boxes(int o, std::vector<Box> &out) const {
    out.clear();
    for (int k = npin_start[o]; k < npin_start[o + 1]; ++k) {
      const int e = pin_net[npin[k]];
      Box bx{kFloatMax, -kFloatMax, kFloatMax, -kFloatMax};
      for (int64_t q = d->net_start[e]; q < d->net_start[e + 1]; ++q) {
        if (d->pin_obj[q] == o) continue;
        const float px = pin_x(q), py = pin_y(q);
        bx.mnx = std::min(bx.mnx, px);
        bx.mxx = std::max(bx.mxx, px);
        bx.mny = std::min(bx.mny, py);
        bx.mxy = std::max(bx.mxy, py);
      }
      out.push_back(bx);
    }
  }
  double moved_hpwl(int o, const std::vector<Box> &boxes) const {
    double s = 0;
    const int k0 = npin_start[o], k1 = npin_start[o + 1];
    for (int k = k0; k < k1; ++k) {
      const int e = pin_net[npin[k]];
      Box bx = boxes[k - k0];
      for (int k2 = k0; k2 < k1; ++k2) {
        const int q = npin[k2];
        if (pin_net[q] != e) continue;
        const float px = pin_x(q), py = pin_y(q);
        bx.mnx = std::min(bx.mnx, px);
        bx.mxx = std::max(bx.mxx, px);
        bx.mny = std::min(bx.mny, py);
        bx.mxy = std::max(bx.mxy, py);
      }
      s += ((bx.mxx - bx.mnx) + (bx.mxy - bx.mny)) + 0.0f;
    }
    return s;
  }

  // PlaceDB::setModuleLocation_2D (placedb.cpp:203-236): clamped into the core.
  void set_ll(int o, float px, float py) {
    if (px < core_llx) px = core_llx;
    if (px + w[o] > core_urx) px = (float)((double)(core_urx - w[o]) - kEps);
    if (py < core_lly) py = core_lly;
    if (py + h[o] > core_ury) py = (float)((double)(core_ury - h[o]) - kEps);
    x[o] = px;
    y[o] = py;
  }

  // PlaceDB::y2RowIndex (placedb.cpp:1577-1584).
  int row_of(float py) const { return (int)((py - core_lly) / rowh); }

  // PlaceDB::isConnected (placedb.cpp:1600-1620).
  bool connected(int a, int b) const {
    for (int i = npin_start[a]; i < npin_start[a + 1]; ++i)
      for (int j = npin_start[b]; j < npin_start[b + 1]; ++j)
        if (pin_net[npin[i]] == pin_net[npin[j]]) return true;
    return false;
  }
};

// ---- one row's book of free space and cells (ISMRow, detailed.cpp:840-1070) ----
struct RowBook {
  double llx = 0, lly = 0;
  std::map<double, double> space;  // start -> length
  std::map<double, int> cells;     // lower-left x -> node

  bool insert_cell(const Board &b, double px, int o) {
    const double cx = std::round(px), cw = std::round((double)b.w[o]);
    auto it = space.upper_bound(cx);
    if (it == space.begin()) return false;
    --it;
    const double s0 = it->first, len = it->second;
    if (len - (cx - s0) < cw) return false;
    if (cx - s0 == 0) {
      if ((s0 + len) - (cx + cw) != 0) space[cx + cw] = (s0 + len) - (cx + cw);
      space.erase(s0);
    } else {
      if (cx + cw < s0 + len) space[cx + cw] = (s0 + len) - (cx + cw);
      it->second = cx - s0;
    }
    cells[cx] = o;
    return true;
  }

  bool remove_tail(double px, double width) {
    px = std::round(px);
    width = std::round(width);
    auto it = space.upper_bound(px);
    if (it == space.begin()) return false;
    --it;
    const double s0 = it->first, len = it->second;
    if (len - (px - s0) < width) return false;
    if (px - s0 == 0) {
      if ((s0 + len) - (px + width) != 0) space[px + width] = (s0 + len) - (px + width);
      space.erase(s0);
    } else {
      if (px + width < s0 + len) space[px + width] = (s0 + len) - (px + width);
      it->second = px - s0;
    }
    return true;
  }

  void insert_space(double px, double width) {
    px = std::round(px);
    width = std::round(width);
    auto nx = space.lower_bound(px);
    bool front = true, end = true;
    if (nx != space.begin()) {
      auto pv = std::prev(nx);
      front = !(pv->first + pv->second == px);
    }
    if (nx != space.end()) end = !(nx->first == px + width);
    if (front && end) {
      space[px] = width;
    } else if (front && !end) {
      const double l = nx->second, k = nx->first;
      space[px] = width + l;
      space.erase(k);
    } else if (!front && end) {
      auto pv = std::prev(space.lower_bound(px));
      pv->second = pv->second + width;
    } else {
      const double k = nx->first, l = nx->second;
      auto pv = std::prev(nx);
      pv->second = pv->second + width + l;
      space.erase(k);
    }
  }

  void remove_cell(const Board &b, int o) {
    insert_space(b.x[o], b.w[o]);
    auto it = cells.find(std::round((double)b.x[o]));
    if (it != cells.end() && it->second == o) cells.erase(it);
  }
};

// Row books of every row: intervals as free space, then every node inserted
// (ISMDP::initializeISMRows, detailed.cpp:776-838; GlobalSwapDP's the same).
std::vector<RowBook> row_books(const Board &b) {
  std::vector<RowBook> rows(b.R);
  for (int r = 0; r < b.R; ++r) {
    rows[r].llx = b.core_llx;
    rows[r].lly = b.row_bottom[r];
    for (const Interval &t : b.iv[r]) rows[r].space[t.start] = (float)(t.end - t.start);
  }
  for (int o = 0; o < b.N; ++o) {
    int covered = 1;
    if (b.h[o] > b.rowh) {
      covered = (int)(b.h[o] / b.rowh);
      if (b.rowh * covered < b.h[o]) ++covered;
    }
    const int r0 = b.row_of(b.y[o]);
    for (int i = 0; i < covered; ++i)
      if (r0 + i >= 0 && r0 + i < b.R) rows[r0 + i].insert_cell(b, b.x[o], o);
  }
  return rows;
}

// ---- linear assignment: Jonker & Volgenant (1987), shortest augmenting
// paths after column reduction, reduction transfer and two rounds of
// augmenting row reduction, on an integer cost matrix (lap2::lapSolve,
// detailed.cpp:1117-1420, same steps and tie rules) ----
struct Assign {
  std::vector<int> rowsol, colsol, v, matches, freer, col, d, pred;
};

// c: n x n row-major; the result (the column of each row) is in w.rowsol.
void assign(const int *cm, int n, Assign &w) {
  auto c = [cm, n](int i, int j) { return cm[(size_t)i * n + j]; };
  const int inf = INT_MAX;
  for (std::vector<int> *a : {&w.rowsol, &w.colsol, &w.v, &w.matches, &w.freer}) a->assign(n, 0);
  w.col.resize(n);
  w.d.resize(n);
  w.pred.resize(n);
  std::vector<int> &rowsol = w.rowsol, &colsol = w.colsol, &v = w.v, &matches = w.matches, &freer = w.freer,
                   &col = w.col, &d = w.d, &pred = w.pred;
  // column reduction, last column first
  for (int j = n - 1; j >= 0; --j) {
    int mn = c(0, j), imin = 0;
    for (int i = 1; i < n; ++i)
      if (c(i, j) < mn) mn = c(i, j), imin = i;
    v[j] = mn;
    if (++matches[imin] == 1) {
      rowsol[imin] = j;
      colsol[j] = imin;
    } else {
      colsol[j] = -1;
    }
  }
  // reduction transfer from rows assigned once
  int nfree = 0;
  for (int i = 0; i < n; ++i) {
    if (matches[i] == 0) {
      freer[nfree++] = i;
    } else if (matches[i] == 1) {
      const int j1 = rowsol[i];
      int mn = inf;
      for (int j = 0; j < n; ++j)
        if (j != j1 && c(i, j) - v[j] < mn) mn = c(i, j) - v[j];
      v[j1] -= mn;
    }
  }
  // augmenting row reduction, twice
  for (int round = 0; round < 2; ++round) {
    int k = 0;
    const int prev = nfree;
    nfree = 0;
    while (k < prev) {
      const int i = freer[k++];
      int umin = c(i, 0) - v[0], j1 = 0, j2 = 0, usub = inf;
      for (int j = 1; j < n; ++j) {
        const int hh = c(i, j) - v[j];
        if (hh < usub) {
          if (hh >= umin) {
            usub = hh;
            j2 = j;
          } else {
            usub = umin;
            umin = hh;
            j2 = j1;
            j1 = j;
          }
        }
      }
      int i0 = colsol[j1];
      if (umin < usub) {
        v[j1] -= usub - umin;
      } else if (i0 >= 0) {
        j1 = j2;
        i0 = colsol[j2];
      }
      rowsol[i] = j1;
      colsol[j1] = i;
      if (i0 >= 0) {
        if (umin < usub) freer[--k] = i0;
        else freer[nfree++] = i0;
      }
    }
  }
  // augmentation: a shortest path from each remaining free row
  for (int f = 0; f < nfree; ++f) {
    const int fr = freer[f];
    for (int j = 0; j < n; ++j) {
      d[j] = c(fr, j) - v[j];
      pred[j] = fr;
      col[j] = j;
    }
    int low = 0, up = 0, last = 0, end = 0, mn = 0;
    bool found = false;
    while (!found) {
      if (up == low) {
        last = low - 1;
        mn = d[col[up++]];
        for (int k = up; k < n; ++k) {
          const int j = col[k], hh = d[j];
          if (hh <= mn) {
            if (hh < mn) {
              up = low;
              mn = hh;
            }
            col[k] = col[up];
            col[up++] = j;
          }
        }
        for (int k = low; k < up; ++k)
          if (colsol[col[k]] < 0) {
            end = col[k];
            found = true;
            break;
          }
      }
      if (!found) {
        const int j1 = col[low++];
        const int i = colsol[j1];
        const int hh = c(i, j1) - v[j1] - mn;
        for (int k = up; k < n; ++k) {
          const int j = col[k];
          const int v2 = c(i, j) - v[j] - hh;
          if (v2 < d[j]) {
            pred[j] = i;
            if (v2 == mn) {
              if (colsol[j] < 0) {
                end = j;
                found = true;
                break;
              }
              col[k] = col[up];
              col[up++] = j;
            }
            d[j] = v2;
          }
        }
      }
    }
    for (int k = 0; k <= last; ++k) {
      const int j1 = col[k];
      v[j1] += d[j1] - mn;
    }
    int i;
    do {
      i = pred[end];
      colsol[end] = i;
      const int j1 = end;
      end = rowsol[i];
      rowsol[i] = j1;
    } while (i != fr);
  }
}

// ---- independent-set matching ----
struct Ism {
  Board &b;
  std::vector<RowBook> rows;
  int max_window = 90, max_cells = 128;  // ISMDP::initializeParams (detailed.cpp:761-774)
  bool double_window = false, independent = false;
  long long matchings = 0;
  std::vector<int> cost;  // work space of the assignment, reused
  Assign lap;
  std::vector<Board::Box> boxes;

  explicit Ism(Board &bb) : b(bb), rows(row_books(bb)) {}

  void sweep(int win, int overlap) {  // ISMSweep (detailed.cpp:264-300)
    if (overlap >= win) overlap = 0;
    const int nx = (int)((b.core_urx - b.core_llx) / b.rowh), ny = (int)((b.core_ury - b.core_lly) / b.rowh);
    for (int i = 0; i < ny; i += win - overlap)
      for (int j = 0; j < nx; j += win - overlap) {
        const float llx = (float)(b.core_llx + j * b.rowh), lly = (float)(b.core_lly + i * b.rowh);
        run(llx, lly, (float)(llx + win * b.rowh), (float)(lly + win * b.rowh));
      }
  }

  // ISMRun (detailed.cpp:336-749) on the window [llx, urx) x [lly, ury).
  void run(float llx, float lly, float urx, float ury) {
    const int r0 = std::max(b.row_of(lly), 0);
    const double top = std::min((double)ury, (double)b.core_ury - b.rowh);
    const int r1 = b.row_of((float)top);
    // cells by width, ascending; equal widths in insertion order
    std::multimap<double, int> by_width;
    for (int i = r0; i <= r1 && i < b.R; ++i) {
      auto a = rows[i].cells.lower_bound(llx), e = rows[i].cells.lower_bound(urx);
      for (; a != e; ++a)
        if (!b.macro[a->second]) by_width.insert({(double)b.w[a->second], a->second});
    }
    if (double_window) {
      const float ww = urx - llx;
      for (int i = r0; i < r1 && i < b.R; ++i) {
        auto a = rows[i].cells.lower_bound(urx + 5 * ww), e = rows[i].cells.lower_bound(urx + 6 * ww);
        for (; a != e; ++a)
          if (!b.macro[a->second]) by_width.insert({(double)b.w[a->second], a->second});
      }
    }
    std::list<int> pending;  // widest first
    for (auto it = by_width.rbegin(); it != by_width.rend(); ++it) pending.push_back(it->second);
    while (!pending.empty()) {
      std::vector<float> sx, sy;  // slots (lower-left)
      std::vector<int> who;       // the cell of each slot, -1 for free space
      std::vector<int> taken_row;
      std::vector<double> tail_x, tail_w;
      std::vector<int> tail_row;
      int taken = 0;
      const double wmax = b.w[pending.front()];
      for (auto it = pending.begin(); it != pending.end();) {
        const int o = *it;
        const float px = b.x[o], py = b.y[o];
        const int r = b.row_of(py);
        bool take = false;
        if (b.w[o] == wmax) {
          take = !independent || !conflicts(who, o);
        } else if (b.w[o] < wmax) {
          auto sp = rows[r].space.upper_bound(px);
          if (sp != rows[r].space.end() && sp->first == (double)(b.w[o] + px) && (b.w[o] + sp->second) >= wmax &&
              (!independent || !conflicts(who, o))) {
            const double tx = b.w[o] + px, tw = wmax - b.w[o];
            tail_x.push_back(tx);
            tail_w.push_back(tw);
            tail_row.push_back(r);
            rows[r].remove_tail(tx, tw);
            take = true;
          }
        }
        if (take) {
          who.push_back(o);
          sx.push_back(px);
          sy.push_back(py);
          taken_row.push_back(r);
          it = pending.erase(it);
          if (++taken >= max_cells) break;
        } else {
          ++it;
        }
      }
      if (taken < max_window) {
        int free_slots = 0;
        for (int i = r0; i < r1 && i < b.R; ++i) {
          auto a = rows[i].space.lower_bound(llx), e = rows[i].space.lower_bound(urx);
          for (; a != e; ++a) {
            const int k = (int)(a->second / wmax);
            for (int j = 0; j < k; ++j) {
              who.push_back(-1);
              sx.push_back((float)(a->first + j * wmax));
              sy.push_back((float)rows[i].lly);
              if (++free_slots + taken >= max_window) break;
            }
          }
        }
      }
      for (size_t i = 0; i < tail_x.size(); ++i) rows[tail_row[i]].insert_space(tail_x[i], tail_w[i]);
      for (int i = 0; i < taken; ++i) rows[taken_row[i]].remove_cell(b, who[i]);
      const int n = (int)who.size();
      cost.assign((size_t)n * n, 0);
      for (int i = 0; i < n; ++i) {
        if (who[i] < 0) continue;
        const int o = who[i];
        float bx = sx[0], by = sy[0];
        double best = kDoubleMax;
        b.other_boxes(o, boxes);
        for (int j = 0; j < n; ++j) {
          b.set_ll(o, sx[j], sy[j]);
          const double wl = b.moved_hpwl(o, boxes);
          if (wl < best) {
            best = wl;
            bx = sx[j];
            by = sy[j];
          }
          cost[(size_t)i * n + j] = (int)wl;
        }
        b.set_ll(o, bx, by);  // later cells are costed with this one at its best slot
      }
      assign(cost.data(), n, lap);
      const std::vector<int> &sol = lap.rowsol;
      ++matchings;
      for (int i = 0; i < n; ++i) {
        if (who[i] < 0) break;
        const int j = sol[i];
        const int r = b.row_of(sy[j]);
        rows[r].insert_cell(b, sx[j], who[i]);
        b.set_ll(who[i], sx[j], sy[j]);
      }
    }
  }

  bool conflicts(const std::vector<int> &chosen, int o) const {
    for (int c : chosen)
      if (c >= 0 && b.connected(c, o)) return true;
    return false;
  }
};

// DetailedPlacer::runISM (detailed.cpp:102-216): sweeps of shrinking windows
// while a sweep gains at least 0.2%, then with independent cells and a
// second window until one gains less than 0.6%.
long long run_ism(Board &b, double time_limit) {
  Ism m(b);
  const auto t0 = std::chrono::steady_clock::now();
  double prev = b.hpwl();
  const double stop = -0.2;
  for (int k = 0;; ++k) {
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!(el < time_limit)) break;
    int win = 20 - k;
    if (win < 15) win = 15 + k % 5;
    int overlap = 2 + k / 2;
    if (overlap > 8) overlap = 8;
    m.sweep(win, overlap);
    const double cur = b.hpwl();
    const double gain = 100.0 * (cur / prev - 1.0);
    if (gain > stop * 3 && m.independent) break;
    if (gain > stop && !m.double_window) {
      m.double_window = true;
      m.independent = true;
    }
    prev = cur;
  }
  return m.matchings;
}

// ---- local reordering ----
struct Segment {
  double bottom, start, end;
  std::vector<int> cells;
};

struct Reorder {
  Board &b;
  std::vector<Segment> segs;
  long long windows = 0;

  explicit Reorder(Board &bb) : b(bb) {
    // LocalReorderingDP::initializeSegments (detailed.cpp:1519-1604)
    for (int r = 0; r < b.R; ++r)
      for (const Interval &t : b.iv[r]) segs.push_back(Segment{b.row_bottom[r], (double)t.start, (double)t.end, {}});
    for (int o = 0; o < b.N; ++o) {
      if (b.macro[o]) continue;
      const double cb = b.y[o], cs = b.x[o], ce = (double)(b.x[o] + b.w[o]);
      // first segment not before the cell (compareLRSegment, detailed.h)
      auto it = std::lower_bound(segs.begin(), segs.end(), 0, [&](const Segment &s, int) {
        if (s.bottom == cb) return s.start <= cs ? s.end < ce : false;
        return s.bottom < cb;
      });
      if (it != segs.end() && it->bottom == cb) it->cells.push_back(o);
    }
    for (Segment &s : segs)
      std::sort(s.cells.begin(), s.cells.end(), [&](int a, int c) { return f_less(b.x[a], b.x[c]); });
  }

  // LocalReorderingDP::solve (detailed.cpp:1436-1517)
  void solve(int win, int overlap) {
    for (Segment &s : segs) {
      const int n = (int)s.cells.size();
      if (n < 1) continue;
      if (n == 1) {
        const int o = s.cells[0];
        const float ox = b.x[o], oy = b.y[o];
        const double wo = b.module_hpwl(o);
        b.set_ll(o, (float)s.start, oy);
        const double wl = b.module_hpwl(o);
        const float rx = (float)(s.end - b.w[o]);
        b.set_ll(o, rx, oy);
        const double wr = b.module_hpwl(o);
        if (wo <= wl && wo <= wr) b.set_ll(o, ox, oy);
        else if (wl <= wo && wl <= wr) b.set_ll(o, (float)s.start, oy);
        else b.set_ll(o, rx, oy);
        continue;
      }
      if (n > win) {
        for (int a = 0; a + win <= n; a += win - overlap) best_order(s.cells, a, a + win);
      } else {
        best_order(s.cells, 0, n);
      }
    }
  }

  // solveForBestOrder + LRSolver (detailed.cpp:1606-1830, 1955-2051): the
  // cells [a, e) of v and, when their span has free space, one block of it.
  struct Search {
    Board &b;
    std::vector<int> cells;         // window cells; -1 = the free block
    std::vector<int> nets;          // nets touching the window (one entry per pin of a window cell)
    std::map<int, int> count0;      // net -> pins of window cells on it
    double space = 0;
    double best = kDoubleMax;
    std::vector<int> best_cells;
    std::vector<double> best_x;
  };

  void best_order(std::vector<int> &v, int a, int e) {
    ++windows;
    Search S{b, {}, {}, {}, 0, kDoubleMax, {}, {}};
    // the current order as the first complete solution
    std::set<int> netset;
    for (int k = a; k < e; ++k)
      for (int p = b.npin_start[v[k]]; p < b.npin_start[v[k] + 1]; ++p) netset.insert(b.pin_net[b.npin[p]]);
    double c0 = 0;
    for (int net : netset) c0 += b.net_hpwl(net);
    S.best = c0;
    for (int k = a; k < e; ++k) {
      S.best_cells.push_back(v[k]);
      S.best_x.push_back(b.x[v[k]]);
    }
    float mnx = kFloatMax, mxx = -kFloatMax, tw = 0;
    for (int k = a; k < e; ++k) {
      const int o = v[k];
      S.cells.push_back(o);
      for (int p = b.npin_start[o]; p < b.npin_start[o + 1]; ++p) S.count0[b.pin_net[b.npin[p]]] += 1;
      mnx = std::min(mnx, b.x[o]);
      mxx = std::max(mxx, b.x[o] + b.w[o]);
      tw += b.w[o];
    }
    if (tw < mxx - mnx) {
      S.space = mxx - mnx - tw;
      S.cells.push_back(-1);
    }
    std::vector<int> placed;
    std::vector<double> xs;
    std::map<int, int> count = S.count0;
    std::vector<int> rest = S.cells;
    dfs(S, rest, placed, xs, count, 0.0, mnx);
    // apply the best order (the current one when nothing beat it)
    int w = a;
    for (size_t k = 0; k < S.best_cells.size(); ++k) {
      const int o = S.best_cells[k];
      if (o < 0) continue;
      b.set_ll(o, (float)S.best_x[k], b.y[o]);
      v[w++] = o;
    }
  }

  void dfs(Search &S, std::vector<int> &rest, std::vector<int> &placed, std::vector<double> &xs,
           std::map<int, int> &count, double cost, double cx) {
    if (rest.empty()) {
      if (cost < S.best) {
        S.best = cost;
        S.best_cells = placed;
        S.best_x = xs;
      }
      return;
    }
    for (size_t k = 0; k < rest.size(); ++k) {
      const int o = rest[k];
      std::map<int, int> cnt = count;
      double c = cost, nx;
      if (o >= 0) {
        b.set_ll(o, (float)cx, b.y[o]);
        nx = cx + b.w[o];
        for (int p = b.npin_start[o]; p < b.npin_start[o + 1]; ++p) {
          const int net = b.pin_net[b.npin[p]];
          if (--cnt[net] == 0) c += b.net_hpwl(net);
        }
      } else {
        nx = cx + S.space;
      }
      if (c < S.best) {
        std::vector<int> r2;
        r2.reserve(rest.size() - 1);
        for (int q : rest)
          if (q != o) r2.push_back(q);
        placed.push_back(o);
        xs.push_back(cx);
        dfs(S, r2, placed, xs, cnt, c, nx);
        placed.pop_back();
        xs.pop_back();
      }
    }
  }
};

// ---- global swap ----
struct Swap {
  Board &b;
  std::vector<RowBook> rows;
  int max_slots = 30;  // GlobalSwapDP::initializeParams (detailed.cpp:2317-2325)
  long long moved = 0;
  std::vector<Board::Box> boxes;

  explicit Swap(Board &bb) : b(bb), rows(row_books(bb)) {}

  // PlaceDB::getOptimialRegion (placedb.cpp:1480-1533); false for a cell
  // without nets.
  bool optimal_region(int o, float &l, float &bt, float &r, float &t) const {
    std::vector<float> xs, ys;
    for (int k = b.npin_start[o]; k < b.npin_start[o + 1]; ++k) {
      const int e = b.pin_net[b.npin[k]];
      double mnx = kDoubleMax, mxx = -kDoubleMax, mny = kDoubleMax, mxy = -kDoubleMax;
      for (int64_t q = b.d->net_start[e]; q < b.d->net_start[e + 1]; ++q) {
        if (b.d->pin_obj[q] == o) continue;
        const double px = b.pin_x(q), py = b.pin_y(q);
        mnx = std::min(mnx, px);
        mxx = std::max(mxx, px);
        mny = std::min(mny, py);
        mxy = std::max(mxy, py);
      }
      xs.push_back((float)mnx);
      xs.push_back((float)mxx);
      ys.push_back((float)mny);
      ys.push_back((float)mxy);
    }
    if (xs.empty()) return false;
    std::sort(xs.begin(), xs.end());
    std::sort(ys.begin(), ys.end());
    l = xs[xs.size() / 2 - 1];
    r = xs[xs.size() / 2];
    bt = ys[ys.size() / 2 - 1];
    t = ys[ys.size() / 2];
    return true;
  }

  // GlobalSwapDP::solve (detailed.cpp:2053-2306)
  void solve() {
    for (int o = 0; o < b.N; ++o) {
      if (b.macro[o]) continue;
      const float cw = b.w[o], ox = b.x[o], oy = b.y[o];
      float l, bt, r, t;
      if (!optimal_region(o, l, bt, r, t)) continue;
      if (f_greaterorequal(bt, b.core_ury) || f_lessorequal(t, b.core_lly) || f_greaterorequal(l, b.core_urx) ||
          f_lessorequal(r, b.core_llx))
        continue;
      l = std::max(l, b.core_llx);
      bt = std::max(bt, b.core_lly);
      r = std::min(r, b.core_urx);
      t = std::min(t, b.core_ury);
      const double st = b.site_step;
      l = (float)(std::floor((l - b.core_llx) / st) * st + b.core_llx);
      r = (float)(std::ceil((r - b.core_llx) / st) * st + b.core_llx);
      bt = (float)(std::floor((bt - b.core_lly) / b.rowh) * b.rowh + b.core_lly);
      t = (float)(std::ceil((t - b.core_lly) / b.rowh) * b.rowh + b.core_lly);
      if (f_equal(l, r) || f_equal(bt, t)) continue;
      if (ox >= l && ox <= r && oy >= bt && oy <= t) continue;  // CRect::inside (global.h:314-317)
      const int r0 = std::max(b.row_of(bt), 0);
      const double top = std::min((double)t, (double)b.core_ury - b.rowh);
      const int r1 = b.row_of((float)top);
      std::vector<float> ex, ey;
      int empty = 0;
      for (int i = r0; i < r1 && i < b.R; ++i) {
        auto a = rows[i].space.lower_bound(l), e = rows[i].space.lower_bound(r);
        for (; a != e; ++a) {
          const int k = (int)(a->second / cw);
          for (int j = 0; j < k; ++j) {
            if (++empty > max_slots) break;
            ex.push_back((float)(a->first + j * cw));
            ey.push_back((float)rows[i].lly);
          }
        }
      }
      std::vector<int> cand;
      std::vector<float> cx, cy;
      if (empty < max_slots) {
        std::vector<int> same;  // equal-width cells, in row and x order
        for (int i = r0; i <= r1 && i < b.R; ++i) {
          auto a = rows[i].cells.lower_bound(l), e = rows[i].cells.lower_bound(r);
          for (; a != e; ++a)
            if (!b.macro[a->second] && f_equal(cw, b.w[a->second])) same.push_back(a->second);
        }
        // a multimap on the (equal) width, read in reverse: last inserted first
        int n_swap = 0;
        for (auto it = same.rbegin(); it != same.rend(); ++it) {
          if (empty + ++n_swap > max_slots) break;
          cand.push_back(*it);
          cx.push_back(b.x[*it]);
          cy.push_back(b.y[*it]);
        }
      }
      const int orow = b.row_of(oy);
      rows[orow].remove_cell(b, o);
      const double old = b.module_hpwl(o);
      bool done = false;
      if (!ex.empty()) b.other_boxes(o, boxes);
      for (size_t s = 0; s < ex.size(); ++s) {
        b.set_ll(o, ex[s], ey[s]);
        if (old > b.moved_hpwl(o, boxes)) {
          done = true;
          rows[b.row_of(ey[s])].insert_cell(b, ex[s], o);
          break;
        }
      }
      if (!done) {
        for (size_t s = 0; s < cand.size(); ++s) {
          const int q = cand[s];
          const double old2 = old + b.module_hpwl(q);
          const int srow = b.row_of(cy[s]);
          rows[srow].remove_cell(b, q);
          b.set_ll(o, cx[s], cy[s]);
          b.set_ll(q, ox, oy);
          if (old2 > b.module_hpwl(o) + b.module_hpwl(q)) {
            done = true;
            rows[srow].insert_cell(b, cx[s], o);
            rows[orow].insert_cell(b, ox, q);
            break;
          }
          rows[srow].insert_cell(b, cx[s], q);
          b.set_ll(q, cx[s], cy[s]);
        }
      }
      if (!done) {
        b.set_ll(o, ox, oy);
        rows[orow].insert_cell(b, ox, o);
      } else {
        ++moved;
      }
    }
  }
};

}  // namespace

extern "C" {

int ep_detailed_place(const ep_design *d, const float *node_xy, const ep_dp_params *p, float *out_xy,
                      ep_dp_stats *stats) {
  std::string &err = host_error();
  err.clear();
  if (!d || !p || !out_xy || d->num_nodes <= 0 || d->num_rows <= 0 || !d->net_start || !d->pin_obj || !d->pin_dx ||
      !d->pin_dy || !d->node_w || !d->node_h || !d->row_bottom || !d->row_height || !d->row_x0 || !d->row_x1)
    return err = "bad arguments", EP_ERR_ARG;
  if (!(p->site_step > 0) || p->passes < 0) return err = "bad site_step or passes", EP_ERR_ARG;
  const auto t0 = std::chrono::steady_clock::now();
  Board b;
  b.d = d;
  b.N = d->num_nodes;
  b.T = d->num_terminals;
  b.E = d->num_nets;
  b.R = d->num_rows;
  Geometry g = geometry(d, node_xy);
  b.x = g.llx;
  b.y = g.lly;
  b.w = g.w;
  b.h = g.h;
  b.macro = g.macro;
  b.core_llx = g.core_llx;
  b.core_lly = g.core_lly;
  b.core_urx = g.core_urx;
  b.core_ury = g.core_ury;
  b.rowh = d->common_row_height;
  b.site_step = p->site_step;
  b.row_bottom.assign(d->row_bottom, d->row_bottom + b.R);
  b.tcx.resize(b.T);
  b.tcy.resize(b.T);
  for (int t = 0; t < b.T; ++t) {
    b.tcx[t] = d->term_x[t];
    b.tcy[t] = d->term_y[t];
  }
  const int64_t P = d->num_pins;
  b.pin_net.resize(P);
  b.npin_start.assign(b.N + 1, 0);
  for (int e = 0; e < b.E; ++e)
    for (int64_t q = d->net_start[e]; q < d->net_start[e + 1]; ++q) {
      const int o = d->pin_obj[q];
      if (o < 0 || o >= b.N + b.T) return err = "pin owner out of range", EP_ERR_ARG;
      b.pin_net[q] = e;
      if (o < b.N) ++b.npin_start[o + 1];
    }
  for (int o = 0; o < b.N; ++o) b.npin_start[o + 1] += b.npin_start[o];
  b.npin.resize(b.npin_start[b.N]);
  {
    std::vector<int> cur(b.npin_start.begin(), b.npin_start.end() - 1);
    for (int64_t q = 0; q < P; ++q)
      if (d->pin_obj[q] < b.N) b.npin[cur[d->pin_obj[q]]++] = (int)q;
  }
  // the moves index rows by each cell's lower edge: the input must be a
  // legal placement (every standard cell inside the core, on a row)
  for (int o = 0; o < b.N; ++o) {
    if (b.macro[o]) continue;
    const int r = b.row_of(b.y[o]);
    if (!(b.x[o] >= b.core_llx && b.x[o] + b.w[o] <= b.core_urx) || r < 0 || r >= b.R ||
        (double)b.y[o] != b.row_bottom[r])
      return err = "node " + std::to_string(o) + " is not on a row inside the core: legalise before detailed placement",
             EP_ERR_ARG;
  }
  // DetailedPlacer::initialization: the interval removal once more after the
  // legaliser's (placedb.cpp:629-849 applied twice)
  if (!row_intervals(d, g, p->site_step, b.iv, err, 2)) return EP_ERR_ARG;
  ep_dp_stats st = {};
  st.hpwl_before = b.hpwl();
  const double limit = p->ism_time_limit > 0 ? p->ism_time_limit : 28800.0;
  auto since = [](std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
  };
  for (int pass = 0; pass < p->passes; ++pass) {
    auto t = std::chrono::steady_clock::now();
    Reorder lr(b);
    for (int k = 0; k < 2; ++k) lr.solve(3, 2);
    st.lr_windows += lr.windows;
    if (pass < EP_DP_MAX_PASSES) st.hpwl_after_lr[pass] = b.hpwl();
    st.seconds_lr += since(t);
    t = std::chrono::steady_clock::now();
    st.ism_matchings += run_ism(b, limit);
    if (pass < EP_DP_MAX_PASSES) st.hpwl_after_ism[pass] = b.hpwl();
    st.seconds_ism += since(t);
    t = std::chrono::steady_clock::now();
    Swap gs(b);
    for (int k = 0; k < 2; ++k) gs.solve();
    st.gs_moves += gs.moved;
    if (pass < EP_DP_MAX_PASSES) st.hpwl_after_gs[pass] = b.hpwl();
    st.seconds_gs += since(t);
  }
  st.hpwl_after = b.hpwl();
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int o = 0; o < b.N; ++o) {
    out_xy[2 * o] = b.x[o] + (float)0.5 * b.w[o];  // Module centre (objects.cpp:518-527)
    out_xy[2 * o + 1] = b.y[o] + (float)0.5 * b.h[o];
  }
  if (stats) *stats = st;
  return EP_OK;
}

}  // extern "C"
