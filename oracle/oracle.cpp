// CPU oracle for the ePlace global-placement hot path.  TEST INFRASTRUCTURE:
// only tests/, the package's smoke check and the bench's CPU-baseline leg load
// it, always as the checker (or the timed CPU baseline), never as the product.
//
// It restates the reference's algorithm (Xinyu-lumos/ePlace) function by
// function; each function cites the reference file:line it follows.  Two
// arithmetic modes, chosen per placer:
//   T = float   "reference" mode: the reference's own float/double mix, same
//               operation order (x86-64 SSE, no FMA contraction), so maps,
//               gradients and scalars match the reference bit for bit except
//               where the reference's Ooura FFT is involved (here: a DCT by
//               FFT in float, different rounding).
//   T = double  "fp64" mode: the same formulas evaluated in double from the
//               same float geometry; the accuracy yardstick for the GPU path.
// Independently of the mode, scatter_fixed() sums the reference's per
// (object, bin) float contribution as int64 fixed point: order-free, so the
// GPU's atomic scatter must match it exactly.
//
// Built with g++ -O2 -ffp-contract=off (oracle/Makefile).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

const double kEps = 1.0e-15;        // global.h:39
const int kNegMaxExp = -300;        // global.h:46

// ---------------------------------------------------------------------------
// glibc rand() restated (TYPE_3 additive feedback generator, srand(seed)).
// The reference never seeds rand() (placedb.cpp:353-379), i.e. seed 1.
// Published algorithm: r0 = seed; r_i = 16807 r_{i-1} mod (2^31-1), i < 31;
// r_i = r_{i-31}, 31 <= i < 34; r_i = r_{i-31} + r_{i-3} (mod 2^32) after;
// output k is r_{k+344} >> 1.
// ---------------------------------------------------------------------------
struct GlibcRand {
  std::vector<uint32_t> r;
  size_t k = 0;
  explicit GlibcRand(uint32_t seed) {
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
    r.push_back(r[i - 31] + r[i - 3]);
    return (int32_t)(r.back() >> 1);
  }
};

// ---------------------------------------------------------------------------
// Geometry helpers (global.h:343-545).
// ---------------------------------------------------------------------------
inline bool float_less(float a, float b) { return (double)(a - b) < -1.0 * kEps; }
inline bool float_greater(float a, float b) { return (double)(a - b) > 1.0 * kEps; }
inline bool float_equal(float a, float b) { return std::fabs((double)(a - b)) < kEps; }
inline bool dless(double a, double b) { return a - b < -1.0 * kEps; }

// getOverlap / getOverlapArea (global.h:477-518): the signed lengths are
// negative when there is overlap; their product is the area.
inline double overlap_len(double x1, double x2, double x3, double x4) {
  double s = std::max(x1, x3), e = std::min(x2, x4);
  return s >= e ? 0.0 : s - e;
}
inline double overlap_area(double l1, double b1, double r1, double t1, double l2, double b2,
                           double r2, double t2) {
  double h = overlap_len(l1, r1, l2, r2);
  if (h == 0) return 0;
  double v = overlap_len(b1, t1, b2, t2);
  if (v == 0) return 0;
  return h * v;
}

// fastExp (global.h:457-471).
template <class T>
inline T fast_exp(T a) {
  a = (T)1 + a / (T)1024;
  for (int i = 0; i < 10; ++i) a *= a;
  return a;
}

// ---------------------------------------------------------------------------
// Radix-2 complex FFT and the DCT-II / DCT-III / DST-III built on it
// (Makhoul's even/odd reordering).  Used for the Poisson solve at sizes where
// the O(N^3) matrix form is too slow.
// ---------------------------------------------------------------------------
// Twiddles e^{-2 pi i k / n}, k < n/2, computed in double, cached per n.
template <class T>
const std::vector<std::complex<T>> &twiddles(size_t n) {
  static thread_local std::vector<std::complex<T>> tw[32];
  int lg = 0;
  while (((size_t)1 << lg) < n) ++lg;
  std::vector<std::complex<T>> &t = tw[lg];
  if (t.size() != n / 2) {
    t.resize(n / 2);
    for (size_t k = 0; k < n / 2; ++k)
      t[k] = std::complex<T>((T)std::cos(2 * M_PI * k / n), (T)-std::sin(2 * M_PI * k / n));
  }
  return t;
}

template <class T>
void fft_inplace(std::vector<std::complex<T>> &a, bool inverse) {
  size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  const std::vector<std::complex<T>> &tw = twiddles<T>(n);
  for (size_t len = 2; len <= n; len <<= 1) {
    size_t step = n / len;
    for (size_t i = 0; i < n; i += len)
      for (size_t j = 0; j < len / 2; ++j) {
        std::complex<T> w = tw[j * step];
        if (inverse) w = std::conj(w);
        std::complex<T> u = a[i + j], v = a[i + j + len / 2] * w;
        a[i + j] = u + v;
        a[i + j + len / 2] = u - v;
      }
  }
}

// X[k] = sum_n x[n] cos(pi k (n + 1/2) / N)
template <class T>
void dct2(const T *x, T *X, int N, int stride_in, int stride_out) {
  std::vector<std::complex<T>> v(N);
  for (int n = 0; n < N / 2; ++n) {
    v[n] = x[(size_t)(2 * n) * stride_in];
    v[N - 1 - n] = x[(size_t)(2 * n + 1) * stride_in];
  }
  fft_inplace(v, false);
  for (int k = 0; k < N; ++k) {
    double a = -M_PI * k / (2.0 * N);
    X[(size_t)k * stride_out] = (T)(v[k].real() * (T)std::cos(a) - v[k].imag() * (T)std::sin(a));
  }
}

// y[n] = sum_k X[k] cos(pi k (n + 1/2) / N)
template <class T>
void dct3(const T *X, T *y, int N, int stride_in, int stride_out) {
  std::vector<std::complex<T>> w(N);
  for (int k = 0; k < N; ++k) {
    T xk = X[(size_t)k * stride_in];
    T xnk = k == 0 ? (T)0 : X[(size_t)(N - k) * stride_in];
    double a = M_PI * k / (2.0 * N);
    std::complex<T> e((T)std::cos(a), (T)std::sin(a));
    w[k] = e * std::complex<T>(xk, -xnk);
  }
  fft_inplace(w, true);
  T x0 = X[0];
  for (int n = 0; n < N / 2; ++n) {
    y[(size_t)(2 * n) * stride_out] = (T)0.5 * w[n].real() + (T)0.5 * x0;
    y[(size_t)(2 * n + 1) * stride_out] = (T)0.5 * w[N - 1 - n].real() + (T)0.5 * x0;
  }
}

// s[n] = sum_{k>=1} A[k] sin(pi k (n + 1/2) / N) = (-1)^n DCT3(B), B[k] = A[N-k]
template <class T>
void dst3(const T *A, T *s, int N, int stride_in, int stride_out) {
  std::vector<T> B(N), y(N);
  B[0] = 0;
  for (int k = 1; k < N; ++k) B[k] = A[(size_t)(N - k) * stride_in];
  dct3(B.data(), y.data(), N, 1, 1);
  for (int n = 0; n < N; ++n) s[(size_t)n * stride_out] = (n & 1) ? -y[n] : y[n];
}

bool is_pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

}  // namespace

// ===========================================================================
// Placer state.  Object order is the reference's ePlaceNodesAndFillers:
// dbNodes then fillers (eplace.cpp:190-191).
// ===========================================================================
extern "C" {

typedef struct orc_design {
  int32_t num_nodes, num_terminals, num_nets, num_rows;
  int64_t num_pins;
  const float *node_w, *node_h, *node_x, *node_y;
  const float *term_w, *term_h, *term_x, *term_y;
  const uint8_t *term_ni;
  const int64_t *net_start;
  const int32_t *pin_obj;
  const float *pin_dx, *pin_dy;
  const double *row_bottom, *row_height, *row_x0, *row_x1;
  double common_row_height;
} orc_design;

}  // extern "C"

namespace {

struct PlacerBase {
  virtual ~PlacerBase() {}
  int mode = 0;  // 0 reference float, 1 fp64
};

template <class T>
struct Placer : PlacerBase {
  // ---- design ----
  int N = 0, F = 0, O = 0, NT = 0, E = 0;
  int64_t P = 0;
  std::vector<float> w, h, area;     // [O]
  std::vector<uint8_t> macro, filler;
  std::vector<int> npins;            // modulePins.size()
  std::vector<float> cx, cy, llx, lly;  // centres and lower-left (Module::center/coor)
  std::vector<float> tcx, tcy, tw, th, tllx, tlly;
  std::vector<uint8_t> tni;
  std::vector<int64_t> net_start;
  std::vector<int32_t> pin_obj;
  std::vector<float> pin_dx, pin_dy;
  std::vector<int64_t> node_pin_start;
  std::vector<int64_t> node_pins;    // pins of each node in modulePins order
  std::vector<double> row_bottom, row_height, row_x0, row_x1;
  float core_llx = 0, core_lly = 0, core_urx = 0, core_ury = 0, total_row_area = 0;
  double common_row_height = 1;
  float td = 1;
  float std_area = 0, macro_area = 0;
  int exp_mode = 0;  // 0 fastExp (reference), 1 exp
  int wl_model = 0;  // 0 WA, 1 LSE
  // ---- bins ----
  int nx = 0, ny = 0;
  float bsx = 0, bsy = 0, bin_area = 0;
  std::vector<float> bllx, burx, blly, bury;
  std::vector<T> term_map, base_map, node_map, filler_map, rho, phi, ex, ey;
  std::vector<int64_t> node_q, filler_q;
  int fixed_shift = 0;
  // ---- gradients and scalars ----
  std::vector<T> gdx, gdy, gwx, gwy, gtx, gty;
  std::vector<T> pin_gx, pin_gy;
  T overflow = 0, invgx = 0, invgy = 0, lambda = 0;
  double last_hpwl = 0, hpwl_bound = 0, wl_value = 0;
  int stage = 0;  // 0 mGP, 1 FILLERONLY, 2 cGP
  int poisson_fast = 1;
  int mgp_iterations = 0;
  // ---- optimiser (nesterov.hpp NSIter, a_k) ----
  std::vector<T> cur_main, cur_ref, cur_grad, last_ref, last_grad;
  T ns_a = 1;
  int iter = 0;
  T last_step = 0;
  int bktrk_count = 0;
  // backtracking acceptance factor (BKTRK_EPS, nesterov.hpp:37); settable only
  // so that tests can plant a wrong value and see the trajectory check fail
  double bktrk_eps = 0.95;
  int bb_short_used = 0;  // the last BB step was the short step
  // DELTA_HPWL_REF (eplace.h:11); settable only for experiments on how the
  // absolute reference throttles lambda on large designs
  double delta_hpwl_ref = 350000;

  // ---------------------------------------------------------------------
  void load(const orc_design *d, float target_density) {
    td = target_density;
    N = d->num_nodes;
    NT = d->num_terminals;
    E = d->num_nets;
    P = d->num_pins;
    common_row_height = d->common_row_height;
    w.assign(d->node_w, d->node_w + N);
    h.assign(d->node_h, d->node_h + N);
    cx.assign(d->node_x, d->node_x + N);
    cy.assign(d->node_y, d->node_y + N);
    for (int i = 0; i < N; ++i) {
      area.push_back(w[i] * h[i]);  // Module ctor: float_mul (objects.h:162-173)
      macro.push_back(float_greater(h[i], (float)common_row_height) ? 1 : 0);  // placedb.cpp:102-111
      filler.push_back(0);
      llx.push_back(cx[i] - (float)0.5 * w[i]);
      lly.push_back(cy[i] - (float)0.5 * h[i]);
    }
    tcx.assign(d->term_x, d->term_x + NT);
    tcy.assign(d->term_y, d->term_y + NT);
    tw.assign(d->term_w, d->term_w + NT);
    th.assign(d->term_h, d->term_h + NT);
    tni.assign(d->term_ni, d->term_ni + NT);
    for (int i = 0; i < NT; ++i) {
      tllx.push_back(tcx[i] - (float)0.5 * tw[i]);
      tlly.push_back(tcy[i] - (float)0.5 * th[i]);
    }
    net_start.assign(d->net_start, d->net_start + E + 1);
    pin_obj.assign(d->pin_obj, d->pin_obj + P);
    pin_dx.assign(d->pin_dx, d->pin_dx + P);
    pin_dy.assign(d->pin_dy, d->pin_dy + P);
    npins.assign(N, 0);
    for (int64_t p = 0; p < P; ++p)
      if (pin_obj[p] < N) npins[pin_obj[p]]++;
    node_pin_start.assign(N + 1, 0);
    for (int i = 0; i < N; ++i) node_pin_start[i + 1] = node_pin_start[i] + npins[i];
    node_pins.assign(node_pin_start[N], 0);
    std::vector<int64_t> fill(node_pin_start.begin(), node_pin_start.end() - 1);
    for (int64_t p = 0; p < P; ++p)
      if (pin_obj[p] < N) node_pins[fill[pin_obj[p]]++] = p;
    row_bottom.assign(d->row_bottom, d->row_bottom + d->num_rows);
    row_height.assign(d->row_height, d->row_height + d->num_rows);
    row_x0.assign(d->row_x0, d->row_x0 + d->num_rows);
    row_x1.assign(d->row_x1, d->row_x1 + d->num_rows);
    // PlaceDB::setCoreRegion (placedb.cpp:52-77); rows hold float start/end.
    float bottom = (float)row_bottom.front();
    float top = (float)(row_bottom.back() + row_height.back());
    float left = (float)row_x0.front(), right = (float)row_x1.front();
    total_row_area = 0;
    for (size_t r = 0; r < row_bottom.size(); ++r) {
      left = std::min(left, (float)row_x0[r]);
      right = std::max(right, (float)row_x1[r]);
      float a = (float)((double)((float)row_x1[r] - (float)row_x0[r]) * row_height[r]);
      total_row_area += a;
    }
    core_llx = left;
    core_lly = bottom;
    core_urx = right;
    core_ury = top;
  }

  // SiteRow::getLL_2D / getUR_2D (objects.cpp:542-553): float start, y + double height.
  float row_lly(size_t r) const { return (float)row_bottom[r]; }
  float row_ury(size_t r) const { return (float)((double)(float)row_bottom[r] + row_height[r]); }

  // PlaceDB::setModuleLocation_2D (placedb.cpp:200-240) for a movable object.
  void set_location(int i, float x, float y) {
    if (x < core_llx) x = core_llx;
    if (x + w[i] > core_urx) x = (float)((double)(core_urx - w[i]) - kEps);
    if (y < core_lly) y = core_lly;
    if (y + h[i] > core_ury) y = (float)((double)(core_ury - h[i]) - kEps);
    llx[i] = x;
    lly[i] = y;
    cx[i] = llx[i] + (float)0.5 * w[i];  // Module::setLocation_2D (objects.cpp:518-527)
    cy[i] = lly[i] + (float)0.5 * h[i];
  }

  // PlaceDB::setModuleCenter_2D (placedb.cpp:253-288) + Module::setCenter_2D.
  void set_center(int i, float x, float y) {
    double hw = 0.5 * (double)w[i], hh = 0.5 * (double)h[i];
    if ((double)x - hw < (double)core_llx) x = (float)((double)core_llx + hw + kEps);
    if ((double)x + hw > (double)core_urx) x = (float)((double)core_urx - hw - kEps);
    if ((double)y - hh < (double)core_lly) y = (float)((double)core_lly + hh + kEps);
    if ((double)y + hh > (double)core_ury) y = (float)((double)core_ury - hh - kEps);
    cx[i] = x;
    cy[i] = y;
    llx[i] = cx[i] - (float)0.5 * w[i];
    lly[i] = cy[i] - (float)0.5 * h[i];
  }

  // EPlacer_2D::fillerInitialization (eplace.cpp:61-208).
  // filler_xy != null: use those centres instead of the rand() sequence.
  void filler_init(uint32_t seed, const float *filler_xy, int filler_count_override) {
    float total_overlap = 0;
    for (int t = 0; t < NT; ++t) {
      if (tni[t]) continue;
      for (size_t r = 0; r < row_bottom.size(); ++r)
        total_overlap = (float)((double)total_overlap +
                                overlap_area(tllx[t], tlly[t], tllx[t] + tw[t], tlly[t] + th[t],
                                             (float)row_x0[r], row_lly(r), (float)row_x1[r], row_ury(r)));
    }
    float ws = total_row_area - total_overlap;
    float sa = 0, ma = 0;
    for (int i = 0; i < N; ++i) {
      if (macro[i]) ma += area[i];
      else sa += area[i];
    }
    std_area = sa;
    macro_area = ma;
    float node_scaled = sa + ma * td;
    float total_filler = ws * td - node_scaled;
    std::vector<float> sorted(area.begin(), area.begin() + N);
    std::sort(sorted.begin(), sorted.end());
    int lo = (int)(0.05 * (float)N), hi = (int)(0.95 * (float)N);
    float tot = 0;
    for (int i = lo; i < hi; ++i) tot += sorted[i];
    float favg = tot / (float)(hi - lo);
    float fh = (float)common_row_height;
    float fw = favg / fh;
    int count = (int)(total_filler / favg + 0.5);
    if (filler_count_override >= 0) count = filler_count_override;
    if (count < 0) count = 0;
    F = count;
    O = N + F;
    GlibcRand rng(seed);
    float rinv = (float)1.0 / (float)2147483647;
    for (int f = 0; f < F; ++f) {
      w.push_back(fw);
      h.push_back(fh);
      area.push_back(fw * fh);
      macro.push_back(0);
      filler.push_back(1);
      npins.push_back(0);
      cx.push_back(0);
      cy.push_back(0);
      llx.push_back(0);
      lly.push_back(0);
      int i = N + f;
      if (filler_xy) {
        set_center(i, filler_xy[2 * f], filler_xy[2 * f + 1]);
      } else {
        // PlaceDB::setModuleLocation_2D_random (placedb.cpp:353-379)
        float x = (float)rng.next();
        float y = (float)rng.next();
        float pw = core_urx - core_llx, ph = core_ury - core_lly;
        ph -= h[i];
        pw -= w[i];
        x = x * rinv * pw;
        y = y * rinv * ph;
        set_location(i, x, y);
      }
    }
  }

  // EPlacer_2D::binInitialization (eplace.cpp:221-412).  bins <= 0: rule.
  void bin_init(int bins_x, int bins_y) {
    float node_area = std_area + macro_area;
    float cw = core_urx - core_llx, ch = core_ury - core_lly;
    float core_area = cw * ch;
    float avg = (float)(1.0 * node_area / N);
    float ideal_area = avg / td;
    int ideal = (int)(1.0 * (core_area / ideal_area) + 0.5f);
    int dim = 1024;
    for (int i = 1; i < 10; ++i)
      if ((2 << i) * (2 << i) <= ideal && (2 << (i + 1)) * (2 << (i + 1)) > ideal) {
        dim = 2 << i;
        break;
      }
    nx = bins_x > 0 ? bins_x : dim;
    ny = bins_y > 0 ? bins_y : dim;
    bsx = cw / (float)nx;
    bsy = ch / (float)ny;
    bin_area = bsx * bsy;
    bllx.resize(nx);
    burx.resize(nx);
    blly.resize(ny);
    bury.resize(ny);
    for (int i = 0; i < nx; ++i) {
      bllx[i] = (float)i * bsx + core_llx;
      burx[i] = bllx[i] + bsx;
    }
    for (int j = 0; j < ny; ++j) {
      blly[j] = (float)j * bsy + core_lly;
      bury[j] = blly[j] + bsy;
    }
    size_t nb = (size_t)nx * ny;
    std::vector<float> tm(nb, 0.f), bm(nb, 0.f);
    // terminal density (eplace.cpp:327-374): fully-inside terminals only.
    for (int t = 0; t < NT; ++t) {
      float tl = tllx[t], tb = tlly[t], tr = tllx[t] + tw[t], tt = tlly[t] + th[t];
      if (tl < core_llx || tl + tw[t] > core_urx) continue;
      if (tb < core_lly || tb + th[t] > core_ury) continue;
      int sx = (int)((tl - core_llx) / bsx), ex_ = (int)((tr - core_llx) / bsx);
      int sy = (int)((tb - core_lly) / bsy), ey_ = (int)((tt - core_lly) / bsy);
      if (ey_ >= ny) ey_ = ny - 1;
      if (ex_ >= nx) ex_ = nx - 1;
      for (int i = sx; i <= ex_; ++i)
        for (int j = sy; j <= ey_; ++j) {
          float &v = tm[(size_t)i * ny + j];
          v = (float)((double)v + (double)td * overlap_area(bllx[i], blly[j], burx[i], bury[j], tl, tb, tr, tt));
        }
    }
    // base density (eplace.cpp:385-409).  Rows that miss the bin add exactly
    // 0.0, which leaves a float sum unchanged, so only rows overlapping the
    // bin in y are visited.
    std::vector<std::vector<int>> rows_of_bin(ny);
    for (size_t r = 0; r < row_bottom.size(); ++r)
      for (int j = 0; j < ny; ++j)
        if (overlap_len(blly[j], bury[j], row_lly(r), row_ury(r)) != 0) rows_of_bin[j].push_back((int)r);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        float avail = 0;
        for (int r : rows_of_bin[j])
          avail = (float)((double)avail + overlap_area(bllx[i], blly[j], burx[i], bury[j], (float)row_x0[r],
                                                       row_lly(r), (float)row_x1[r], row_ury(r)));
        float ba = bsx * bsy;
        bm[(size_t)i * ny + j] = float_equal(ba, avail) ? 0.f : td * (ba - avail);
      }
    term_map.assign(tm.begin(), tm.end());
    base_map.assign(bm.begin(), bm.end());
    node_map.assign(nb, 0);
    filler_map.assign(nb, 0);
    node_q.assign(nb, 0);
    filler_q.assign(nb, 0);
    rho.assign(nb, 0);
    phi.assign(nb, 0);
    ex.assign(nb, 0);
    ey.assign(nb, 0);
    gdx.assign(O, 0);
    gdy.assign(O, 0);
    gtx.assign(O, 0);
    gty.assign(O, 0);
    gwx.assign(N, 0);
    gwy.assign(N, 0);
    pin_gx.assign(P, 0);
    pin_gy.assign(P, 0);
  }

  // Inflated rectangle, scales and bin range of one object; shared by the
  // scatter (eplace.cpp:1042-1104) and the gather (eplace.cpp:601-659).
  struct Footprint {
    float l, b, r, t, sx, sy;
    int i0, i1, j0, j1;
  };
  Footprint footprint(int o) const {
    Footprint fp;
    fp.sx = 1;
    fp.sy = 1;
    fp.l = llx[o];
    fp.b = lly[o];
    fp.r = llx[o] + w[o];
    fp.t = lly[o] + h[o];
    if (float_less(w[o], bsx)) {
      fp.sx = w[o] / bsx;
      fp.l = (float)((double)cx[o] - 0.5 * (double)bsx);
      fp.r = (float)((double)cx[o] + 0.5 * (double)bsx);
    }
    if (float_less(h[o], bsy)) {
      fp.sy = h[o] / bsy;
      fp.b = (float)((double)cy[o] - 0.5 * (double)bsy);
      fp.t = (float)((double)cy[o] + 0.5 * (double)bsy);
    }
    fp.i0 = (int)((fp.l - core_llx) / bsx);
    fp.i1 = (int)((fp.r - core_llx) / bsx);
    fp.j0 = (int)((fp.b - core_lly) / bsy);
    fp.j1 = (int)((fp.t - core_lly) / bsy);
    if (fp.j1 >= ny) fp.j1 = ny - 1;
    if (fp.i1 >= nx) fp.i1 = nx - 1;
    return fp;
  }

  // EPlacer_2D::binNodeDensityUpdate (eplace.cpp:1029-1133), summed in T.
  void scatter() {
    std::fill(node_map.begin(), node_map.end(), (T)0);
    std::fill(filler_map.begin(), filler_map.end(), (T)0);
    for (int o = 0; o < O; ++o) {
      Footprint fp = footprint(o);
      for (int i = fp.i0; i <= fp.i1; ++i)
        for (int j = fp.j0; j <= fp.j1; ++j) {
          double ovd = overlap_area(bllx[i], blly[j], burx[i], bury[j], fp.l, fp.b, fp.r, fp.t);
          size_t k = (size_t)i * ny + j;
          if (mode == 0) {
            float ov = (float)ovd;
            if (macro[o]) node_map[k] += fp.sx * fp.sy * td * ov;
            else if (filler[o]) filler_map[k] += fp.sx * fp.sy * ov;
            else node_map[k] += fp.sx * fp.sy * ov;
          } else {
            double s = float_less(w[o], bsx) ? (double)w[o] / (double)bsx : 1.0;
            double t = float_less(h[o], bsy) ? (double)h[o] / (double)bsy : 1.0;
            double c = s * t * ovd * (macro[o] ? (double)td : 1.0);
            if (filler[o]) filler_map[k] += (T)c;
            else node_map[k] += (T)c;
          }
        }
    }
  }

  // The reference's float contribution of object o to bin (i, j) (the
  // addend of eplace.cpp:1116/1123/1127), before accumulation.
  static float contribution(const Footprint &fp, double ovd, bool is_macro, float tdv) {
    float ov = (float)ovd;
    if (is_macro) return fp.sx * fp.sy * tdv * ov;
    return fp.sx * fp.sy * ov;
  }

  // Same scatter, each contribution rounded to int64 units of 2^-shift and
  // summed exactly.
  void scatter_fixed(int shift) {
    fixed_shift = shift;
    std::fill(node_q.begin(), node_q.end(), 0);
    std::fill(filler_q.begin(), filler_q.end(), 0);
    double scale = std::ldexp(1.0, shift);
    for (int o = 0; o < O; ++o) {
      Footprint fp = footprint(o);
      for (int i = fp.i0; i <= fp.i1; ++i)
        for (int j = fp.j0; j <= fp.j1; ++j) {
          double ovd = overlap_area(bllx[i], blly[j], burx[i], bury[j], fp.l, fp.b, fp.r, fp.t);
          float c = contribution(fp, ovd, macro[o] != 0, td);
          int64_t q = (int64_t)std::llrint((double)c * scale);
          size_t k = (size_t)i * ny + j;
          if (filler[o]) filler_q[k] += q;
          else node_q[k] += q;
        }
    }
  }

  // Maps from the fixed-point sums (what the GPU's density finaliser uses).
  void maps_from_fixed() {
    double inv = std::ldexp(1.0, -fixed_shift);
    for (size_t k = 0; k < node_q.size(); ++k) {
      node_map[k] = (T)(float)((double)node_q[k] * inv);
      filler_map[k] = (T)(float)((double)filler_q[k] * inv);
    }
  }

  // EPlacer_2D::densityOverflowUpdate (eplace.cpp:438-453).
  void overflow_update() {
    T nas = (T)std_area + (T)macro_area * (T)td;
    T inv_ba = (T)(1.0 / (double)(bsx * bsy));
    T sum = 0;
    for (size_t k = 0; k < node_map.size(); ++k) {
      T d = (node_map[k] + term_map[k] + base_map[k]) * inv_ba - (T)td;
      sum += std::max((T)0, d) * (T)bin_area;
    }
    overflow = sum / nas;
  }

  // Charge density fed to the Poisson solve (eplace.cpp:571-580).
  void density_to_rho() {
    T inv_ba = (T)(1.0 / (double)(bsx * bsy));
    for (size_t k = 0; k < rho.size(); ++k) {
      T e = node_map[k] + base_map[k] + filler_map[k] + term_map[k];
      rho[k] = e * inv_ba;
    }
  }

  // Wave numbers (fft.cpp:129-147): long double product, stored as float.
  void wave_numbers(std::vector<T> &wx, std::vector<T> &wy) const {
    const long double PI = 3.141592653589793238462L;
    wx.resize(nx);
    wy.resize(ny);
    for (int i = 0; i < nx; ++i) {
      if (mode == 0) wx[i] = (T)(float)(PI * (float)i / (float)nx);
      else wx[i] = (T)(M_PI * i / nx);
    }
    for (int j = 0; j < ny; ++j) {
      if (mode == 0) wy[j] = (T)(float)(PI * (float)j / (float)ny * (float)bsy / (float)bsx);
      else wy[j] = (T)(M_PI * j / ny * (double)bsy / (double)bsx);
    }
  }

  // FFT_2D::doFFT (fft.cpp:190-284) in closed form:
  //   rho^ = (4/(NxNy)) c_u c_v sum rho cos cos,   c_0 = 1/2
  //   phi^ = rho^ / (wx^2 + wy^2), phi^[0][0] = 0; Ex^ = wx phi^, Ey^ = wy phi^
  //   phi = sum phi^ cos cos;  Ex = sum Ex^ sin_x cos_y;  Ey = sum Ey^ cos_x sin_y
  void poisson() {
    std::vector<T> wx, wy;
    wave_numbers(wx, wy);
    size_t nb = (size_t)nx * ny;
    std::vector<T> a(nb), t(nb);
    bool fast = poisson_fast && is_pow2(nx) && is_pow2(ny);
    // forward DCT-II along y then x
    if (fast) {
      for (int i = 0; i < nx; ++i) dct2(&rho[(size_t)i * ny], &t[(size_t)i * ny], ny, 1, 1);
      for (int j = 0; j < ny; ++j) dct2(&t[j], &a[j], nx, ny, ny);
    } else {
      matrix_transform(rho, a, 0, 0);
    }
    T norm = (T)(4.0 / nx / ny);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        T v = a[(size_t)i * ny + j];
        if (i == 0) v *= (T)0.5;
        if (j == 0) v *= (T)0.5;
        a[(size_t)i * ny + j] = v * norm;
      }
    std::vector<T> sp(nb), sx(nb), sy(nb);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        size_t k = (size_t)i * ny + j;
        if (i == 0 && j == 0) {
          sp[k] = sx[k] = sy[k] = 0;
          continue;
        }
        T wx2 = wx[i] * wx[i], wy2 = wy[j] * wy[j];
        T p = a[k] / (wx2 + wy2);
        sp[k] = p;
        sx[k] = p * wx[i];
        sy[k] = p * wy[j];
      }
    if (fast) {
      // phi: DCT-III both;  Ex: DST-III in x, DCT-III in y;  Ey: DCT-III x, DST-III y
      for (int i = 0; i < nx; ++i) dct3(&sp[(size_t)i * ny], &t[(size_t)i * ny], ny, 1, 1);
      for (int j = 0; j < ny; ++j) dct3(&t[j], &phi[j], nx, ny, ny);
      for (int i = 0; i < nx; ++i) dct3(&sx[(size_t)i * ny], &t[(size_t)i * ny], ny, 1, 1);
      for (int j = 0; j < ny; ++j) dst3(&t[j], &ex[j], nx, ny, ny);
      for (int i = 0; i < nx; ++i) dst3(&sy[(size_t)i * ny], &t[(size_t)i * ny], ny, 1, 1);
      for (int j = 0; j < ny; ++j) dct3(&t[j], &ey[j], nx, ny, ny);
    } else {
      matrix_transform(sp, phi, 1, 1);
      matrix_transform(sx, ex, 2, 1);
      matrix_transform(sy, ey, 1, 2);
    }
  }

  // Direct O(N^3) separable transform.  kind 0: forward cos sum over index;
  // 1: inverse cos; 2: inverse sin.  kx applies along x, ky along y.
  void matrix_transform(const std::vector<T> &in, std::vector<T> &out, int kx, int ky) const {
    auto basis = [](int kind, int k, int x, int n) -> double {
      if (kind == 0) return std::cos(M_PI * k * (x + 0.5) / n);   // [k][x]
      if (kind == 1) return std::cos(M_PI * x * (k + 0.5) / n);   // spectral x -> spatial k
      return std::sin(M_PI * x * (k + 0.5) / n);
    };
    size_t nb = (size_t)nx * ny;
    std::vector<T> t(nb, 0);
    std::vector<T> By((size_t)ny * ny), Bx((size_t)nx * nx);
    for (int a = 0; a < ny; ++a)
      for (int b = 0; b < ny; ++b) By[(size_t)a * ny + b] = (T)basis(ky, a, b, ny);
    for (int a = 0; a < nx; ++a)
      for (int b = 0; b < nx; ++b) Bx[(size_t)a * nx + b] = (T)basis(kx, a, b, nx);
    for (int i = 0; i < nx; ++i)
      for (int a = 0; a < ny; ++a) {
        T s = 0;
        for (int b = 0; b < ny; ++b) s += By[(size_t)a * ny + b] * in[(size_t)i * ny + b];
        t[(size_t)i * ny + a] = s;
      }
    for (int a = 0; a < nx; ++a)
      for (int j = 0; j < ny; ++j) {
        T s = 0;
        for (int b = 0; b < nx; ++b) s += Bx[(size_t)a * nx + b] * t[(size_t)b * ny + j];
        out[(size_t)a * ny + j] = s;
      }
  }

  // Density gradient gather (eplace.cpp:599-674).
  void gather() {
    for (int o = 0; o < O; ++o) {
      Footprint fp = footprint(o);
      T gx = 0, gy = 0;
      for (int i = fp.i0; i <= fp.i1; ++i)
        for (int j = fp.j0; j <= fp.j1; ++j) {
          double ovd = overlap_area(bllx[i], blly[j], burx[i], bury[j], fp.l, fp.b, fp.r, fp.t);
          size_t k = (size_t)i * ny + j;
          T ov;
          if (mode == 0) ov = (T)(float)((double)(fp.sx * fp.sy) * ovd);
          else {
            double s = float_less(w[o], bsx) ? (double)w[o] / (double)bsx : 1.0;
            double t = float_less(h[o], bsy) ? (double)h[o] / (double)bsy : 1.0;
            ov = (T)(s * t * ovd);
          }
          gx += ov * ex[k];
          gy += ov * ey[k];
        }
      gdx[o] = gx;
      gdy[o] = gy;
    }
  }

  // gamma schedule (eplace.cpp:475-500); returns 1/gamma per dimension.
  void gamma_update() {
    float cx_ = 0.125f / bsx, cy_ = 0.125f / bsy;
    if (mode == 0) {
      float tau = (float)overflow;
      if (tau > 1.0) {
        cx_ = (float)(cx_ * 0.1);
        cy_ = (float)(cy_ * 0.1);
      } else if (tau < 0.1) {
        cx_ = (float)(cx_ * 10.0);
        cy_ = (float)(cy_ * 10.0);
      } else {
        float temp = (float)(1.0 / std::pow(10.0, (tau - 0.1) * 20 / 9.0 - 1.0));
        cx_ *= temp;
        cy_ *= temp;
      }
      invgx = cx_;
      invgy = cy_;
    } else {
      double tau = (double)overflow;
      double bx = 0.125 / (double)bsx, by = 0.125 / (double)bsy, m;
      if (tau > 1.0) m = 0.1;
      else if (tau < 0.1) m = 10.0;
      else m = 1.0 / std::pow(10.0, (tau - 0.1) * 20 / 9.0 - 1.0);
      invgx = (T)(bx * m);
      invgy = (T)(by * m);
    }
  }

  T pin_x(int64_t p) const {
    int o = pin_obj[p];
    float c = o < N ? cx[o] : tcx[o - N];
    return (T)(c + pin_dx[p]);  // Pin::calculateAbsolutePos (objects.cpp:454-459), float add
  }
  T pin_y(int64_t p) const {
    int o = pin_obj[p];
    float c = o < N ? cy[o] : tcy[o - N];
    return (T)(c + pin_dy[p]);
  }

  T expf_(T a) const { return exp_mode == 0 ? fast_exp(a) : (T)std::exp(a); }

  // calcNetBoundPins + calcWA_Wirelength_2D + per-node gradient
  // (placedb.cpp:434-470, objects.cpp:75-144, 248-415, eplace.cpp:531-549).
  void wirelength() {
    hpwl_bound = 0;
    wl_value = 0;
    for (int e = 0; e < E; ++e) {
      int64_t s = net_start[e], t = net_start[e + 1];
      if (s == t) continue;
      double maxx = -1.79769313486231570815e+308, minx = 1.79769313486231570815e+308;
      double maxy = maxx, miny = minx;
      int64_t bxmax = s, bxmin = s, bymax = s, bymin = s;
      for (int64_t p = s; p < t; ++p) {
        double x = (double)pin_x(p), y = (double)pin_y(p);
        if (x < minx) { minx = x; bxmin = p; }
        if (x > maxx) { maxx = x; bxmax = p; }
        if (y < miny) { miny = y; bymin = p; }
        if (y > maxy) { maxy = y; bymax = p; }
      }
      hpwl_bound += (maxx - minx) + (maxy - miny);
      T pmaxx = pin_x(bxmax), pminx = pin_x(bxmin), pmaxy = pin_y(bymax), pminy = pin_y(bymin);
      if (wl_model == 0) wa_net(s, t, pmaxx, pminx, pmaxy, pminy);
      else lse_net(s, t, pmaxx, pminx, pmaxy, pminy);
    }
    for (int i = 0; i < N; ++i) {
      T gx = 0, gy = 0;
      for (int64_t q = node_pin_start[i]; q < node_pin_start[i + 1]; ++q) {
        gx += pin_gx[node_pins[q]];
        gy += pin_gy[node_pins[q]];
      }
      gwx[i] = gx;
      gwy[i] = gy;
    }
  }

  void wa_net(int64_t s, int64_t t, T pmaxx, T pminx, T pmaxy, T pminy) {
    // Net::calcWirelengthWA_2D (objects.cpp:248-345)
    int64_t n = t - s;
    std::vector<T> emx(n), emnx(n), emy(n), emny(n);
    std::vector<char> zmx(n), zmnx(n), zmy(n), zmny(n);
    T nmx = 0, dmx = 0, nmnx = 0, dmnx = 0, nmy = 0, dmy = 0, nmny = 0, dmny = 0;
    for (int64_t p = s; p < t; ++p) {
      int64_t q = p - s;
      T x = pin_x(p), y = pin_y(p);
      T amx = (x - pmaxx) * invgx, amnx = (pminx - x) * invgx;
      T amy = (y - pmaxy) * invgy, amny = (pminy - y) * invgy;
      zmx[q] = !(amx > (T)kNegMaxExp);
      zmnx[q] = !(amnx > (T)kNegMaxExp);
      zmy[q] = !(amy > (T)kNegMaxExp);
      zmny[q] = !(amny > (T)kNegMaxExp);
      if (!zmx[q]) { emx[q] = expf_(amx); nmx += x * emx[q]; dmx += emx[q]; }
      if (!zmnx[q]) { emnx[q] = expf_(amnx); nmnx += x * emnx[q]; dmnx += emnx[q]; }
      if (!zmy[q]) { emy[q] = expf_(amy); nmy += y * emy[q]; dmy += emy[q]; }
      if (!zmny[q]) { emny[q] = expf_(amny); nmny += y * emny[q]; dmny += emny[q]; }
    }
    wl_value += (double)((nmx / dmx - nmnx / dmnx) + (nmy / dmy - nmny / dmny));
    // Net::getWirelengthGradientWA_2D (objects.cpp:359-415)
    for (int64_t p = s; p < t; ++p) {
      int64_t q = p - s;
      T x = pin_x(p), y = pin_y(p);
      T gmx = 0, gmy = 0, gnx = 0, gny = 0;
      if (!zmx[q]) {
        T gd = invgx * emx[q];
        T gn = emx[q] + x * gd;
        gmx = (gn * dmx - gd * nmx) / (dmx * dmx);
      }
      if (!zmy[q]) {
        T gd = invgy * emy[q];
        T gn = emy[q] + y * gd;
        gmy = (gn * dmy - gd * nmy) / (dmy * dmy);
      }
      if (!zmnx[q]) {
        T gd = invgx * emnx[q];
        T gn = emnx[q] - x * gd;
        gnx = (gn * dmnx + gd * nmnx) / (dmnx * dmnx);
      }
      if (!zmny[q]) {
        T gd = invgy * emny[q];
        T gn = emny[q] - y * gd;
        gny = (gn * dmny + gd * nmny) / (dmny * dmny);
      }
      pin_gx[p] = gmx - gnx;
      pin_gy[p] = gmy - gny;
    }
  }

  void lse_net(int64_t s, int64_t t, T pmaxx, T pminx, T pmaxy, T pminy) {
    // Net::calcWirelengthLSE_2D / getWirelengthGradientLSE_2D (objects.cpp:161-241, 423-437)
    int64_t n = t - s;
    std::vector<T> emx(n), emnx(n), emy(n), emny(n);
    T smx = 0, smnx = 0, smy = 0, smny = 0;
    for (int64_t p = s; p < t; ++p) {
      int64_t q = p - s;
      T x = pin_x(p), y = pin_y(p);
      T amx = (x - pmaxx) * invgx, amnx = (pminx - x) * invgx;
      T amy = (y - pmaxy) * invgy, amny = (pminy - y) * invgy;
      emx[q] = amx > (T)kNegMaxExp ? expf_(amx) : (T)0;
      emnx[q] = amnx > (T)kNegMaxExp ? expf_(amnx) : (T)0;
      emy[q] = amy > (T)kNegMaxExp ? expf_(amy) : (T)0;
      emny[q] = amny > (T)kNegMaxExp ? expf_(amny) : (T)0;
      smx += emx[q];
      smnx += emnx[q];
      smy += emy[q];
      smny += emny[q];
    }
    wl_value += (double)(pmaxx - pminx + std::log(smx) / invgx + std::log(smnx) / invgx) +
                (double)(pmaxy - pminy + std::log(smy) / invgy + std::log(smny) / invgy);
    for (int64_t p = s; p < t; ++p) {
      int64_t q = p - s;
      pin_gx[p] = emx[q] / smx - emnx[q] / smnx;
      pin_gy[p] = emy[q] / smy - emny[q] / smny;
    }
  }

  // EPlacer_2D::wirelengthGradientUpdate (eplace.cpp:466-550).
  void wirelength_gradient_update() {
    overflow_update();
    gamma_update();
    wirelength();
  }

  // totalGradientUpdate combine (eplace.cpp:687-738).
  void combine() {
    for (int o = 0; o < O; ++o) {
      T pre;
      T cnt = (T)npins[o];
      T charge = (T)area[o];
      pre = (T)1 / std::max((T)1, cnt + lambda * charge);
      if (filler[o]) {
        gtx[o] = pre * lambda * gdx[o];
        gty[o] = pre * lambda * gdy[o];
      } else {
        gtx[o] = pre * (lambda * gdx[o] - gwx[o]);
        gty[o] = pre * (lambda * gdy[o] - gwy[o]);
      }
    }
  }

  void total_gradient_update(bool fixed_density, int shift) {
    if (fixed_density) {
      scatter_fixed(shift);
      maps_from_fixed();
    } else {
      scatter();
    }
    density_to_rho();
    poisson();
    gather();
    wirelength_gradient_update();
    combine();
  }

  // PlaceDB::calcHPWL (placedb.cpp:424-432; objects.cpp:35-73).
  double calc_hpwl() const {
    double sum = 0;
    for (int e = 0; e < E; ++e) {
      if (net_start[e] == net_start[e + 1]) continue;
      double maxx = -1.79769313486231570815e+308, minx = 1.79769313486231570815e+308;
      double maxy = maxx, miny = minx;
      for (int64_t p = net_start[e]; p < net_start[e + 1]; ++p) {
        double x = (double)(float)pin_x(p), y = (double)(float)pin_y(p);
        minx = std::min(minx, x);
        maxx = std::max(maxx, x);
        miny = std::min(miny, y);
        maxy = std::max(maxy, y);
      }
      sum += (maxx - minx) + (maxy - miny);
    }
    return sum;
  }

  // penaltyFactorInitilization (eplace.cpp:812-835).
  void penalty_init() {
    last_hpwl = calc_hpwl();
    T num = 0, den = 0;
    for (int i = 0; i < N; ++i) {
      num += std::fabs(gwx[i]);
      num += std::fabs(gwy[i]);
      den += std::fabs(gdx[i]);
      den += std::fabs(gdy[i]);
    }
    for (int i = N; i < O; ++i) {
      den += std::fabs(gdx[i]);
      den += std::fabs(gdy[i]);
    }
    lambda = num / den;
  }

  // updatePenaltyFactor (eplace.cpp:846-883).
  void update_penalty() {
    double hp = calc_hpwl();
    if (mode == 0) {
      float cur = (float)hp;
      float mult;
      double delta = cur - last_hpwl;
      if (delta < 0.0) mult = (float)1.05;
      else mult = (float)std::pow(1.05, (-(delta) / delta_hpwl_ref + 1.0));
      if (float_greater(mult, (float)1.05)) mult = (float)1.05;
      if (float_less(mult, (float)0.95)) mult = (float)0.95;
      lambda *= mult;
      last_hpwl = cur;
    } else {
      double delta = hp - last_hpwl;
      double mult = delta < 0.0 ? 1.05 : std::pow(1.05, -delta / delta_hpwl_ref + 1.0);
      mult = std::min(1.05, std::max(0.95, mult));
      lambda = (T)((double)lambda * mult);
      last_hpwl = hp;
    }
  }

  // ---- stage subsets (eplace.cpp:190-203, 742-804) ----
  bool in_stage(int o) const {
    if (stage == 0) return true;
    if (stage == 1) return filler[o] != 0;
    return filler[o] || !macro[o];
  }
  std::vector<int> stage_objects() const {
    std::vector<int> v;
    for (int o = 0; o < O; ++o)
      if (in_stage(o)) v.push_back(o);
    return v;
  }
  void get_position(std::vector<T> &xy) const {
    xy.clear();
    for (int o : stage_objects()) {
      xy.push_back((T)cx[o]);
      xy.push_back((T)cy[o]);
    }
  }
  void get_gradient(std::vector<T> &g) const {
    g.clear();
    for (int o : stage_objects()) {
      g.push_back(gtx[o]);
      g.push_back(gty[o]);
    }
  }
  void set_position(const std::vector<T> &xy) {
    std::vector<int> so = stage_objects();
    for (size_t k = 0; k < so.size(); ++k) set_center(so[k], (float)xy[2 * k], (float)xy[2 * k + 1]);
  }

  // calc_lipschitz_constant (optutils.hpp:19-35).
  static T lipschitz(const std::vector<T> &xt, const std::vector<T> &xt1, const std::vector<T> &gt,
                     const std::vector<T> &gt1) {
    T num = 0, den = 0;
    for (size_t k = 0; k < xt.size(); k += 2) {
      T dgx = gt[k] - gt1[k], dgy = gt[k + 1] - gt1[k + 1];
      T dxx = xt[k] - xt1[k], dxy = xt[k + 1] - xt1[k + 1];
      num += dgx * dgx + dgy * dgy + (T)0 * (T)0;
      den += dxx * dxx + dxy * dxy + (T)0 * (T)0;
    }
    return std::sqrt(num) / std::sqrt(den);
  }

  // Barzilai-Borwein step of opt_step_bb (nesterov.hpp:245-281): with
  // s = y_k - y_{k-1} and y = g_k - g_{k-1} over the stage vector (z = 0),
  // the short step s.y / y.y when positive, else min(|s| / |y|, 1/L).  The
  // reference forms the norms and the dot product with Eigen's float
  // reductions; here they are summed in order in T.
  static T bb_step(const std::vector<T> &xt, const std::vector<T> &xt1, const std::vector<T> &gt,
                   const std::vector<T> &gt1, T lip_step, int *used_short) {
    T ss = 0, yy = 0, sy = 0;
    for (size_t k = 0; k < xt.size(); ++k) {
      T sk = xt[k] - xt1[k], yk = gt[k] - gt1[k];
      ss += sk * sk;
      yy += yk * yk;
      sy += sk * yk;
    }
    T s_norm = std::sqrt(ss), y_norm = std::sqrt(yy);
    T short_step = sy / (y_norm * y_norm);
    *used_short = short_step > 0;
    if (short_step > 0) return short_step;
    return std::min(s_norm / y_norm, lip_step);
  }

  // ---- EplaceNesterovOpt (nesterov.hpp:86-415) ----
  void opt_init() {
    iter = 0;
    ns_a = 1;
    get_position(cur_main);
  }
  bool stop_condition(float target_overflow) {
    if (stage == 0) {
      bool j = ((float)overflow < target_overflow) || (iter > 1000);
      if (j) mgp_iterations = iter;
      return j;
    }
    if (stage == 1) return iter >= 20;
    return ((float)overflow < 0.07f) || (iter > 1000);
  }
  void nesterov_move(const std::vector<T> &grad, T step, T a_new, std::vector<T> &nmain,
                     std::vector<T> &nref) {
    size_t n = cur_main.size();
    nmain.resize(n);
    nref.resize(n);
    T coef = (ns_a - 1) / a_new;
    for (size_t k = 0; k < n; ++k) {
      T np = cur_ref[k] + grad[k] * step;
      nmain[k] = np;
      nref[k] = np + (np - cur_main[k]) * coef;
    }
  }
  // One opt_step (nesterov.hpp:151-200) with vanilla (375-415), bktrk
  // (310-369) or bb (231-302): method 0, 1 or 2.
  void opt_step(int bktrk, bool fixed_density, int shift) {
    get_position(cur_ref);
    total_gradient_update(fixed_density, shift);
    get_gradient(cur_grad);
    T step = iter == 0 ? (T)1 : (T)1 / lipschitz(cur_ref, last_ref, cur_grad, last_grad);
    if (bktrk == 2) {
      if (iter > 0) step = bb_step(cur_ref, last_ref, cur_grad, last_grad, step, &bb_short_used);
      bktrk = 0;  // one move per iteration, as vanilla
    }
    T a_new = ((T)1 + std::sqrt((T)4 * (ns_a * ns_a) + (T)1)) / (T)2;
    std::vector<T> nmain, nref, ngrad;
    if (!bktrk) {
      nesterov_move(cur_grad, step, a_new, nmain, nref);
      set_position(nref);
    } else {
      while (true) {
        nesterov_move(cur_grad, step, a_new, nmain, nref);
        set_position(nref);
        total_gradient_update(fixed_density, shift);
        get_gradient(ngrad);
        T nstep = (T)1 / lipschitz(nref, cur_ref, ngrad, cur_grad);
        // BKTRK_EPS is a double literal: the product and the comparison are
        // double in the reference too (nesterov.hpp:355)
        if (bktrk_eps * (double)step <= (double)nstep) {
          step = nstep;
          break;
        }
        ++bktrk_count;
        step = nstep;
      }
    }
    last_step = step;
    last_ref = cur_ref;
    last_grad = cur_grad;
    cur_main = nmain;
    ns_a = a_new;
    update_penalty();
    ++iter;
  }
  void wrap_up() { set_position(cur_main); }
};

template <class T>
Placer<T> *as(void *p) {
  return static_cast<Placer<T> *>(static_cast<PlacerBase *>(p));
}

}  // namespace

// ===========================================================================
// C API (ctypes)
// ===========================================================================
extern "C" {

int orc_glibc_rand(uint32_t seed, int n, int32_t *out) {
  GlibcRand r(seed);
  for (int i = 0; i < n; ++i) out[i] = r.next();
  return 0;
}

// mode 0 reference float, 1 fp64.  filler_count < 0: reference count.
void *orc_create(const orc_design *d, float target_density, int bins_x, int bins_y, int mode, int exp_mode,
                 int wl_model, const float *filler_xy, int filler_count, uint32_t seed) {
  PlacerBase *b;
  if (mode == 0) {
    auto *p = new Placer<float>();
    p->mode = 0;
    p->exp_mode = exp_mode;
    p->wl_model = wl_model;
    p->load(d, target_density);
    p->filler_init(seed, filler_xy, filler_count);
    p->bin_init(bins_x, bins_y);
    b = p;
  } else {
    auto *p = new Placer<double>();
    p->mode = 1;
    p->exp_mode = exp_mode;
    p->wl_model = wl_model;
    p->load(d, target_density);
    p->filler_init(seed, filler_xy, filler_count);
    p->bin_init(bins_x, bins_y);
    b = p;
  }
  return b;
}

void orc_destroy(void *p) { delete static_cast<PlacerBase *>(p); }

#define ORC_DISPATCH(p, ...)                          \
  do {                                                \
    if (static_cast<PlacerBase *>(p)->mode == 0) {    \
      auto *P = as<float>(p);                         \
      __VA_ARGS__;                                    \
    } else {                                          \
      auto *P = as<double>(p);                        \
      __VA_ARGS__;                                    \
    }                                                 \
  } while (0)

// ints: N F O NT E nx ny stage iter mgp_iterations bktrk_count fixed_shift P
void orc_info_i(void *p, int64_t *out) {
  ORC_DISPATCH(p, {
    int64_t v[] = {P->N, P->F, P->O, P->NT, P->E, P->nx, P->ny, P->stage, P->iter, P->mgp_iterations,
                   P->bktrk_count, P->fixed_shift, P->P};
    memcpy(out, v, sizeof(v));
  });
}

// doubles: bsx bsy bin_area core(4) total_row_area std_area macro_area
//          overflow invgx invgy lambda last_hpwl hpwl_bound wl_value td last_step ns_a
//          filler_w filler_h
void orc_info_d(void *p, double *out) {
  ORC_DISPATCH(p, {
    double fw = P->F ? P->w[P->N] : 0, fh = P->F ? P->h[P->N] : 0;
    double v[] = {P->bsx, P->bsy, P->bin_area, P->core_llx, P->core_lly, P->core_urx, P->core_ury,
                  P->total_row_area, P->std_area, P->macro_area, (double)P->overflow, (double)P->invgx,
                  (double)P->invgy, (double)P->lambda, P->last_hpwl, P->hpwl_bound, P->wl_value, P->td,
                  (double)P->last_step, (double)P->ns_a, fw, fh};
    memcpy(out, v, sizeof(v));
  });
}

void orc_set_scalar(void *p, const char *name, double v) {
  std::string n(name);
  ORC_DISPATCH(p, {
    if (n == "lambda") P->lambda = v;
    else if (n == "overflow") P->overflow = v;
    else if (n == "invgx") P->invgx = v;
    else if (n == "invgy") P->invgy = v;
    else if (n == "last_hpwl") P->last_hpwl = v;
    else if (n == "stage") P->stage = (int)v;
    else if (n == "poisson_fast") P->poisson_fast = (int)v;
    else if (n == "exp_mode") P->exp_mode = (int)v;
    else if (n == "bktrk_eps") P->bktrk_eps = v;
    else if (n == "delta_hpwl_ref") P->delta_hpwl_ref = v;
  });
}

// Copy a named array out as double (or int64 for *_q).  Returns its length.
int64_t orc_get(void *p, const char *name, void *out) {
  std::string n(name);
  int64_t len = -1;
  ORC_DISPATCH(p, {
    auto put = [&](const auto &vec) {
      len = (int64_t)vec.size();
      if (out)
        for (int64_t i = 0; i < len; ++i) ((double *)out)[i] = (double)vec[i];
    };
    auto put2 = [&](const auto &a, const auto &b) {
      len = (int64_t)a.size() * 2;
      if (out)
        for (size_t i = 0; i < a.size(); ++i) {
          ((double *)out)[2 * i] = (double)a[i];
          ((double *)out)[2 * i + 1] = (double)b[i];
        }
    };
    auto puti = [&](const std::vector<int64_t> &vec) {
      len = (int64_t)vec.size();
      if (out) memcpy(out, vec.data(), vec.size() * 8);
    };
    if (n == "node_map") put(P->node_map);
    else if (n == "filler_map") put(P->filler_map);
    else if (n == "term_map") put(P->term_map);
    else if (n == "base_map") put(P->base_map);
    else if (n == "rho") put(P->rho);
    else if (n == "phi") put(P->phi);
    else if (n == "ex") put(P->ex);
    else if (n == "ey") put(P->ey);
    else if (n == "node_q") puti(P->node_q);
    else if (n == "filler_q") puti(P->filler_q);
    else if (n == "grad_density") put2(P->gdx, P->gdy);
    else if (n == "grad_wirelength") put2(P->gwx, P->gwy);
    else if (n == "grad_total") put2(P->gtx, P->gty);
    else if (n == "pin_grad") put2(P->pin_gx, P->pin_gy);
    else if (n == "positions") put2(P->cx, P->cy);
    else if (n == "ll") put2(P->llx, P->lly);
    else if (n == "wh") put2(P->w, P->h);
    else if (n == "bin_llx") put(P->bllx);
    else if (n == "bin_urx") put(P->burx);
    else if (n == "bin_lly") put(P->blly);
    else if (n == "bin_ury") put(P->bury);
    else if (n == "cur_main") put(P->cur_main);
    else if (n == "stage_positions") { std::vector<decltype(P->lambda)> v; P->get_position(v); put(v); }
    else if (n == "stage_gradient") { std::vector<decltype(P->lambda)> v; P->get_gradient(v); put(v); }
  });
  return len;
}

// Set a named array from double (or int64 for *_q).
int orc_set(void *p, const char *name, const void *in) {
  std::string n(name);
  int ok = 0;
  ORC_DISPATCH(p, {
    using TT = decltype(P->lambda);
    const double *d = (const double *)in;
    if (n == "positions") {  // centres, raw (no clamp); lower-left re-derived
      for (int o = 0; o < P->O; ++o) {
        P->cx[o] = (float)d[2 * o];
        P->cy[o] = (float)d[2 * o + 1];
        P->llx[o] = P->cx[o] - (float)0.5 * P->w[o];
        P->lly[o] = P->cy[o] - (float)0.5 * P->h[o];
      }
      ok = 1;
    } else if (n == "ll") {
      for (int o = 0; o < P->O; ++o) {
        P->llx[o] = (float)d[2 * o];
        P->lly[o] = (float)d[2 * o + 1];
      }
      ok = 1;
    } else if (n == "stage_positions_clamped") {
      std::vector<TT> v(d, d + 2 * P->stage_objects().size());
      P->set_position(v);
      ok = 1;
    } else if (n == "rho") {
      for (size_t k = 0; k < P->rho.size(); ++k) P->rho[k] = (TT)d[k];
      ok = 1;
    } else if (n == "ex") {
      for (size_t k = 0; k < P->ex.size(); ++k) P->ex[k] = (TT)d[k];
      ok = 1;
    } else if (n == "ey") {
      for (size_t k = 0; k < P->ey.size(); ++k) P->ey[k] = (TT)d[k];
      ok = 1;
    } else if (n == "grad_density") {
      for (int o = 0; o < P->O; ++o) { P->gdx[o] = (TT)d[2 * o]; P->gdy[o] = (TT)d[2 * o + 1]; }
      ok = 1;
    } else if (n == "grad_wirelength") {
      for (int o = 0; o < P->N; ++o) { P->gwx[o] = (TT)d[2 * o]; P->gwy[o] = (TT)d[2 * o + 1]; }
      ok = 1;
    }
  });
  return ok;
}

// Individual steps of totalGradientUpdate and the optimiser.
int orc_run(void *p, const char *what, int arg0, int arg1) {
  std::string n(what);
  int ok = 1;
  ORC_DISPATCH(p, {
    if (n == "scatter") P->scatter();
    else if (n == "scatter_fixed") { P->scatter_fixed(arg0); P->maps_from_fixed(); }
    else if (n == "overflow") P->overflow_update();
    else if (n == "rho") P->density_to_rho();
    else if (n == "poisson") P->poisson();
    else if (n == "gather") P->gather();
    else if (n == "gamma") P->gamma_update();
    else if (n == "wirelength") P->wirelength();
    else if (n == "wirelength_gradient_update") P->wirelength_gradient_update();
    else if (n == "combine") P->combine();
    else if (n == "total_gradient_update") P->total_gradient_update(arg0 != 0, arg1);
    else if (n == "penalty_init") P->penalty_init();
    else if (n == "update_penalty") P->update_penalty();
    else if (n == "opt_init") P->opt_init();
    else if (n == "opt_step") P->opt_step(arg0, arg1 >= 0, arg1);
    else if (n == "wrap_up") P->wrap_up();
    else ok = 0;
  });
  return ok;
}

double orc_hpwl(void *p) {
  double v = 0;
  ORC_DISPATCH(p, v = P->calc_hpwl());
  return v;
}

int orc_stop_condition(void *p, float target_overflow) {
  int v = 0;
  ORC_DISPATCH(p, v = P->stop_condition(target_overflow) ? 1 : 0);
  return v;
}

// Standalone Poisson solve on a given rho (x-major nx*ny), fp64 arithmetic.
// method 0: O(N^3) matrix form, 1: FFT form (powers of two).
int orc_poisson_f64(int nx, int ny, float bsx, float bsy, const double *rho, double *phi, double *ex,
                    double *ey, int method) {
  Placer<double> p;
  p.mode = 1;
  p.nx = nx;
  p.ny = ny;
  p.bsx = bsx;
  p.bsy = bsy;
  p.poisson_fast = method;
  size_t nb = (size_t)nx * ny;
  p.rho.assign(rho, rho + nb);
  p.phi.assign(nb, 0);
  p.ex.assign(nb, 0);
  p.ey.assign(nb, 0);
  p.poisson();
  memcpy(phi, p.phi.data(), nb * 8);
  memcpy(ex, p.ex.data(), nb * 8);
  memcpy(ey, p.ey.data(), nb * 8);
  return 0;
}

}  // extern "C"
