// Quadratic initial placement (QPlace/qplace.cpp:20-295) on the GPU.
//
// Per outer iteration (qplace.cpp:59-102):
//   1. bound pins of every net, first strict minimum / maximum in pin order
//      (Net::calcBoundPin, objects.cpp:75-140), and the HPWL;
//   2. the X and Y systems of QPPlacer::createSparseMatrix
//      (qplace.cpp:123-280), built node-centrically: each movable node walks
//      its pins and, per pin, the partners the reference pairs it with, so
//      every row is written by one thread in a fixed order (no sort, no
//      atomics).  Rows 0..N-1 are X, N..2N-1 are Y; the diagonal and b are
//      kept apart from the off-diagonal CSR entries;
//   3. Eigen's BiCGSTAB with the identity preconditioner
//      (Eigen/src/IterativeLinearSolvers/BiCGSTAB.h, bicgstab(), the
//      van der Vorst algorithm with Eigen's restart rule) on both systems at
//      once (blockIdx.y = direction), float vectors as in the reference,
//      dot products and row sums accumulated in double with a fixed order;
//   4. PlaceDB::setModuleCenter_2D's clamp of the solution
//      (placedb.cpp:253-290).
// The host reads two errors and one HPWL per outer iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "ep_internal.h"
#include "reduce.h"

namespace ep {
namespace {

constexpr int kQpThreads = 256;
constexpr int kQpLanes = 4;  // lanes per matrix row in the solver's row products
constexpr int kQpMaxBlocks = 4096;
constexpr int kQpLargeNet = 64;  // nets above this get a block each for the bound pins
constexpr double kQpMinDistance = 25.0;  // MIN_DISTANCE, qplace.h:20
constexpr int kScanPerThread = 4;
constexpr int kScanTile = 1024 * kScanPerThread;

struct QpView {
  int n;  // movable nodes
  const int32_t *net_start, *pin_obj, *pin_net, *npin_start, *npin_list;
  const float *pin_dx, *pin_dy, *cx, *cy, *tx, *ty, *c1;
  const int4 *bound;  // per net: pin ids of x min, x max, y min, y max (-1: no pins)
};

__device__ __forceinline__ float qp_abs(const QpView &v, int q, int dim) {
  const int o = v.pin_obj[q];
  if (dim == 0) return (o < v.n ? v.cx[o] : v.tx[o - v.n]) + v.pin_dx[q];
  return (o < v.n ? v.cy[o] : v.ty[o - v.n]) + v.pin_dy[q];
}

// weightX / weightY (qplace.cpp:174-186): constant1 / max(d, MIN_DISTANCE),
// the second form divided in double as the reference's double literal does.
__device__ __forceinline__ float qp_weight(float c1, float d) {
  if (d - 25.0f > (float)kEps || fabsf(d - 25.0f) < (float)kEps) return c1 / d;
  return (float)((double)c1 / kQpMinDistance);
}

// Visits the partners of pin a (on movable node i) in direction dim, in
// pin order, as qplace.cpp:153-270 pairs them: X pairs contain a bound pin;
// Y pairs pass the reference's condition pin2 == ymin || pin1 == ymax ||
// pin2 == ymax (pin1 the earlier pin of the pair).  Partners on node i are
// skipped (qplace.cpp:163-166).
template <class F>
__device__ __forceinline__ void qp_partners(const QpView &v, int i, int a, int dim, F &&visit) {
  const int net = v.pin_net[a];
  const int s = v.net_start[net], e = v.net_start[net + 1];
  const int4 B = v.bound[net];
  auto go = [&](int q) {
    if (q >= 0 && v.pin_obj[q] != i) visit(q, net);  // q < 0: no bound pin (non-finite positions)
  };
  if (dim == 0) {
    if (a == B.x || a == B.y) {
      for (int q = s; q < e; ++q)
        if (q != a) go(q);
    } else {
      go(min(B.x, B.y));
      if (B.y != B.x) go(max(B.x, B.y));
    }
  } else {
    const int lo = B.z, hi = B.w;
    if (a == hi) {
      for (int q = s; q < e; ++q)
        if (q != a) go(q);
    } else if (a == lo) {
      for (int q = s; q < e; ++q)
        if (q != a && (q == hi || q < a)) go(q);
    } else {
      // q < hi-pin or later: the pair (a, hi) always passes; (a, lo) passes
      // when lo is the later pin.  Visit in pin order.
      if (lo > a && lo != hi && lo < hi) go(lo);
      go(hi);
      if (lo > a && lo != hi && lo > hi) go(lo);
    }
  }
}

// moveNodesCenterToCenter (placedb.cpp:591-598) / setModuleCenter_2D clamp
// (placedb.cpp:253-290): the float argument is replaced in place, so the
// second test sees the first's result.
__device__ __forceinline__ float qp_clamp(float x, float w, float lo, float hi) {
  if ((double)x - 0.5 * (double)w < (double)lo) x = (float)((double)lo + 0.5 * (double)w + kEps);
  if ((double)x + 0.5 * (double)w > (double)hi) x = (float)((double)hi - 0.5 * (double)w - kEps);
  return x;
}

__global__ void qp_place_kernel(int n, const float *__restrict__ xs, const float *__restrict__ ys, int from_center,
                                float center_x, float center_y, const float *__restrict__ w,
                                const float *__restrict__ h, float llx, float lly, float urx, float ury, float *cx,
                                float *cy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = from_center ? center_x : xs[i], y = from_center ? center_y : ys[i];
  cx[i] = qp_clamp(x, w[i], llx, urx);
  cy[i] = qp_clamp(y, h[i], lly, ury);
}

// Net::calcBoundPin for nets of <= kQpLargeNet pins, one thread each; the
// HPWL sum per block (PlaceDB::calcNetBoundPins, placedb.cpp:462-470).
__global__ void __launch_bounds__(kQpThreads) qp_bound_kernel(QpView v, int nets, int4 *bound, double *partials) {
  const int net = blockIdx.x * blockDim.x + threadIdx.x;
  double hp = 0;
  if (net < nets) {
    const int s = v.net_start[net], e = v.net_start[net + 1];
    if (e - s <= kQpLargeNet) {
      int4 B = make_int4(-1, -1, -1, -1);
      double minx = 1.7976931348623157e308, maxx = -1.7976931348623157e308, miny = minx, maxy = maxx;
      for (int q = s; q < e; ++q) {
        const double x = qp_abs(v, q, 0), y = qp_abs(v, q, 1);
        if (x < minx) minx = x, B.x = q;
        if (x > maxx) maxx = x, B.y = q;
        if (y < miny) miny = y, B.z = q;
        if (y > maxy) maxy = y, B.w = q;
      }
      bound[net] = B;
      if (e > s) hp = (maxx - minx) + (maxy - miny);
    }
  }
  hp = block_sum(hp);
  if (threadIdx.x == 0) partials[blockIdx.x] = hp;
}

// Larger nets: one block each; ties go to the lowest pin id, which is the
// first strict extremum of the sequential scan.
__global__ void __launch_bounds__(kQpThreads) qp_bound_large_kernel(QpView v, const int32_t *__restrict__ large,
                                                                    int4 *bound, double *hpwl_large) {
  const int net = large[blockIdx.x];
  const int s = v.net_start[net], e = v.net_start[net + 1];
  __shared__ double sv[4][kQpThreads];
  __shared__ int si[4][kQpThreads];
  double val[4] = {1.7976931348623157e308, -1.7976931348623157e308, 1.7976931348623157e308,
                   -1.7976931348623157e308};
  int idx[4] = {-1, -1, -1, -1};
  for (int q = s + threadIdx.x; q < e; q += blockDim.x) {
    const double x = qp_abs(v, q, 0), y = qp_abs(v, q, 1);
    if (x < val[0]) val[0] = x, idx[0] = q;
    if (x > val[1]) val[1] = x, idx[1] = q;
    if (y < val[2]) val[2] = y, idx[2] = q;
    if (y > val[3]) val[3] = y, idx[3] = q;
  }
  for (int k = 0; k < 4; ++k) sv[k][threadIdx.x] = val[k], si[k][threadIdx.x] = idx[k];
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int k = 0; k < 4; ++k) {
        const double a = sv[k][threadIdx.x], b = sv[k][threadIdx.x + off];
        const int ia = si[k][threadIdx.x], ib = si[k][threadIdx.x + off];
        const bool lo = (k & 1) == 0;
        bool take_b;
        if (ib < 0) take_b = false;
        else if (ia < 0) take_b = true;
        else take_b = lo ? (b < a || (b == a && ib < ia)) : (b > a || (b == a && ib < ia));
        if (take_b) sv[k][threadIdx.x] = b, si[k][threadIdx.x] = ib;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bound[net] = make_int4(si[0][0], si[1][0], si[2][0], si[3][0]);
    hpwl_large[blockIdx.x] = (sv[1][0] - sv[0][0]) + (sv[3][0] - sv[2][0]);
  }
}

__global__ void __launch_bounds__(kFinish) qp_hpwl_finish_kernel(const double *__restrict__ partials, int np,
                                                                   const double *__restrict__ large, int nl,
                                                                   double *out) {
  double s = block_sum_partials(partials, np);
  double l = block_sum_partials(large, nl);
  if (threadIdx.x == 0) *out = s + l;
}

// Row lengths: off-diagonal entries (movable partners) of row i (X) and
// N + i (Y).
__global__ void qp_count_kernel(QpView v, int32_t *cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * v.n) return;
  const int dim = t >= v.n, i = t - dim * v.n;
  int c = 0;
  for (int k = v.npin_start[i]; k < v.npin_start[i + 1]; ++k)
    qp_partners(v, i, v.npin_list[k], dim, [&](int q, int) { c += v.pin_obj[q] < v.n; });
  cnt[t] = c;
}

// Rows, diagonal and b (qplace.cpp:190-218 and 244-268) in a fixed order.
__global__ void qp_fill_kernel(QpView v, const int32_t *__restrict__ rowptr, int32_t *col, float *val, float *diag,
                               float *rhs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * v.n) return;
  const int dim = t >= v.n, i = t - dim * v.n;
  int w = rowptr[t];
  double dg = 0, bb = 0;
  for (int k = v.npin_start[i]; k < v.npin_start[i + 1]; ++k) {
    const int a = v.npin_list[k];
    const float xa = qp_abs(v, a, dim);
    const float oa = dim ? v.pin_dy[a] : v.pin_dx[a];
    qp_partners(v, i, a, dim, [&](int q, int net) {
      const float xq = qp_abs(v, q, dim);
      const float wt = qp_weight(v.c1[net], fabsf(xa - xq));
      const int mq = v.pin_obj[q];
      dg += wt;
      if (mq < v.n) {
        const float oq = dim ? v.pin_dy[q] : v.pin_dx[q];
        col[w] = mq;
        val[w] = -wt;
        ++w;
        bb += -1.0 * (double)wt * (double)(oa - oq);
      } else {
        bb += (double)(wt * (xq - oa));
      }
    });
  }
  diag[t] = (float)dg;
  rhs[t] = (float)bb;
}

// --- exclusive scan of int32 counts (n + 1 entries, the last one 0) -------
__device__ __forceinline__ int wave_incl_scan(int x) {
  const int lane = threadIdx.x & 63;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  return x;
}

__device__ __forceinline__ int block_excl_scan(int x, int *total) {
  __shared__ int wsum[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int incl = wave_incl_scan(x);
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int k = 0; k < nw; ++k) {
      const int s = wsum[k];
      wsum[k] = run;
      run += s;
    }
    *total = run;
  }
  __syncthreads();
  const int r = incl - x + wsum[wave];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(1024) scan_tiles_kernel(int n, const int32_t *__restrict__ in, int32_t *out,
                                                          int32_t *tile_sum) {
  __shared__ int total;
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
  int v[kScanPerThread], s = 0;
  for (int k = 0; k < kScanPerThread; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  int run = block_excl_scan(s, &total);
  for (int k = 0; k < kScanPerThread; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) scan_top_kernel(int tiles, int32_t *tile_sum) {
  __shared__ int total;
  const int per = (tiles + blockDim.x - 1) / blockDim.x;
  const int t0 = threadIdx.x * per, t1 = min(tiles, t0 + per);
  int s = 0;
  for (int t = t0; t < t1; ++t) s += tile_sum[t];
  int run = block_excl_scan(s, &total);
  for (int t = t0; t < t1; ++t) {
    const int x = tile_sum[t];
    tile_sum[t] = run;
    run += x;
  }
}

__global__ void scan_add_kernel(int n, const int32_t *__restrict__ tile_sum, int32_t *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_sum[i / kScanTile];
}

// --- BiCGSTAB ---------------------------------------------------------------
struct QpSolveScalars {
  float rho, rho_old, alpha, w, beta, r_sq, r0_sq, rhs_sq, tol2, error, rho_next;
  int iter, restarts, active, restart, zero, max_iter;
};

struct QpSys {
  int n;
  const int32_t *rowptr, *col;  // rows d*n .. d*n+n-1 belong to direction d
  const float *val, *diag, *b;
  float *x, *r, *r0, *p, *v, *s, *t;
};

// Sum over the kQpLanes lanes of one row of val * vec[col]; lane 0 of the
// group gets the result.  Fixed order: lane-strided partial sums, then a
// fixed shuffle tree.
__device__ __forceinline__ double qp_row(const QpSys &S, int row, int lane, const float *__restrict__ vec) {
  double acc0 = 0, acc1 = 0;
  int k = S.rowptr[row] + lane;
  const int e = S.rowptr[row + 1];
  for (; k + kQpLanes < e; k += 2 * kQpLanes) {
    const int c0 = S.col[k], c1 = S.col[k + kQpLanes];
    const float v0 = S.val[k], v1 = S.val[k + kQpLanes];
    acc0 += (double)v0 * (double)vec[c0];
    acc1 += (double)v1 * (double)vec[c1];
  }
  if (k < e) acc0 += (double)S.val[k] * (double)vec[S.col[k]];
  double acc = acc0 + acc1;
  for (int off = kQpLanes / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, kQpLanes);
  return acc;
}

__device__ __forceinline__ double qp_diag(const QpSys &S, int row) { return S.diag ? (double)S.diag[row] : 0.0; }

// r = b - A x, r0 = r; partial r.r and b.b (BiCGSTAB.h: the set-up).  When
// `restart_only`, only directions flagged for a restart do it.
__global__ void __launch_bounds__(kQpThreads) qp_residual_kernel(QpSys S, QpSolveScalars *sc, int restart_only,
                                                                 double *partials) {
  const int d = blockIdx.y;
  const QpSolveScalars s = sc[d];
  if (restart_only && !s.restart) return;
  const int n = S.n, off = d * n;
  const int lane = threadIdx.x % kQpLanes;
  double rr = 0, bb = 0;
  for (int row = (blockIdx.x * blockDim.x + threadIdx.x) / kQpLanes; row < n;
       row += gridDim.x * blockDim.x / kQpLanes) {
    const double ax = qp_row(S, off + row, lane, S.x + off);
    if (lane == 0) {
      const float bi = S.b[off + row];
      const float ri = bi - (float)(qp_diag(S, off + row) * (double)S.x[off + row] + ax);
      S.r[off + row] = ri;
      S.r0[off + row] = ri;
      rr += (double)ri * ri;
      bb += (double)bi * bi;
    }
  }
  rr = block_sum(rr);
  bb = block_sum(bb);
  if (threadIdx.x == 0) {
    partials[(2 * d) * kQpMaxBlocks + blockIdx.x] = rr;
    partials[(2 * d + 1) * kQpMaxBlocks + blockIdx.x] = bb;
  }
}

// First loop test of bicgstab() and the first rho / beta.
__global__ void __launch_bounds__(kFinish) qp_setup_finish_kernel(QpSolveScalars *sc, const double *partials,
                                                                   int nblocks, float tol, int max_iter) {
  const int d = blockIdx.x;
  const double *const src[2] = {partials + (2 * d) * kQpMaxBlocks, partials + (2 * d + 1) * kQpMaxBlocks};
  double sum[2];
  block_sum_partials_strided_n(src, nblocks, 1, sum);
  const float rr = (float)sum[0], bb = (float)sum[1];
  if (threadIdx.x) return;
  QpSolveScalars s = QpSolveScalars{};
  s.max_iter = max_iter;
  s.r0_sq = rr;
  s.rhs_sq = bb;
  s.r_sq = rr;
  s.error = tol;
  s.rho = s.alpha = s.w = 1.f;
  if (bb == 0.f) {
    s.zero = 1;  // x = 0, error stays the tolerance
  } else {
    s.tol2 = tol * tol * bb;
    const float eps = 1.1920928955078125e-07f;
    s.active = s.r_sq > s.tol2 && 0 < max_iter;
    if (s.active) {
      s.rho_old = s.rho;
      s.rho = rr;  // r0 . r with r0 = r
      if (fabsf(s.rho) < eps * eps * s.r0_sq) s.restart = 1;
      else s.beta = (s.rho / s.rho_old) * (s.alpha / s.w);
    }
  }
  sc[d] = s;
}

// p = r + beta (p - w v), in place.
__global__ void __launch_bounds__(kQpThreads) qp_pvec_kernel(QpSys S, const QpSolveScalars *sc) {
  const int d = blockIdx.y;
  const QpSolveScalars s = sc[d];
  if (!s.active || s.restart) return;
  const int n = S.n, off = d * n;
  const float beta = s.beta, w = s.w;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    S.p[off + i] = S.r[off + i] + beta * (S.p[off + i] - w * S.v[off + i]);
}

// s = r - alpha v.
__global__ void __launch_bounds__(kQpThreads) qp_svec_kernel(QpSys S, const QpSolveScalars *sc) {
  const int d = blockIdx.y;
  const QpSolveScalars s = sc[d];
  if (!s.active || s.restart) return;
  const int n = S.n, off = d * n;
  const float alpha = s.alpha;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    S.s[off + i] = S.r[off + i] - alpha * S.v[off + i];
}

// which 0: v = A p, partial r0.v;  which 1: t = A s, partials t.t and t.s.
__global__ void __launch_bounds__(kQpThreads) qp_spmv_kernel(QpSys S, const QpSolveScalars *sc, int which,
                                                             double *partials) {
  const int d = blockIdx.y;
  const QpSolveScalars s = sc[d];
  if (!s.active || s.restart) return;
  const int n = S.n, off = d * n, lane = threadIdx.x % kQpLanes;
  const float *in = (which ? S.s : S.p) + off;
  float *out = (which ? S.t : S.v) + off;
  const float *dot = (which ? S.s : S.r0) + off;
  double d0 = 0, d1 = 0;
  for (int row = (blockIdx.x * blockDim.x + threadIdx.x) / kQpLanes; row < n;
       row += gridDim.x * blockDim.x / kQpLanes) {
    const double a = qp_row(S, off + row, lane, in);
    if (lane == 0) {
      const float yi = (float)(qp_diag(S, off + row) * (double)in[row] + a);
      out[row] = yi;
      if (which) {
        d0 += (double)yi * yi;
        d1 += (double)yi * dot[row];
      } else {
        d0 += (double)dot[row] * yi;
      }
    }
  }
  d0 = block_sum(d0);
  if (which) d1 = block_sum(d1);
  if (threadIdx.x == 0) {
    partials[(2 * d) * kQpMaxBlocks + blockIdx.x] = d0;
    if (which) partials[(2 * d + 1) * kQpMaxBlocks + blockIdx.x] = d1;
  }
}

__global__ void __launch_bounds__(kFinish) qp_alpha_kernel(QpSolveScalars *sc, const double *partials, int nblocks) {
  const int d = blockIdx.x;
  if (!sc[d].active || sc[d].restart) return;
  const float r0v = (float)block_sum_partials(partials + (2 * d) * kQpMaxBlocks, nblocks);
  if (threadIdx.x == 0) sc[d].alpha = sc[d].rho / r0v;
}

__global__ void __launch_bounds__(kFinish) qp_omega_kernel(QpSolveScalars *sc, const double *partials, int nblocks) {
  const int d = blockIdx.x;
  if (!sc[d].active || sc[d].restart) return;
  const double *const src[2] = {partials + (2 * d) * kQpMaxBlocks, partials + (2 * d + 1) * kQpMaxBlocks};
  double sum[2];
  block_sum_partials_strided_n(src, nblocks, 1, sum);
  const float tt = (float)sum[0], ts = (float)sum[1];
  if (threadIdx.x == 0) sc[d].w = tt > 0.f ? ts / tt : 0.f;
}

// x += alpha p + w s; r = s - w t; partial r.r and r0.r.
__global__ void __launch_bounds__(kQpThreads) qp_update_kernel(QpSys S, const QpSolveScalars *sc, double *partials) {
  const int d = blockIdx.y;
  const QpSolveScalars s = sc[d];
  if (!s.active || s.restart) return;
  const int n = S.n, off = d * n;
  const float alpha = s.alpha, w = s.w;
  double rr = 0, r0r = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float si = S.s[off + i];
    S.x[off + i] = S.x[off + i] + (alpha * S.p[off + i] + w * si);
    const float ri = si - w * S.t[off + i];
    S.r[off + i] = ri;
    rr += (double)ri * ri;
    r0r += (double)S.r0[off + i] * ri;
  }
  rr = block_sum(rr);
  r0r = block_sum(r0r);
  if (threadIdx.x == 0) {
    partials[(2 * d) * kQpMaxBlocks + blockIdx.x] = rr;
    partials[(2 * d + 1) * kQpMaxBlocks + blockIdx.x] = r0r;
  }
}

// End of one bicgstab() iteration and the head of the next: the loop test,
// rho, the restart test, beta.
__global__ void __launch_bounds__(kFinish) qp_iter_finish_kernel(QpSolveScalars *sc, const double *partials,
                                                                  int nblocks) {
  const int d = blockIdx.x;
  if (!sc[d].active || sc[d].restart) return;
  const double *const src[2] = {partials + (2 * d) * kQpMaxBlocks, partials + (2 * d + 1) * kQpMaxBlocks};
  double sum[2];
  block_sum_partials_strided_n(src, nblocks, 1, sum);
  const float rr = (float)sum[0], r0r = (float)sum[1];
  if (threadIdx.x) return;
  QpSolveScalars s = sc[d];
  s.iter += 1;
  s.r_sq = rr;
  s.active = s.r_sq > s.tol2 && s.iter < s.max_iter;
  if (s.active) {
    const float eps = 1.1920928955078125e-07f;
    s.rho_old = s.rho;
    s.rho = r0r;
    if (fabsf(s.rho) < eps * eps * s.r0_sq) s.restart = 1;
    else s.beta = (s.rho / s.rho_old) * (s.alpha / s.w);
  }
  sc[d] = s;
}

// After a restart's residual: rho = r0_sq = |r|^2, the first restart resets
// the iteration count, then beta.
__global__ void __launch_bounds__(kFinish) qp_restart_finish_kernel(QpSolveScalars *sc, const double *partials,
                                                                     int nblocks) {
  const int d = blockIdx.x;
  if (!sc[d].restart) return;
  const float rr = (float)block_sum_partials(partials + (2 * d) * kQpMaxBlocks, nblocks);
  if (threadIdx.x) return;
  QpSolveScalars s = sc[d];
  s.rho = s.r0_sq = rr;
  if (s.restarts++ == 0) s.iter = 0;
  s.beta = (s.rho / s.rho_old) * (s.alpha / s.w);
  s.restart = 0;
  sc[d] = s;
}

__global__ void qp_zero_kernel(int64_t n, float *a, float *b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.f, b[i] = 0.f;
}

// Final error (tol_error = sqrt(|r|^2 / |b|^2)) and x = 0 for b = 0.
__global__ void qp_solve_end_kernel(int n, int ndim, QpSolveScalars *sc, float *x) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ndim * n) return;
  const int d = t / n;
  if (sc[d].zero) x[t] = 0.f;
  if (t % n == 0 && !sc[d].zero) sc[d].error = sqrtf(sc[d].r_sq / sc[d].rhs_sq);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
#define QP_CHECK(x)                                                                   \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) throw std::string(#x ": ") + hipGetErrorString(e_);       \
  } while (0)

template <class T>
T *dalloc(std::vector<void *> &owned, size_t n) {
  void *p = nullptr;
  QP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  owned.push_back(p);
  return (T *)p;
}

template <class T>
T *dupload(std::vector<void *> &owned, const T *src, size_t n, hipStream_t st) {
  T *p = dalloc<T>(owned, n);
  if (n) QP_CHECK(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, st));
  return p;
}

int grid_for(int64_t n, int per_block) { return (int)std::max<int64_t>(1, (n + per_block - 1) / per_block); }

// BiCGSTAB state on the device for ndim systems of n rows.
struct Solver {
  QpSys S{};
  QpSolveScalars *sc = nullptr;
  double *partials = nullptr;
  int nblocks = 1, ndim = 1;

  void create(std::vector<void *> &owned, int n, int nd) {
    ndim = nd;
    const size_t m = (size_t)n * nd;
    S.n = n;
    S.x = dalloc<float>(owned, m);
    S.r = dalloc<float>(owned, m);
    S.r0 = dalloc<float>(owned, m);
    S.p = dalloc<float>(owned, m);
    S.v = dalloc<float>(owned, m);
    S.s = dalloc<float>(owned, m);
    S.t = dalloc<float>(owned, m);
    sc = dalloc<QpSolveScalars>(owned, nd);
    partials = dalloc<double>(owned, (size_t)2 * nd * kQpMaxBlocks);
    nblocks = std::min(kQpMaxBlocks, grid_for(n, kQpThreads / kQpLanes));
  }

  // Solves from the guess in S.x; out[ndim] receives the final scalars of
  // each direction (iterations, error).
  void solve(hipStream_t st, int max_iter, float tol, QpSolveScalars *out) {
    const int n = S.n;
    const dim3 g(nblocks, ndim), gu(std::min(kQpMaxBlocks, grid_for(n, kQpThreads)), ndim);
    const int64_t m = (int64_t)n * ndim;
    hipLaunchKernelGGL(qp_zero_kernel, dim3(grid_for(m, 256)), dim3(256), 0, st, m, S.p, S.v);
    hipLaunchKernelGGL(qp_residual_kernel, g, dim3(kQpThreads), 0, st, S, sc, 0, partials);
    hipLaunchKernelGGL(qp_setup_finish_kernel, dim3(ndim), dim3(kFinish), 0, st, sc, partials, nblocks, tol,
                       max_iter);
    // Restarts reset the count once, so a direction may run up to
    // 2 max_iter iterations; the host checks every kCheck iterations.
    constexpr int kCheck = 16;
    for (int d = 0; d < ndim; ++d) out[d] = QpSolveScalars{};
    for (int it = 0;; ++it) {
      hipLaunchKernelGGL(qp_residual_kernel, g, dim3(kQpThreads), 0, st, S, sc, 1, partials);
      hipLaunchKernelGGL(qp_restart_finish_kernel, dim3(ndim), dim3(kFinish), 0, st, sc, partials, nblocks);
      hipLaunchKernelGGL(qp_pvec_kernel, gu, dim3(kQpThreads), 0, st, S, sc);
      hipLaunchKernelGGL(qp_spmv_kernel, g, dim3(kQpThreads), 0, st, S, sc, 0, partials);
      hipLaunchKernelGGL(qp_alpha_kernel, dim3(ndim), dim3(kFinish), 0, st, sc, partials, nblocks);
      hipLaunchKernelGGL(qp_svec_kernel, gu, dim3(kQpThreads), 0, st, S, sc);
      hipLaunchKernelGGL(qp_spmv_kernel, g, dim3(kQpThreads), 0, st, S, sc, 1, partials);
      hipLaunchKernelGGL(qp_omega_kernel, dim3(ndim), dim3(kFinish), 0, st, sc, partials, nblocks);
      hipLaunchKernelGGL(qp_update_kernel, gu, dim3(kQpThreads), 0, st, S, sc, partials);
      hipLaunchKernelGGL(qp_iter_finish_kernel, dim3(ndim), dim3(kFinish), 0, st, sc, partials, gu.x);
      if ((it + 1) % kCheck == 0 || it + 1 >= 2 * max_iter) {
        QP_CHECK(hipMemcpyAsync(out, sc, ndim * sizeof(QpSolveScalars), hipMemcpyDeviceToHost, st));
        QP_CHECK(hipStreamSynchronize(st));
        bool any = false;
        for (int d = 0; d < ndim; ++d) any |= out[d].active != 0;
        if (!any || it + 1 >= 2 * max_iter) break;
      }
    }
    hipLaunchKernelGGL(qp_solve_end_kernel, dim3(grid_for(m, 256)), dim3(256), 0, st, n, ndim, sc, S.x);
    QP_CHECK(hipMemcpyAsync(out, sc, ndim * sizeof(QpSolveScalars), hipMemcpyDeviceToHost, st));
    QP_CHECK(hipStreamSynchronize(st));
  }
};

// Design on the device plus the assembled systems.
struct QpCtx {
  std::vector<void *> owned;
  hipStream_t st = nullptr;
  int n = 0, T = 0, E = 0;
  int64_t P = 0;
  QpView v{};
  float *cx = nullptr, *cy = nullptr, *w = nullptr, *h = nullptr;
  int4 *bound = nullptr;
  int32_t *large = nullptr;
  int nlarge = 0;
  double *bpart = nullptr, *blarge = nullptr, *hpwl = nullptr;
  int32_t *cnt = nullptr, *rowptr = nullptr, *col = nullptr, *tiles = nullptr;
  float *val = nullptr, *diag = nullptr, *rhs = nullptr;
  int64_t cap = 0;
  float core[4] = {0, 0, 0, 0};

  ~QpCtx() {
    if (st) (void)hipStreamSynchronize(st);
    for (void *p : owned) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }

  void create(const ep_design *d, int device) {
    QP_CHECK(hipSetDevice(device));
    QP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    n = d->num_nodes;
    T = d->num_terminals;
    E = d->num_nets;
    P = d->num_pins;
    if (P >= (int64_t)1 << 28) throw std::string("too many pins for the quadratic placer");
    std::vector<int32_t> ns(E + 1), pnet(P), nps(n + 1, 0), c1i;
    std::vector<float> c1(E);
    std::vector<int32_t> lg;
    cap = 0;
    for (int e = 0; e <= E; ++e) ns[e] = (int32_t)d->net_start[e];
    for (int e = 0; e < E; ++e) {
      const int64_t p = d->net_start[e + 1] - d->net_start[e];
      // constant1 = 1.0 / ((float)pinCount - 1.0) (qplace.cpp:151)
      c1[e] = (float)(1.0 / ((double)(float)p - 1.0));
      if (p > kQpLargeNet) lg.push_back(e);
      for (int64_t q = d->net_start[e]; q < d->net_start[e + 1]; ++q) pnet[q] = e;
      if (p >= 2) cap += 2 * (2 * p - 3);  // entries of one direction: 2 per pair
    }
    cap *= 2;
    if (cap >= (int64_t)INT32_MAX) throw std::string("quadratic placer matrix too large");
    for (int64_t q = 0; q < P; ++q)
      if (d->pin_obj[q] < n) ++nps[d->pin_obj[q] + 1];
    for (int i = 0; i < n; ++i) nps[i + 1] += nps[i];
    std::vector<int32_t> npl(nps[n]);
    {
      std::vector<int32_t> cur(nps.begin(), nps.end() - 1);
      for (int64_t q = 0; q < P; ++q)
        if (d->pin_obj[q] < n) npl[cur[d->pin_obj[q]]++] = (int32_t)q;
    }
    std::vector<float> tx(T), ty(T);
    for (int t = 0; t < T; ++t) tx[t] = d->term_x[t], ty[t] = d->term_y[t];
    // coreRegion (placedb.cpp:52-77)
    float llx = (float)d->row_x0[0], urx = (float)d->row_x1[0];
    for (int r = 0; r < d->num_rows; ++r) llx = std::min(llx, (float)d->row_x0[r]), urx = std::max(urx, (float)d->row_x1[r]);
    core[0] = llx;
    core[1] = (float)d->row_bottom[0];
    core[2] = urx;
    core[3] = (float)d->row_bottom[d->num_rows - 1] + (float)d->row_height[d->num_rows - 1];

    v.n = n;
    v.net_start = dupload(owned, ns.data(), ns.size(), st);
    v.pin_obj = dupload(owned, d->pin_obj, (size_t)P, st);
    v.pin_net = dupload(owned, pnet.data(), pnet.size(), st);
    v.npin_start = dupload(owned, nps.data(), nps.size(), st);
    v.npin_list = dupload(owned, npl.data(), npl.size(), st);
    v.pin_dx = dupload(owned, d->pin_dx, (size_t)P, st);
    v.pin_dy = dupload(owned, d->pin_dy, (size_t)P, st);
    v.tx = dupload(owned, tx.data(), tx.size(), st);
    v.ty = dupload(owned, ty.data(), ty.size(), st);
    v.c1 = dupload(owned, c1.data(), c1.size(), st);
    w = dupload(owned, d->node_w, (size_t)n, st);
    h = dupload(owned, d->node_h, (size_t)n, st);
    cx = dalloc<float>(owned, n);
    cy = dalloc<float>(owned, n);
    v.cx = cx;
    v.cy = cy;
    bound = dalloc<int4>(owned, E);
    v.bound = bound;
    nlarge = (int)lg.size();
    large = dupload(owned, lg.data(), lg.size(), st);
    bpart = dalloc<double>(owned, grid_for(E, kQpThreads));
    blarge = dalloc<double>(owned, std::max(nlarge, 1));
    hpwl = dalloc<double>(owned, 1);
    cnt = dalloc<int32_t>(owned, 2 * (size_t)n + 1);
    rowptr = dalloc<int32_t>(owned, 2 * (size_t)n + 1);
    tiles = dalloc<int32_t>(owned, grid_for(2 * (int64_t)n + 1, kScanTile));
    col = dalloc<int32_t>(owned, cap);
    val = dalloc<float>(owned, cap);
    diag = dalloc<float>(owned, 2 * (size_t)n);
    rhs = dalloc<float>(owned, 2 * (size_t)n);
    QP_CHECK(hipMemsetAsync(cnt, 0, (2 * (size_t)n + 1) * sizeof(int32_t), st));
  }

  // Node centres from host (or the core centre) through the clamp.
  void place(const float *xs, const float *ys, bool from_center) {
    const float ccx = (float)(0.5 * (double)(core[2] + core[0])), ccy = (float)(0.5 * (double)(core[3] + core[1]));
    hipLaunchKernelGGL(qp_place_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, n, xs, ys, from_center ? 1 : 0,
                       ccx, ccy, w, h, core[0], core[1], core[2], core[3], cx, cy);
  }

  void set_centres(const float *xy) {
    std::vector<float> x(n), y(n);
    for (int i = 0; i < n; ++i) x[i] = xy[2 * i], y[i] = xy[2 * i + 1];
    QP_CHECK(hipMemcpyAsync(cx, x.data(), n * 4, hipMemcpyHostToDevice, st));
    QP_CHECK(hipMemcpyAsync(cy, y.data(), n * 4, hipMemcpyHostToDevice, st));
    QP_CHECK(hipStreamSynchronize(st));
  }

  double bounds() {
    const int nb = grid_for(E, kQpThreads);
    hipLaunchKernelGGL(qp_bound_kernel, dim3(nb), dim3(kQpThreads), 0, st, v, E, bound, bpart);
    if (nlarge)
      hipLaunchKernelGGL(qp_bound_large_kernel, dim3(nlarge), dim3(kQpThreads), 0, st, v, large, bound, blarge);
    hipLaunchKernelGGL(qp_hpwl_finish_kernel, dim3(1), dim3(kFinish), 0, st, bpart, nb, blarge, nlarge, hpwl);
    double h = 0;
    QP_CHECK(hipMemcpyAsync(&h, hpwl, sizeof(double), hipMemcpyDeviceToHost, st));
    QP_CHECK(hipStreamSynchronize(st));
    return h;
  }

  // createSparseMatrix; returns the number of off-diagonal entries.
  int64_t assemble() {
    const int rows = 2 * n;
    hipLaunchKernelGGL(qp_count_kernel, dim3(grid_for(rows, 256)), dim3(256), 0, st, v, cnt);
    const int tiles_n = grid_for(rows + 1, kScanTile);
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(tiles_n), dim3(1024), 0, st, rows + 1, cnt, rowptr, tiles);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(1024), 0, st, tiles_n, tiles);
    hipLaunchKernelGGL(scan_add_kernel, dim3(grid_for(rows + 1, 256)), dim3(256), 0, st, rows + 1, tiles, rowptr);
    int32_t nnz = 0;
    QP_CHECK(hipMemcpyAsync(&nnz, rowptr + rows, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    QP_CHECK(hipStreamSynchronize(st));
    if (nnz > cap) throw std::string("quadratic placer matrix exceeds its bound");
    hipLaunchKernelGGL(qp_fill_kernel, dim3(grid_for(rows, 256)), dim3(256), 0, st, v, rowptr, col, val, diag, rhs);
    return nnz;
  }
};

// ---------------------------------------------------------------------------
// Work space of ep_qp_solve_noalloc: the six BiCGSTAB work vectors (r, r0,
// p, v, s, t) of up to EP_QP_NOALLOC_MAX_ROWS rows, the per-block partial
// sums, the per-direction scalars and a validation flag, in one static
// device array.  Each device that runs the library holds its own copy until
// the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device.
constexpr size_t kNaRows = EP_QP_NOALLOC_MAX_ROWS;
constexpr size_t kNaVecBytes = kNaRows * sizeof(float);
constexpr size_t kNaPartialOff = 6 * kNaVecBytes;
constexpr size_t kNaScalarOff = kNaPartialOff + (size_t)2 * 2 * kQpMaxBlocks * sizeof(double);
constexpr size_t kNaFlagOff = kNaScalarOff + 256;
constexpr size_t kNaBytes = kNaFlagOff + 256;
static_assert(2 * sizeof(QpSolveScalars) <= 256, "scalar slot");
static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
unsigned char *g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on `device` (the current device), looked up
// once per device.
unsigned char *scratch_on(int device) {
  if (device < 0 || device >= kNaMaxDevices) throw std::string("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_na_mu);
  if (!g_na_ptr[device]) {
    void *p = nullptr;
    QP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::string("hipGetSymbolAddress returned no address for the work space");
    g_na_ptr[device] = static_cast<unsigned char *>(p);
  }
  return g_na_ptr[device];
}

__global__ void qp_flag_clear_kernel(int *flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0;
}

// Checks a device CSR before the solve reads it through its column
// indices: rowptr[0] = 0, rowptr[rows] = nnz, rows non-decreasing, columns in
// [0, n).  Reads only rowptr[0..rows] and col[k] for 0 <= k < nnz.  Bits of
// *flag: 1 bad row bounds, 2 decreasing rows, 4 column out of range.
__global__ void qp_check_system_kernel(int rows, int n, int64_t nnz, const int32_t *__restrict__ rowptr,
                                       const int32_t *__restrict__ col, int *flag) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int64_t a = rowptr[r], b = rowptr[r + 1];
  int bad = 0;
  if (r == 0 && a != 0) bad |= 1;
  if (r == rows - 1 && b != nnz) bad |= 1;
  if (b < a) {
    bad |= 2;
  } else if (a < 0 || b > nnz) {
    bad |= 1;
  } else {
    for (int64_t k = a; k < b; ++k)
      if (col[k] < 0 || col[k] >= n) {
        bad |= 4;
        break;
      }
  }
  if (bad) atomicOr(flag, bad);
}

thread_local std::string g_qp_err;

int qp_fail(const std::string &m, int code) {
  g_qp_err = m;
  return code;
}

bool design_ok(const ep_design *d) {
  return d && d->num_nodes > 0 && d->num_nets >= 0 && d->num_rows > 0 && d->num_pins >= 0 && d->net_start &&
         d->pin_obj && d->pin_dx && d->pin_dy && d->node_w && d->node_h && d->row_bottom && d->row_height &&
         d->row_x0 && d->row_x1 && (d->num_terminals == 0 || (d->term_x && d->term_y));
}

}  // namespace
}  // namespace ep

using namespace ep;

extern "C" {

const char *ep_qp_last_error(void) { return g_qp_err.c_str(); }

int ep_qplace(const ep_design *design, const ep_qp_params *params, float *node_xy, double *log,
              ep_qp_stats *stats) {
  g_qp_err.clear();
  if (!design_ok(design) || !params || !node_xy) return qp_fail("bad arguments", EP_ERR_ARG);
  if (params->max_iterations < 0 || params->solver_iterations < 0) return qp_fail("negative iteration count", EP_ERR_ARG);
  try {
    auto t0 = std::chrono::steady_clock::now();
    QpCtx c;
    c.create(design, params->device);
    const int n = c.n;
    Solver sv;
    sv.create(c.owned, n, 2);
    if (params->start_at_center) {
      c.place(nullptr, nullptr, true);
    } else {
      c.set_centres(node_xy);
    }
    // moveNodesCenterToCenter / calcHPWL / calcNetBoundPins (qplace.cpp:23-29)
    const double h0 = c.bounds();
    ep_qp_stats s{};
    s.initial_hpwl = h0;
    int it = 0;
    QpSolveScalars res[2];
    for (;; ++it) {
      const int64_t nnz = c.assemble();
      (void)nnz;
      sv.S.rowptr = c.rowptr;
      sv.S.col = c.col;
      sv.S.val = c.val;
      sv.S.diag = c.diag;
      sv.S.b = c.rhs;
      // solveWithGuess: the guess is the current centre (qplace.cpp:139-140)
      QP_CHECK(hipMemcpyAsync(sv.S.x, c.cx, n * 4, hipMemcpyDeviceToDevice, c.st));
      QP_CHECK(hipMemcpyAsync(sv.S.x + n, c.cy, n * 4, hipMemcpyDeviceToDevice, c.st));
      sv.solve(c.st, params->solver_iterations, params->solver_tolerance, res);
      // updateModuleLocation (qplace.cpp:283-295)
      c.place(sv.S.x, sv.S.x + n, false);
      const double hp = c.bounds();
      s.solver_iterations += res[0].iter + res[1].iter;
      s.restarts += res[0].restarts + res[1].restarts;
      s.error_x = res[0].error;
      s.error_y = res[1].error;
      s.final_hpwl = hp;
      if (log) {
        log[4 * it + 0] = hp;
        log[4 * it + 1] = res[0].error;
        log[4 * it + 2] = res[1].error;
        log[4 * it + 3] = (double)(res[0].iter + res[1].iter);
      }
      // qplace.cpp:94-101
      if (fabsf(res[0].error) < params->target_error && fabsf(res[1].error) < params->target_error && it > 4) break;
      if (it >= params->max_iterations) break;
    }
    s.iterations = it + 1;
    std::vector<float> x(n), y(n);
    QP_CHECK(hipMemcpyAsync(x.data(), c.cx, n * 4, hipMemcpyDeviceToHost, c.st));
    QP_CHECK(hipMemcpyAsync(y.data(), c.cy, n * 4, hipMemcpyDeviceToHost, c.st));
    QP_CHECK(hipStreamSynchronize(c.st));
    for (int i = 0; i < n; ++i) node_xy[2 * i] = x[i], node_xy[2 * i + 1] = y[i];
    s.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = s;
    return EP_OK;
  } catch (const std::string &e) {
    return qp_fail(e, EP_ERR_HIP);
  } catch (const std::bad_alloc &) {
    return qp_fail("host allocation failed", EP_ERR_NOMEM);
  }
}

int ep_qp_matrix(const ep_design *design, const float *node_xy, int32_t device, ep_qp_matrix_out *out) {
  g_qp_err.clear();
  if (!design_ok(design) || !node_xy || !out) return qp_fail("bad arguments", EP_ERR_ARG);
  try {
    QpCtx c;
    c.create(design, device);
    const int n = c.n;
    c.set_centres(node_xy);
    out->hpwl = c.bounds();
    const int64_t nnz = c.assemble();
    out->nnz = nnz;
    if (out->rowptr && out->col && out->val && out->diag && out->b) {
      if (out->capacity < nnz) return qp_fail("capacity below nnz", EP_ERR_ARG);
      QP_CHECK(hipMemcpyAsync(out->rowptr, c.rowptr, (2 * (size_t)n + 1) * 4, hipMemcpyDeviceToHost, c.st));
      QP_CHECK(hipMemcpyAsync(out->col, c.col, nnz * 4, hipMemcpyDeviceToHost, c.st));
      QP_CHECK(hipMemcpyAsync(out->val, c.val, nnz * 4, hipMemcpyDeviceToHost, c.st));
      QP_CHECK(hipMemcpyAsync(out->diag, c.diag, 2 * (size_t)n * 4, hipMemcpyDeviceToHost, c.st));
      QP_CHECK(hipMemcpyAsync(out->b, c.rhs, 2 * (size_t)n * 4, hipMemcpyDeviceToHost, c.st));
      QP_CHECK(hipStreamSynchronize(c.st));
    }
    return EP_OK;
  } catch (const std::string &e) {
    return qp_fail(e, EP_ERR_HIP);
  } catch (const std::bad_alloc &) {
    return qp_fail("host allocation failed", EP_ERR_NOMEM);
  }
}

int ep_qp_solve(const ep_qp_system *sys, float *x, int32_t max_iterations, float tolerance, int32_t device,
                float *error, int32_t *iterations) {
  g_qp_err.clear();
  if (!sys || !x || sys->n <= 0 || sys->ndim < 1 || sys->ndim > 2 || !sys->rowptr || !sys->b ||
      (sys->nnz > 0 && (!sys->col || !sys->val)) || max_iterations < 0)
    return qp_fail("bad arguments", EP_ERR_ARG);
  const int n = sys->n, nd = sys->ndim;
  const size_t rows = (size_t)n * nd;
  if (sys->rowptr[0] != 0 || sys->rowptr[rows] != sys->nnz) return qp_fail("rowptr does not match nnz", EP_ERR_ARG);
  for (size_t r = 0; r < rows; ++r)
    if (sys->rowptr[r + 1] < sys->rowptr[r]) return qp_fail("rowptr not monotone", EP_ERR_ARG);
  for (int64_t k = 0; k < sys->nnz; ++k)
    if (sys->col[k] < 0 || sys->col[k] >= n) return qp_fail("column index out of range", EP_ERR_ARG);
  try {
    QP_CHECK(hipSetDevice(device));
    std::vector<void *> owned;
    hipStream_t st = nullptr;
    QP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct Guard {
      std::vector<void *> &o;
      hipStream_t s;
      ~Guard() {
        (void)hipStreamSynchronize(s);
        for (void *p : o) (void)hipFree(p);
        (void)hipStreamDestroy(s);
      }
    } guard{owned, st};
    Solver sv;
    sv.create(owned, n, nd);
    sv.S.rowptr = dupload(owned, sys->rowptr, rows + 1, st);
    sv.S.col = dupload(owned, sys->col, (size_t)sys->nnz, st);
    sv.S.val = dupload(owned, sys->val, (size_t)sys->nnz, st);
    sv.S.diag = sys->diag ? dupload(owned, sys->diag, rows, st) : nullptr;
    sv.S.b = dupload(owned, sys->b, rows, st);
    QP_CHECK(hipMemcpyAsync(sv.S.x, x, rows * 4, hipMemcpyHostToDevice, st));
    QpSolveScalars res[2];
    sv.solve(st, max_iterations, tolerance, res);
    QP_CHECK(hipMemcpyAsync(x, sv.S.x, rows * 4, hipMemcpyDeviceToHost, st));
    QP_CHECK(hipStreamSynchronize(st));
    for (int d = 0; d < nd; ++d) {
      if (error) error[d] = res[d].error;
      if (iterations) iterations[d] = res[d].iter;
    }
    return EP_OK;
  } catch (const std::string &e) {
    return qp_fail(e, EP_ERR_HIP);
  } catch (const std::bad_alloc &) {
    return qp_fail("host allocation failed", EP_ERR_NOMEM);
  }
}

int ep_qp_solve_noalloc(const ep_qp_system *sys, float *x, int32_t max_iterations, float tolerance, int32_t device,
                        void *stream, float *error, int32_t *iterations) {
  g_qp_err.clear();
  if (!sys || !x || sys->n <= 0 || sys->ndim < 1 || sys->ndim > 2 || !sys->rowptr || !sys->b ||
      (sys->nnz > 0 && (!sys->col || !sys->val)) || sys->nnz < 0 || max_iterations < 0)
    return qp_fail("bad arguments", EP_ERR_ARG);
  const int n = sys->n, nd = sys->ndim;
  const int64_t rows = (int64_t)n * nd;
  if (rows > (int64_t)EP_QP_NOALLOC_MAX_ROWS)
    return qp_fail("system of " + std::to_string(rows) + " rows exceeds the allocation-free work space of " +
                       std::to_string((long long)EP_QP_NOALLOC_MAX_ROWS) + " rows",
                   EP_ERR_ARG);
  if (sys->nnz > (int64_t)INT32_MAX) return qp_fail("too many entries", EP_ERR_ARG);
  try {
    QP_CHECK(hipSetDevice(device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned char *ws = scratch_on(device);
    int *flag = reinterpret_cast<int *>(ws + kNaFlagOff);
    hipLaunchKernelGGL(qp_flag_clear_kernel, dim3(1), dim3(64), 0, st, flag);
    hipLaunchKernelGGL(qp_check_system_kernel, dim3(grid_for(rows, 256)), dim3(256), 0, st, (int)rows, n, sys->nnz,
                       sys->rowptr, sys->col, flag);
    int bad = 0;
    QP_CHECK(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    QP_CHECK(hipStreamSynchronize(st));
    if (bad & 4) return qp_fail("column index out of range", EP_ERR_ARG);
    if (bad & 2) return qp_fail("rowptr not monotone", EP_ERR_ARG);
    if (bad) return qp_fail("rowptr does not match nnz", EP_ERR_ARG);
    Solver sv;
    sv.ndim = nd;
    sv.nblocks = std::min(kQpMaxBlocks, grid_for(n, kQpThreads / kQpLanes));
    sv.S.n = n;
    sv.S.x = x;
    float *vec[6];
    for (int i = 0; i < 6; ++i) vec[i] = reinterpret_cast<float *>(ws + i * kNaVecBytes);
    sv.S.r = vec[0];
    sv.S.r0 = vec[1];
    sv.S.p = vec[2];
    sv.S.v = vec[3];
    sv.S.s = vec[4];
    sv.S.t = vec[5];
    sv.partials = reinterpret_cast<double *>(ws + kNaPartialOff);
    sv.sc = reinterpret_cast<QpSolveScalars *>(ws + kNaScalarOff);
    sv.S.rowptr = sys->rowptr;
    sv.S.col = sys->col;
    sv.S.val = sys->val;
    sv.S.diag = sys->diag;
    sv.S.b = sys->b;
    QpSolveScalars res[2];
    sv.solve(st, max_iterations, tolerance, res);
    for (int d = 0; d < nd; ++d) {
      if (error) error[d] = res[d].error;
      if (iterations) iterations[d] = res[d].iter;
    }
    return EP_OK;
  } catch (const std::string &e) {
    return qp_fail(e, EP_ERR_HIP);
  }
}

}  // extern "C"
