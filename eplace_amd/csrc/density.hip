// Bin-density scatter, density finaliser and field gather + combine.
//
// Compiled with -ffp-contract=off: the footprint and the per-(object, bin)
// contribution reproduce the reference's float/double arithmetic exactly, so
// the int64 fixed-point maps are bit-identical to the CPU oracle's.

#include <algorithm>

#include "ep_internal.h"
#include "reduce.h"

namespace ep {

struct Foot {
  float l, b, r, t, sx, sy;
  int i0, i1, j0, j1;
};

// float_less (global.h:350-353): float difference compared in double.
__device__ __forceinline__ bool f_less(float a, float b) { return (double)(a - b) < -kEps; }

// Inflated rectangle, smoothing scales and clamped bin range of one object
// (eplace.cpp:1042-1104 for the scatter, 601-659 for the gather).
__device__ __forceinline__ Foot footprint(float cx, float cy, float w, float h, const Grid &g) {
  Foot f;
  f.sx = 1.f;
  f.sy = 1.f;
  f.l = cx - 0.5f * w;  // Module::setCenter_2D (objects.cpp:530-539)
  f.b = cy - 0.5f * h;
  f.r = f.l + w;        // Module::getUR_2D (objects.cpp:497-509)
  f.t = f.b + h;
  if (f_less(w, g.bsx)) {
    f.sx = w / g.bsx;
    f.l = (float)((double)cx - 0.5 * (double)g.bsx);
    f.r = (float)((double)cx + 0.5 * (double)g.bsx);
  }
  if (f_less(h, g.bsy)) {
    f.sy = h / g.bsy;
    f.b = (float)((double)cy - 0.5 * (double)g.bsy);
    f.t = (float)((double)cy + 0.5 * (double)g.bsy);
  }
  f.i0 = (int)((f.l - g.core_llx) / g.bsx);
  f.i1 = (int)((f.r - g.core_llx) / g.bsx);
  f.j0 = (int)((f.b - g.core_lly) / g.bsy);
  f.j1 = (int)((f.t - g.core_lly) / g.bsy);
  // hi clamp as the reference; lo clamp only guards non-finite inputs
  f.i0 = max(f.i0, 0);
  f.j0 = max(f.j0, 0);
  f.i1 = min(f.i1, g.nx - 1);
  f.j1 = min(f.j1, g.ny - 1);
  return f;
}

// getOverlapArea (global.h:477-518).
__device__ __forceinline__ double overlap_len(double x1, double x2, double x3, double x4) {
  double s = fmax(x1, x3), e = fmin(x2, x4);
  return s >= e ? 0.0 : s - e;
}
// Bin boundaries are recomputed the way the grid built them (i * bsx + llx,
// then + bsx, float arithmetic without contraction), so they are the stored
// values without a memory read.
__device__ __forceinline__ double bin_overlap(const Grid &g, int i, int j, const Foot &f) {
  const float lx = (float)i * g.bsx + g.core_llx;
  double hl = overlap_len(lx, lx + g.bsx, f.l, f.r);
  if (hl == 0) return 0;
  const float ly = (float)j * g.bsy + g.core_lly;
  double vl = overlap_len(ly, ly + g.bsy, f.b, f.t);
  if (vl == 0) return 0;
  return hl * vl;
}

// Addend of eplace.cpp:1116/1123/1127 for one bin, as int64 units of 2^-shift.
__device__ __forceinline__ unsigned long long fixed_contribution(float s, double ov, double scale) {
  float c = s * (float)ov;
  return (unsigned long long)__double2ll_rn((double)c * scale);
}

// ---------------------------------------------------------------------------
// Tile-sorted scatter.  Hot bins (clustered placements put hundreds of
// contributions on one bin) serialise HBM/L2 atomics, so objects are sorted
// by (kind, 32x32-bin tile of their footprint's lower-left bin) every
// evaluation, each chunk of <= kTileChunk objects of one tile accumulates in
// an LDS window of the tile plus a kTileHalo-bin high-side halo, and only the
// window's nonzero bins are added to HBM.  Footprint bins outside the window
// (objects wider than the halo) go straight to HBM.  Integer sums make the
// result independent of the order.
// ---------------------------------------------------------------------------
// Chunks of 512 objects: the field gather (a block per chunk, grid of up to
// 4096 blocks) gains more from the extra blocks than the scatter loses to the
// extra window clears and flushes; 1024 and 256 were slower overall at 1M
// cells.
constexpr int kTile = 32, kTileHalo = 3, kTileWin = kTile + kTileHalo, kTileChunk = 512;

constexpr int kSortThreads = 1024;

__device__ __forceinline__ uint32_t tile_key(const float2 p, float w, float h, uint8_t kind, uint8_t big,
                                             const Grid &g, int tyc, int ntiles) {
  if (big) return 3u * ntiles;
  const Foot f = footprint(p.x, p.y, w, h, g);
  const int ti = min(f.i0, g.nx - 1) / kTile, tj = min(f.j0, g.ny - 1) / kTile;
  return (uint32_t)kind * ntiles + (uint32_t)(ti * tyc + tj);
}

// Pass 1: block b counts the keys of objects [b per, (b+1) per) in LDS and
// reserves its run inside each key with one atomic on the key's running
// total (bcount[b][k] = the run's offset inside the key).  The grid also
// clears the nb-bin fixed-point maps the scatter accumulates into.
__global__ void __launch_bounds__(kSortThreads) tile_hist_kernel(
    int O, int per, const float2 *__restrict__ pos, const float *__restrict__ w, const float *__restrict__ h,
    const uint8_t *__restrict__ kind, const uint8_t *__restrict__ is_big, Grid g, int tyc, int ntiles, int nkeys,
    uint32_t *okey, uint32_t *bcount, uint32_t *ktotal, int64_t nb, unsigned long long *node_q,
    unsigned long long *filler_q) {
  extern __shared__ uint32_t hist[];
  for (int k = threadIdx.x; k < nkeys; k += blockDim.x) hist[k] = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x) {
    node_q[i] = 0;
    filler_q[i] = 0;
  }
  __syncthreads();
  const int o0 = blockIdx.x * per, o1 = min(O, o0 + per);
  for (int o = o0 + threadIdx.x; o < o1; o += blockDim.x) {
    const uint32_t k = tile_key(pos[o], w[o], h[o], kind[o], is_big[o], g, tyc, ntiles);
    okey[o] = k;
    atomicAdd(&hist[k], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nkeys; k += blockDim.x) {
    const uint32_t n = hist[k];
    bcount[(size_t)blockIdx.x * nkeys + k] = n ? atomicAdd(&ktotal[k], n) : 0u;
  }
}

// Pass 1 fused with the Nesterov move that produces the positions: block b
// moves the stage's objects of [b per, (b+1) per) (unless the run is done)
// and counts the keys of all of them, as tile_hist_kernel does from the
// stored positions.  The evaluation that follows starts at pass 2.
__global__ void __launch_bounds__(kSortThreads) move_hist_kernel(
    MoveArgs a, const DevScalars *sc, int O, int per, float2 *pos, const float *__restrict__ w,
    const float *__restrict__ h, const uint8_t *__restrict__ kind, const uint8_t *__restrict__ is_big, Grid g,
    int tyc, int ntiles, int nkeys, uint32_t *okey, uint32_t *bcount, uint32_t *ktotal, int64_t nb,
    unsigned long long *node_q, unsigned long long *filler_q) {
  extern __shared__ uint32_t hist[];
  for (int k = threadIdx.x; k < nkeys; k += blockDim.x) hist[k] = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x) {
    node_q[i] = 0;
    filler_q[i] = 0;
  }
  const bool moving = !sc->done;
  const float step = (float)sc->step, coef = (float)sc->coef;
  __syncthreads();
  const int o0 = blockIdx.x * per, o1 = min(O, o0 + per);
  for (int o = o0 + threadIdx.x; o < o1; o += blockDim.x) {
    const float wo = w[o], ho = h[o];
    const float2 p = moving && a.in_stage[o] ? move_object(a, o, pos, wo, ho, step, coef) : pos[o];
    const uint32_t k = tile_key(p, wo, ho, kind[o], is_big[o], g, tyc, ntiles);
    okey[o] = k;
    atomicAdd(&hist[k], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nkeys; k += blockDim.x) {
    const uint32_t n = hist[k];
    bcount[(size_t)blockIdx.x * nkeys + k] = n ? atomicAdd(&ktotal[k], n) : 0u;
  }
}

// Pass 2: every block scans the key totals into the first slot of each key
// (the same result in every block) and places its objects at that slot + its
// offset inside the key + an LDS cursor.  Block 0 also writes the chunks of
// <= kTileChunk objects per key (the sentinel key makes none).  The totals
// are cleared by the scatter, after every block has read them.
__global__ void __launch_bounds__(kSortThreads) tile_place_kernel(
    int O, int per, int nkeys, const uint32_t *__restrict__ okey, const uint32_t *__restrict__ bcount,
    const uint32_t *__restrict__ ktotal, const float2 *__restrict__ pos, const float *__restrict__ w,
    const float *__restrict__ h, uint32_t *sobj, float4 *svals, int4 *chunks, int *nchunks) {
  extern __shared__ uint32_t cursor[];
  __shared__ int pslot[kSortThreads], pchunk[kSortThreads];
  const int kper = (nkeys + blockDim.x - 1) / blockDim.x;
  const int k0 = threadIdx.x * kper, k1 = min(k0 + kper, nkeys);
  int slots = 0, cnt = 0;
  for (int k = k0; k < k1; ++k) {
    const int n = (int)ktotal[k];
    slots += n;
    if (k < nkeys - 1) cnt += (n + kTileChunk - 1) / kTileChunk;
  }
  pslot[threadIdx.x] = slots;
  pchunk[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {  // inclusive scans
    const int a = threadIdx.x >= (unsigned)off ? pslot[threadIdx.x - off] : 0;
    const int b = threadIdx.x >= (unsigned)off ? pchunk[threadIdx.x - off] : 0;
    __syncthreads();
    pslot[threadIdx.x] += a;
    pchunk[threadIdx.x] += b;
    __syncthreads();
  }
  int slot = pslot[threadIdx.x] - slots, cpos = pchunk[threadIdx.x] - cnt;
  const uint32_t *bc = bcount + (size_t)blockIdx.x * nkeys;
  for (int k = k0; k < k1; ++k) {
    const int n = (int)ktotal[k];
    cursor[k] = (uint32_t)slot + bc[k];
    if (blockIdx.x == 0 && k < nkeys - 1)
      for (int c = 0; c < n; c += kTileChunk) chunks[cpos++] = make_int4(k, slot + c, slot + min(c + kTileChunk, n), 0);
    slot += n;
  }
  if (blockIdx.x == 0 && threadIdx.x == blockDim.x - 1) *nchunks = pchunk[threadIdx.x];
  __syncthreads();
  const int o0 = blockIdx.x * per, o1 = min(O, o0 + per);
  for (int o = o0 + threadIdx.x; o < o1; o += blockDim.x) {
    const uint32_t sl = atomicAdd(&cursor[okey[o]], 1u);
    const float2 p = pos[o];
    sobj[sl] = (uint32_t)o;
    svals[sl] = make_float4(p.x, p.y, w[o], h[o]);
  }
}

// Chunks of one key (kind, tile) accumulate in an LDS window of the tile plus
// its high-side halo; footprint bins outside the window go to global memory
// directly.
__global__ void __launch_bounds__(kBlock) tile_scatter_kernel(const int4 *__restrict__ chunks,
                                                               const int *__restrict__ nchunks,
                                                               const float4 *__restrict__ vals, Grid g, int tyc,
                                                               int ntiles,
                                                               double scale, unsigned long long *node_q,
                                                               unsigned long long *filler_q, int nkeys,
                                                               uint32_t *ktotal) {
  __shared__ unsigned long long win[kTileWin * kTileWin];
  // the key totals of this sort are read; zero them for the next one
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nkeys; k += gridDim.x * blockDim.x) ktotal[k] = 0;
  const int total = *nchunks;
  for (int ci = blockIdx.x; ci < total; ci += gridDim.x) {
    const int4 ch = chunks[ci];
    const int kind = ch.x / ntiles, tile = ch.x % ntiles;
    const int ib = (tile / tyc) * kTile, jb = (tile % tyc) * kTile;
    unsigned long long *dst = kind == KIND_FILLER ? filler_q : node_q;
    for (int t = threadIdx.x; t < kTileWin * kTileWin; t += blockDim.x) win[t] = 0;
    __syncthreads();
    for (int q = ch.y + threadIdx.x; q < ch.z; q += blockDim.x) {
      float4 v = vals[q];
      Foot f = footprint(v.x, v.y, v.z, v.w, g);
      float s = f.sx * f.sy;
      if (kind == KIND_MACRO) s = s * g.td;
      for (int i = f.i0; i <= f.i1; ++i)
        for (int j = f.j0; j <= f.j1; ++j) {
          double ov = bin_overlap(g, i, j, f);
          if (ov == 0) continue;
          unsigned long long c = fixed_contribution(s, ov, scale);
          int li = i - ib, lj = j - jb;
          if (li >= 0 && li < kTileWin && lj >= 0 && lj < kTileWin)
            atomicAdd(&win[li * kTileWin + lj], c);
          else
            atomicAdd(dst + (size_t)i * g.ny + j, c);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kTileWin * kTileWin; t += blockDim.x) {
      unsigned long long c = win[t];
      if (c == 0) continue;
      int i = ib + t / kTileWin, j = jb + t % kTileWin;
      atomicAdd(dst + (size_t)i * g.ny + j, c);
    }
    __syncthreads();
  }
}

// Objects covering many bins (macros): one block each, threads over bins.
__global__ void __launch_bounds__(kBlock) scatter_big_kernel(const int32_t *__restrict__ big,
                                                              const float2 *__restrict__ pos,
                                                              const float *__restrict__ w,
                                                              const float *__restrict__ h,
                                                              const uint8_t *__restrict__ kind, Grid g,
                                                              double scale, unsigned long long *node_q,
                                                              unsigned long long *filler_q) {
  int o = big[blockIdx.x];
  uint8_t k = kind[o];
  Foot f = footprint(pos[o].x, pos[o].y, w[o], h[o], g);
  float s = f.sx * f.sy;
  if (k == KIND_MACRO) s = s * g.td;
  unsigned long long *dst = k == KIND_FILLER ? filler_q : node_q;
  int ni = f.i1 - f.i0 + 1, nj = f.j1 - f.j0 + 1;
  if (ni <= 0 || nj <= 0) return;
  int64_t total = (int64_t)ni * nj;
  for (int64_t t = threadIdx.x; t < total; t += blockDim.x) {
    int i = f.i0 + (int)(t / nj), j = f.j0 + (int)(t % nj);
    double ov = bin_overlap(g, i, j, f);
    if (ov == 0) continue;
    atomicAdd(dst + (size_t)i * g.ny + j, fixed_contribution(s, ov, scale));
  }
}

// Fixed-point maps -> float maps -> rho (eplace.cpp:571-580) and per-block
// partial sums of the overflow area (eplace.cpp:438-453) and of rho.
__global__ void __launch_bounds__(kBlock) finalize_density_kernel(
    int nb, const unsigned long long *__restrict__ node_q, const unsigned long long *__restrict__ filler_q,
    const float *__restrict__ term, const float *__restrict__ base, Grid g, int ldp, double inv_scale,
    float *__restrict__ rho, double *partials, const DevScalars *sc) {
  if (sc->done) return;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  double ovf = 0, rsum = 0, rmax = 0;
  if (k < nb) {
    float node = (float)((double)(long long)node_q[k] * inv_scale);
    float fill = (float)((double)(long long)filler_q[k] * inv_scale);
    float t = term[k], b = base[k];
    float e = node + b + fill + t;
    int i = k / g.ny, j = k - i * g.ny;
    float r = e * g.inv_bin_area;
    rho[(size_t)i * ldp + j] = r;
    rsum = r;
    rmax = fabs((double)r);
    float d = (node + t + b) * g.inv_bin_area - g.td;
    ovf = (double)(fmaxf(0.f, d) * g.bin_area);
  }
  double s = block_sum(ovf);
  double rs = block_sum(rsum);
  double rm = block_max(rmax);
  if (threadIdx.x == 0) {
    partials[kPartialTau + blockIdx.x] = s;
    partials[kPartialRho + blockIdx.x] = rs;
    partials[kPartialRhoMax + blockIdx.x] = rm;
  }
}

// tau = overflow area / node area, mean rho, and the gamma schedule
// (eplace.cpp:475-500) in the reference's float arithmetic; counts the
// evaluation.
__global__ void __launch_bounds__(kFinish) finish_density_kernel(int nblocks, int nb, Grid g,
                                                                 double node_area_scaled, const double *partials,
                                                                 DevScalars *sc) {
  if (sc->done) return;
  const double *const src[2] = {partials + kPartialTau, partials + kPartialRho};
  double sum[2];
  block_sum_partials_strided_n(src, nblocks, 1, sum);
  const double tot = sum[0], rtot = sum[1];
  double rmax = 0;
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x) rmax = fmax(rmax, partials[kPartialRhoMax + i]);
  rmax = block_max(rmax);
  if (threadIdx.x == 0) {
    double tau = tot / node_area_scaled;
    sc->overflow = tau;
    sc->rho_mean = rtot / (double)nb;
    sc->rho_max = rmax;
    float cx = 0.125f / g.bsx, cy = 0.125f / g.bsy;
    float tf = (float)tau;
    if (tf > 1.0) {
      cx = (float)(cx * 0.1);
      cy = (float)(cy * 0.1);
    } else if (tf < 0.1) {
      cx = (float)(cx * 10.0);
      cy = (float)(cy * 10.0);
    } else {
      float temp = (float)(1.0 / pow(10.0, ((double)tf - 0.1) * 20 / 9.0 - 1.0));
      cx *= temp;
      cy *= temp;
    }
    sc->inv_gamma_x = cx;
    sc->inv_gamma_y = cy;
    sc->evals += 1;
  }
}

// Density gradient of one object: sum over its bins of sx*sy*overlap*E
// (eplace.cpp:661-671; the product is rounded to float as there).
__device__ __forceinline__ float2 gather_one(const Foot &f, const Grid &g, int ldp, const float *ex,
                                             const float *ey, int i, int j) {
  double ov = bin_overlap(g, i, j, f);
  float a = (float)((double)(f.sx * f.sy) * ov);
  size_t k = (size_t)i * ldp + j;
  return make_float2(a * ex[k], a * ey[k]);
}

// Tile-sorted gather: the chunks of the last scatter (objects of one kind
// whose footprint starts in one 32x32-bin tile) read Ex/Ey from an LDS copy
// of the tile plus its high-side halo; footprint bins outside it come from
// global memory.  Same arithmetic and order as the per-object gather, so
// the result is identical.
__global__ void __launch_bounds__(kBlock) gather_tile_kernel(const int4 *__restrict__ chunks,
                                                              const int *__restrict__ nchunks,
                                                              const uint32_t *__restrict__ sobj,
                                                              const float4 *__restrict__ vals, Grid g, int tyc,
                                                              int ntiles, int ldp, const float *__restrict__ ex,
                                                              const float *__restrict__ ey, float2 *gd,
                                                              const DevScalars *sc) {
  if (sc->done) return;
  __shared__ float wex[kTileWin * kTileWin], wey[kTileWin * kTileWin];
  const int total = *nchunks;
  for (int ci = blockIdx.x; ci < total; ci += gridDim.x) {
    const int4 ch = chunks[ci];
    const int tile = ch.x % ntiles;
    const int ib = (tile / tyc) * kTile, jb = (tile % tyc) * kTile;
    __syncthreads();  // the previous chunk is done with the window
    for (int t = threadIdx.x; t < kTileWin * kTileWin; t += blockDim.x) {
      const int i = ib + t / kTileWin, j = jb + t % kTileWin;
      float a = 0.f, b = 0.f;
      if (i < g.nx && j < g.ny) {
        const size_t k = (size_t)i * ldp + j;
        a = ex[k];
        b = ey[k];
      }
      wex[t] = a;
      wey[t] = b;
    }
    __syncthreads();
    for (int q = ch.y + threadIdx.x; q < ch.z; q += blockDim.x) {
      const float4 v = vals[q];
      const Foot f = footprint(v.x, v.y, v.z, v.w, g);
      float gx = 0.f, gy = 0.f;
      for (int i = f.i0; i <= f.i1; ++i)
        for (int j = f.j0; j <= f.j1; ++j) {
          const double ov = bin_overlap(g, i, j, f);
          const float a = (float)((double)(f.sx * f.sy) * ov);
          const int li = i - ib, lj = j - jb;
          float e0, e1;
          if (li >= 0 && li < kTileWin && lj >= 0 && lj < kTileWin) {
            e0 = wex[li * kTileWin + lj];
            e1 = wey[li * kTileWin + lj];
          } else {
            const size_t k = (size_t)i * ldp + j;
            e0 = ex[k];
            e1 = ey[k];
          }
          gx += a * e0;
          gy += a * e1;
        }
      gd[sobj[q]] = make_float2(gx, gy);
    }
  }
}

__global__ void __launch_bounds__(kBlock) gather_big_kernel(const int32_t *__restrict__ big,
                                                             const float2 *__restrict__ pos,
                                                             const float *__restrict__ w,
                                                             const float *__restrict__ h, Grid g, int ldp,
                                                             const float *__restrict__ ex,
                                                             const float *__restrict__ ey, float2 *gd,
                                                             const DevScalars *sc) {
  if (sc->done) return;
  int o = big[blockIdx.x];
  Foot f = footprint(pos[o].x, pos[o].y, w[o], h[o], g);
  int ni = f.i1 - f.i0 + 1, nj = f.j1 - f.j0 + 1;
  double gx = 0, gy = 0;
  if (ni > 0 && nj > 0) {
    int64_t total = (int64_t)ni * nj;
    for (int64_t t = threadIdx.x; t < total; t += blockDim.x) {
      float2 v = gather_one(f, g, ldp, ex, ey, f.i0 + (int)(t / nj), f.j0 + (int)(t % nj));
      gx += v.x;
      gy += v.y;
    }
  }
  gx = block_sum(gx);
  gy = block_sum(gy);
  if (threadIdx.x == 0) gd[o] = make_float2((float)gx, (float)gy);
}

// Node wirelength-gradient reduction over the node's pins in modulePins
// order (eplace.cpp:531-549) when recompute is set, preconditioned
// combine (eplace.cpp:694-737) and, optionally, the Lipschitz partial sums
// |g - g_prev|^2 and |y - y_prev|^2 over the stage's objects
// (optutils.hpp:19-35) with, for the step controller, |grad WL|^2 and
// |grad D|^2, and for the Barzilai-Borwein step (y - y_prev).(g - g_prev)
// (nesterov.hpp:245-268) over the stage.  Blocks walk the objects XCD-contiguously: with
// nets stored by degree class and then by their lowest node, the pin
// gradients a range of nodes reads lie in one narrow band per class.
__global__ void __launch_bounds__(kBlock) gather_combine_kernel(
    int O, int N, const float2 *__restrict__ pos, const uint8_t *__restrict__ kind,
    const int32_t *__restrict__ npins, const float *__restrict__ area, const int32_t *__restrict__ node_pin_start,
    const int32_t *__restrict__ node_pin_lane, const float2 *__restrict__ pin_grad, const float2 *gd, float2 *gwl,
    float2 *gtot, int recompute, int lipschitz, const uint8_t *__restrict__ in_stage,
    const float2 *__restrict__ ref_now, const float2 *__restrict__ ref_prev,
    const float2 *__restrict__ grad_prev, double *partials, float2 *pos_out, float2 *gtot_out,
    int at_next, const DevScalars *sc) {
  if (sc->done) return;
  int o = xcd_block() * blockDim.x + threadIdx.x;
  double ln = 0, ld = 0, nw = 0, nd = 0, sy = 0, ln2 = 0, ld2 = 0;
  if (o < O) {
    uint8_t k = kind[o];
    const float2 d = gd[o];  // from the tile / big-object gather
    float2 wl = make_float2(0.f, 0.f);
    if (o < N) {
      if (recompute) {
        int s = node_pin_start[o], e = node_pin_start[o + 1];
        for (int q = s; q < e; ++q) {
          float2 pg = pin_grad[node_pin_lane[q]];
          wl.x += pg.x;
          wl.y += pg.y;
        }
        gwl[o] = wl;
      } else {
        wl = gwl[o];
      }
    }
    float lam = (float)sc->lambda;
    float pre = 1.f / fmaxf(1.0f, (float)npins[o] + lam * area[o]);
    float2 t;
    if (k == KIND_FILLER) {
      t.x = pre * lam * d.x;
      t.y = pre * lam * d.y;
    } else {
      t.x = pre * (lam * d.x - wl.x);
      t.y = pre * (lam * d.y - wl.y);
    }
    gtot[o] = t;
    // at_next: the same positions' gradient at the staged lambda
    // (updatePenaltyFactor of this trial), which the combine-only pass after
    // an accepted trial would compute: it becomes g_{k+1}, and its Lipschitz
    // sums against y_k, g_k at the clamped positions give the next step
    float2 t2 = t;
    if (at_next) {
      const float lam2 = (float)sc->lambda_next;
      const float pre2 = 1.f / fmaxf(1.0f, (float)npins[o] + lam2 * area[o]);
      if (k == KIND_FILLER) {
        t2.x = pre2 * lam2 * d.x;
        t2.y = pre2 * lam2 * d.y;
      } else {
        t2.x = pre2 * (lam2 * d.x - wl.x);
        t2.y = pre2 * (lam2 * d.y - wl.y);
      }
    }
    // the optimiser's copies of y_k and g_k for the next trials
    if (pos_out) pos_out[o] = pos[o];
    if (gtot_out) gtot_out[o] = t2;
    if (lipschitz && in_stage[o]) {
      float2 gp = grad_prev[o];
      float2 rn = ref_now ? ref_now[o] : pos[o];
      float2 rp = ref_prev[o];
      double dgx = (double)t.x - gp.x, dgy = (double)t.y - gp.y;
      double dxx = (double)rn.x - rp.x, dxy = (double)rn.y - rp.y;
      ln = dgx * dgx + dgy * dgy;
      ld = dxx * dxx + dxy * dxy;
      nw = (double)wl.x * wl.x + (double)wl.y * wl.y;
      nd = (double)d.x * d.x + (double)d.y * d.y;
      sy = dxx * dgx + dxy * dgy;
      if (at_next) {
        const float2 pc = pos[o];
        double egx = (double)t2.x - gp.x, egy = (double)t2.y - gp.y;
        double exx = (double)pc.x - rp.x, exy = (double)pc.y - rp.y;
        ln2 = egx * egx + egy * egy;
        ld2 = exx * exx + exy * exy;
      }
    }
  }
  if (lipschitz && at_next) {
    double v[7] = {ln, ld, nw, nd, sy, ln2, ld2};
    block_sum_n(v);
    if (threadIdx.x == 0) {
      partials[kPartialLip + 2 * blockIdx.x] = v[0];
      partials[kPartialLip + 2 * blockIdx.x + 1] = v[1];
      partials[kPartialNorm + 2 * blockIdx.x] = v[2];
      partials[kPartialNorm + 2 * blockIdx.x + 1] = v[3];
      partials[kPartialSY + 2 * blockIdx.x] = v[4];
      partials[kPartialLipNext + 2 * blockIdx.x] = v[5];
      partials[kPartialLipNext + 2 * blockIdx.x + 1] = v[6];
    }
  } else if (lipschitz) {
    double v[5] = {ln, ld, nw, nd, sy};
    block_sum_n(v);
    if (threadIdx.x == 0) {
      partials[kPartialLip + 2 * blockIdx.x] = v[0];
      partials[kPartialLip + 2 * blockIdx.x + 1] = v[1];
      partials[kPartialNorm + 2 * blockIdx.x] = v[2];
      partials[kPartialNorm + 2 * blockIdx.x + 1] = v[3];
      partials[kPartialSY + 2 * blockIdx.x] = v[4];
    }
  }
}

// lambda0 = sum |gwl| / sum |gd| (eplace.cpp:812-835): partial sums here,
// the quotient in lambda_init_finish_kernel.
__global__ void __launch_bounds__(kBlock) lambda_init_kernel(int O, int N, const float2 *__restrict__ gd,
                                                              const float2 *__restrict__ gwl,
                                                              double *partials) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0, b = 0;
  if (o < O) {
    float2 d = gd[o];
    b = fabs((double)d.x) + fabs((double)d.y);
    if (o < N) {
      float2 wv = gwl[o];
      a = fabs((double)wv.x) + fabs((double)wv.y);
    }
  }
  a = block_sum(a);
  b = block_sum(b);
  if (threadIdx.x == 0) {
    partials[kPartialLambda + 2 * blockIdx.x] = a;
    partials[kPartialLambda + 2 * blockIdx.x + 1] = b;
  }
}

__global__ void __launch_bounds__(kFinish) lambda_init_finish_kernel(int nblocks, const double *partials,
                                                                     DevScalars *sc) {
  double na = block_sum_partials_strided(partials + kPartialLambda, nblocks, 2);
  double nb = block_sum_partials_strided(partials + kPartialLambda + 1, nblocks, 2);
  if (threadIdx.x == 0) {
    sc->lambda = na / nb;
    sc->last_hpwl = sc->hpwl;
  }
}

// ---------------------------------------------------------------------------
static int blocks_for(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

__global__ void clear_u32_kernel(uint32_t *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

void launch_clear_key_totals(Ctx &c) {
  TileSort &t = c.tiles;
  hipLaunchKernelGGL(clear_u32_kernel, dim3((t.nkeys + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream, t.ktotal,
                     t.nkeys);
}

void launch_move_hist(Ctx &c, const MoveArgs &a) {
  TileSort &t = c.tiles;
  const int64_t nb = (int64_t)c.nx * c.ny;
  const size_t lds = (size_t)t.nkeys * sizeof(uint32_t);
  hipLaunchKernelGGL(move_hist_kernel, dim3(t.nblocks), dim3(kSortThreads), lds, c.stream, a, c.sc, c.O, t.per,
                     c.pos, c.w, c.h, c.kind, c.is_big, c.grid(), t.tyc, t.ntiles, t.nkeys, t.okey, t.bcount,
                     t.ktotal, nb, c.node_q, c.filler_q);
}

void launch_scatter(Ctx &c, bool counted) {
  Grid g = c.grid();
  double scale = ldexp(1.0, c.shift);
  TileSort &t = c.tiles;
  const int64_t nb = (int64_t)c.nx * c.ny;
  const size_t lds = (size_t)t.nkeys * sizeof(uint32_t);
  if (!counted)
    hipLaunchKernelGGL(tile_hist_kernel, dim3(t.nblocks), dim3(kSortThreads), lds, c.stream, c.O, t.per, c.pos,
                       c.w, c.h, c.kind, c.is_big, g, t.tyc, t.ntiles, t.nkeys, t.okey, t.bcount, t.ktotal, nb,
                       c.node_q, c.filler_q);
  hipLaunchKernelGGL(tile_place_kernel, dim3(t.nblocks), dim3(kSortThreads), lds, c.stream, c.O, t.per, t.nkeys,
                     t.okey, t.bcount, t.ktotal, c.pos, c.w, c.h, t.sobj, t.svals, t.chunks, t.nchunks);
  hipLaunchKernelGGL(tile_scatter_kernel, dim3(t.grid), dim3(kBlock), 0, c.stream, t.chunks, t.nchunks, t.svals, g,
                     t.tyc, t.ntiles, scale, c.node_q, c.filler_q, t.nkeys, t.ktotal);
  if (c.num_big)
    hipLaunchKernelGGL(scatter_big_kernel, dim3(c.num_big), dim3(kBlock), 0, c.stream, c.big, c.pos, c.w,
                       c.h, c.kind, g, scale, c.node_q, c.filler_q);
}

int poisson_ld(const Ctx &c);

bool tile_sort_create(Ctx &c, std::string &err) {
  TileSort &t = c.tiles;
  t.tyc = (c.ny + kTile - 1) / kTile;
  t.ntiles = ((c.nx + kTile - 1) / kTile) * t.tyc;
  t.nkeys = 3 * t.ntiles + 1;
  if ((size_t)t.nkeys * sizeof(uint32_t) > 64 * 1024) {
    err = "grid too large for the tile sort";
    return false;
  }
  const int O = std::max(c.O, 1);
  t.nblocks = std::max(1, std::min(256, (O + kSortThreads - 1) / kSortThreads));
  t.per = (O + t.nblocks - 1) / t.nblocks;
  const size_t max_chunks = (size_t)O / kTileChunk + 3 * (size_t)t.ntiles + 1;
  bool ok = hipMalloc(&t.okey, (size_t)O * 4) == hipSuccess &&
            hipMalloc(&t.bcount, (size_t)t.nkeys * t.nblocks * 4) == hipSuccess &&
            hipMalloc(&t.ktotal, (size_t)t.nkeys * 4) == hipSuccess &&
            hipMalloc(&t.sobj, (size_t)O * 4) == hipSuccess && hipMalloc(&t.svals, (size_t)O * 16) == hipSuccess &&
            hipMalloc(&t.chunks, max_chunks * sizeof(int4)) == hipSuccess &&
            hipMalloc(&t.nchunks, sizeof(int)) == hipSuccess;
  ok = ok && hipMemset(t.ktotal, 0, (size_t)t.nkeys * 4) == hipSuccess;
  t.grid = (int)std::min<size_t>(max_chunks, 4096);
  if (!ok) err = "tile sort allocation failed";
  return ok;
}

void tile_sort_destroy(Ctx &c) {
  TileSort &t = c.tiles;
  void *p[] = {t.okey, t.bcount, t.ktotal, t.sobj, t.svals, t.chunks, t.nchunks};
  for (void *q : p)
    if (q) (void)hipFree(q);
  t = TileSort();
}

void launch_finalize_density(Ctx &c) {
  Grid g = c.grid();
  int nb = c.nx * c.ny;
  double nas = (double)(c.std_area + c.macro_area * c.td);
  hipLaunchKernelGGL(finalize_density_kernel, dim3(blocks_for(nb)), dim3(kBlock), 0, c.stream, nb, c.node_q,
                     c.filler_q, c.term_map, c.base_map, g, poisson_ld(c), ldexp(1.0, -c.shift), c.rho,
                     c.partials, c.sc);
  hipLaunchKernelGGL(finish_density_kernel, dim3(1), dim3(kFinish), 0, c.stream, blocks_for(nb), nb, g, nas,
                     c.partials, c.sc);
}

void launch_field_gather(Ctx &c) {
  Grid g = c.grid();
  int ldp = poisson_ld(c);
  TileSort &t = c.tiles;
  hipLaunchKernelGGL(gather_tile_kernel, dim3(t.grid), dim3(kBlock), 0, c.stream, t.chunks, t.nchunks, t.sobj,
                     t.svals, g, t.tyc, t.ntiles, ldp, c.ex, c.ey, c.gd, c.sc);
  if (c.num_big)
    hipLaunchKernelGGL(gather_big_kernel, dim3(c.num_big), dim3(kBlock), 0, c.stream, c.big, c.pos, c.w, c.h, g,
                       ldp, c.ex, c.ey, c.gd, c.sc);
}

void launch_combine(Ctx &c, bool recompute, bool lipschitz, const float2 *ref_now, const float2 *ref_prev,
                    const float2 *grad_prev, float2 *pos_out, float2 *gtot_out, bool at_next_lambda) {
  hipLaunchKernelGGL(gather_combine_kernel, dim3(object_blocks(c.O)), dim3(kBlock), 0, c.stream, c.O, c.N,
                     c.pos, c.kind, c.npins, c.area, c.node_pin_start, c.node_pin_lane, c.pin_grad, c.gd, c.gwl,
                     c.gtot, recompute ? 1 : 0, lipschitz ? 1 : 0, c.in_stage, ref_now, ref_prev, grad_prev,
                     c.partials, pos_out, gtot_out, at_next_lambda ? 1 : 0, c.sc);
}

void launch_gather_combine(Ctx &c, bool recompute, bool lipschitz, const float2 *ref_now,
                           const float2 *ref_prev, const float2 *grad_prev, float2 *pos_out, float2 *gtot_out) {
  if (recompute) launch_field_gather(c);
  launch_combine(c, recompute, lipschitz, ref_now, ref_prev, grad_prev, pos_out, gtot_out);
}

void launch_lambda_init(Ctx &c) {
  hipLaunchKernelGGL(lambda_init_kernel, dim3(blocks_for(c.O)), dim3(kBlock), 0, c.stream, c.O, c.N, c.gd,
                     c.gwl, c.partials);
  hipLaunchKernelGGL(lambda_init_finish_kernel, dim3(1), dim3(kFinish), 0, c.stream, blocks_for(c.O), c.partials,
                     c.sc);
}

}  // namespace ep
