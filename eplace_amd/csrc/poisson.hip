// Poisson solve for the bin charge density as dense transform GEMMs on MFMA.
//
// Reference: FFT_2D::doFFT (FFT/fft.cpp:190-284) on Ooura's ddct2d /
// ddsct2d / ddcst2d (FFT/fftsg2d.cpp:105-297).  Closed form, x-major maps
// (map[x][y], as the reference's bins[x][y]):
//   rho^ = F_X rho F_Y^T,   F[k][x] = (2/N) c_k cos(pi k (x+1/2)/N), c_0 = 1/2
//   phi^ = rho^ / (wx^2 + wy^2), phi^[0][0] = 0
//   P = phi^ C_Y,  Q = phi^ W_Y S_Y
//   Ex = (S_X^T W_X) P,  Ey = C_X^T Q,  phi = C_X^T P
// with C[k][x] = cos(pi k (x+1/2)/N), S[k][x] = sin(pi k (x+1/2)/N),
// wx = pi u / Nx, wy = pi v / Ny * (bin_h / bin_w) (fft.cpp:129-147).
//
// Parity folding.  Every transform matrix T[k][x] of length N satisfies
// T[k][N-1-x] = +-(-1)^k T[k][x], so each GEMM is split by the parity of the
// frequency index and runs on half-length operands:
//   forward (sum over space):  even k take a[x] + a[N-1-x], odd k take
//                              a[x] - a[N-1-x], x < N/2;
//   inverse (sum over frequency): out[x] = e + o, out[N-1-x] = +-e +- o,
//                              e/o the sums over even/odd frequencies.
// That halves the multiply-adds.  Spectra are stored parity-split (even
// frequencies first, then odd; each half padded to a multiple of 64).
//
// Operand layout.  Each stage is C[M x N] = sum_k A[m][k] B[n][k] with both
// operands k-contiguous: A is a constant transform matrix, B is the data,
// and the stages alternate which side is transposed so no transpose is
// ever materialised:
//   1. T1t[v][x] = sum_y F_Y[v][y] rho[x][y]        (B folded over y)
//   2. spec[u][v] = sum_x F_X[u][x] T1t[v][x] / w^2  (B folded over x)
//   3. Pt[y][u] = sum_v C_Y[v][y] spec[u][v]         (output mirrored in y)
//      Qt[y][u] = sum_v wy S_Y[v][y] spec[u][v]
//   4. Ex[x][y] = sum_u wx S_X[u][x] Pt[y][u]        (output mirrored in x)
//      Ey = C_X^T Qt^T, phi = C_X^T Pt^T
// Blocks of 64x64 outputs (four waves of 32x32) stage 64 x 32 slices of
// both operands in a double-buffered LDS, one barrier per k step, with the
// next slice's global loads in flight during the multiply.
//
// Arithmetic modes (ep_params.poisson_mode):
//   0  FP32 MFMA (16x16x4);
//   1  split-BF16: each FP32 value is a_hi + a_lo (two bf16; the constant
//      matrices are stored split), a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi on
//      bf16 MFMA (16x16x32), accumulated in FP32; the operands keep 16
//      significant bits;
//   2  split-FP16: the same three products on fp16 MFMA (16x16x32).  fp16
//      halves keep 11 bits each, so a_hi + a_lo carries 22; its exponent
//      range is narrow, so every operand is multiplied by a power of two
//      that puts its largest magnitude near 2^13 (exact in FP32, undone in
//      the epilogue).  The constant matrices have fixed scales; the data
//      operand of a stage takes its scale from the largest |value| the
//      previous stage wrote (per-wave maxima written by its epilogue; for
//      stage 1 the largest charge density).  Values far below the largest
//      lose only absolute precision below 2^-27 of it, whether or not fp16
//      subnormals are flushed;
//   3  split-BF16, three planes: a = a_0 + a_1 + a_2 (24 bits), six products
//      (every pair of plane orders summing to at most 2).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cmath>
#include <type_traits>
#include <vector>

#include "ep_internal.h"

namespace ep {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int kMaxPlanes = 3;
// Largest |operand| after scaling in mode 2 (folded data adds two values,
// so the stored halves stay below 2^14 < 65504).
constexpr int kF16ScaleExp = 13;

// A constant transform matrix: FP32, its bf16 planes and its scaled fp16
// halves (value * 2^f16_exp).
struct Mat {
  float *f = nullptr;
  __bf16 *bp[kMaxPlanes] = {nullptr, nullptr, nullptr};
  _Float16 *hp[2] = {nullptr, nullptr};
  int f16_exp = 0;
};

struct Poisson {
  int mode = 0;                // poisson_mode
  int nxp = 0, nyp = 0;        // padded map sizes (multiples of 64)
  int hx = 0, hy = 0;          // nx/2, ny/2
  int hxp = 0, hyp = 0;        // halves padded to multiples of 64
  Mat FYA;   // [2hyp x hyp]   F_Y[v][y], rows parity-split v
  Mat FXA;   // [2hxp x hxp]   F_X[u][x], rows parity-split u
  Mat CYT;   // [hyp x 2hyp]   C_Y[v][y], row y, columns parity-split v
  Mat SYT;   // [hyp x 2hyp]   wy_v S_Y[v][y]
  Mat CX2;   // [hxp x 2hxp]   C_X[u][x], row x, columns parity-split u
  Mat SX2;   // [hxp x 2hxp]   wx_u S_X[u][x]
  float *wx2p = nullptr, *wy2p = nullptr;  // squared wave numbers in parity-split order, -1 = padding
  float *T1t = nullptr;   // [2hyp x nxp]
  float *spec = nullptr;  // [2hxp x 2hyp]
  float *PQt = nullptr;   // [2 nyp x 2hxp]: Pt rows 0..nyp, Qt rows nyp..2nyp
  float *wmax = nullptr;  // mode 2: per-wave max |output| of stages 1-3 [3][kMaxWaveMax]
  void *bsplit = nullptr; // split modes: the current stage's data operand as planes
  float *bscale = nullptr; // mode 2: the scale split_kernel applied, per stage [4]
};

constexpr int kMaxWaveMax = 1 << 16;

int poisson_ld(const Ctx &c) { return c.poisson ? c.poisson->nyp : c.ny; }

// One launch of up to three GEMMs C = A B^T (blockIdx.z selects).
struct PGemm {
  const float *Af[3];                 // FP32 A (mode 0)
  const void *Ap[3][kMaxPlanes];      // A planes (modes 1-3)
  float a_inv[3];                     // mode 2: 2^-f16_exp of A
  const float *B[3];
  float *C[3];
  float sb[3];         // fold sign of B along k (0: no fold)
  float se[3], so[3];  // mirrored output row = se * even + so * odd
  int row_off[3];      // spectral row of the batch's first row (epilogue)
  int lda, ldb, ldc;
  int M, N, K;         // multiples of 64, 64, 32
  int nfb, hb;         // B fold: mirror k -> nfb-1-k, valid k < hb
  int out_fold;        // 1: rows split even/odd over K, row m < ho also writes nfo-1-m
  int nfo, ho;
  int khalf;           // with out_fold: K of the even part (the odd part follows)
  int epi;             // 1: divide by wx2p[row] + wy2p[col], zero the DC term
  const double *sub;   // B := B - *sub before folding (removes the mean of rho)
  const float *wx2p, *wy2p;
  // split modes: B as planes written by split_kernel, rows of ldbp
  const void *Bp[3][kMaxPlanes];
  int ldbp;
  const float *bscale;  // mode 2: the power of two split_kernel scaled B by
  float *out_max;       // mode 2 (optional): one max |C| per wave of the launch
  unsigned long long *stamp;  // optional: per-block (start, end) on the wall clock, this launch's
  const int *evals;           // slots of parity 0; parity 1 at stamp_set further
  int stamp_set;
};

constexpr int PBM = 64, PBN = 64, PK = 32;
constexpr int kGroups = 2;  // k-groups per block: each runs half the k steps, then they add up
constexpr int kGroupThreads = 256;

// Thread index inside the thread's k-group.
__device__ __forceinline__ int gtid() { return threadIdx.x & (kGroupThreads - 1); }
constexpr int kLd32 = PK + 4;  // FP32 LDS row stride (floats), rows 16-byte aligned
// 16-bit LDS tiles: 64-byte rows (32 k) of four 16-byte chunks, chunk c of
// row r stored at c ^ ((r >> 1 ^ r >> 3) & 3).  The stores (whole rows) and
// the MFMA operand reads (16 consecutive rows, one chunk) then touch
// distinct 16-byte bank groups per 8 lanes with 32 banks and per 16 lanes
// with 64.
__device__ __forceinline__ int swz16(int r, int c) { return 8 * (c ^ ((r >> 1 ^ r >> 3) & 3)); }

// Element type, vector type and plane count of each split mode.
template <int MODE> struct Split;
template <> struct Split<1> { typedef __bf16 T; typedef bf16x8 V; static constexpr int NP = 2; };
template <> struct Split<2> { typedef _Float16 T; typedef f16x8 V; static constexpr int NP = 2; };
template <> struct Split<3> { typedef __bf16 T; typedef bf16x8 V; static constexpr int NP = 3; };

__device__ __forceinline__ floatx4 mfma16(bf16x8 a, bf16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mfma16(f16x8 a, f16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Per-thread share of one k step: row tid / 4 of the 64-row A and B tiles,
// eight consecutive k at 8 (tid % 4).
struct Stage32 {
  float a[8], b[8];
};
template <int MODE> struct StageS {
  typename Split<MODE>::V a[Split<MODE>::NP], b[Split<MODE>::NP];
};
struct Smem32 {
  float A[2][PBM][kLd32], B[2][PBN][kLd32];
};
template <int MODE> struct SmemS {
  typename Split<MODE>::T A[2][Split<MODE>::NP][PBM][PK], B[2][Split<MODE>::NP][PBN][PK];
};

__device__ __forceinline__ void load_b(const PGemm &p, const float *__restrict__ B, float sb, float sub, int col0,
                                       int k0, float (&v)[8]) {
  const int r = gtid() >> 2, k = k0 + 8 * (gtid() & 3);
  const float *row = B + (size_t)(col0 + r) * p.ldb;
  const float4 a0 = *reinterpret_cast<const float4 *>(row + k);
  const float4 a1 = *reinterpret_cast<const float4 *>(row + k + 4);
  v[0] = a0.x; v[1] = a0.y; v[2] = a0.z; v[3] = a0.w;
  v[4] = a1.x; v[5] = a1.y; v[6] = a1.z; v[7] = a1.w;
  if (sb != 0.f) {
    if ((p.nfb & 3) == 0 && k + 8 <= p.hb) {
      // mirror k + i -> nfb - 1 - k - i: two aligned float4, reversed
      const float *m = row + (p.nfb - 8 - k);
      const float4 m0 = *reinterpret_cast<const float4 *>(m);
      const float4 m1 = *reinterpret_cast<const float4 *>(m + 4);
      const float mv[8] = {m1.w, m1.z, m1.y, m1.x, m0.w, m0.z, m0.y, m0.x};
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (v[i] - sub) + sb * (mv[i] - sub);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = k + i < p.hb ? (v[i] - sub) + sb * (row[p.nfb - 1 - k - i] - sub) : 0.f;
    }
  } else if (sub != 0.f) {
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] -= sub;
  }
}

__device__ __forceinline__ void load_stage(const PGemm &p, int z, const float *B, float sb, float sub, int row0,
                                           int col0, int k0, Stage32 &st) {
  const int r = gtid() >> 2, k = k0 + 8 * (gtid() & 3);
  const float *a = p.Af[z] + (size_t)(row0 + r) * p.lda + k;
  const float4 a0 = *reinterpret_cast<const float4 *>(a);
  const float4 a1 = *reinterpret_cast<const float4 *>(a + 4);
  st.a[0] = a0.x; st.a[1] = a0.y; st.a[2] = a0.z; st.a[3] = a0.w;
  st.a[4] = a1.x; st.a[5] = a1.y; st.a[6] = a1.z; st.a[7] = a1.w;
  load_b(p, B, sb, sub, col0, k0, st.b);
}

// Split modes: both operands arrive as planes (A stored split at set-up, B
// split once per stage by split_kernel), one 16-byte load per plane.
template <int MODE>
__device__ __forceinline__ void load_stage(const PGemm &p, int z, const float *, float, float, int row0, int col0,
                                           int k0, StageS<MODE> &st) {
  typedef typename Split<MODE>::T T;
  typedef typename Split<MODE>::V V;
  const int r = gtid() >> 2, k = k0 + 8 * (gtid() & 3);
  const size_t offa = (size_t)(row0 + r) * p.lda + k, offb = (size_t)(col0 + r) * p.ldbp + k;
#pragma unroll
  for (int q = 0; q < Split<MODE>::NP; ++q) {
    st.a[q] = *reinterpret_cast<const V *>(static_cast<const T *>(p.Ap[z][q]) + offa);
    st.b[q] = *reinterpret_cast<const V *>(static_cast<const T *>(p.Bp[z][q]) + offb);
  }
}

__device__ __forceinline__ void store_stage(Smem32 &s, int buf, const Stage32 &st, float) {
  const int r = gtid() >> 2, k = 8 * (gtid() & 3);
  *reinterpret_cast<float4 *>(&s.A[buf][r][k]) = make_float4(st.a[0], st.a[1], st.a[2], st.a[3]);
  *reinterpret_cast<float4 *>(&s.A[buf][r][k + 4]) = make_float4(st.a[4], st.a[5], st.a[6], st.a[7]);
  *reinterpret_cast<float4 *>(&s.B[buf][r][k]) = make_float4(st.b[0], st.b[1], st.b[2], st.b[3]);
  *reinterpret_cast<float4 *>(&s.B[buf][r][k + 4]) = make_float4(st.b[4], st.b[5], st.b[6], st.b[7]);
}

template <int MODE>
__device__ __forceinline__ void store_stage(SmemS<MODE> &s, int buf, const StageS<MODE> &st, float) {
  typedef typename Split<MODE>::V V;
  const int r = gtid() >> 2, k = swz16(r, gtid() & 3);
#pragma unroll
  for (int q = 0; q < Split<MODE>::NP; ++q) {
    *reinterpret_cast<V *>(&s.A[buf][q][r][k]) = st.a[q];
    *reinterpret_cast<V *>(&s.B[buf][q][r][k]) = st.b[q];
  }
}

// FP32: lane group g = lane / 16 takes k = 8 g + s in sub-step s for both
// operands, so every 16x16x4 sub-step pairs the same k values.
__device__ __forceinline__ void mma(const Smem32 &s, int buf, floatx4 (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, wave = gtid() >> 6;
  const int kq = 8 * (lane >> 4);
  float a[2][8], b[2][8];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi) {
    const float *q = &s.A[buf][(wave >> 1) * 32 + mi * 16 + (lane & 15)][kq];
    const float4 x0 = *reinterpret_cast<const float4 *>(q), x1 = *reinterpret_cast<const float4 *>(q + 4);
    a[mi][0] = x0.x; a[mi][1] = x0.y; a[mi][2] = x0.z; a[mi][3] = x0.w;
    a[mi][4] = x1.x; a[mi][5] = x1.y; a[mi][6] = x1.z; a[mi][7] = x1.w;
  }
#pragma unroll
  for (int ni = 0; ni < 2; ++ni) {
    const float *q = &s.B[buf][(wave & 1) * 32 + ni * 16 + (lane & 15)][kq];
    const float4 x0 = *reinterpret_cast<const float4 *>(q), x1 = *reinterpret_cast<const float4 *>(q + 4);
    b[ni][0] = x0.x; b[ni][1] = x0.y; b[ni][2] = x0.z; b[ni][3] = x0.w;
    b[ni][4] = x1.x; b[ni][5] = x1.y; b[ni][6] = x1.z; b[ni][7] = x1.w;
  }
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni)
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mi][t], b[ni][t], acc[mi][ni], 0, 0, 0);
}

// Split modes: lane group g supplies the eight k values 8 g .. 8 g + 7 of
// its row (A) and column (B); both operands use the same slots.  The plane
// products are issued smallest first: (qa, qb) with qa + qb = NP - 1 down
// to (0, 0).
template <int MODE>
__device__ __forceinline__ void mma(const SmemS<MODE> &s, int buf, floatx4 (&acc)[2][2]) {
  typedef typename Split<MODE>::V V;
  constexpr int NP = Split<MODE>::NP;
  const int lane = threadIdx.x & 63, wave = gtid() >> 6;
  V a[2][NP], b[2][NP];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi) {
    const int r = (wave >> 1) * 32 + mi * 16 + (lane & 15);
#pragma unroll
    for (int q = 0; q < NP; ++q) a[mi][q] = *reinterpret_cast<const V *>(&s.A[buf][q][r][swz16(r, lane >> 4)]);
  }
#pragma unroll
  for (int ni = 0; ni < 2; ++ni) {
    const int r = (wave & 1) * 32 + ni * 16 + (lane & 15);
#pragma unroll
    for (int q = 0; q < NP; ++q) b[ni][q] = *reinterpret_cast<const V *>(&s.B[buf][q][r][swz16(r, lane >> 4)]);
  }
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int ord = NP - 1; ord >= 0; --ord)
#pragma unroll
        for (int qa = ord; qa >= 0; --qa) acc[mi][ni] = mfma16(a[mi][qa], b[ni][ord - qa], acc[mi][ni]);
}

// Register stages per k-group: the global loads of this many rounds are in
// flight while a round multiplies.
#ifndef EP_PGEMM_DEPTH
#define EP_PGEMM_DEPTH 2
#endif
constexpr int kDepth = EP_PGEMM_DEPTH;

// acc += A[rows, k_begin..k_end) B[cols, k_begin..k_end)^T, the k steps
// split between the block's k-groups (group g takes steps g, g + kGroups,
// ...).  Each group has its own LDS double buffer and kDepth register
// stages: while round r multiplies from LDS, the global loads of rounds
// r + 1 .. r + kDepth are in flight (round r + 1's go to LDS at the end of
// round r, and round r's stage then takes round r + kDepth's loads).  Every
// group passes the same number of block barriers (one per round), also when
// it has no step left.
template <class ST, class SM>
__device__ __forceinline__ void run_k(const PGemm &p, int z, const float *B, float sb, float sub, float bscale,
                                      int row0, int col0, int k_begin, int k_end, SM &s, floatx4 (&acc)[2][2]) {
  // K spans of a run are multiples of kGroups * PK (the host checks), so
  // every group runs the same rounds and every test below depends only on
  // kernel arguments: branches stay scalar, and the loads of the next
  // rounds stay in flight across them.
  const int grp = __builtin_amdgcn_readfirstlane(threadIdx.x / kGroupThreads);
  const int rounds = (k_end - k_begin) / (PK * kGroups);
  auto kstep = [&](int r) { return k_begin + (r * kGroups + grp) * PK; };
  ST st[kDepth];  // st[r % kDepth] holds round r's loads until they go to LDS
  if (rounds >= 2 * kDepth && rounds % kDepth == 0) {
    // Branch-free pipeline: the wait before each LDS store then covers
    // exactly the loads of the round it stores, with kDepth - 1 later
    // rounds still in flight.
#pragma unroll
    for (int i = 0; i < kDepth; ++i) load_stage(p, z, B, sb, sub, row0, col0, kstep(i), st[i]);
    __syncthreads();  // the previous pass may still read buffer 0
    store_stage(s, 0, st[0], bscale);
    __syncthreads();
    int r0 = 0;
    for (; r0 < rounds - kDepth; r0 += kDepth) {
#pragma unroll
      for (int j = 0; j < kDepth; ++j) {
        const int r = r0 + j;
        load_stage(p, z, B, sb, sub, row0, col0, kstep(r + kDepth), st[j]);
        // keep the loads ahead of the multiply (the scheduler would sink
        // them below the next LDS store, leaving one round to cover them)
        __builtin_amdgcn_sched_barrier(0);
        mma(s, r & 1, acc);
        store_stage(s, (r + 1) & 1, st[(j + 1) % kDepth], bscale);
        __syncthreads();
      }
    }
    // the last kDepth rounds: nothing left to load
#pragma unroll
    for (int j = 0; j < kDepth; ++j) {
      const int r = r0 + j;
      mma(s, r & 1, acc);
      if (j + 1 < kDepth) store_stage(s, (r + 1) & 1, st[(j + 1) % kDepth], bscale);
      __syncthreads();
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < kDepth; ++i)
    if (i < rounds) load_stage(p, z, B, sb, sub, row0, col0, kstep(i), st[i]);
  __syncthreads();  // the previous pass may still read buffer 0
  if (rounds > 0) store_stage(s, 0, st[0], bscale);
  __syncthreads();
  for (int r0 = 0; r0 < rounds; r0 += kDepth) {
#pragma unroll
    for (int j = 0; j < kDepth; ++j) {
      const int r = r0 + j;
      if (r >= rounds) break;
      // LDS buffer r & 1 holds round r; its stage st[j] is free again
      if (r + kDepth < rounds) load_stage(p, z, B, sb, sub, row0, col0, kstep(r + kDepth), st[j]);
      mma(s, r & 1, acc);
      if (r + 1 < rounds) store_stage(s, (r + 1) & 1, st[(j + 1) % kDepth], bscale);
      __syncthreads();
    }
  }
}

// Adds the other k-groups' accumulators into group 0's through LDS (the
// groups' stage buffers are free after the last round's barrier).
template <class SM>
__device__ __forceinline__ void reduce_groups(SM (&s)[kGroups], floatx4 (&acc)[2][2]) {
  float *red = reinterpret_cast<float *>(&s[0]);
  const int grp = threadIdx.x / kGroupThreads, t = gtid();
  for (int g = 1; g < kGroups; ++g) {
    if (grp == g)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) red[((i * 2 + j) * 4 + r) * kGroupThreads + t] = acc[i][j][r];
    __syncthreads();
    if (grp == 0)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] += red[((i * 2 + j) * 4 + r) * kGroupThreads + t];
    __syncthreads();
  }
}

// Power of two 2^e with |v| 2^e <= 2^kF16ScaleExp for every |v| <= m
// (1 for m = 0 or not finite).
__device__ __forceinline__ float f16_scale(float m) {
  if (!(m > 0.f) || !isfinite(m)) return 1.f;
  int e;
  frexpf(m, &e);  // m < 2^e
  e = kF16ScaleExp - e;
  e = e < -100 ? -100 : (e > 100 ? 100 : e);
  return ldexpf(1.f, e);
}

template <int MODE>
__global__ void __launch_bounds__(kGroups * kGroupThreads) pgemm_kernel(PGemm p) {
  typedef typename std::conditional<MODE != 0, StageS<MODE == 0 ? 1 : MODE>, Stage32>::type ST;
  typedef typename std::conditional<MODE != 0, SmemS<MODE == 0 ? 1 : MODE>, Smem32>::type SM;
  static_assert(sizeof(SM) >= 16 * kGroupThreads * sizeof(float), "reduction buffer");
  __shared__ __attribute__((aligned(16))) SM s[kGroups];
  const int blk = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  unsigned long long *stamp = nullptr;
  if (p.stamp && threadIdx.x == 0) {
    stamp = p.stamp + (size_t)(*p.evals & 1) * p.stamp_set + 2 * blk;
    stamp[0] = wall_clock64();
  }
  const int z = blockIdx.z;
  const float *__restrict__ B = p.B[z];
  float *__restrict__ C = p.C[z];
  const float sb = p.sb[z];
  const float sub = p.sub ? (float)*p.sub : 0.f;
  const int grp = __builtin_amdgcn_readfirstlane(threadIdx.x / kGroupThreads);
  const int lane = threadIdx.x & 63, wave = gtid() >> 6;
  const int row0 = blockIdx.y * PBM, col0 = blockIdx.x * PBN;
  const int roww = row0 + (wave >> 1) * 32, colw = col0 + (wave & 1) * 32;
  // mode 2: undo the operands' power-of-two scales in the epilogue
  float bscale = 1.f, out_scale = 1.f;
  if (MODE == 2) {
    bscale = *p.bscale;
    out_scale = p.a_inv[z] / bscale;
  }
  floatx4 acc[2][2], odd[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
      odd[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
  // the spectral divisor's wave numbers of this thread's outputs, loaded
  // before the k loop so that their latency is not the epilogue's
  float wx2[2][4] = {}, wy2[2] = {};
  if (p.epi == 1 && grp == 0) {
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int r = 0; r < 4; ++r) wx2[mi][r] = p.wx2p[p.row_off[z] + roww + mi * 16 + 4 * (lane >> 4) + r];
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) wy2[ni] = p.wy2p[colw + ni * 16 + (lane & 15)];
  }
  if (p.out_fold) {
    run_k<ST>(p, z, B, sb, sub, bscale, row0, col0, 0, p.khalf, s[grp], acc);
    run_k<ST>(p, z, B, sb, sub, bscale, row0, col0, p.khalf, p.K, s[grp], odd);
    reduce_groups(s, acc);
    reduce_groups(s, odd);
  } else {
    run_k<ST>(p, z, B, sb, sub, bscale, row0, col0, 0, p.K, s[grp], acc);
    reduce_groups(s, acc);
  }
  if (grp != 0) return;
  const float se = p.se[z], so = p.so[z];
  float wmax = 0.f;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = roww + mi * 16 + 4 * (lane >> 4) + r;
        const int col = colw + ni * 16 + (lane & 15);
        float e = acc[mi][ni][r], o = odd[mi][ni][r];
        if (MODE == 2) {
          e *= out_scale;
          o *= out_scale;
        }
        if (p.out_fold) {
          if (row < p.ho) {
            const float v0 = e + o, v1 = se * e + so * o;
            C[(size_t)row * p.ldc + col] = v0;
            C[(size_t)(p.nfo - 1 - row) * p.ldc + col] = v1;
            wmax = fmaxf(wmax, fmaxf(fabsf(v0), fabsf(v1)));
          }
        } else if (p.epi == 1) {
          const float a = wx2[mi][r], b = wy2[ni];
          const bool dc = p.row_off[z] + row == 0 && col == 0;
          const float v = (a < 0.f || b < 0.f || dc) ? 0.f : e / (a + b);
          C[(size_t)row * p.ldc + col] = v;
          wmax = fmaxf(wmax, fabsf(v));
        } else {
          C[(size_t)row * p.ldc + col] = e;
          wmax = fmaxf(wmax, fabsf(e));
        }
      }
  if (MODE == 2 && p.out_max) {
    for (int off = 32; off > 0; off >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, off, 64));
    if (lane == 0) p.out_max[blk * (kGroupThreads / 64) + wave] = wmax;
  }
  if (stamp) stamp[1] = wall_clock64();
}

// One stage's data operand split into planes for the split-mode GEMMs:
// batch z, row n, columns k < K of
//   v = (src[n][k] - sub) + sb_z (src[n][nfb-1-k] - sub) for k < hb, 0 beyond  (fold)
//   v = src[n][k]                                                              (no fold)
// times the mode-2 scale, stored as plane q at out + ((z NP + q) rows + n) K + k.
// Each element is split once here instead of once per GEMM block that reads
// it.  Mode 2 takes the scale from the bound on |src| (the largest of
// in_max[0..in_n), or *in_max_d) and stores it for the GEMM's epilogue.
struct SplitJob {
  const float *src;
  int ld, rows, K, nbatch;
  float sb[2];
  int fold, nfb, hb;
  const double *sub;
  const float *in_max;
  int in_n;
  const double *in_max_d;
  float *scale_out;
  void *out;
};

constexpr int kSplitThreads = 256;

template <int MODE>
__global__ void __launch_bounds__(kSplitThreads) split_kernel(SplitJob j) {
  typedef typename Split<MODE>::T T;
  typedef typename Split<MODE>::V V;
  constexpr int NP = Split<MODE>::NP;
  float scale = 1.f;
  if (MODE == 2) {
    float m = 0.f;
    if (j.in_max_d) {
      m = (float)*j.in_max_d;
    } else {
      __shared__ float wm[kSplitThreads / 64];
      for (int i = threadIdx.x; i < j.in_n; i += blockDim.x) m = fmaxf(m, j.in_max[i]);
      for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
      if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
      __syncthreads();
      m = wm[0];
#pragma unroll
      for (int i = 1; i < kSplitThreads / 64; ++i) m = fmaxf(m, wm[i]);
    }
    scale = f16_scale(m);  // folded values add two of at most m
    if (blockIdx.x == 0 && threadIdx.x == 0) *j.scale_out = scale;
  }
  const float sub = j.sub ? (float)*j.sub : 0.f;
  const int per_row = j.K / 8;
  const int64_t per_batch = (int64_t)j.rows * per_row, total = per_batch * j.nbatch;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(t / per_batch);
    const int64_t rem = t - z * per_batch;
    const int n = (int)(rem / per_row), k = 8 * (int)(rem - (int64_t)n * per_row);
    const float *row = j.src + (size_t)n * j.ld;
    float v[8];
    if (j.fold) {
      const float sb = j.sb[z];
      if ((j.nfb & 3) == 0 && k + 8 <= j.hb) {
        // k + i and its mirror nfb - 1 - k - i: two aligned float4 each
        const float4 a0 = *reinterpret_cast<const float4 *>(row + k);
        const float4 a1 = *reinterpret_cast<const float4 *>(row + k + 4);
        const float4 m0 = *reinterpret_cast<const float4 *>(row + j.nfb - 8 - k);
        const float4 m1 = *reinterpret_cast<const float4 *>(row + j.nfb - 4 - k);
        const float a[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
        const float mv[8] = {m1.w, m1.z, m1.y, m1.x, m0.w, m0.z, m0.y, m0.x};
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = (a[i] - sub) + sb * (mv[i] - sub);
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i)
          v[i] = k + i < j.hb ? (row[k + i] - sub) + sb * (row[j.nfb - 1 - k - i] - sub) : 0.f;
      }
    } else {
      const float4 a0 = *reinterpret_cast<const float4 *>(row + k);
      const float4 a1 = *reinterpret_cast<const float4 *>(row + k + 4);
      v[0] = a0.x; v[1] = a0.y; v[2] = a0.z; v[3] = a0.w;
      v[4] = a1.x; v[5] = a1.y; v[6] = a1.z; v[7] = a1.w;
    }
    V pl[NP];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      float r = MODE == 2 ? v[i] * scale : v[i];
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const T x = (T)r;
        pl[q][i] = x;
        r -= (float)x;
      }
    }
    T *out = static_cast<T *>(j.out);
#pragma unroll
    for (int q = 0; q < NP; ++q)
      *reinterpret_cast<V *>(out + ((size_t)(z * NP + q) * j.rows + n) * j.K + k) = pl[q];
  }
}

static void split(Ctx &c, const SplitJob &j) {
  const int64_t chunks = (int64_t)j.nbatch * j.rows * (j.K / 8);
  const int blocks = (int)std::min<int64_t>(256, (chunks + kSplitThreads - 1) / kSplitThreads);
  switch (c.poisson->mode) {
    case 1: hipLaunchKernelGGL(split_kernel<1>, dim3(blocks), dim3(kSplitThreads), 0, c.stream, j); break;
    case 2: hipLaunchKernelGGL(split_kernel<2>, dim3(blocks), dim3(kSplitThreads), 0, c.stream, j); break;
    default: hipLaunchKernelGGL(split_kernel<3>, dim3(blocks), dim3(kSplitThreads), 0, c.stream, j); break;
  }
}

// Number of per-wave maxima a launch of this shape writes.
static int wave_max_count(const PGemm &p, int batch) {
  return (p.N / PBN) * (p.M / PBM) * batch * (kGroupThreads / 64);
}

// Every K span is a multiple of the 64-element padding of the halves, so
// run_k's rounds are whole for both k-groups.
static_assert(64 % (PK * kGroups) == 0, "K spans must hold whole rounds");

static void gemm(Ctx &c, const PGemm &p, int batch) {
  dim3 grid(p.N / PBN, p.M / PBM, batch);
  const dim3 block(kGroups * kGroupThreads);
  switch (c.poisson->mode) {
    case 1: hipLaunchKernelGGL(pgemm_kernel<1>, grid, block, 0, c.stream, p); break;
    case 2: hipLaunchKernelGGL(pgemm_kernel<2>, grid, block, 0, c.stream, p); break;
    case 3: hipLaunchKernelGGL(pgemm_kernel<3>, grid, block, 0, c.stream, p); break;
    default: hipLaunchKernelGGL(pgemm_kernel<0>, grid, block, 0, c.stream, p); break;
  }
}

static int round_up(int v, int m) { return (v + m - 1) / m * m; }

template <class T>
static bool put(T *&d, const std::vector<T> &h) {
  return hipMalloc(&d, h.size() * sizeof(T)) == hipSuccess &&
         hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

// Uploads a constant matrix in the form the mode reads: FP32 always, bf16
// planes (modes 1 and 3), or fp16 halves of the matrix times 2^f16_exp
// (mode 2; the exponent puts the largest |entry| in [2^12, 2^13)).
static bool upload(Mat &m, const std::vector<float> &h, int mode) {
  if (!put(m.f, h)) return false;
  if (mode == 1 || mode == 3) {
    const int np = mode == 3 ? 3 : 2;
    std::vector<float> rem(h);
    for (int q = 0; q < np; ++q) {
      std::vector<__bf16> pl(h.size());
      for (size_t i = 0; i < h.size(); ++i) {
        pl[i] = (__bf16)rem[i];
        rem[i] -= (float)pl[i];
      }
      if (!put(m.bp[q], pl)) return false;
    }
  } else if (mode == 2) {
    float mx = 0.f;
    for (float v : h) mx = std::max(mx, std::fabs(v));
    int e = 0;
    if (mx > 0.f) std::frexp(mx, &e);
    m.f16_exp = kF16ScaleExp - e;
    std::vector<_Float16> hi(h.size()), lo(h.size());
    for (size_t i = 0; i < h.size(); ++i) {
      const float v = std::ldexp(h[i], m.f16_exp);
      hi[i] = (_Float16)v;
      lo[i] = (_Float16)(v - (float)hi[i]);
    }
    if (!put(m.hp[0], hi) || !put(m.hp[1], lo)) return false;
  }
  return true;
}

static void release(Mat &m) {
  void *p[] = {m.f, m.bp[0], m.bp[1], m.bp[2], m.hp[0], m.hp[1]};
  for (void *q : p)
    if (q) (void)hipFree(q);
  m = Mat();
}

bool poisson_create(Ctx &c, std::string &err) {
  if (c.nx % 2 || c.ny % 2) {
    err = "the transform needs an even number of bins per side";
    return false;
  }
  Poisson *q = new Poisson();
  c.poisson = q;
  q->mode = c.poisson_mode;
  const int nx = c.nx, ny = c.ny;
  q->nxp = round_up(nx, 64);
  q->nyp = round_up(ny, 64);
  q->hx = nx / 2;
  q->hy = ny / 2;
  q->hxp = round_up(q->hx, 64);
  q->hyp = round_up(q->hy, 64);
  const int X = q->nxp, Y = q->nyp, Hx = q->hx, Hy = q->hy, Hxp = q->hxp, Hyp = q->hyp;
  const double pi = M_PI;
  std::vector<double> wx(nx), wy(ny);
  for (int u = 0; u < nx; ++u) wx[u] = pi * u / nx;
  for (int v = 0; v < ny; ++v) wy[v] = pi * v / ny * ((double)c.bsy / (double)c.bsx);
  // frequency of slot k of a parity-split axis whose halves are hp long
  auto freq = [](int k, int hp) { return k < hp ? 2 * k : 2 * (k - hp) + 1; };
  auto F = [&](int k, int x, int n) { return (2.0 / n) * (k == 0 ? 0.5 : 1.0) * std::cos(pi * k * (x + 0.5) / n); };
  std::vector<float> FYA((size_t)2 * Hyp * Hyp, 0.f), FXA((size_t)2 * Hxp * Hxp, 0.f);
  std::vector<float> CYT((size_t)Hyp * 2 * Hyp, 0.f), SYT((size_t)Hyp * 2 * Hyp, 0.f);
  std::vector<float> CX2((size_t)Hxp * 2 * Hxp, 0.f), SX2((size_t)Hxp * 2 * Hxp, 0.f);
  std::vector<float> wx2p(2 * Hxp, -1.f), wy2p(2 * Hyp, -1.f);
  for (int k = 0; k < 2 * Hyp; ++k) {
    const int v = freq(k, Hyp);
    if (v >= ny) continue;
    wy2p[k] = (float)(wy[v] * wy[v]);
    for (int y = 0; y < Hy; ++y) {
      FYA[(size_t)k * Hyp + y] = (float)F(v, y, ny);
      CYT[(size_t)y * 2 * Hyp + k] = (float)std::cos(pi * v * (y + 0.5) / ny);
      SYT[(size_t)y * 2 * Hyp + k] = (float)(wy[v] * std::sin(pi * v * (y + 0.5) / ny));
    }
  }
  for (int k = 0; k < 2 * Hxp; ++k) {
    const int u = freq(k, Hxp);
    if (u >= nx) continue;
    wx2p[k] = (float)(wx[u] * wx[u]);
    for (int x = 0; x < Hx; ++x) {
      FXA[(size_t)k * Hxp + x] = (float)F(u, x, nx);
      CX2[(size_t)x * 2 * Hxp + k] = (float)std::cos(pi * u * (x + 0.5) / nx);
      SX2[(size_t)x * 2 * Hxp + k] = (float)(std::sin(pi * u * (x + 0.5) / nx) * wx[u]);
    }
  }
  auto up = [&](float *&d, const std::vector<float> &h) -> bool {
    if (hipMalloc(&d, h.size() * sizeof(float)) != hipSuccess) return false;
    return hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
  };
  auto zeros = [&](float *&d, size_t n) -> bool {
    if (hipMalloc(&d, n * sizeof(float)) != hipSuccess) return false;
    return hipMemset(d, 0, n * sizeof(float)) == hipSuccess;
  };
  const int md = q->mode;
  // planes of the largest stage operand: stage 1 (2 batches x X rows x Hyp),
  // stage 2 (2 x 2Hyp x Hxp), stage 3 (2Hxp x 2Hyp), stage 4 (2Y x 2Hxp)
  const size_t split_elems = std::max(std::max((size_t)2 * X * Hyp, (size_t)4 * Hyp * Hxp),
                                      std::max((size_t)4 * Hxp * Hyp, (size_t)4 * Y * Hxp));
  const int np = md == 3 ? 3 : 2;
  bool ok = upload(q->FYA, FYA, md) && upload(q->FXA, FXA, md) && upload(q->CYT, CYT, md) &&
            upload(q->SYT, SYT, md) && upload(q->CX2, CX2, md) && upload(q->SX2, SX2, md) && up(q->wx2p, wx2p) &&
            up(q->wy2p, wy2p) && (md != 2 || zeros(q->wmax, (size_t)3 * kMaxWaveMax)) &&
            (md == 0 || (hipMalloc(&q->bsplit, split_elems * np * 2) == hipSuccess &&
                         hipMemset(q->bsplit, 0, split_elems * np * 2) == hipSuccess)) &&
            (md == 0 || zeros(q->bscale, 4)) &&
            zeros(q->T1t, (size_t)2 * Hyp * X) && zeros(q->spec, (size_t)2 * Hxp * 2 * Hyp) &&
            zeros(q->PQt, (size_t)2 * Y * 2 * Hxp) && zeros(c.rho, (size_t)X * Y) && zeros(c.phi, (size_t)X * Y) &&
            zeros(c.ex, (size_t)X * Y) && zeros(c.ey, (size_t)X * Y);
  // wall-clock stamps: the blocks of each stage's launch (batch 2; the
  // standalone solve with phi adds a third batch to stage 4)
  GemmStamps &st = c.gemm_stamps;
  st.nblk[0] = (X / PBN) * (Hyp / PBM) * 2;
  st.nblk[1] = (2 * Hyp / PBN) * (Hxp / PBM) * 2;
  st.nblk[2] = (2 * Hxp / PBN) * (Hyp / PBM) * 2;
  st.nblk[3] = (Y / PBN) * (Hxp / PBM) * 2;
  st.stride = std::max(std::max(st.nblk[0], st.nblk[1]), std::max(st.nblk[2], st.nblk[3] / 2 * 3));
  const size_t nstamp = (size_t)4 * kGemmStamped * st.stride;
  st.evals = &c.sc->evals;
  ok = ok && hipMalloc(&st.t, nstamp * sizeof(unsigned long long)) == hipSuccess &&
       hipMemset(st.t, 0, nstamp * sizeof(unsigned long long)) == hipSuccess;
  if (!ok) err = "poisson_create: device allocation failed";
  // per-wave maxima of the largest launch (stages 1 and 3: (X/64)(Hyp/64) and
  // (2Hxp/64)(Hyp/64) blocks per batch, two batches, four waves each)
  const long waves = 2L * 4 * (Hyp / 64) * std::max(X / 64, 2 * Hxp / 64);
  if (ok && md == 2 && waves > kMaxWaveMax) {
    err = "poisson_create: grid too large for the split-FP16 transforms";
    ok = false;
  }
  return ok;
}

void poisson_destroy(Ctx &c) {
  Poisson *q = c.poisson;
  if (!q) return;
  Mat *mats[] = {&q->FYA, &q->FXA, &q->CYT, &q->SYT, &q->CX2, &q->SX2};
  for (Mat *m : mats) release(*m);
  float *ptrs[] = {q->wx2p, q->wy2p, q->T1t, q->spec, q->PQt, q->wmax, q->bscale};
  if (q->bsplit) (void)hipFree(q->bsplit);
  for (float *p : ptrs)
    if (p) (void)hipFree(p);
  if (c.gemm_stamps.t) (void)hipFree(c.gemm_stamps.t);
  c.gemm_stamps = GemmStamps{};
  delete q;
  c.poisson = nullptr;
}

static void set_a(PGemm &p, int z, const Mat &m, size_t off) {
  p.Af[z] = m.f + off;
  for (int q = 0; q < kMaxPlanes; ++q) p.Ap[z][q] = m.bp[q] ? m.bp[q] + off : nullptr;
  if (m.hp[0]) {
    p.Ap[z][0] = m.hp[0] + off;
    p.Ap[z][1] = m.hp[1] + off;
    p.Ap[z][2] = nullptr;
  }
  p.a_inv[z] = std::ldexp(1.f, -m.f16_exp);
}

void launch_poisson(Ctx &c, bool with_phi, hipEvent_t *gev) {
  Poisson *q = c.poisson;
  int g_launch = 0;
  auto timed_gemm = [&](PGemm &p, int batch) {
    p.stamp = c.gemm_stamps.t + (size_t)2 * g_launch * c.gemm_stamps.stride;
    p.evals = c.gemm_stamps.evals;
    p.stamp_set = 2 * kGemmStamped * c.gemm_stamps.stride;
    if (gev) (void)hipEventRecord(gev[2 * g_launch], c.stream);
    gemm(c, p, batch);
    if (gev) (void)hipEventRecord(gev[2 * g_launch + 1], c.stream);
    ++g_launch;
  };
  const int X = q->nxp, Y = q->nyp, Hxp = q->hxp, Hyp = q->hyp;
  const bool split_mode = q->mode != 0;
  const int np = q->mode == 3 ? 3 : 2;
  // split modes: plane qq of batch z of a split of `rows` rows of K, from row r0
  auto plane = [&](int z, int qq, int rows, int K, int r0) -> const void * {
    return static_cast<const char *>(q->bsplit) + (((size_t)(z * np + qq) * rows + r0) * K) * 2;
  };
  auto set_b = [&](PGemm &p, int z, int bz, int rows, int K, int r0) {
    for (int qq = 0; qq < np; ++qq) p.Bp[z][qq] = plane(bz, qq, rows, K, r0);
    p.ldbp = K;
  };
  PGemm p = {};
  SplitJob j = {};
  // 1. T1t[v][x] = sum_y F_Y[v][y] (rho[x][y] - mean), even v on fold +,
  // odd v on fold -.  A constant only moves rho^[0][0], which the solve
  // discards; removing it keeps its float rounding out of the low
  // frequencies.  Mode 2 scales rho by its largest value.
  p.sub = &c.sc->rho_mean;
  set_a(p, 0, q->FYA, 0);
  set_a(p, 1, q->FYA, (size_t)Hyp * Hyp);
  p.B[0] = p.B[1] = c.rho;
  p.C[0] = q->T1t;
  p.C[1] = q->T1t + (size_t)Hyp * X;
  p.sb[0] = 1.f;
  p.sb[1] = -1.f;
  p.lda = Hyp; p.ldb = Y; p.ldc = X;
  p.M = Hyp; p.N = X; p.K = Hyp;
  p.nfb = c.ny; p.hb = q->hy;
  p.out_max = q->wmax;
  if (split_mode) {
    j = SplitJob{};
    j.src = c.rho; j.ld = Y; j.rows = X; j.K = Hyp; j.nbatch = 2;
    j.sb[0] = 1.f; j.sb[1] = -1.f; j.fold = 1; j.nfb = c.ny; j.hb = q->hy;
    j.sub = &c.sc->rho_mean; j.in_max_d = &c.sc->rho_max; j.scale_out = q->bscale; j.out = q->bsplit;
    split(c, j);
    set_b(p, 0, 0, X, Hyp, 0);
    set_b(p, 1, 1, X, Hyp, 0);
    p.bscale = q->bscale;
  }
  int n_prev = wave_max_count(p, 2);
  timed_gemm(p, 2);
  // 2. spec[u][v] = sum_x F_X[u][x] T1t[v][x] / (wx^2 + wy^2)
  p = PGemm{};
  set_a(p, 0, q->FXA, 0);
  set_a(p, 1, q->FXA, (size_t)Hxp * Hxp);
  p.B[0] = p.B[1] = q->T1t;
  p.C[0] = q->spec;
  p.C[1] = q->spec + (size_t)Hxp * 2 * Hyp;
  p.sb[0] = 1.f;
  p.sb[1] = -1.f;
  p.row_off[0] = 0;
  p.row_off[1] = Hxp;
  p.lda = Hxp; p.ldb = X; p.ldc = 2 * Hyp;
  p.M = Hxp; p.N = 2 * Hyp; p.K = Hxp;
  p.nfb = c.nx; p.hb = q->hx;
  p.epi = 1;
  p.wx2p = q->wx2p;
  p.wy2p = q->wy2p;
  p.out_max = q->wmax + kMaxWaveMax;
  if (split_mode) {
    j = SplitJob{};
    j.src = q->T1t; j.ld = X; j.rows = 2 * Hyp; j.K = Hxp; j.nbatch = 2;
    j.sb[0] = 1.f; j.sb[1] = -1.f; j.fold = 1; j.nfb = c.nx; j.hb = q->hx;
    j.in_max = q->wmax; j.in_n = n_prev; j.scale_out = q->bscale + 1; j.out = q->bsplit;
    split(c, j);
    set_b(p, 0, 0, 2 * Hyp, Hxp, 0);
    set_b(p, 1, 1, 2 * Hyp, Hxp, 0);
    p.bscale = q->bscale + 1;
  }
  n_prev = wave_max_count(p, 2);
  timed_gemm(p, 2);
  // 3. Pt[y][u] = sum_v C_Y[v][y] spec[u][v] (mirror: e - o),
  //    Qt[y][u] = sum_v wy S_Y[v][y] spec[u][v] (mirror: o - e)
  p = PGemm{};
  set_a(p, 0, q->CYT, 0);
  set_a(p, 1, q->SYT, 0);
  p.B[0] = p.B[1] = q->spec;
  p.C[0] = q->PQt;
  p.C[1] = q->PQt + (size_t)Y * 2 * Hxp;
  p.se[0] = 1.f;  p.so[0] = -1.f;
  p.se[1] = -1.f; p.so[1] = 1.f;
  p.lda = 2 * Hyp; p.ldb = 2 * Hyp; p.ldc = 2 * Hxp;
  p.M = Hyp; p.N = 2 * Hxp; p.K = 2 * Hyp; p.khalf = Hyp;
  p.out_fold = 1; p.nfo = c.ny; p.ho = q->hy;
  p.out_max = q->wmax + 2 * kMaxWaveMax;
  if (split_mode) {
    j = SplitJob{};
    j.src = q->spec; j.ld = 2 * Hyp; j.rows = 2 * Hxp; j.K = 2 * Hyp; j.nbatch = 1;
    j.in_max = q->wmax + kMaxWaveMax; j.in_n = n_prev; j.scale_out = q->bscale + 2; j.out = q->bsplit;
    split(c, j);
    set_b(p, 0, 0, 2 * Hxp, 2 * Hyp, 0);
    set_b(p, 1, 0, 2 * Hxp, 2 * Hyp, 0);
    p.bscale = q->bscale + 2;
  }
  n_prev = wave_max_count(p, 2);
  timed_gemm(p, 2);
  // 4. Ex[x][y] = sum_u wx S_X[u][x] Pt[y][u] (mirror: o - e),
  //    Ey = sum_u C_X[u][x] Qt[y][u], phi = sum_u C_X[u][x] Pt[y][u] (e - o)
  p = PGemm{};
  set_a(p, 0, q->SX2, 0);
  set_a(p, 1, q->CX2, 0);
  set_a(p, 2, q->CX2, 0);
  p.B[0] = q->PQt;
  p.B[1] = q->PQt + (size_t)Y * 2 * Hxp;
  p.B[2] = q->PQt;
  p.C[0] = c.ex;
  p.C[1] = c.ey;
  p.C[2] = c.phi;
  p.se[0] = -1.f; p.so[0] = 1.f;
  p.se[1] = 1.f;  p.so[1] = -1.f;
  p.se[2] = 1.f;  p.so[2] = -1.f;
  p.lda = 2 * Hxp; p.ldb = 2 * Hxp; p.ldc = Y;
  p.M = Hxp; p.N = Y; p.K = 2 * Hxp; p.khalf = Hxp;
  p.out_fold = 1; p.nfo = c.nx; p.ho = q->hx;
  if (split_mode) {
    j = SplitJob{};
    j.src = q->PQt; j.ld = 2 * Hxp; j.rows = 2 * Y; j.K = 2 * Hxp; j.nbatch = 1;
    j.in_max = q->wmax + 2 * kMaxWaveMax; j.in_n = n_prev; j.scale_out = q->bscale + 3; j.out = q->bsplit;
    split(c, j);
    set_b(p, 0, 0, 2 * Y, 2 * Hxp, 0);
    set_b(p, 1, 0, 2 * Y, 2 * Hxp, Y);
    set_b(p, 2, 0, 2 * Y, 2 * Hxp, 0);
    p.bscale = q->bscale + 3;
  }
  timed_gemm(p, with_phi ? 3 : 2);
}

}  // namespace ep
