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
// Two arithmetic modes: FP32 MFMA (16x16x4), and split-BF16: each FP32
// value is a_hi + a_lo (two bf16; the constant matrices are stored split),
// and a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi on bf16 MFMA (16x16x32),
// accumulated in FP32.
#include <hip/hip_bf16.h>

#include <cmath>
#include <vector>

#include "ep_internal.h"

namespace ep {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// A constant transform matrix: FP32, plus its bf16 hi/lo split.
struct Mat {
  float *f = nullptr;
  __bf16 *hi = nullptr, *lo = nullptr;
};

struct Poisson {
  int mode = 0;                // 0 FP32, 1 split-BF16
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
};

int poisson_ld(const Ctx &c) { return c.poisson ? c.poisson->nyp : c.ny; }

// One launch of up to three GEMMs C = A B^T (blockIdx.z selects).
struct PGemm {
  const float *Af[3];          // FP32 A (mode 0)
  const __bf16 *Ah[3], *Al[3];  // split A (mode 1)
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
};

constexpr int PBM = 64, PBN = 64, PK = 32;
constexpr int kGroups = 2;  // k-groups per block: each runs half the k steps, then they add up
constexpr int kGroupThreads = 256;

// Thread index inside the thread's k-group.
__device__ __forceinline__ int gtid() { return threadIdx.x & (kGroupThreads - 1); }
constexpr int kLd32 = PK + 4;  // FP32 LDS row stride (floats), rows 16-byte aligned
// bf16 LDS tiles: 64-byte rows (32 k) of four 16-byte chunks, chunk c of row
// r stored at c ^ ((r >> 1 ^ r >> 3) & 3).  The stores (whole rows) and the
// MFMA operand reads (16 consecutive rows, one chunk) then touch distinct
// 16-byte bank groups per 8 lanes with 32 banks and per 16 lanes with 64.
__device__ __forceinline__ int swz16(int r, int c) { return 8 * (c ^ ((r >> 1 ^ r >> 3) & 3)); }

// Per-thread share of one k step: row tid / 4 of the 64-row A and B tiles,
// eight consecutive k at 8 (tid % 4).
struct Stage32 {
  float a[8], b[8];
};
struct Stage16 {
  bf16x8 ah, al;
  float b[8];
};
struct Smem32 {
  float A[2][PBM][kLd32], B[2][PBN][kLd32];
};
struct Smem16 {
  __bf16 Ah[2][PBM][PK], Al[2][PBM][PK], Bh[2][PBN][PK], Bl[2][PBN][PK];
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

__device__ __forceinline__ void load_stage(const PGemm &p, int z, const float *B, float sb, float sub, int row0,
                                           int col0, int k0, Stage16 &st) {
  const int r = gtid() >> 2, k = k0 + 8 * (gtid() & 3);
  const size_t off = (size_t)(row0 + r) * p.lda + k;
  st.ah = *reinterpret_cast<const bf16x8 *>(p.Ah[z] + off);
  st.al = *reinterpret_cast<const bf16x8 *>(p.Al[z] + off);
  load_b(p, B, sb, sub, col0, k0, st.b);
}

__device__ __forceinline__ void store_stage(Smem32 &s, int buf, const Stage32 &st) {
  const int r = gtid() >> 2, k = 8 * (gtid() & 3);
  *reinterpret_cast<float4 *>(&s.A[buf][r][k]) = make_float4(st.a[0], st.a[1], st.a[2], st.a[3]);
  *reinterpret_cast<float4 *>(&s.A[buf][r][k + 4]) = make_float4(st.a[4], st.a[5], st.a[6], st.a[7]);
  *reinterpret_cast<float4 *>(&s.B[buf][r][k]) = make_float4(st.b[0], st.b[1], st.b[2], st.b[3]);
  *reinterpret_cast<float4 *>(&s.B[buf][r][k + 4]) = make_float4(st.b[4], st.b[5], st.b[6], st.b[7]);
}

__device__ __forceinline__ void store_stage(Smem16 &s, int buf, const Stage16 &st) {
  const int r = gtid() >> 2, k = swz16(r, gtid() & 3);
  *reinterpret_cast<bf16x8 *>(&s.Ah[buf][r][k]) = st.ah;
  *reinterpret_cast<bf16x8 *>(&s.Al[buf][r][k]) = st.al;
  bf16x8 h, l;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const __bf16 x = (__bf16)st.b[i];
    h[i] = x;
    l[i] = (__bf16)(st.b[i] - (float)x);
  }
  *reinterpret_cast<bf16x8 *>(&s.Bh[buf][r][k]) = h;
  *reinterpret_cast<bf16x8 *>(&s.Bl[buf][r][k]) = l;
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

// split-BF16: lane group g supplies the eight k values 8 g .. 8 g + 7 of its
// row (A) and column (B); both operands use the same slots.
__device__ __forceinline__ void mma(const Smem16 &s, int buf, floatx4 (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, wave = gtid() >> 6;
  bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi) {
    const int r = (wave >> 1) * 32 + mi * 16 + (lane & 15);
    ah[mi] = *reinterpret_cast<const bf16x8 *>(&s.Ah[buf][r][swz16(r, lane >> 4)]);
    al[mi] = *reinterpret_cast<const bf16x8 *>(&s.Al[buf][r][swz16(r, lane >> 4)]);
  }
#pragma unroll
  for (int ni = 0; ni < 2; ++ni) {
    const int r = (wave & 1) * 32 + ni * 16 + (lane & 15);
    bh[ni] = *reinterpret_cast<const bf16x8 *>(&s.Bh[buf][r][swz16(r, lane >> 4)]);
    bl[ni] = *reinterpret_cast<const bf16x8 *>(&s.Bl[buf][r][swz16(r, lane >> 4)]);
  }
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) {
      acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mi], bh[ni], acc[mi][ni], 0, 0, 0);
      acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bl[ni], acc[mi][ni], 0, 0, 0);
      acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mi], bh[ni], acc[mi][ni], 0, 0, 0);
    }
}

// acc += A[rows, k_begin..k_end) B[cols, k_begin..k_end)^T, the k steps
// split between the block's k-groups (group g takes steps g, g + kGroups,
// ...).  Each group has its own LDS double buffer and two register stages:
// while round r multiplies from LDS, the global loads of rounds r + 1 and
// r + 2 are in flight (round r + 1's go to LDS at the end of round r).
// Every group passes the same number of block barriers (one per round),
// also when it has no step left.
template <class ST, class SM>
__device__ __forceinline__ void run_k(const PGemm &p, int z, const float *B, float sb, float sub, int row0,
                                      int col0, int k_begin, int k_end, SM &s, floatx4 (&acc)[2][2]) {
  const int grp = threadIdx.x / kGroupThreads;
  const int steps = (k_end - k_begin + PK - 1) / PK;
  const int rounds = (steps + kGroups - 1) / kGroups;
  auto kstep = [&](int r) { return k_begin + (r * kGroups + grp) * PK; };
  auto have = [&](int r) { return kstep(r) < k_end; };
  ST even, odd;  // register stages of the even and the odd rounds
  if (have(0)) load_stage(p, z, B, sb, sub, row0, col0, kstep(0), even);
  if (have(1)) load_stage(p, z, B, sb, sub, row0, col0, kstep(1), odd);
  __syncthreads();  // the previous pass may still read buffer 0
  if (have(0)) store_stage(s, 0, even);
  __syncthreads();
  for (int r = 0; r < rounds; r += 2) {
    // round r: buffer 0 holds step r, `odd` step r + 1, `even` is free
    if (have(r + 2)) load_stage(p, z, B, sb, sub, row0, col0, kstep(r + 2), even);
    if (have(r)) mma(s, 0, acc);
    if (have(r + 1)) store_stage(s, 1, odd);
    __syncthreads();
    if (r + 1 >= rounds) break;
    // round r + 1: buffer 1 holds step r + 1, `even` step r + 2, `odd` is free
    if (have(r + 3)) load_stage(p, z, B, sb, sub, row0, col0, kstep(r + 3), odd);
    if (have(r + 1)) mma(s, 1, acc);
    if (have(r + 2)) store_stage(s, 0, even);
    __syncthreads();
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

template <int BF16>
__global__ void __launch_bounds__(kGroups * kGroupThreads) pgemm_kernel(PGemm p) {
  typedef typename std::conditional<BF16 != 0, Stage16, Stage32>::type ST;
  typedef typename std::conditional<BF16 != 0, Smem16, Smem32>::type SM;
  static_assert(sizeof(SM) >= 16 * kGroupThreads * sizeof(float), "reduction buffer");
  __shared__ __attribute__((aligned(16))) SM s[kGroups];
  const int z = blockIdx.z;
  const float *__restrict__ B = p.B[z];
  float *__restrict__ C = p.C[z];
  const float sb = p.sb[z];
  const float sub = p.sub ? (float)*p.sub : 0.f;
  const int grp = threadIdx.x / kGroupThreads;
  const int lane = threadIdx.x & 63, wave = gtid() >> 6;
  const int row0 = blockIdx.y * PBM, col0 = blockIdx.x * PBN;
  const int roww = row0 + (wave >> 1) * 32, colw = col0 + (wave & 1) * 32;
  floatx4 acc[2][2], odd[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
      odd[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
  if (p.out_fold) {
    run_k<ST>(p, z, B, sb, sub, row0, col0, 0, p.khalf, s[grp], acc);
    run_k<ST>(p, z, B, sb, sub, row0, col0, p.khalf, p.K, s[grp], odd);
    reduce_groups(s, acc);
    reduce_groups(s, odd);
  } else {
    run_k<ST>(p, z, B, sb, sub, row0, col0, 0, p.K, s[grp], acc);
    reduce_groups(s, acc);
  }
  if (grp != 0) return;
  const float se = p.se[z], so = p.so[z];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = roww + mi * 16 + 4 * (lane >> 4) + r;
        const int col = colw + ni * 16 + (lane & 15);
        const float e = acc[mi][ni][r], o = odd[mi][ni][r];
        if (p.out_fold) {
          if (row < p.ho) {
            C[(size_t)row * p.ldc + col] = e + o;
            C[(size_t)(p.nfo - 1 - row) * p.ldc + col] = se * e + so * o;
          }
        } else if (p.epi == 1) {
          const float a = p.wx2p[p.row_off[z] + row], b = p.wy2p[col];
          const bool dc = p.row_off[z] + row == 0 && col == 0;
          C[(size_t)row * p.ldc + col] = (a < 0.f || b < 0.f || dc) ? 0.f : e / (a + b);
        } else {
          C[(size_t)row * p.ldc + col] = e;
        }
      }
}

static void gemm(Ctx &c, const PGemm &p, int batch) {
  dim3 grid(p.N / PBN, p.M / PBM, batch);
  const dim3 block(kGroups * kGroupThreads);
  if (c.poisson->mode == 1) hipLaunchKernelGGL(pgemm_kernel<1>, grid, block, 0, c.stream, p);
  else hipLaunchKernelGGL(pgemm_kernel<0>, grid, block, 0, c.stream, p);
}

static int round_up(int v, int m) { return (v + m - 1) / m * m; }

static bool upload(Mat &m, const std::vector<float> &h) {
  std::vector<__bf16> hi(h.size()), lo(h.size());
  for (size_t i = 0; i < h.size(); ++i) {
    hi[i] = (__bf16)h[i];
    lo[i] = (__bf16)(h[i] - (float)hi[i]);
  }
  return hipMalloc(&m.f, h.size() * 4) == hipSuccess &&
         hipMemcpy(m.f, h.data(), h.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMalloc(&m.hi, h.size() * 2) == hipSuccess &&
         hipMemcpy(m.hi, hi.data(), h.size() * 2, hipMemcpyHostToDevice) == hipSuccess &&
         hipMalloc(&m.lo, h.size() * 2) == hipSuccess &&
         hipMemcpy(m.lo, lo.data(), h.size() * 2, hipMemcpyHostToDevice) == hipSuccess;
}

static void release(Mat &m) {
  void *p[] = {m.f, m.hi, m.lo};
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
  bool ok = upload(q->FYA, FYA) && upload(q->FXA, FXA) && upload(q->CYT, CYT) && upload(q->SYT, SYT) &&
            upload(q->CX2, CX2) && upload(q->SX2, SX2) && up(q->wx2p, wx2p) && up(q->wy2p, wy2p) &&
            zeros(q->T1t, (size_t)2 * Hyp * X) && zeros(q->spec, (size_t)2 * Hxp * 2 * Hyp) &&
            zeros(q->PQt, (size_t)2 * Y * 2 * Hxp) && zeros(c.rho, (size_t)X * Y) && zeros(c.phi, (size_t)X * Y) &&
            zeros(c.ex, (size_t)X * Y) && zeros(c.ey, (size_t)X * Y);
  if (!ok) err = "poisson_create: device allocation failed";
  return ok;
}

void poisson_destroy(Ctx &c) {
  Poisson *q = c.poisson;
  if (!q) return;
  Mat *mats[] = {&q->FYA, &q->FXA, &q->CYT, &q->SYT, &q->CX2, &q->SX2};
  for (Mat *m : mats) release(*m);
  float *ptrs[] = {q->wx2p, q->wy2p, q->T1t, q->spec, q->PQt};
  for (float *p : ptrs)
    if (p) (void)hipFree(p);
  delete q;
  c.poisson = nullptr;
}

static void set_a(PGemm &p, int z, const Mat &m, size_t off) {
  p.Af[z] = m.f + off;
  p.Ah[z] = m.hi + off;
  p.Al[z] = m.lo + off;
}

void launch_poisson(Ctx &c, bool with_phi) {
  Poisson *q = c.poisson;
  const int X = q->nxp, Y = q->nyp, Hxp = q->hxp, Hyp = q->hyp;
  PGemm p = {};
  // 1. T1t[v][x] = sum_y F_Y[v][y] (rho[x][y] - mean), even v on fold +,
  // odd v on fold -.  A constant only moves rho^[0][0], which the solve
  // discards; removing it keeps its float rounding out of the low
  // frequencies.
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
  gemm(c, p, 2);
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
  gemm(c, p, 2);
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
  gemm(c, p, 2);
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
  gemm(c, p, with_phi ? 3 : 2);
}

}  // namespace ep
