// Poisson solve for the bin charge density as dense transform GEMMs on MFMA.
//
// Reference: FFT_2D::doFFT (FFT/fft.cpp:190-284) on Ooura's ddct2d /
// ddsct2d / ddcst2d (FFT/fftsg2d.cpp:105-297).  Closed form, x-major maps
// (map[x][y], as the reference's bins[x][y]):
//   rho^ = F_X rho F_Y^T,   F[k][x] = (2/N) c_k cos(pi k (x+1/2)/N), c_0 = 1/2
//   phi^ = rho^ / (wx^2 + wy^2), phi^[0][0] = 0         (GEMM 2 epilogue)
//   [P | Q] = phi^ [C_Y | W_Y S_Y]                         (GEMM 3)
//   Ex = (S_X^T W_X) P,  Ey = C_X^T Q,  phi = C_X^T P       (GEMM 4, batched)
// with C[k][x] = cos(pi k (x+1/2)/N), S[k][x] = sin(pi k (x+1/2)/N),
// wx = pi u / Nx, wy = pi v / Ny * (bin_h / bin_w) (fft.cpp:129-147).
// Grids are zero-padded to multiples of 64 so every GEMM tile is full.
#include <cmath>
#include <vector>

#include "ep_internal.h"

namespace ep {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct Poisson {
  int nxp = 0, nyp = 0;
  float *FYt = nullptr, *FX = nullptr, *CYWS = nullptr, *SXtWX = nullptr, *CXt = nullptr;
  float *T1 = nullptr, *spec = nullptr, *PQ = nullptr;
  float *wx2 = nullptr, *wy2 = nullptr;
};

int poisson_ld(const Ctx &c) { return c.poisson ? c.poisson->nyp : c.ny; }

struct GemmBatch {
  const float *A[3];
  const float *B[3];
  float *C[3];
  int lda, ldb, ldc;
  int M, N, K;
  int epi;  // 0 store, 1 spectral divide
  const double *sub;  // if set, A := A - *sub (removes the DC of rho)
  const float *wx2, *wy2;
  int nx, ny;
};

constexpr int BM = 64, BN = 64, BK = 16;

__device__ __forceinline__ float epilogue(const GemmBatch &p, int row, int col, float v) {
  if (p.epi == 1) {
    if (row >= p.nx || col >= p.ny || (row == 0 && col == 0)) return 0.f;
    return v / (p.wx2[row] + p.wy2[col]);
  }
  return v;
}

// C[M x N] = A[M x K] B[K x N], row-major, FP32 MFMA 16x16x4.  Block tile
// 64x64 (4 waves of 32x32), K tile 16; LDS rows padded against bank
// conflicts.  Requires M, N multiples of 64, K of 16, 16-byte aligned rows.
__global__ void __launch_bounds__(256) gemm_f32_kernel(GemmBatch p) {
  const int bz = blockIdx.z;
  const float *__restrict__ A = p.A[bz];
  const float *__restrict__ B = p.B[bz];
  float *__restrict__ C = p.C[bz];
  const int row0 = blockIdx.y * BM, col0 = blockIdx.x * BN;
  __shared__ float As[BM][BK + 1];
  __shared__ __attribute__((aligned(16))) float Bs[BK][BN + 16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int ar = tid >> 2, ac = (tid & 3) * 4;
  const int br = tid >> 4, bc = (tid & 15) * 4;
  const float sub = p.sub ? (float)*p.sub : 0.f;
  for (int k0 = 0; k0 < p.K; k0 += BK) {
    float4 va = *reinterpret_cast<const float4 *>(A + (size_t)(row0 + ar) * p.lda + k0 + ac);
    va.x -= sub;
    va.y -= sub;
    va.z -= sub;
    va.w -= sub;
    float4 vb = *reinterpret_cast<const float4 *>(B + (size_t)(k0 + br) * p.ldb + col0 + bc);
    As[ar][ac] = va.x;
    As[ar][ac + 1] = va.y;
    As[ar][ac + 2] = va.z;
    As[ar][ac + 3] = va.w;
    *reinterpret_cast<float4 *>(&Bs[br][bc]) = vb;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float a[2], b[2];
#pragma unroll
      for (int mi = 0; mi < 2; ++mi) a[mi] = As[wm * 32 + mi * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) b[ni] = Bs[kk + (lane >> 4)][wn * 32 + ni * 16 + (lane & 15)];
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mi], b[ni], acc[mi][ni], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = row0 + wm * 32 + mi * 16 + 4 * (lane >> 4) + r;
        int col = col0 + wn * 32 + ni * 16 + (lane & 15);
        C[(size_t)row * p.ldc + col] = epilogue(p, row, col, acc[mi][ni][r]);
      }
}

static void gemm(Ctx &c, GemmBatch p, int batch) {
  dim3 grid(p.N / BN, p.M / BM, batch);
  hipLaunchKernelGGL(gemm_f32_kernel, grid, dim3(256), 0, c.stream, p);
}

static int round_up(int v, int m) { return (v + m - 1) / m * m; }

bool poisson_create(Ctx &c, std::string &err) {
  Poisson *q = new Poisson();
  c.poisson = q;
  const int nx = c.nx, ny = c.ny;
  q->nxp = round_up(nx, 64);
  q->nyp = round_up(ny, 64);
  const int X = q->nxp, Y = q->nyp;
  const double pi = M_PI;
  std::vector<float> FYt((size_t)Y * Y, 0.f), FX((size_t)X * X, 0.f), CYWS((size_t)Y * 2 * Y, 0.f),
      SXtWX((size_t)X * X, 0.f), CXt((size_t)X * X, 0.f), wx2(X, 0.f), wy2(Y, 0.f);
  std::vector<double> wx(nx), wy(ny);
  for (int u = 0; u < nx; ++u) wx[u] = pi * u / nx;
  for (int v = 0; v < ny; ++v) wy[v] = pi * v / ny * ((double)c.bsy / (double)c.bsx);
  for (int u = 0; u < nx; ++u) wx2[u] = (float)(wx[u] * wx[u]);
  for (int v = 0; v < ny; ++v) wy2[v] = (float)(wy[v] * wy[v]);
  for (int v = 0; v < ny; ++v)
    for (int y = 0; y < ny; ++y) {
      double cs = std::cos(pi * v * (y + 0.5) / ny), sn = std::sin(pi * v * (y + 0.5) / ny);
      FYt[(size_t)y * Y + v] = (float)((2.0 / ny) * (v == 0 ? 0.5 : 1.0) * cs);
      CYWS[(size_t)v * 2 * Y + y] = (float)cs;
      CYWS[(size_t)v * 2 * Y + Y + y] = (float)(wy[v] * sn);
    }
  for (int u = 0; u < nx; ++u)
    for (int x = 0; x < nx; ++x) {
      double cs = std::cos(pi * u * (x + 0.5) / nx), sn = std::sin(pi * u * (x + 0.5) / nx);
      FX[(size_t)u * X + x] = (float)((2.0 / nx) * (u == 0 ? 0.5 : 1.0) * cs);
      SXtWX[(size_t)x * X + u] = (float)(sn * wx[u]);
      CXt[(size_t)x * X + u] = (float)cs;
    }
  auto up = [&](float *&d, const std::vector<float> &h) -> bool {
    if (hipMalloc(&d, h.size() * sizeof(float)) != hipSuccess) return false;
    return hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
  };
  auto zeros = [&](float *&d, size_t n) -> bool {
    if (hipMalloc(&d, n * sizeof(float)) != hipSuccess) return false;
    return hipMemset(d, 0, n * sizeof(float)) == hipSuccess;
  };
  bool ok = up(q->FYt, FYt) && up(q->FX, FX) && up(q->CYWS, CYWS) && up(q->SXtWX, SXtWX) && up(q->CXt, CXt) &&
            up(q->wx2, wx2) && up(q->wy2, wy2) && zeros(q->T1, (size_t)X * Y) && zeros(q->spec, (size_t)X * Y) &&
            zeros(q->PQ, (size_t)X * 2 * Y) && zeros(c.rho, (size_t)X * Y) && zeros(c.phi, (size_t)X * Y) &&
            zeros(c.ex, (size_t)X * Y) && zeros(c.ey, (size_t)X * Y);
  if (!ok) err = "poisson_create: device allocation failed";
  return ok;
}

void poisson_destroy(Ctx &c) {
  Poisson *q = c.poisson;
  if (!q) return;
  float *ptrs[] = {q->FYt, q->FX, q->CYWS, q->SXtWX, q->CXt, q->T1, q->spec, q->PQ, q->wx2, q->wy2};
  for (float *p : ptrs)
    if (p) (void)hipFree(p);
  delete q;
  c.poisson = nullptr;
}

void launch_poisson(Ctx &c, bool with_phi) {
  Poisson *q = c.poisson;
  const int X = q->nxp, Y = q->nyp;
  GemmBatch p = {};
  p.nx = c.nx;
  p.ny = c.ny;
  p.wx2 = q->wx2;
  p.wy2 = q->wy2;
  // 1. T1 = (rho - mean) F_Y^T.  A constant only moves rho^[0][0], which
  // the solve discards; removing it keeps its float rounding out of the
  // low frequencies.  Padding rows/columns meet zero rows of F_Y^T / F_X.
  p.sub = &c.sc->rho_mean;
  p.A[0] = c.rho;
  p.B[0] = q->FYt;
  p.C[0] = q->T1;
  p.lda = Y;
  p.ldb = Y;
  p.ldc = Y;
  p.M = X;
  p.N = Y;
  p.K = Y;
  p.epi = 0;
  gemm(c, p, 1);
  // 2. phi^ = (F_X T1) / (wx^2 + wy^2)
  p.sub = nullptr;
  p.A[0] = q->FX;
  p.B[0] = q->T1;
  p.C[0] = q->spec;
  p.lda = X;
  p.ldb = Y;
  p.ldc = Y;
  p.M = X;
  p.N = Y;
  p.K = X;
  p.epi = 1;
  gemm(c, p, 1);
  // 3. [P | Q] = phi^ [C_Y | W_Y S_Y]
  p.A[0] = q->spec;
  p.B[0] = q->CYWS;
  p.C[0] = q->PQ;
  p.lda = Y;
  p.ldb = 2 * Y;
  p.ldc = 2 * Y;
  p.M = X;
  p.N = 2 * Y;
  p.K = Y;
  p.epi = 0;
  gemm(c, p, 1);
  // 4. Ex = (S_X^T W_X) P,  Ey = C_X^T Q,  [phi = C_X^T P]
  p.A[0] = q->SXtWX;
  p.B[0] = q->PQ;
  p.C[0] = c.ex;
  p.A[1] = q->CXt;
  p.B[1] = q->PQ + Y;
  p.C[1] = c.ey;
  p.A[2] = q->CXt;
  p.B[2] = q->PQ;
  p.C[2] = c.phi;
  p.lda = X;
  p.ldb = 2 * Y;
  p.ldc = Y;
  p.M = X;
  p.N = Y;
  p.K = X;
  p.epi = 0;
  gemm(c, p, with_phi ? 3 : 2);
}

}  // namespace ep
