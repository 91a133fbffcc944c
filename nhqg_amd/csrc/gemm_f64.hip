// FP64 vertical contractions on the CDNA matrix cores.
//
// Every vertical operator of the step (Chebyshev <-> CGL transforms,
// Galerkin stencils and left inverses, tau projections, d/dZ, SBP transfers)
// is a small real matrix A (at most 257 x 257) applied to all modes of a
// level-major field: C = A * B with B a (K x 2*nm_pad) real matrix.  One
// workgroup owns 64 columns and ALL rows of C, so B is streamed from HBM
// exactly once; A chunks come from L2.  Each wave computes the full column
// height (MT tiles of 16 rows) for 16 columns with
// v_mfma_f64_16x16x4_f64, and the result passes through a linear epilogue
// that fuses the stage-assembly arithmetic.
#include "common.h"

#define KC 8    // k-chunk staged in LDS per iteration (double buffered)

// Wave w owns row tiles w, w+4, w+8, ... and all four 16-column tiles of
// the block, so each A value read from LDS feeds 4 MFMAs and each B value
// feeds up to MTW.  LDS holds k-major chunks (rows contiguous); the next
// chunk is fetched into registers while the current one is multiplied.
template <int MT>
__global__ void __launch_bounds__(256)
gemm_f64_kernel(int M, int K, const double* __restrict__ A, int lda,
                const double* __restrict__ B, double* C, int ld, Epi e) {
  constexpr int MP = MT * 16;
  constexpr int SA = MP + 2;          // padded row length of the A chunk
  constexpr int SB = 64 + 2;
  constexpr int MTW = (MT + 3) / 4;   // row tiles per wave
  constexpr int NA = (MP * KC + 255) / 256;
  constexpr int NB = (KC * 64) / 256;
  __shared__ double As[2][KC][SA];
  __shared__ double Bs[2][KC][SB];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int n0 = blockIdx.x * 64;
  const int li = lane & 15;
  const int lk = lane >> 4;

  v4d acc[MTW][4];
#pragma unroll
  for (int i = 0; i < MTW; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[i][c] = v4d{0.0, 0.0, 0.0, 0.0};

  double ra[NA], rb[NB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int idx = tid + u * 256;        // idx = r * KC + kc (kc fastest)
      const int r = idx / KC, kc = idx % KC;
      ra[u] = (idx < MP * KC && r < M && k0 + kc < K) ? A[(size_t)r * lda + k0 + kc] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int idx = tid + u * 256;        // idx = kr * 64 + c
      const int kr = idx >> 6, cc = idx & 63;
      rb[u] = (k0 + kr < K) ? B[(size_t)(k0 + kr) * ld + n0 + cc] : 0.0;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int idx = tid + u * 256;
      if (idx < MP * KC) As[buf][idx % KC][idx / KC] = ra[u];
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int idx = tid + u * 256;
      Bs[buf][idx >> 6][idx & 63] = rb[u];
    }
  };

  const int nchunks = (K + KC - 1) / KC;
  fetch(0);
  stash(0);
  __syncthreads();
  for (int ch = 0; ch < nchunks; ++ch) {
    const int cur = ch & 1;
    if (ch + 1 < nchunks) fetch((ch + 1) * KC);
#pragma unroll
    for (int kk = 0; kk < KC / 4; ++kk) {
      const int kr = kk * 4 + lk;
      double b[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[cur][kr][c * 16 + li];
#pragma unroll
      for (int i = 0; i < MTW; ++i) {
        const int t = wave + 4 * i;
        if (t < MT) {
          const double a = As[cur][kr][t * 16 + li];
#pragma unroll
          for (int c = 0; c < 4; ++c)
            acc[i][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[c], acc[i][c], 0, 0, 0);
        }
      }
    }
    if (ch + 1 < nchunks) stash(cur ^ 1);
    __syncthreads();
  }

  // v_mfma_f64_16x16x4_f64 output: D[i][j] of a tile sits in lane
  // j + 16*(i%4), register i/4 (one row group per register).
#pragma unroll
  for (int i = 0; i < MTW; ++i) {
    const int t = wave + 4 * i;
    if (t >= MT) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int col = n0 + c * 16 + li;
      const int mode = col >> 1;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = t * 16 + r * 4 + lk;
        if (row < M) {
          const size_t idx = (size_t)row * ld + col;
          C[idx] = epi_apply(e, acc[i][c][r], idx, mode);
        }
      }
    }
  }
}

template <int MT>
static hipError_t launch_mt(int M, int N, int K, const double* A, int lda, const double* B,
                            double* C, int ld, const Epi& e, hipStream_t s) {
  dim3 grid(N / 64);
  hipLaunchKernelGGL(gemm_f64_kernel<MT>, grid, dim3(256), 0, s, M, K, A, lda, B, C, ld, e);
  return hipGetLastError();
}

hipError_t launch_gemm_f64(int M, int N, int K, const double* A, int lda, const double* B,
                           double* C, int ld, const Epi& e, hipStream_t s) {
  if (N % 64 != 0 || M <= 0 || K <= 0) return hipErrorInvalidValue;
  const int mt = (M + 15) / 16;
  switch (mt) {
#define NHQG_MT(n) case n: return launch_mt<n>(M, N, K, A, lda, B, C, ld, e, s);
    NHQG_MT(1) NHQG_MT(2) NHQG_MT(3) NHQG_MT(4) NHQG_MT(5) NHQG_MT(6) NHQG_MT(7) NHQG_MT(8)
    NHQG_MT(9) NHQG_MT(10) NHQG_MT(11) NHQG_MT(12) NHQG_MT(13) NHQG_MT(14) NHQG_MT(15)
    NHQG_MT(16) NHQG_MT(17)
#undef NHQG_MT
    default: return hipErrorInvalidValue;
  }
}

__global__ void __launch_bounds__(256)
combine_kernel(size_t n, int ld, double* out, Epi e) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= n) return;
  const int mode = (int)((idx % (size_t)ld) >> 1);
  out[idx] = epi_apply(e, 0.0, idx, mode);
}

hipError_t launch_combine(int nl, int ld, double* out, const Epi& e, hipStream_t s) {
  const size_t n = (size_t)nl * ld;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(combine_kernel, dim3(blocks), dim3(256), 0, s, n, ld, out, e);
  return hipGetLastError();
}
