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

#define KC 16   // k-chunk staged in LDS per iteration

template <int MT>
__global__ void __launch_bounds__(256)
gemm_f64_kernel(int M, int K, const double* __restrict__ A, int lda,
                const double* __restrict__ B, double* C, int ld, Epi e) {
  __shared__ double As[MT * 16][KC + 1];
  __shared__ double Bs[KC][64 + 1];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int n0 = blockIdx.x * 64;
  const int li = lane & 15;   // row of A / column of B and D held by this lane
  const int lk = lane >> 4;   // k index (A, B) and row quad (D)

  v4d acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};

  for (int k0 = 0; k0 < K; k0 += KC) {
    for (int idx = tid; idx < MT * 16 * KC; idx += 256) {
      const int r = idx / KC, c = idx % KC;
      As[r][c] = (r < M && k0 + c < K) ? A[(size_t)r * lda + k0 + c] : 0.0;
    }
    for (int idx = tid; idx < KC * 64; idx += 256) {
      const int r = idx >> 6, c = idx & 63;
      Bs[r][c] = (k0 + r < K) ? B[(size_t)(k0 + r) * ld + n0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC / 4; ++kk) {
      const double b = Bs[kk * 4 + lk][wave * 16 + li];
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        const double a = As[t * 16 + li][kk * 4 + lk];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }

  // v_mfma_f64_16x16x4_f64 output: D[i][j] of tile t sits in lane
  // j + 16*(i%4), register i/4 (one row group per register).
  const int col = n0 + wave * 16 + li;
  const int mode = col >> 1;
#pragma unroll
  for (int t = 0; t < MT; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = t * 16 + r * 4 + lk;
      if (row < M) {
        const size_t idx = (size_t)row * ld + col;
        C[idx] = epi_apply(e, acc[t][r], idx, mode);
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
