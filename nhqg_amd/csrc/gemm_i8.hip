// Mixed-precision vertical contractions: FP64 results from int8 slice
// products on the CDNA4 matrix cores (v_mfma_i32_16x16x64_i8).
//
// A (the fixed operator, at most 257 x 257) and each column c of B (the
// field) are written in fixed point with S signed 7-bit slices:
//   A[r][k] = 2^ea_r  sum_p 2^-7p a_p[r][k],   B[k][c] = 2^eb_c  sum_q 2^-7q b_q[k][c]
// with |a_p|, |b_q| <= 127 and a remainder below 2^-7S of the row / column
// maximum.  Every slice product sum_k a_p b_q is exact in int32 (256 * 127^2
// * S < 2^31), so the only roundings are the dropped remainder and slice
// pairs (p + q > S + 1) and the final FP64 sums:
//   C[r][c] ~ 2^(ea_r + eb_c) sum_{l=2}^{S+1} 2^-7l sum_{p+q=l} (a_p b_q)[r][c]
// The pairs of one level l share an int32 accumulator.  S = 6 keeps ~42 bits
// (21 products, each 1/64 of an FP64 MFMA's cost per flop).
//
// k = 0 is added exactly in FP64 (A[r][0] * B[0][c]) and the slices cover
// k = 1 .. min(K, 257) - 1 (at most 256, four k-steps of 64).  The first
// Chebyshev coefficient carries the vertically uniform (barotropic) part of
// q, often far larger than the rest of its column; left in the slices it
// would set the column's scale, and operators that annihilate it (d/dZ,
// whose other entries reach ~Nz^2) would amplify the remainder of the small
// coefficients.
//
// Per wave: 16 columns.  The wave loads its columns of B in FP64 once (the
// same HBM traffic as the FP64 kernel), finds each column's exponent, and
// keeps all S slices of all k in registers.  A's slices, pre-split on the
// host, stream tile by tile (16 rows) through LDS, shared by the 4 waves of
// a workgroup.  Each tile's accumulators are turned back into FP64 and go
// through the same epilogue arithmetic as the FP64 kernel.
#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>

#include "common.h"

typedef int v4i __attribute__((ext_vector_type(4)));

#define GI_NW 4   // waves per workgroup

// 4 consecutive rows (row0 + r) of one column: epi_apply's arithmetic with
// each extra field read for the 4 elements at once.
NHQG_DEV void epi_rows4(const Epi& e, double* C, const double (&v)[4], int row0, int M, int ld,
                        int col, int mode) {
  double o[4];
  const double sa = e.s_acc ? e.s_acc[mode] : 1.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    o[r] = e.a_acc * v[r];
    if (e.s_acc) o[r] *= sa;
  }
  for (int x = 0; x < e.nX; ++x) {
    const double* X = e.X[x];
    const double ax = e.aX[x];
    const bool hs = e.sX[x] != nullptr;
    const double sx = hs ? e.sX[x][mode] : 1.0;
    double w[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) w[r] = X[(size_t)min(row0 + r, M - 1) * ld + col];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double t = ax * w[r];
      if (hs) t *= sx;
      o[r] += t;
    }
  }
  const double so = e.s_out ? e.s_out[mode] : 1.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = row0 + r;
    if (row < M) {
      double t = o[r];
      if (e.s_out) t *= so;
      if (e.d_tab) t *= e.d_tab[(size_t)row * e.d_ld + mode];
      C[(size_t)row * ld + col] = t;
    }
  }
}

template <int S>
__global__ void __launch_bounds__(GI_NW * 64)
gemm_i8_kernel(GemmBatch bt) {
  constexpr int NT = GI_NW * 64;
  __shared__ v4i As[2][S * 4 * 64];   // [buf][(p * nk + kk) * 64 + lane]
  __shared__ double sa_s[17 * 16], af_s[17 * 16];   // row scales and A[r][0] (nti <= 17)

  const GemmJob& J = bt.job[blockIdx.y];
  const v4i* __restrict__ A8 = (const v4i*)J.op.A8;
  const int M = J.op.M, K = J.op.K, nk = J.op.nk, nti = J.op.nti;
  const int KI = K - 1;   // sliced k = 1 .. K-1 (K <= 257)
  const int ld = bt.ld;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int col = blockIdx.x * (GI_NW * 16) + wave * 16 + li;
  // this lane's column; row offsets fit in 32 bits (257 * ld < 2^31 for Nx <= 1024)
  const double* __restrict__ Bc = J.B + col;
  const int mode = col >> 1;
  const int tile_v = S * nk * 64;   // v4i per A tile

  // A tile 0 in flight while B is sliced
  v4i ra[S];
  auto fetchA = [&](int t) {
#pragma unroll
    for (int u = 0; u < S; ++u) {
      const int idx = tid + u * NT;
      ra[u] = A8[(size_t)t * tile_v + min(idx, tile_v - 1)];
    }
  };
  auto stashA = [&](int buf) {
#pragma unroll
    for (int u = 0; u < S; ++u) {
      const int idx = tid + u * NT;
      if (idx < tile_v) As[buf][idx] = ra[u];
    }
  };
  fetchA(0);
  for (int i = tid; i < nti * 16; i += NT) {
    sa_s[i] = J.op.sa[i];
    af_s[i] = J.op.afirst[i];
  }

  // this lane's 64 values of its column (sliced index kl -> k = kl + 1),
  // all loads in flight at once; rows beyond K-1 are clamped and zeroed
  double bv[4][16];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int kl = kk * 64 + lk * 16 + j;
      bv[kk][j] = Bc[(min(kl, KI - 1) + 1) * ld];
    }
  const double bfirst = Bc[0];
  // column exponent (the 4 lanes of a column share it)
  double mx = 0.0;
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int kl = kk * 64 + lk * 16 + j;
      if (kl >= KI) bv[kk][j] = 0.0;
      mx = fmax(mx, fabs(bv[kk][j]));
    }
  mx = fmax(mx, __shfl_xor(mx, 16));
  mx = fmax(mx, __shfl_xor(mx, 32));
  int eb = 0;
  if (mx > 0.0) (void)frexp(mx, &eb);
  const double inv = ldexp(1.0, -eb);

  // slices: bs[kk][q] holds bytes kl = kk*64 + lk*16 + (0..15) of slice q
  v4i bs[4][S];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      int pk[S];
#pragma unroll
      for (int q = 0; q < S; ++q) pk[q] = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        double x = bv[kk][g * 4 + b] * inv;
#pragma unroll
        for (int q = 0; q < S; ++q) {
          const double y = x * 128.0;
          const double d = fmin(fmax(rint(y), -127.0), 127.0);
          x = y - d;
          pk[q] |= ((int)d & 0xff) << (8 * b);
        }
      }
#pragma unroll
      for (int q = 0; q < S; ++q) bs[kk][q][g] = pk[q];
    }
  const double scale_c = ldexp(1.0, eb - 14);

  stashA(0);
  __syncthreads();

  for (int t = 0; t < nti; ++t) {
    const int cur = t & 1;
    if (t + 1 < nti) fetchA(t + 1);
    v4i acc[S];
#pragma unroll
    for (int l = 0; l < S; ++l) acc[l] = v4i{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      if (kk < nk) {
#pragma unroll
        for (int p = 0; p < S; ++p) {
          const v4i a = As[cur][(p * nk + kk) * 64 + lane];
#pragma unroll
          for (int q = 0; q < S - p; ++q)
            acc[p + q] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, bs[kk][q], acc[p + q], 0, 0, 0);
        }
      }
    }
    // output rows t*16 + 4*lk + r, column li (v_mfma_i32_16x16x64_i8 layout)
    const int row0 = t * 16 + lk * 4;
    double v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double s = (double)acc[S - 1][r];
#pragma unroll
      for (int l = S - 2; l >= 0; --l) s = (double)acc[l][r] + s * 0.0078125;
      v[r] = fma(af_s[row0 + r], bfirst, s * scale_c * sa_s[row0 + r]);
    }
    if (J.C2) epi_rows4(J.e2, J.C2, v, row0, M, ld, col, mode);
    epi_rows4(J.e, J.C, v, row0, M, ld, col, mode);
    if (t + 1 < nti) stashA(cur ^ 1);
    __syncthreads();
  }
}

bool gemm_i8_supported(int S) { return S >= 4 && S <= 7; }

hipError_t launch_gemm_i8(const GemmBatch& b, int nj, hipStream_t s) {
  if (nj < 1 || nj > GEMM_MAX_JOBS || b.ld % (GI_NW * 16) != 0 || !gemm_i8_supported(b.slices))
    return hipErrorInvalidValue;
  for (int j = 0; j < nj; ++j) {
    const GemmJob& J = b.job[j];
    const GemmOp& op = J.op;
    if (!op.A8 || !op.sa || !op.afirst || op.slices != b.slices || op.M < 1 || op.M > 257 ||
        op.K < 2 ||
        op.K > 257 || op.nk != (op.K - 1 + 63) / 64 ||
        op.nti != (op.M + 15) / 16 || J.W || !J.B || !J.C)
      return hipErrorInvalidValue;
  }
  dim3 grid(b.ld / (GI_NW * 16), nj);
  switch (b.slices) {
    case 4: hipLaunchKernelGGL(gemm_i8_kernel<4>, grid, dim3(GI_NW * 64), 0, s, b); break;
    case 5: hipLaunchKernelGGL(gemm_i8_kernel<5>, grid, dim3(GI_NW * 64), 0, s, b); break;
    case 6: hipLaunchKernelGGL(gemm_i8_kernel<6>, grid, dim3(GI_NW * 64), 0, s, b); break;
    default: hipLaunchKernelGGL(gemm_i8_kernel<7>, grid, dim3(GI_NW * 64), 0, s, b); break;
  }
  return hipGetLastError();
}

// Host split of A (M x K row-major, 2 <= K <= 257) into S slices of
// k = 1 .. K-1 in the kernel's tile order: byte ((((t * S + p) * nk + kk) *
// 64 + lane) * 16 + j) holds slice p of A[t*16 + lane%16][1 + kk*64 +
// (lane/16)*16 + j]; sa[r] = 2^ea_r over those k; afirst[r] = A[r][0].  Rows
// up to nti*16 padded with zeros.
void gemm_i8_split(const double* A, int M, int K, int S, std::vector<int8_t>& a8,
                   std::vector<double>& sa, std::vector<double>& afirst) {
  const int KI = K - 1, nk = (KI + 63) / 64, nti = (M + 15) / 16;
  a8.assign((size_t)S * nti * nk * 64 * 16, 0);
  sa.assign((size_t)nti * 16, 0.0);
  afirst.assign((size_t)nti * 16, 0.0);
  for (int r = 0; r < M; ++r) {
    const double* row = A + (size_t)r * K;
    double mx = 0.0;
    for (int kl = 0; kl < KI; ++kl) mx = std::fmax(mx, std::fabs(row[kl + 1]));
    int e = 0;
    if (mx > 0.0) (void)std::frexp(mx, &e);
    sa[r] = std::ldexp(1.0, e);
    const double inv = std::ldexp(1.0, -e);
    const int t = r / 16, li = r % 16;
    for (int kl = 0; kl < KI; ++kl) {
      const int kk = kl / 64, lk = (kl % 64) / 16, j = kl % 16;
      const int lane = li + 16 * lk;
      double x = row[kl + 1] * inv;
      for (int p = 0; p < S; ++p) {
        const double y = x * 128.0;
        const double d = std::fmin(std::fmax(std::nearbyint(y), -127.0), 127.0);
        x = y - d;
        a8[((((size_t)t * S + p) * nk + kk) * 64 + lane) * 16 + j] = (int8_t)d;
      }
    }
    afirst[r] = row[0];
  }
}
