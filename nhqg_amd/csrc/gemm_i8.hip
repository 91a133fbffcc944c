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
#include <algorithm>
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

// ---------------------------------------------------------------------------
// Split operators (gemm_split.hip): the two half-size products E and O of a
// reflection-symmetric operator as int8 slice products.  Each half has its
// own FP64 first-k term, column exponent and row scales, and at most 128
// sliced k (two k-steps of 64); a tile holds both halves' 16-row slices, the
// same LDS footprint as a dense tile, for about half the dense MFMA work.
// The inputs of the halves (strided rows, or the butterfly B[k] +- B[N-k])
// and the outputs (strided rows, or E + O and E - O on reflected rows)
// follow the FP64 split kernel.
// ---------------------------------------------------------------------------
namespace {
constexpr int GI_MAXT = 9;      // row tiles per half
}  // namespace

// GI_NKH: k-steps of 64 per half (1 for halves at most 65 deep, else 2)
template <int S, int GI_NKH, bool IB, bool OB>
__global__ void __launch_bounds__(GI_NW * 64)
gemm_i8_split_kernel(GemmBatch bt) {
  constexpr int NT = GI_NW * 64;
  constexpr int TV = 2 * S * GI_NKH * 64;   // v4i per tile, both halves
  constexpr int NA = (TV + NT - 1) / NT;
  __shared__ v4i As[2][TV];
  __shared__ double sa_s[2][GI_MAXT * 16], af_s[2][GI_MAXT * 16];

  const GemmJob& J = bt.job[blockIdx.y];
  const v4i* __restrict__ A8 = (const v4i*)J.op.sA8;
  const int K = J.op.K, nti = J.op.snti;
  const int nk0 = J.op.sp_nk[0], nk1 = J.op.sp_nk[1];
  const int nbi = J.op.sp_nbi, hbi = nbi >> 1;
  const int istr = J.op.sp_istr, io0 = J.op.sp_ioff[0], io1 = J.op.sp_ioff[1];
  const int ld = bt.ld;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int col = blockIdx.x * (GI_NW * 16) + wave * 16 + li;
  const double* __restrict__ Bc = J.B + col;
  const int mode = col >> 1;

  v4i ra[NA];
  auto fetchA = [&](int t) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int idx = tid + u * NT;
      ra[u] = A8[(size_t)t * TV + min(idx, TV - 1)];
    }
  };
  auto stashA = [&](int buf) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int idx = tid + u * NT;
      if (idx < TV) As[buf][idx] = ra[u];
    }
  };
  fetchA(0);
  for (int i = tid; i < nti * 16; i += NT) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      sa_s[h][i] = J.op.ssa[h * nti * 16 + i];
      af_s[h][i] = J.op.safirst[h * nti * 16 + i];
    }
  }

  // input k of the two halves (k < nk_h), all loads in flight at once
  auto rows_of = [&](int k, int& r0, int& r1) {
    if constexpr (IB) {
      r0 = min(k, hbi);
      r1 = nbi - r0;
    } else {
      r0 = min(istr * k + io0, K - 1);
      r1 = min(istr * k + io1, K - 1);
    }
  };
  auto combine = [&](int k, double lo, double hi, double& e, double& o) {
    if constexpr (IB) {
      e = k < hbi ? lo + hi : (k == hbi ? lo : 0.0);
      o = k < hbi ? lo - hi : 0.0;
    } else {
      e = lo;
      o = hi;
    }
    if (k >= nk0) e = 0.0;
    if (k >= nk1) o = 0.0;
  };
  double lo[GI_NKH][16], hi[GI_NKH][16];
#pragma unroll
  for (int kk = 0; kk < GI_NKH; ++kk)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      int r0, r1;
      rows_of(kk * 64 + lk * 16 + j + 1, r0, r1);   // sliced index kl -> k = kl + 1
      lo[kk][j] = Bc[(size_t)r0 * ld];
      hi[kk][j] = Bc[(size_t)r1 * ld];
    }
  double bf[2];
  {
    int r0, r1;
    rows_of(0, r0, r1);
    combine(0, Bc[(size_t)r0 * ld], Bc[(size_t)r1 * ld], bf[0], bf[1]);
  }
  double mx[2] = {0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < GI_NKH; ++kk)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      double e, o;
      combine(kk * 64 + lk * 16 + j + 1, lo[kk][j], hi[kk][j], e, o);
      lo[kk][j] = e;   // from here on: the half inputs
      hi[kk][j] = o;
      mx[0] = fmax(mx[0], fabs(e));
      mx[1] = fmax(mx[1], fabs(o));
    }
  double inv[2], scale_c[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    mx[h] = fmax(mx[h], __shfl_xor(mx[h], 16));
    mx[h] = fmax(mx[h], __shfl_xor(mx[h], 32));
    int eb = 0;
    if (mx[h] > 0.0) (void)frexp(mx[h], &eb);
    inv[h] = ldexp(1.0, -eb);
    scale_c[h] = ldexp(1.0, eb - 14);
  }
  // slices: bs[h][kk][q] holds bytes kl = kk*64 + lk*16 + (0..15) of slice q
  v4i bs[2][GI_NKH][S];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int kk = 0; kk < GI_NKH; ++kk)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        int pk[S];
#pragma unroll
        for (int q = 0; q < S; ++q) pk[q] = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          double x = (h == 0 ? lo[kk][g * 4 + b] : hi[kk][g * 4 + b]) * inv[h];
#pragma unroll
          for (int q = 0; q < S; ++q) {
            const double y = x * 128.0;
            const double d = fmin(fmax(rint(y), -127.0), 127.0);
            x = y - d;
            pk[q] |= ((int)d & 0xff) << (8 * b);
          }
        }
#pragma unroll
        for (int q = 0; q < S; ++q) bs[h][kk][q][g] = pk[q];
      }

  stashA(0);
  __syncthreads();

  const int nbo = J.op.sp_nbo, hbo = nbo >> 1;
  for (int t = 0; t < nti; ++t) {
    const int cur = t & 1;
    if (t + 1 < nti) fetchA(t + 1);
    double v[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      v4i acc[S];
#pragma unroll
      for (int l = 0; l < S; ++l) acc[l] = v4i{0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < GI_NKH; ++kk)
#pragma unroll
        for (int p = 0; p < S; ++p) {
          const v4i a = As[cur][((h * S + p) * GI_NKH + kk) * 64 + lane];
#pragma unroll
          for (int q = 0; q < S - p; ++q)
            acc[p + q] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, bs[h][kk][q], acc[p + q], 0, 0, 0);
        }
      // half rows t*16 + 4*lk + r, column li (v_mfma_i32_16x16x64_i8 layout)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = t * 16 + lk * 4 + r;
        double s = (double)acc[S - 1][r];
#pragma unroll
        for (int l = S - 2; l >= 0; --l) s = (double)acc[l][r] + s * 0.0078125;
        v[h][r] = fma(af_s[h][rr], bf[h], s * scale_c[h] * sa_s[h][rr]);
      }
    }
    double val[8];
    int row[8];
    bool ok[8];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = t * 16 + lk * 4 + r;
      if constexpr (OB) {
        val[r] = v[0][r] + v[1][r];
        row[r] = rr;
        ok[r] = rr <= hbo;
        val[4 + r] = v[0][r] - v[1][r];
        row[4 + r] = nbo - rr;
        ok[4 + r] = rr < hbo;
      } else {
        val[r] = v[0][r];
        row[r] = J.op.sp_ostr * rr + J.op.sp_ooff[0];
        ok[r] = rr < J.op.sp_nr[0];
        val[4 + r] = v[1][r];
        row[4 + r] = J.op.sp_ostr * rr + J.op.sp_ooff[1];
        ok[4 + r] = rr < J.op.sp_nr[1];
      }
    }
    if (J.C2) epi_rows<8>(J.e2, J.C2, val, row, ok, ld, col, mode);
    epi_rows<8>(J.e, J.C, val, row, ok, ld, col, mode);
    if (t + 1 < nti) stashA(cur ^ 1);
    __syncthreads();
  }
}

bool gemm_i8_split_batch_ok(const GemmBatch& b, int nj) {
  if (nj < 1 || nj > GEMM_MAX_JOBS || b.ld % (GI_NW * 16) != 0 || !gemm_i8_supported(b.slices))
    return false;
  const GemmOp& o = b.job[0].op;
  for (int j = 0; j < nj; ++j) {
    const GemmJob& J = b.job[j];
    const GemmOp& p = J.op;
    if (!p.sA8 || !p.ssa || !p.safirst || p.slices != b.slices || p.sp_ib != o.sp_ib ||
        p.sp_ob != o.sp_ob || p.snkh != o.snkh || p.snkh < 1 || p.snkh > 2 || p.snti < 1 ||
        p.snti > GI_MAXT || std::max(p.sp_nk[0], p.sp_nk[1]) > p.snkh * 64 + 1 || J.W ||
        !J.B || !J.C)
      return false;
  }
  return true;
}

template <int S, int NKH>
static void launch_i8s(const GemmBatch& b, int nj, hipStream_t s) {
  const GemmOp& o = b.job[0].op;
  dim3 grid(b.ld / (GI_NW * 16), nj);
  if (o.sp_ib && o.sp_ob)
    hipLaunchKernelGGL((gemm_i8_split_kernel<S, NKH, true, true>), grid, dim3(GI_NW * 64), 0, s,
                       b);
  else if (o.sp_ib)
    hipLaunchKernelGGL((gemm_i8_split_kernel<S, NKH, true, false>), grid, dim3(GI_NW * 64), 0, s,
                       b);
  else if (o.sp_ob)
    hipLaunchKernelGGL((gemm_i8_split_kernel<S, NKH, false, true>), grid, dim3(GI_NW * 64), 0, s,
                       b);
  else
    hipLaunchKernelGGL((gemm_i8_split_kernel<S, NKH, false, false>), grid, dim3(GI_NW * 64), 0,
                       s, b);
}

template <int S>
static void launch_i8s_k(const GemmBatch& b, int nj, hipStream_t s) {
  if (b.job[0].op.snkh == 1) launch_i8s<S, 1>(b, nj, s);
  else launch_i8s<S, 2>(b, nj, s);
}

hipError_t launch_gemm_i8_split(const GemmBatch& b, int nj, hipStream_t s) {
  if (!gemm_i8_split_batch_ok(b, nj)) return hipErrorInvalidValue;
  switch (b.slices) {
    case 4: launch_i8s_k<4>(b, nj, s); break;
    case 5: launch_i8s_k<5>(b, nj, s); break;
    case 6: launch_i8s_k<6>(b, nj, s); break;
    default: launch_i8s_k<7>(b, nj, s); break;
  }
  return hipGetLastError();
}

// Slices of the two halves E (sp_nr[0] x sp_nk[0]) and O of a split
// operator, read back from its packed FP64 halves (E rows, O rows, then E's
// row 16*sp_mth when sp_xrow), each sliced like gemm_i8_split: k = 0 kept in
// FP64, k >= 1 in S slices with one exponent per half row.
void gemm_i8_split_halves(const std::vector<double>& packed, const GemmOp& op, int S,
                          std::vector<int8_t>& a8, std::vector<double>& sa,
                          std::vector<double>& afirst, int& snti, int& snkh) {
  const int MP = 16 * op.sp_mth, MW = 2 * MP + 16;
  snti = (std::max(op.sp_nr[0], op.sp_nr[1]) + 15) / 16;
  snkh = std::max(1, (std::max(op.sp_nk[0], op.sp_nk[1]) - 1 + 63) / 64);
  const int GI_NKH = snkh;
  a8.assign((size_t)snti * 2 * S * GI_NKH * 64 * 16, 0);
  sa.assign((size_t)2 * snti * 16, 0.0);
  afirst.assign((size_t)2 * snti * 16, 0.0);
  for (int h = 0; h < 2; ++h)
    for (int r = 0; r < op.sp_nr[h]; ++r) {
      const size_t c = (h == 0 && r == MP) ? 2 * MP : (size_t)h * MP + r;
      auto a = [&](int k) { return packed[(size_t)k * MW + c]; };
      const int nk = op.sp_nk[h];
      double mx = 0.0;
      for (int k = 1; k < nk; ++k) mx = std::fmax(mx, std::fabs(a(k)));
      int e = 0;
      if (mx > 0.0) (void)std::frexp(mx, &e);
      sa[(size_t)h * snti * 16 + r] = std::ldexp(1.0, e);
      afirst[(size_t)h * snti * 16 + r] = a(0);
      const double inv = std::ldexp(1.0, -e);
      const int t = r / 16, li = r % 16;
      for (int k = 1; k < nk; ++k) {
        const int kl = k - 1, kk = kl / 64, lk = (kl % 64) / 16, j = kl % 16;
        const int lane = li + 16 * lk;
        double x = a(k) * inv;
        for (int p = 0; p < S; ++p) {
          const double y = x * 128.0;
          const double d = std::fmin(std::fmax(std::nearbyint(y), -127.0), 127.0);
          x = y - d;
          a8[(((((size_t)t * 2 + h) * S + p) * GI_NKH + kk) * 64 + lane) * 16 + j] = (int8_t)d;
        }
      }
    }
}
