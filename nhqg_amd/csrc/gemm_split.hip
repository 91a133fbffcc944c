// Vertical contractions split by the reflection symmetry of the Chebyshev
// grid.
//
// The CGL levels z_i = cos(pi i / N) satisfy z_{N-i} = -z_i, and
// T_k(-z) = (-1)^k T_k(z), so the vertical operators of the step are
// structured under z -> -z:
//  * coefficient -> coefficient maps (d/dZ, the projections and Galerkin
//    stencils, B0 and its eigenvectors taken per parity) are parity blocks:
//    A[i][j] = 0 unless row i and column j fall in matching halves;
//  * coefficient -> value maps satisfy A[N-i][k] = +-(-1)^k A[i][k] (a
//    butterfly on the output rows); value -> coefficient maps
//    A[k][N-i] = +-(-1)^k A[k][i] (a butterfly on the input rows); the
//    value -> value SBP cycle satisfies A[N-i][N-j] = A[i][j] (both).
// Each such product is two products of about half the rows and half the
// depth: half the flops of the dense one.  The structure holds to the
// roundoff of the operators' construction (<= 1e-13 relative); the halves
// are built from the symmetrised operator, so a split product differs from
// the dense one by that roundoff.
//
// Kernel: per wave 16 columns of B and both halves (E rows in acc0, O rows
// in acc1, MTH tiles of 16 each), A's halves staged through LDS in k-chunks
// as in the dense kernel (gemm_f64.hip), B in a register ring of raw rows
// that are combined (butterfly, substep update) when used, so the ring's
// loads stay in flight.  Chunks are 16 rows deep, 8 with the fused substep
// update (its extra ring registers would not fit 2 waves per SIMD); each
// operator's halves are padded to the chunk of the kernel it runs in.
#include "common.h"

#include <algorithm>
#include <cmath>

#ifndef SP_NW
#define SP_NW 4   // waves per workgroup (two workgroups per CU)
#endif
#ifndef SP_KC
#define SP_KC 16  // k-chunk staged in LDS per iteration (double buffered)
#endif
#ifndef SP_KCU
#define SP_KCU 8  // the same with the fused substep update
#endif
#ifndef SP_R
#define SP_R 2    // B ring depth in chunks
#endif

namespace {

constexpr int SP_MAXH = 9;   // row tiles per half: M <= 2 * 144

}  // namespace

// Epilogue of tile T and up.  v_mfma_f64_16x16x4_f64 output: D[i][j] of a
// tile sits in lane j + 16*(i%4), register i/4, so a half's row is
// rr = T*16 + r*4 + lk.  Tiles go two at a time (16 values per epi_rows, so
// each fused term costs one memory round trip per tile pair) by
// compile-time recursion: an epilogue loop too large to unroll would index
// the accumulators at run time and move them to scratch memory.
template <int T, int MTH, bool OB>
NHQG_DEV void split_epilogue(const GemmJob& J, const v4d (&acc0)[MTH], const v4d (&acc1)[MTH],
                             int ld, int col, int lk) {
  if constexpr (T < MTH) {
    constexpr int NT = (T + 1 < MTH) ? 2 : 1;   // tiles in this call
    constexpr int NV = 8 * NT;
    const int mode = col >> 1;
    const int nbo = J.op.sp_nbo, hbo = nbo >> 1;
    double val[NV];
    int row[NV];
    bool ok[NV];
#pragma unroll
    for (int u = 0; u < NT; ++u) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = (T + u) * 16 + r * 4 + lk;
        const double E = acc0[T + u][r], O = acc1[T + u][r];
        const int a = 8 * u + r, b = 8 * u + 4 + r;
        if constexpr (OB) {
          val[a] = E + O;
          row[a] = rr;
          ok[a] = rr <= hbo;
          val[b] = E - O;
          row[b] = nbo - rr;
          ok[b] = rr < hbo;
        } else {
          val[a] = E;
          row[a] = J.op.sp_ostr * rr + J.op.sp_ooff[0];
          ok[a] = rr < J.op.sp_nr[0];
          val[b] = O;
          row[b] = J.op.sp_ostr * rr + J.op.sp_ooff[1];
          ok[b] = rr < J.op.sp_nr[1];
        }
      }
    }
    if (J.C2) epi_rows<NV>(J.e2, J.C2, val, row, ok, ld, col, mode);
    epi_rows<NV>(J.e, J.C, val, row, ok, ld, col, mode);
    split_epilogue<T + NT, MTH, OB>(J, acc0, acc1, ld, col, lk);
  }
}

template <int MTH, bool IB, bool OB, bool UPD>
__global__ void __launch_bounds__(SP_NW * 64) __attribute__((amdgpu_waves_per_eu(2)))
gemm_split_kernel(GemmBatch bt) {
  constexpr int NT = SP_NW * 64;
  constexpr int KC = UPD ? SP_KCU : SP_KC;
  constexpr int R = SP_R;
  constexpr int KS = KC / 4;
  constexpr int MP = MTH * 16;       // rows per half
  constexpr int MW = 2 * MP + 16;    // one packed row of A: E, O, E's extra row (+ pad)
  constexpr int SA = MW;             // LDS pitch, an odd multiple of 128 bytes: the 4 k
                                     // rows of a read fall on distinct banks
  constexpr int NA = (MW * KC + NT - 1) / NT;
  constexpr bool FULL = (NA * NT == MW * KC);
  __shared__ double As[2][KC][SA];

  const GemmJob& J = bt.job[blockIdx.y];
  const double* __restrict__ At = J.op.sp_At;
  const int K = J.op.K;
  const int nchunks = J.op.sp_kp / KC;
  const double* __restrict__ B = J.B;
  const int ld = bt.ld;
  const int nbi = J.op.sp_nbi, hbi = nbi >> 1;
  const int istr = J.op.sp_istr, io0 = J.op.sp_ioff[0], io1 = J.op.sp_ioff[1];

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int col = blockIdx.x * (SP_NW * 16) + wave * 16 + li;

  v4d acc0[MTH], acc1[MTH];
#pragma unroll
  for (int t = 0; t < MTH; ++t) {
    acc0[t] = v4d{0.0, 0.0, 0.0, 0.0};
    acc1[t] = v4d{0.0, 0.0, 0.0, 0.0};
  }
  double ra[NA];
  double xr = 0.0;   // E's row MP (sp_xrow): partial dot over this lane's k
  // raw B rows of the ring: (lo, hi) = (B[k], B[N-k]) for an input
  // butterfly, else the rows of half 0 and half 1
  double lo[R][KS], hi[R][KS];
  double wl[UPD ? R : 1][KS], wh[UPD ? R : 1][KS], gl[UPD ? R : 1][KS], gh[UPD ? R : 1][KS];

  auto fetchA = [&](int ch) {
    const double* src = At + (size_t)ch * KC * MW;
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      int idx = tid + u * NT;   // idx = kc * MW + r
      if (!FULL) idx = min(idx, MW * KC - 1);
      ra[u] = src[idx];
    }
  };
  auto stashA = [&](int buf) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int idx = tid + u * NT;
      if (FULL || idx < MW * KC) As[buf][idx / MW][idx % MW] = ra[u];
    }
  };
  auto fetchB = [&](int slot, int ch) {
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int k = ch * KC + 4 * s + lk;
      int r0, r1;
      if constexpr (IB) {
        r0 = min(k, hbi);
        r1 = nbi - r0;
      } else {
        r0 = min(istr * k + io0, K - 1);   // rows past the half meet zero columns of A
        r1 = min(istr * k + io1, K - 1);
      }
      lo[slot][s] = B[(size_t)r0 * ld + col];
      hi[slot][s] = B[(size_t)r1 * ld + col];
      if constexpr (UPD) {
        wl[slot][s] = J.W[(size_t)r0 * ld + col];
        wh[slot][s] = J.W[(size_t)r1 * ld + col];
        gl[slot][s] = J.g[r0];
        gh[slot][s] = J.g[r1];
      }
    }
  };
  // the two half inputs of k-row k from ring slot j
  auto inputs = [&](int j, int s, int k, double& ve, double& vo) {
    if constexpr (IB) {
      const int r0 = min(k, hbi), r1 = nbi - r0;
      double a = lo[j][s], b = hi[j][s];
      if constexpr (UPD) {
        // SBP substep update of the input: th - c * w * g, boundary rows zero
        a = (r0 == 0 || r0 >= K - 1) ? 0.0 : a - J.bc * wl[j][s] * gl[j][s];
        b = (r1 == 0 || r1 >= K - 1) ? 0.0 : b - J.bc * wh[j][s] * gh[j][s];
      }
      ve = k < hbi ? a + b : (k == hbi ? a : 0.0);
      vo = k < hbi ? a - b : 0.0;
    } else {
      ve = lo[j][s];
      vo = hi[j][s];
    }
  };

  fetchA(0);
  stashA(0);
#pragma unroll
  for (int p = 0; p < R - 1; ++p) fetchB(p, min(p, nchunks - 1));
  __syncthreads();

  for (int ch0 = 0; ch0 < nchunks; ch0 += R) {
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int ch = ch0 + j;
      if (ch < nchunks) {
        const int cur = ch & 1;
        fetchA(min(ch + 1, nchunks - 1));
        fetchB((j + R - 1) % R, min(ch + R - 1, nchunks - 1));
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          const int kr = s * 4 + lk;
          double ve, vo;
          inputs(j, s, ch * KC + kr, ve, vo);
#pragma unroll
          for (int t = 0; t < MTH; ++t) {
            const double a = As[cur][kr][t * 16 + li];
            acc0[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, ve, acc0[t], 0, 0, 0);
          }
#pragma unroll
          for (int t = 0; t < MTH; ++t) {
            const double a = As[cur][kr][MP + t * 16 + li];
            acc1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, vo, acc1[t], 0, 0, 0);
          }
          xr = fma(As[cur][kr][2 * MP], ve, xr);
        }
        stashA(cur ^ 1);
        __syncthreads();
      }
    }
  }

  split_epilogue<0, MTH, OB>(J, acc0, acc1, ld, col, lk);
  if (J.op.sp_xrow) {
    // E's row MP: the four k-residue partial sums sit in lanes li + 16*lk.
    // With an output butterfly it is the middle level (O's is zero there).
    xr += __shfl_xor(xr, 16);
    xr += __shfl_xor(xr, 32);
    if (lk == 0) {
      const double v[1] = {xr};
      const int row[1] = {OB ? MP : J.op.sp_ostr * MP + J.op.sp_ooff[0]};
      const bool ok[1] = {true};
      if (J.C2) epi_rows<1>(J.e2, J.C2, v, row, ok, ld, col, col >> 1);
      epi_rows<1>(J.e, J.C, v, row, ok, ld, col, col >> 1);
    }
  }
}

// ---------------------------------------------------------------------------
// host side: structure detection and packing
// ---------------------------------------------------------------------------

namespace {

// Rows (or columns) of a half: index str*r + off[h] for r < n[h].
struct Side {
  int str, off[2], n[2];
};

Side side_parity(int L, int p) {   // half 0: indices of parity p
  return Side{2, {p, 1 - p}, {(L - p + 1) / 2, (L - (1 - p) + 1) / 2}};
}

Side side_contig(int L) {          // half 0: the first ceil(L/2) indices
  const int h = (L + 1) / 2;
  return Side{1, {0, h}, {h, L - h}};
}

int half_of(const Side& s, int i) {
  return s.str == 2 ? ((i & 1) == s.off[0] ? 0 : 1) : (i < s.off[1] ? 0 : 1);
}

}  // namespace

int gemm_split_kc(bool upd) { return upd ? SP_KCU : SP_KC; }

bool gemm_split_plan(const double* A, int M, int K, bool prefer_cols, double tol, int min_mth,
                     int kc, GemmOp& op, std::vector<double>& packed) {
  if (kc < 1) return false;
  if (M < 8 || K < 8 || M > 2 * 16 * SP_MAXH) return false;
  double mx = 0.0;
  for (size_t i = 0; i < (size_t)M * K; ++i) mx = std::max(mx, std::fabs(A[i]));
  if (mx == 0.0) return false;
  const double th = tol * mx;
  auto a = [&](int i, int j) { return A[(size_t)i * K + j]; };

  // the plan: E/O matrices as functions of (half row r, half k)
  int ib = 0, ob = 0, nr[2], nk[2];
  Side rs{}, cs{};
  int sigma = 0;
  int kind = 0;   // 1 blocks, 2 output butterfly, 3 input butterfly, 4 both

  // parity / contiguous blocks
  {
    const Side rows[2] = {side_parity(M, 0), side_contig(M)};
    const Side cols[3] = {side_parity(K, 0), side_parity(K, 1), side_contig(K)};
    for (int ri = 0; ri < 2 && !kind; ++ri)
      for (int ci = 0; ci < 3 && !kind; ++ci) {
        bool ok = true;
        for (int i = 0; i < M && ok; ++i) {
          const int hr = half_of(rows[ri], i);
          for (int j = 0; j < K; ++j)
            if (half_of(cols[ci], j) != hr && std::fabs(a(i, j)) > th) {
              ok = false;
              break;
            }
        }
        if (ok) {
          kind = 1;
          rs = rows[ri];
          cs = cols[ci];
        }
      }
  }
  // A[N-i][k] = s_k A[i][k], s_k = (-1)^(k + sigma)
  auto rowsym = [&](int sg) {
    if (!(M & 1)) return false;
    const int N = M - 1;
    for (int i = 0; i < M; ++i)
      for (int k = 0; k < K; ++k) {
        const double s = ((k + sg) & 1) ? -1.0 : 1.0;
        if (std::fabs(a(N - i, k) - s * a(i, k)) > th) return false;
      }
    return true;
  };
  // A[k][N-i] = s_k A[k][i]
  auto colsym = [&](int sg) {
    if (!(K & 1)) return false;
    const int N = K - 1;
    for (int k = 0; k < M; ++k) {
      const double s = ((k + sg) & 1) ? -1.0 : 1.0;
      for (int i = 0; i < K; ++i)
        if (std::fabs(a(k, N - i) - s * a(k, i)) > th) return false;
    }
    return true;
  };
  auto bothsym = [&]() {
    if (M != K || !(M & 1)) return false;
    const int N = M - 1;
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < K; ++j)
        if (std::fabs(a(N - i, N - j) - a(i, j)) > th) return false;
    return true;
  };
  for (int pass = 0; pass < 2 && !kind; ++pass) {
    const bool cols_now = (pass == 0) == prefer_cols;
    for (int sg = 0; sg < 2 && !kind; ++sg) {
      if (cols_now ? colsym(sg) : rowsym(sg)) {
        kind = cols_now ? 3 : 2;
        sigma = sg;
      }
    }
  }
  if (!kind && bothsym()) kind = 4;
  if (!kind) return false;

  // half sizes
  if (kind == 1) {
    for (int h = 0; h < 2; ++h) {
      nr[h] = rs.n[h];
      nk[h] = cs.n[h];
    }
  } else {
    ob = (kind == 2 || kind == 4);
    ib = (kind == 3 || kind == 4);
    if (ob) {
      // O's middle row is zero: the two butterfly outputs of row N/2 coincide
      nr[0] = (M - 1) / 2 + 1;
      nr[1] = (M - 1) / 2;
    } else {
      rs = side_parity(M, sigma);   // E: rows whose column reflection sign is +1
      nr[0] = rs.n[0];
      nr[1] = rs.n[1];
    }
    if (ib) {
      nk[0] = (K - 1) / 2 + 1;
      nk[1] = (K - 1) / 2;
    } else {
      cs = side_parity(K, sigma);   // E: columns whose row reflection sign is +1
      nk[0] = cs.n[0];
      nk[1] = cs.n[1];
    }
  }
  if (std::min(nr[0], nr[1]) < 1 || std::min(nk[0], nk[1]) < 1) return false;
  // E's last row may go to the VALU (the dense kernel's 257th row): with
  // nr[0] = 16*mth + 1 the halves then take mth tiles, not mth + 1.  Under an
  // output butterfly that row must be the middle level.
  int mth = std::max((std::max(nr[0], nr[1]) + 15) / 16, min_mth);
  const int mx_ = std::max((std::max(nr[0] - 1, nr[1]) + 15) / 16, min_mth);
  int xrow = 0;
  if (nr[0] == 16 * mx_ + 1 && (!ob || 16 * mx_ == (M - 1) / 2)) {
    mth = mx_;
    xrow = 1;
  }
  if (mth > SP_MAXH) return false;
  const int MP = 16 * mth, MW = 2 * MP + 16;
  const int kp = (std::max(nk[0], nk[1]) + kc - 1) / kc * kc;

  // E/O entries
  const int Nr = M - 1, Nc = K - 1;
  auto abar = [&](int r, int j) {   // operator symmetrised under both reflections
    return kind == 4 ? 0.5 * (a(r, j) + a(Nr - r, Nc - j)) : a(r, j);
  };
  auto entry = [&](int h, int r, int k) -> double {
    if (kind == 1) return a(rs.str * r + rs.off[h], cs.str * k + cs.off[h]);
    if (kind == 2) {
      // output butterfly; half 0 takes the columns with A[N-r][j] = +A[r][j]
      const int j = cs.str * k + cs.off[h];
      return 0.5 * (h == 0 ? a(r, j) + a(Nr - r, j) : a(r, j) - a(Nr - r, j));
    }
    // input butterfly (kinds 3 and 4): column k pairs with Nc - k
    const int i = (kind == 3) ? rs.str * r + rs.off[h] : r;
    if (h == 0 && k == Nc / 2) return abar(i, k);
    return 0.5 * (h == 0 ? abar(i, k) + abar(i, Nc - k) : abar(i, k) - abar(i, Nc - k));
  };
  packed.assign((size_t)kp * MW, 0.0);
  for (int h = 0; h < 2; ++h)
    for (int r = 0; r < nr[h]; ++r)
      for (int k = 0; k < nk[h]; ++k) {
        // E's row MP (xrow) sits after the O rows
        const size_t c = (h == 0 && r == MP) ? 2 * MP : (size_t)h * MP + r;
        packed[(size_t)k * MW + c] = entry(h, r, k);
      }

  op.sp_mth = mth;
  op.sp_xrow = xrow;
  op.sp_kp = kp;
  op.sp_kc = kc;
  op.sp_ib = ib;
  op.sp_ob = ob;
  op.sp_istr = ib ? 0 : cs.str;
  op.sp_ioff[0] = ib ? 0 : cs.off[0];
  op.sp_ioff[1] = ib ? 0 : cs.off[1];
  op.sp_ostr = ob ? 0 : rs.str;
  op.sp_ooff[0] = ob ? 0 : rs.off[0];
  op.sp_ooff[1] = ob ? 0 : rs.off[1];
  op.sp_nr[0] = nr[0];
  op.sp_nr[1] = nr[1];
  op.sp_nk[0] = nk[0];
  op.sp_nk[1] = nk[1];
  op.sp_nbi = ib ? K - 1 : 0;
  op.sp_nbo = ob ? M - 1 : 0;
  return true;
}

bool gemm_split_batch_ok(const GemmBatch& b, int nj) {
  if (nj < 1 || nj > GEMM_MAX_JOBS || b.ld % (SP_NW * 16) != 0) return false;
  const GemmOp& o = b.job[0].op;
  const bool upd = b.job[0].W != nullptr;
  for (int j = 0; j < nj; ++j) {
    const GemmOp& p = b.job[j].op;
    if (!p.sp_At || p.sp_mth != o.sp_mth || p.sp_ib != o.sp_ib || p.sp_ob != o.sp_ob ||
        (b.job[j].W != nullptr) != upd || !b.job[j].B || !b.job[j].C)
      return false;
    if (upd && (!p.sp_ib || !b.job[j].g)) return false;
    if (p.sp_kp % gemm_split_kc(upd) != 0) return false;   // padded for the other kernel
  }
  return o.sp_mth >= 1 && o.sp_mth <= SP_MAXH;
}

template <int MTH, bool IB, bool OB, bool UPD>
static hipError_t launch_sp(const GemmBatch& b, int nj, hipStream_t s) {
  dim3 grid(b.ld / (SP_NW * 16), nj);
  hipLaunchKernelGGL((gemm_split_kernel<MTH, IB, OB, UPD>), grid, dim3(SP_NW * 64), 0, s, b);
  return hipGetLastError();
}

template <int MTH>
static hipError_t launch_sp_mth(const GemmBatch& b, int nj, bool ib, bool ob, bool upd,
                                hipStream_t s) {
  if (!ib) {
    if (upd) return hipErrorInvalidValue;
    return ob ? launch_sp<MTH, false, true, false>(b, nj, s)
              : launch_sp<MTH, false, false, false>(b, nj, s);
  }
  if (ob)
    return upd ? launch_sp<MTH, true, true, true>(b, nj, s)
               : launch_sp<MTH, true, true, false>(b, nj, s);
  return upd ? launch_sp<MTH, true, false, true>(b, nj, s)
             : launch_sp<MTH, true, false, false>(b, nj, s);
}

hipError_t launch_gemm_split(const GemmBatch& b, int nj, hipStream_t s) {
  if (!gemm_split_batch_ok(b, nj)) return hipErrorInvalidValue;
  const GemmOp& o = b.job[0].op;
  const bool ib = o.sp_ib, ob = o.sp_ob, upd = b.job[0].W != nullptr;
  switch (o.sp_mth) {
#define NHQG_SP(n) case n: return launch_sp_mth<n>(b, nj, ib, ob, upd, s);
    NHQG_SP(1) NHQG_SP(2) NHQG_SP(3) NHQG_SP(4) NHQG_SP(5) NHQG_SP(6) NHQG_SP(7) NHQG_SP(8)
    NHQG_SP(9)
#undef NHQG_SP
    default: return hipErrorInvalidValue;
  }
}
