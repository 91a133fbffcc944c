// FP64 vertical contractions on the CDNA matrix cores.
//
// Every vertical operator of the step (Chebyshev <-> CGL transforms,
// Galerkin stencils and left inverses, tau projections, d/dZ, SBP transfers)
// is a small real matrix A (at most 257 x 257) applied to all modes of a
// level-major field: C = A * B with B a (K x 2*nm_pad) real matrix.  A
// launch runs up to three independent contractions of the same row-tile
// count (blockIdx.y picks the job), so one launch carries enough work to
// keep every CU busy while other workgroups are in their prologue or
// writing results.
//
// Per wave: 16 columns of B and the full column height of C -- MT row
// tiles of 16 on v_mfma_f64_16x16x4_f64 plus, for 257-row operators
// (MT = 16, M = 257), the last row as a VALU dot product -- so each B
// value is read from HBM once and each A fragment (from LDS) feeds one MFMA.  A is stored k-major and zero
// padded (GemmOp), so chunk loads are coalesced, the LDS stash is
// conflict-free and no load needs a bounds branch: B rows beyond K are
// clamped to row K-1 and meet zero columns of A.  B is prefetched R chunks
// ahead in a statically indexed register ring, and the A chunk for the next
// iteration is loaded before the B ring entry, so the in-order vmcnt wait
// for the A stash leaves the ring in flight.
#include "common.h"

#ifndef GK_KC
#define GK_KC 8   // k-chunk staged in LDS per iteration (double buffered)
#endif
#ifndef GK_NW
#define GK_NW 4   // waves per workgroup; 16 columns each
#endif
#ifndef GK_R
#define GK_R 2    // B ring depth in chunks
#endif

static_assert(GK_KC % 4 == 0, "k-chunk must be a multiple of the MFMA depth");

// Epilogue of tiles [t0, t0 + G) of one lane (16 columns per wave, rows
// t*16 + r*4 + lk): the same per-element arithmetic as epi_apply, but each
// extra field is read for all 4G elements at once, so a term costs one
// memory latency rather than one per element.  Rows beyond M are clamped
// for the loads and not stored.
template <int MT, int G>
NHQG_DEV void epi_store(const Epi& e, double* C, const v4d (&acc)[MT], int t0, int M, int ld,
                        int col, int lk, int mode) {
  double o[G][4];
  const double sa = e.s_acc ? e.s_acc[mode] : 1.0;
#pragma unroll
  for (int t = 0; t < G; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      o[t][r] = e.a_acc * (t0 + t < MT ? acc[t0 + t < MT ? t0 + t : 0][r] : 0.0);
      if (e.s_acc) o[t][r] *= sa;
    }
  for (int x = 0; x < e.nX; ++x) {
    const double* X = e.X[x];
    const double ax = e.aX[x];
    const bool hs = e.sX[x] != nullptr;
    const double sx = hs ? e.sX[x][mode] : 1.0;
    double v[G][4];
#pragma unroll
    for (int t = 0; t < G; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = min((t0 + t) * 16 + r * 4 + lk, M - 1);
        v[t][r] = X[(size_t)row * ld + col];
      }
#pragma unroll
    for (int t = 0; t < G; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double w = ax * v[t][r];
        if (hs) w *= sx;
        o[t][r] += w;
      }
  }
  const double so = e.s_out ? e.s_out[mode] : 1.0;
#pragma unroll
  for (int t = 0; t < G; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = (t0 + t) * 16 + r * 4 + lk;
      if (t0 + t < MT && row < M) {
        double w = o[t][r];
        if (e.s_out) w *= so;
        if (e.d_tab) w *= e.d_tab[(size_t)row * e.d_ld + mode];
        C[(size_t)row * ld + col] = w;
      }
    }
}

// Operators of at most 96 rows (Nz <= 95) have short k loops, so a launch is
// dominated by load latency and the epilogue: their kernels keep one row
// tile per epilogue group and ask for 4-5 waves per SIMD (no spills at
// these bounds), where the 4-tile epilogue needs ~160 VGPRs (3 waves).  On
// the open-top 512^2 x 64 step the class went from 2.28 to 1.90 ms with
// this and 4-wave workgroups (3 -> 5 resident per CU).
template <int MT, bool UPD>
constexpr int gk_min_waves() { return MT <= 5 ? (UPD ? 4 : 5) : MT == 6 ? 4 : 1; }
template <int MT>
constexpr int gk_epi_group() { return MT <= 6 ? 1 : 4; }

template <int MT, bool UPD>
__global__ void __launch_bounds__(GK_NW * 64)
__attribute__((amdgpu_waves_per_eu(gk_min_waves<MT, UPD>(), 8)))
gemm_f64_kernel(GemmBatch bt) {
  constexpr int NT = GK_NW * 64;
  constexpr int KC = GK_KC;
  constexpr int R = GK_R;
  constexpr int MP = MT * 16;
  constexpr int SA = MP + ((MT % 2 == 0) ? 16 : 0);   // row pitch: lk rows land on disjoint banks
  constexpr int NA = (MP * KC + NT - 1) / NT;
  constexpr bool FULL = (NA * NT == MP * KC);
  constexpr int KS = KC / 4;
  // the VALU last row exists only for 257-row operators (MT = 16); below
  // that the launcher rejects M > 16 * MT
  constexpr bool LAST = MT == 16;
  __shared__ double As[2][KC][SA];

  const GemmJob& J = bt.job[blockIdx.y];
  const double* __restrict__ At = J.op.At;
  const int lda = J.op.lda;
  const int M = J.op.M, K = J.op.K;
  const int nchunks = J.op.kp / KC;
  const double* __restrict__ B = J.B;
  const int ld = bt.ld;

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int col = blockIdx.x * (GK_NW * 16) + wave * 16 + li;

  v4d acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  double last = 0.0;
  double ra[NA];
  double bq[R][KS], lq[LAST ? R : 1][KS];
  double wq[UPD ? R : 1][KS], gq[UPD ? R : 1][KS];

  auto fetchA = [&](int ch) {
    const double* src = At + (size_t)ch * KC * lda;
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      int idx = tid + u * NT;                // idx = kc * MP + r
      if (!FULL) idx = min(idx, MP * KC - 1);
      ra[u] = src[(size_t)(idx / MP) * lda + idx % MP];
    }
  };
  auto stashA = [&](int buf) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int idx = tid + u * NT;
      if (FULL || idx < MP * KC) As[buf][idx / MP][idx % MP] = ra[u];
    }
  };
  auto fetchB = [&](int slot, int ch) {
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int kk = ch * KC + 4 * s + lk;
      const int kb = min(kk, K - 1);
      bq[slot][s] = B[(size_t)kb * ld + col];
      if constexpr (LAST) lq[slot][s] = At[(size_t)kk * lda + MP];
      if constexpr (UPD) {
        wq[slot][s] = J.W[(size_t)kb * ld + col];
        gq[slot][s] = J.g[kb];
      }
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
          double bv = bq[j][s];
          if constexpr (UPD) {
            // SBP substep update of the input: th - c * w * g, boundary rows zero
            const int kk = ch * KC + kr;
            bv = (kk == 0 || kk >= K - 1) ? 0.0 : bv - J.bc * wq[j][s] * gq[j][s];
          }
#pragma unroll
          for (int t = 0; t < MT; ++t) {
            const double a = As[cur][kr][t * 16 + li];
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[t], 0, 0, 0);
          }
          if constexpr (LAST) last = fma(lq[j][s], bv, last);
        }
        stashA(cur ^ 1);
        __syncthreads();
      }
    }
  }

  // v_mfma_f64_16x16x4_f64 output: D[i][j] of a tile sits in lane
  // j + 16*(i%4), register i/4 (one row group per register).
  const int mode = col >> 1;
  constexpr int G = gk_epi_group<MT>();
  if (J.C2) {
#pragma unroll
    for (int t0 = 0; t0 < MT; t0 += G) epi_store<MT, G>(J.e2, J.C2, acc, t0, M, ld, col, lk, mode);
  }
#pragma unroll
  for (int t0 = 0; t0 < MT; t0 += G) epi_store<MT, G>(J.e, J.C, acc, t0, M, ld, col, lk, mode);
  if (LAST && M > MP) {
    // the four k-residue partial sums sit in lanes li, li+16, li+32, li+48
    last += __shfl_xor(last, 16);
    last += __shfl_xor(last, 32);
    if (lk == 0) {
      const size_t idx = (size_t)MP * ld + col;
      J.C[idx] = epi_apply(J.e, last, idx, MP, mode);
      if (J.C2) J.C2[idx] = epi_apply(J.e2, last, idx, MP, mode);
    }
  }
}

int gemm_row_tiles(int M) { return (M > 256 ? 256 : M) / 16 + ((M > 256 ? 256 : M) % 16 != 0); }

int gemm_op_lda(int mt) { return mt * 16 + 8; }

int gemm_op_kp(int K) { return (K + GK_KC - 1) / GK_KC * GK_KC; }

void gemm_pack_op(const double* A, int M, int K, int mt, double* out) {
  const int lda = gemm_op_lda(mt), kp = gemm_op_kp(K), mp = mt * 16;
  for (size_t i = 0; i < (size_t)lda * kp; ++i) out[i] = 0.0;
  for (int r = 0; r < M && r <= mp; ++r)
    for (int k = 0; k < K; ++k) out[(size_t)k * lda + r] = A[(size_t)r * K + k];
}

template <int MT>
static hipError_t launch_mt(const GemmBatch& b, int nj, bool upd, hipStream_t s) {
  dim3 grid(b.ld / (GK_NW * 16), nj);
  if (upd)
    hipLaunchKernelGGL((gemm_f64_kernel<MT, true>), grid, dim3(GK_NW * 64), 0, s, b);
  else
    hipLaunchKernelGGL((gemm_f64_kernel<MT, false>), grid, dim3(GK_NW * 64), 0, s, b);
  return hipGetLastError();
}

hipError_t launch_gemm_batch(const GemmBatch& b, int nj, hipStream_t s) {
  if (nj < 1 || nj > GEMM_MAX_JOBS || b.ld % (GK_NW * 16) != 0) return hipErrorInvalidValue;
  const int mt = b.job[0].op.mt;
  bool upd = false;
  for (int j = 0; j < nj; ++j) {
    const GemmOp& op = b.job[j].op;
    if (op.mt != mt || op.M < 1 || op.M > mt * 16 + (mt == 16) || op.K < 1 || op.kp % GK_KC != 0 ||
        op.kp < op.K || op.lda != gemm_op_lda(mt) || !op.At || !b.job[j].B || !b.job[j].C)
      return hipErrorInvalidValue;
    const bool u = b.job[j].W != nullptr;
    if (j > 0 && u != upd) return hipErrorInvalidValue;
    if (u && !b.job[j].g) return hipErrorInvalidValue;
    upd = u;
  }
  switch (mt) {
#define NHQG_MT(n) case n: return launch_mt<n>(b, nj, upd, s);
    NHQG_MT(1) NHQG_MT(2) NHQG_MT(3) NHQG_MT(4) NHQG_MT(5) NHQG_MT(6) NHQG_MT(7) NHQG_MT(8)
    NHQG_MT(9) NHQG_MT(10) NHQG_MT(11) NHQG_MT(12) NHQG_MT(13) NHQG_MT(14) NHQG_MT(15)
    NHQG_MT(16)
#undef NHQG_MT
    default: return hipErrorInvalidValue;
  }
}

__global__ void __launch_bounds__(256)
combine_kernel(size_t n, int ld, double* out, Epi e) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= n) return;
  const int mode = (int)((idx % (size_t)ld) >> 1);
  out[idx] = epi_apply(e, 0.0, idx, (int)(idx / (size_t)ld), mode);
}

hipError_t launch_combine(int nl, int ld, double* out, const Epi& e, hipStream_t s) {
  const size_t n = (size_t)nl * ld;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(combine_kernel, dim3(blocks), dim3(256), 0, s, n, ld, out, e);
  return hipGetLastError();
}
