// Per-mode and per-level kernels of the NHQG step: the batched |k|^2-shell
// w solve, deterministic Parseval horizontal means, the SBP2 thermal
// corrector's serial 1-D pieces, post-step sanitation and the layout
// conversions at the host boundary.
#include "common.h"

// ---------------------------------------------------------------------------
// Shell solve (nhqg/solver.py:648-675, 720-736): for every active |k|^2
// shell s, w = inv_s * rhs for all modes of the shell, fused with the theta
// back-substitution theta = (R_th + alpha * w) * inv_alpha_th.
//
// One workgroup per shell, 8 modes (16 real columns = one MFMA tile width)
// per pass: W (np x 16) = inv_s (np x np) * X (np x 16) on
// v_mfma_f64_16x16x4_f64, wave w owning row tiles 4w .. 4w+3 of the
// 64*TW-padded order np.  The shell matrices (packed by shell_pack_inv:
// inv_s[j][i] at s*np*np + j*np + i, zero padded) stream from HBM straight
// into the A operand through a register ring SH_PD k-steps deep; X is
// gathered into LDS once per pass.  The kernel is bound by streaming the
// matrices once per call (~np^2 * 8 bytes per shell).
// ---------------------------------------------------------------------------
#define SH_CH 8
#define SH_PD 8

template <int TW>
__global__ void __launch_bounds__(256)
shell_solve_kernel(int Ng, int ld, const int* __restrict__ shell_ptr,
                   const int* __restrict__ shell_modes, const double* __restrict__ invP,
                   const double* __restrict__ rhs, double* __restrict__ w_out,
                   const double* __restrict__ Rth, double* __restrict__ th_out,
                   const double* __restrict__ iath, double alpha) {
  constexpr int NP = 64 * TW;       // padded order; 4 waves x TW row tiles of 16
  constexpr int KSTEPS = NP / 4;
  static_assert(KSTEPS % SH_PD == 0, "prefetch depth must divide the k-steps");
  __shared__ double xs[NP][16];
  __shared__ int ms[SH_CH];
  const int sh = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int p0 = shell_ptr[sh], p1 = shell_ptr[sh + 1];
  // A fragment of row tile t at k-step k: inv[t*16 + li][4k + lk]
  const double* Ms = invP + (size_t)sh * NP * NP + (size_t)lk * NP + wave * TW * 16 + li;
  for (int c0 = p0; c0 < p1; c0 += SH_CH) {
    const int nc = min(SH_CH, p1 - c0);
    if (tid < SH_CH) ms[tid] = tid < nc ? shell_modes[c0 + tid] : -1;
    __syncthreads();
    for (int idx = tid; idx < NP * 16; idx += 256) {
      const int j = idx >> 4, c = idx & 15;
      const int m = ms[c >> 1];
      xs[j][c] = (m >= 0 && j < Ng) ? rhs[(size_t)j * ld + 2 * m + (c & 1)] : 0.0;
    }
    __syncthreads();
    v4d acc[TW];
#pragma unroll
    for (int t = 0; t < TW; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    double a[SH_PD][TW];
#pragma unroll
    for (int p = 0; p < SH_PD; ++p)
#pragma unroll
      for (int t = 0; t < TW; ++t) a[p][t] = Ms[(size_t)4 * p * NP + t * 16];
#pragma unroll 1
    for (int k0 = 0; k0 < KSTEPS; k0 += SH_PD) {
#pragma unroll
      for (int p = 0; p < SH_PD; ++p) {
        const int k = k0 + p;
        const double b = xs[4 * k + lk][li];
#pragma unroll
        for (int t = 0; t < TW; ++t)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[p][t], b, acc[t], 0, 0, 0);
        const int kn = min(k + SH_PD, KSTEPS - 1);
#pragma unroll
        for (int t = 0; t < TW; ++t) a[p][t] = Ms[(size_t)4 * kn * NP + t * 16];
        // compiler-only fence: keeps each refill SH_PD k-steps ahead of its
        // use (otherwise the loads are re-formed at the top of the consuming
        // iteration and their latency is exposed)
        asm volatile("" ::: "memory");
      }
    }
    // D[i][n] of a tile sits in lane n + 16*(i%4), register i/4
    const int m = ms[li >> 1];
    if (m >= 0) {
      const double ia = iath[m];
#pragma unroll
      for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = (wave * TW + t) * 16 + r * 4 + lk;
          if (row < Ng) {
            const size_t idx = (size_t)row * ld + 2 * m + (li & 1);
            const double w = acc[t][r];
            w_out[idx] = w;
            th_out[idx] = (Rth[idx] + alpha * w) * ia;
          }
        }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Shell factors on the device (nhqg/grid.py:349-432, q_boundary='none'):
// one workgroup per shell writes M = A_s^T = (alpha I - beta K - c B0)^T
// into its padded slot (zeros outside Ng x Ng) and inverts it in place by
// Gauss-Jordan elimination with partial pivoting in FP64.  Row swaps of M
// are undone as column swaps at the end, so the slot ends up holding
// (A_s^-1)^T, the layout shell_solve_kernel reads.  B0T and KT are the
// transposed matrices (coalesced reads); a zero pivot sets *status.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024)
shell_build_kernel(int Ng, int np, const double* __restrict__ B0T, const double* __restrict__ KT,
                   const double* __restrict__ coef, double* __restrict__ invP,
                   int* __restrict__ status) {
  __shared__ double rowk[256], colk[256];
  __shared__ int piv[256];
  __shared__ double wv[16];
  __shared__ int wi[16];
  const int sh = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  double* M = invP + (size_t)sh * np * np;
  const double al = coef[3 * sh], be = coef[3 * sh + 1], cc = coef[3 * sh + 2];
  // M[j][i] = A[i][j] = alpha d_ij - beta K[i][j] - c B0[i][j]
  for (int j = wave; j < np; j += 16)
    for (int i = lane; i < np; i += 64) {
      double v = 0.0;
      if (i < Ng && j < Ng) {
        double a = (i == j) ? al : 0.0;
        if (KT) a -= be * KT[(size_t)j * Ng + i];
        v = a - cc * B0T[(size_t)j * Ng + i];
      }
      M[(size_t)j * np + i] = v;
    }
  __syncthreads();
  for (int k = 0; k < Ng; ++k) {
    // pivot: largest |M[r][k]|, r >= k (lowest index on ties)
    double best = -1.0;
    int bi = k;
    for (int r = k + tid; r < Ng; r += 1024) {
      const double v = fabs(M[(size_t)r * np + k]);
      if (v > best) {
        best = v;
        bi = r;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off);
      const int oi = __shfl_xor(bi, off);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if (lane == 0) {
      wv[wave] = best;
      wi[wave] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double b = wv[0];
      int i0 = wi[0];
      for (int w = 1; w < 16; ++w)
        if (wv[w] > b || (wv[w] == b && wi[w] < i0)) {
          b = wv[w];
          i0 = wi[w];
        }
      piv[k] = i0;
      if (!(b > 0.0)) *status = 1;
    }
    __syncthreads();
    const int p = piv[k];
    if (p != k)
      for (int c = tid; c < Ng; c += 1024) {
        const double t = M[(size_t)k * np + c];
        M[(size_t)k * np + c] = M[(size_t)p * np + c];
        M[(size_t)p * np + c] = t;
      }
    __syncthreads();
    const double pv = M[(size_t)k * np + k];
    for (int c = tid; c < Ng; c += 1024)
      rowk[c] = (c == k) ? 1.0 / pv : M[(size_t)k * np + c] / pv;
    for (int r = tid; r < Ng; r += 1024) colk[r] = (r == k) ? 0.0 : M[(size_t)r * np + k];
    __syncthreads();
    for (int r = wave; r < Ng; r += 16) {
      const double f = colk[r];
      for (int c = lane; c < Ng; c += 64) {
        double* e = M + (size_t)r * np + c;
        if (r == k)
          *e = rowk[c];
        else
          *e = ((c == k) ? 0.0 : *e) - f * rowk[c];
      }
    }
    __syncthreads();
  }
  for (int k = Ng - 1; k >= 0; --k) {
    const int p = piv[k];
    if (p != k)
      for (int r = tid; r < Ng; r += 1024) {
        const double t = M[(size_t)r * np + k];
        M[(size_t)r * np + k] = M[(size_t)r * np + p];
        M[(size_t)r * np + p] = t;
      }
    __syncthreads();
  }
}

hipError_t launch_shell_build(int n_shells, int Ng, const double* B0T, const double* KT,
                              const double* coef, double* invP, int* status, hipStream_t s) {
  if (Ng > 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(shell_build_kernel, dim3(n_shells), dim3(1024), 0, s, Ng,
                     shell_pad_order(Ng), B0T, KT, coef, invP, status);
  return hipGetLastError();
}

// Diagonal of the shell solves in the eigenbasis of B0, built once per
// context: tab[r][m] = 1 / (alpha_m - c_m lam_r).
__global__ void __launch_bounds__(256)
eigen_diag_kernel(int nl, int nmp, const double* __restrict__ alpha, const double* __restrict__ c,
                  const double* __restrict__ lam, double* __restrict__ tab) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)nl * nmp) return;
  const int r = (int)(idx / nmp), m = (int)(idx % nmp);
  tab[idx] = 1.0 / (alpha[m] - c[m] * lam[r]);
}

hipError_t launch_eigen_diag(int nl, int nmp, const double* alpha, const double* c,
                             const double* lam, double* tab, hipStream_t s) {
  const size_t n = (size_t)nl * nmp;
  hipLaunchKernelGGL(eigen_diag_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nl,
                     nmp, alpha, c, lam, tab);
  return hipGetLastError();
}

int shell_pad_order(int Ng) { return 64 * ((Ng + 63) / 64); }

void shell_pack_inv(const double* invT, int n_shells, int Ng, int s0, int s1, double* out) {
  const int np = shell_pad_order(Ng);
  for (size_t i = 0; i < (size_t)(s1 - s0) * np * np; ++i) out[i] = 0.0;
  for (int s = s0; s < s1; ++s) {
    const double* src = invT + (size_t)s * Ng * Ng;
    double* dst = out + (size_t)(s - s0) * np * np;
    for (int j = 0; j < Ng; ++j)
      for (int i = 0; i < Ng; ++i) dst[(size_t)j * np + i] = src[(size_t)j * Ng + i];
  }
  (void)n_shells;
}

hipError_t launch_shell_solve(int n_shells, int Ng, int ld, const int* shell_ptr,
                              const int* shell_modes, const double* invP, const double* rhs,
                              double* w_out, const double* Rth, double* th_out,
                              const double* iath, double alpha, hipStream_t s) {
  const int tw = shell_pad_order(Ng) / 64;
  switch (tw) {
#define NHQG_TW(n)                                                                          \
  case n:                                                                                   \
    hipLaunchKernelGGL(shell_solve_kernel<n>, dim3(n_shells), dim3(256), 0, s, Ng, ld,      \
                       shell_ptr, shell_modes, invP, rhs, w_out, Rth, th_out, iath, alpha); \
    break;
    NHQG_TW(1) NHQG_TW(2) NHQG_TW(3) NHQG_TW(4)
#undef NHQG_TW
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Horizontal means <a b>_xy per level by discrete Parseval:
//   sum_k pw_k Re(Ha_k conj(Hb_k)),   pw = {1 on the ky=0 and ky=Nx/2
//   columns, 2 elsewhere} / Nx^4,
// where H is the Hermitian part on the self-conjugate columns (what irfft2
// keeps).  This equals mean(irfft2(a) * irfft2(b)) of
// nhqg/solver.py:231-245 exactly in exact arithmetic; the products are
// summed in a fixed tree (no atomics), partials[l][p] per block.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
parseval_partial_kernel(int nm, int ld, int P, const double* __restrict__ a,
                        const double* __restrict__ b, const double* __restrict__ pw,
                        const int* __restrict__ partner, double* __restrict__ partial) {
  __shared__ double red[256];
  const int l = blockIdx.y, p = blockIdx.x, tid = threadIdx.x;
  const int chunk = (nm + P - 1) / P;
  const int m0 = p * chunk, m1 = min(nm, m0 + chunk);
  const double* ar = a + (size_t)l * ld;
  const double* br = b + (size_t)l * ld;
  double sum = 0.0;
  for (int m = m0 + tid; m < m1; m += 256) {
    double xr = ar[2 * m], xi = ar[2 * m + 1];
    double yr = br[2 * m], yi = br[2 * m + 1];
    const int q = partner ? partner[m] : -1;
    if (q >= 0) {
      xr = 0.5 * (xr + ar[2 * q]);
      xi = 0.5 * (xi - ar[2 * q + 1]);
      yr = 0.5 * (yr + br[2 * q]);
      yi = 0.5 * (yi - br[2 * q + 1]);
    }
    sum += pw[m] * (xr * yr + xi * yi);
  }
  red[tid] = sum;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) partial[(size_t)l * P + p] = red[0];
}

hipError_t launch_parseval(int nl, int nm, int ld, int P, const double* a, const double* b,
                           const double* pw, const int* partner, double* partial,
                           hipStream_t s) {
  hipLaunchKernelGGL(parseval_partial_kernel, dim3(P, nl), dim3(256), 0, s, nm, ld, P, a, b, pw,
                     partner, partial);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// SBP2 corrector, serial 1-D part (nhqg/solver.py:929-942, 885-905).
// The corrector matrix A = I - (c1 L + c2 D1 diag(<w^2>) D1) with identity
// boundary rows is pentadiagonal and fixed for all substeps of a call, so
// the preparation kernel factors it in LDS (no pivoting: A is I plus a
// positive semidefinite, diagonally dominant band) and forms A^{-1}
// column by column (one column per thread, banded substitutions with the
// recurrences in registers).  Each substep is then three small parallel
// matrix-vector products.  d1b/lb hold the tridiagonal SBP D1 and
// Laplacian as [i*3 + (j-i+1)].
// ---------------------------------------------------------------------------
#define SBP_NMAX 272
#define SBP_MV_ROWS 2   // rows per wavefront of small_matvec_kernel

// y[i] = sum_j M[i][j] x[j] for a small dense row-major M, one wavefront per
// row: lanes stride the row (coalesced) and a fixed xor-shuffle tree sums
// the lane partials, so the result does not depend on scheduling.
NHQG_DEV double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

NHQG_DEV void block_matvec(int n_out, int n_in, const double* __restrict__ M, const double* x,
                           double* y) {
  // SBP_MV_ROWS rows per wavefront at a time, so that the row loads of a
  // round are all in flight together (the matrices come from L2 / HBM);
  // the grid's wavefronts share the rows
  constexpr int R = SBP_MV_ROWS;
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  const int gw = blockIdx.x * nw + (threadIdx.x >> 6), ngw = gridDim.x * nw;
  for (int i0 = gw * R; i0 < n_out; i0 += ngw * R) {
    double t[R];
#pragma unroll
    for (int r = 0; r < R; ++r) t[r] = 0.0;
#pragma unroll
    for (int jj = 0; jj < (SBP_NMAX + 63) / 64; ++jj) {
      const int j = lane + 64 * jj;
      if (j < n_in) {
        const double xj = x[j];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = min(i0 + r, n_out - 1);
          t[r] += M[(size_t)i * n_in + j] * xj;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double v = wave_sum(t[r]);
      if (lane == 0 && i0 + r < n_out) y[i0 + r] = v;
    }
  }
}


NHQG_DEV double band3(const double* b3, int i, int j, int n) {
  const int o = j - i;
  if (o < -1 || o > 1 || j < 0 || j >= n) return 0.0;
  return b3[i * 3 + o + 1];
}

// Corrector preamble (nhqg/solver.py:930-942) in one block: <w_sbp^2> per
// level from the Parseval partials, the band matrices B = I + half and
// A = I - half (boundary rows of A set to identity) and the pentadiagonal
// LU of A without pivoting (one thread, register recurrence).  LUg gets the
// bands (a, b multipliers; d, e, f of U) and then the reciprocal pivots.
__global__ void __launch_bounds__(512)
sbp_prepare_kernel(int n, int P, const double* __restrict__ partial, const double* __restrict__ d1b,
                   const double* __restrict__ lb, double c1, double c2, double* __restrict__ Bg,
                   double* __restrict__ LUg) {
  __shared__ double w2[SBP_NMAX];
  __shared__ double d1s[3 * SBP_NMAX], lbs[3 * SBP_NMAX];
  __shared__ double LU[SBP_NMAX * 5];   // per row: a (-2), b (-1), d, e (+1), f (+2)
  __shared__ double rd[SBP_NMAX];       // 1 / pivot
  const int tid = threadIdx.x;
  for (int i = tid; i < 3 * n; i += 512) {
    d1s[i] = d1b[i];
    lbs[i] = lb[i];
  }
  for (int i = tid; i < n; i += 512) {
    double sum = 0.0;
    for (int p = 0; p < P; ++p) sum += partial[(size_t)i * P + p];
    w2[i] = sum;
  }
  __syncthreads();
  for (int i = tid; i < n; i += 512) {
    for (int o = -2; o <= 2; ++o) {
      const int j = i + o;
      double h = 0.0;
      if (j >= 0 && j < n) {
        h = c1 * band3(lbs, i, j, n);
        for (int k = i - 1; k <= i + 1; ++k) {
          if (k < 0 || k >= n) continue;
          h += c2 * band3(d1s, i, k, n) * w2[k] * band3(d1s, k, j, n);
        }
      }
      const double id = (o == 0) ? 1.0 : 0.0;
      Bg[i * 5 + o + 2] = id + h;
      LU[i * 5 + o + 2] = (i == 0 || i == n - 1) ? id : id - h;
    }
  }
  __syncthreads();
  if (tid == 0) {
    // row k+1 (b1, d1, e1) and row k+2 (a2, b2, d2) carried in registers
    double d0 = LU[2], e0 = LU[3], f0 = LU[4];
    double b1 = LU[5 + 1], d1 = LU[5 + 2], e1 = LU[5 + 3];
    double a2 = n > 2 ? LU[10 + 0] : 0.0, b2 = n > 2 ? LU[10 + 1] : 0.0;
    double d2 = n > 2 ? LU[10 + 2] : 0.0;
    // chunks of 8 steps: the original band entries a step consumes (f of
    // row k+1, e of row k+2, a, b, d of row k+3; none is stored before that
    // step) are loaded before the chunk's dependent chain
    constexpr int C = 8;
    for (int k0 = 0; k0 < n - 1; k0 += C) {
      double fk1[C], ek2[C], ak3[C], bk3[C], dk3[C];
#pragma unroll
      for (int u = 0; u < C; ++u) {
        const int k = k0 + u;
        const int r1 = min(k + 1, n - 1), r2 = min(k + 2, n - 1), r3 = min(k + 3, n - 1);
        fk1[u] = LU[r1 * 5 + 4];
        ek2[u] = LU[r2 * 5 + 3];
        ak3[u] = LU[r3 * 5 + 0];
        bk3[u] = LU[r3 * 5 + 1];
        dk3[u] = LU[r3 * 5 + 2];
      }
      // one elimination step; chk = false where k + 3 < n is known
      auto step = [&](int u, int k, bool chk) {
        const double r = 1.0 / d0;
        rd[k] = r;
        const double m1 = b1 * r;
        LU[(k + 1) * 5 + 1] = m1;
        const double nd1 = d1 - m1 * e0;
        const double ne1 = e1 - m1 * f0;
        double nb2 = 0.0, nd2 = 0.0;
        if (!chk || k + 2 < n) {
          const double m2 = a2 * r;
          LU[(k + 2) * 5 + 0] = m2;
          nb2 = b2 - m2 * e0;
          nd2 = d2 - m2 * f0;
        }
        LU[(k + 1) * 5 + 2] = nd1;
        LU[(k + 1) * 5 + 3] = ne1;
        // shift the window: row k+1 becomes the pivot row
        d0 = nd1;
        e0 = ne1;
        f0 = fk1[u];
        b1 = nb2;
        d1 = nd2;
        e1 = (!chk || k + 2 < n) ? ek2[u] : 0.0;
        if (!chk || k + 3 < n) {
          a2 = ak3[u];
          b2 = bk3[u];
          d2 = dk3[u];
        }
      };
      // chunks away from the last rows run without bound tests
      if (k0 + C + 3 <= n) {
#pragma unroll
        for (int u = 0; u < C; ++u) step(u, k0 + u, false);
      } else {
#pragma unroll
        for (int u = 0; u < C; ++u)
          if (k0 + u < n - 1) step(u, k0 + u, true);
      }
    }
    rd[n - 1] = 1.0 / d0;
  }
  __syncthreads();
  for (int i = tid; i < 5 * n; i += 512) LUg[i] = LU[i];
  for (int i = tid; i < n; i += 512) LUg[5 * n + i] = rd[i];
}

// ---------------------------------------------------------------------------
// SBP substeps as one affine map per corrector call.  Within a call the
// pentadiagonal A (and B, D1, mu dt_s) are fixed, and a substep
// (nhqg/solver.py:885-905)
//   tb = thb2s th, rh = Zr (B tb - mu dt_s D1 F), nw = A^-1 rh,
//   g = D1 (tb + nw) / 2, th' = s2thb Z nw
// (Zr, Z zero the two boundary rows) is affine in (th, F):
//   th' = M1 th + M2 F,  g = G1 th + G2 F
// with X = A^-1 Zr B thb2s, Y = A^-1 Zr D1 and
//   M1 = s2thb Z X, M2 = -mu dt_s s2thb Z Y,
//   G1 = (D1 thb2s + D1 X) / 2, G2 = -mu dt_s D1 Y / 2.
// sbp_cols_kernel solves the 2n banded systems for X and Y (one thread per
// column, the LU recurrences of the preamble), sbp_compose_kernel forms the
// 2n x 2n map [M1 M2; G1 G2] row-major, and each substep is one many-block
// matvec (sbp_apply_kernel) instead of a single-workgroup chain.
// ---------------------------------------------------------------------------
// Right-hand sides of the column solve, one element per thread, into X|Y
// (n x 2n row-major): Zr (B thb2s) in columns [0, n), Zr D1 in [n, 2n).
__global__ void __launch_bounds__(256)
sbp_rhs_kernel(int n, const double* __restrict__ Bg, const double* __restrict__ thb2s,
               const double* __restrict__ d1b, double* __restrict__ XY) {
  const int ld = 2 * n;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= n * ld) return;
  const int i = idx / ld, col = idx % ld;
  double r = 0.0;
  if (i > 0 && i < n - 1) {
    if (col < n) {
#pragma unroll
      for (int o = -2; o <= 2; ++o) {
        const int j = i + o;
        if (j >= 0 && j < n) r += Bg[i * 5 + o + 2] * thb2s[(size_t)j * n + col];
      }
    } else {
      r = band3(d1b, i, col - n, n);
    }
  }
  XY[idx] = r;
}

// Banded solve of every column of X|Y with the preamble's LU, one column
// per lane (64 per block).  The recurrences are serial in the row, so the
// block's four waves first stage its 64 right-hand-side columns and the LU
// bands in LDS (wave w loads rows w, w + 4, ..., two batches of SBP_LB rows
// in flight per lane), wave 0 alone runs the forward and backward
// substitutions from LDS (the next chunk of SBP_CH rows read ahead of the
// current chunk's chain), and the four waves write the columns back.  Same
// operations in the same order as a plain row loop: the result does not
// depend on the staging.
#define SBP_CH 8
#define SBP_LB 16
#define SBP_CW 4   // waves per block
__global__ void __launch_bounds__(SBP_CW * 64)
sbp_cols_kernel(int n, const double* __restrict__ LUg, double* __restrict__ XY) {
  __shared__ double cs[SBP_NMAX * 64];
  __shared__ double ba[SBP_NMAX], bb[SBP_NMAX], be[SBP_NMAX], bf[SBP_NMAX], brd[SBP_NMAX];
  const int t = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = blockIdx.x * 64 + t;
  const int ld = 2 * n;
  const bool on = col < ld;
  const double* rdp = LUg + 5 * n;
  const double* cp = XY + min(col, ld - 1);
  for (int i = threadIdx.x; i < n; i += SBP_CW * 64) {
    ba[i] = LUg[i * 5 + 0];
    bb[i] = LUg[i * 5 + 1];
    be[i] = LUg[i * 5 + 3];
    bf[i] = LUg[i * 5 + 4];
    brd[i] = rdp[i];
  }
  {
    // this wave's rows i = wv + SBP_CW * j; batch b + 1 in flight while
    // batch b is stored
    const int nr = (n - wv + SBP_CW - 1) / SBP_CW;
    double v0[SBP_LB], v1[SBP_LB];
    auto load = [&](double (&v)[SBP_LB], int j0) {
#pragma unroll
      for (int u = 0; u < SBP_LB; ++u)
        v[u] = cp[(size_t)(wv + SBP_CW * min(j0 + u, nr - 1)) * ld];
    };
    auto store = [&](const double (&v)[SBP_LB], int j0) {
#pragma unroll
      for (int u = 0; u < SBP_LB; ++u)
        if (j0 + u < nr) cs[(wv + SBP_CW * (j0 + u)) * 64 + t] = v[u];
    };
    const int nb = (nr + SBP_LB - 1) / SBP_LB;
    if (nb > 0) load(v0, 0);
    for (int b = 0; b < nb; b += 2) {
      if (b + 1 < nb) load(v1, (b + 1) * SBP_LB);
      store(v0, b * SBP_LB);
      if (b + 2 < nb) load(v0, (b + 2) * SBP_LB);
      if (b + 1 < nb) store(v1, (b + 1) * SBP_LB);
    }
  }
  __syncthreads();
  if (wv == 0) {
    const int nch = (n + SBP_CH - 1) / SBP_CH;
    // forward: y_i = r_i - (b_i y_{i-1} + a_i y_{i-2})
    double rv[SBP_CH], av[SBP_CH], bv[SBP_CH], rn[SBP_CH], an[SBP_CH], bn[SBP_CH];
    auto fload = [&](int c) {
#pragma unroll
      for (int u = 0; u < SBP_CH; ++u) {
        const int i = min(c * SBP_CH + u, n - 1);
        rn[u] = cs[i * 64 + t];
        an[u] = ba[i];
        bn[u] = bb[i];
      }
    };
    fload(0);
    double y1 = 0.0, y2 = 0.0;
    for (int c = 0; c < nch; ++c) {
#pragma unroll
      for (int u = 0; u < SBP_CH; ++u) {
        rv[u] = rn[u];
        av[u] = an[u];
        bv[u] = bn[u];
      }
      if (c + 1 < nch) fload(c + 1);
      auto fstep = [&](int u, int i) {
        const double v = rv[u] - (bv[u] * y1 + av[u] * y2);
        cs[i * 64 + t] = v;
        y2 = y1;
        y1 = v;
      };
      // full chunks without a per-row bound test
      if ((c + 1) * SBP_CH <= n) {
#pragma unroll
        for (int u = 0; u < SBP_CH; ++u) fstep(u, c * SBP_CH + u);
      } else {
#pragma unroll
        for (int u = 0; u < SBP_CH; ++u)
          if (c * SBP_CH + u < n) fstep(u, c * SBP_CH + u);
      }
    }
    // backward: x_i = (y_i - e_i x_{i+1} - f_i x_{i+2}) / d_i
    double ev[SBP_CH], fv[SBP_CH], dv[SBP_CH], en[SBP_CH], fn[SBP_CH], dn[SBP_CH];
    auto bload = [&](int c) {
#pragma unroll
      for (int u = 0; u < SBP_CH; ++u) {
        const int i = max(n - 1 - (c * SBP_CH + u), 0);
        rn[u] = cs[i * 64 + t];   // y from the forward pass (all rows written)
        en[u] = be[i];
        fn[u] = bf[i];
        dn[u] = brd[i];
      }
    };
    bload(0);
    double x1 = 0.0, x2 = 0.0;
    for (int c = 0; c < nch; ++c) {
#pragma unroll
      for (int u = 0; u < SBP_CH; ++u) {
        rv[u] = rn[u];
        ev[u] = en[u];
        fv[u] = fn[u];
        dv[u] = dn[u];
      }
      if (c + 1 < nch) bload(c + 1);
      auto bstep = [&](int u, int i) {
        const double v = (rv[u] - ev[u] * x1 - fv[u] * x2) * dv[u];
        cs[i * 64 + t] = v;
        x2 = x1;
        x1 = v;
      };
      if ((c + 1) * SBP_CH <= n) {
#pragma unroll
        for (int u = 0; u < SBP_CH; ++u) bstep(u, n - 1 - (c * SBP_CH + u));
      } else {
#pragma unroll
        for (int u = 0; u < SBP_CH; ++u)
          if (n - 1 - (c * SBP_CH + u) >= 0) bstep(u, n - 1 - (c * SBP_CH + u));
      }
    }
  }
  __syncthreads();
  if (on)
    for (int i = wv; i < n; i += SBP_CW) XY[(size_t)i * ld + col] = cs[i * 64 + t];
}

// Mall (2n x 2n, row-major) = [M1 M2; G1 G2].  Rows [0, n): s2thb Z XY as a
// 32 x 32-tiled product through LDS (4 outputs per thread, the next k tile
// in flight in registers while the current one is used), the Y half
// scaled by -mu dt_s; rows [n, 2n): the banded D1 rows, one element per
// thread.
#define SBP_TILE 32
__global__ void __launch_bounds__(256)
sbp_compose_kernel(int n, const double* __restrict__ XY, const double* __restrict__ s2thb,
                   const double* __restrict__ thb2s, const double* __restrict__ d1b, double mu_sdt,
                   double* __restrict__ Mall) {
  const int ld = 2 * n;
  const int tid = threadIdx.x;
  const int ntr = (n + SBP_TILE - 1) / SBP_TILE, ntc = (ld + SBP_TILE - 1) / SBP_TILE;
  if ((int)blockIdx.x < ntr * ntc) {
    __shared__ double As[SBP_TILE][SBP_TILE + 1], Bs[SBP_TILE][SBP_TILE + 1];
    const int r0 = (blockIdx.x / ntc) * SBP_TILE, c0 = (blockIdx.x % ntc) * SBP_TILE;
    const int tc = tid % SBP_TILE, tr = tid / SBP_TILE;   // 8 row groups of 4 rows
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // the next k tile is loaded into registers while the current one is used
    constexpr int EPT = SBP_TILE * SBP_TILE / 256;
    double pa[EPT], pb[EPT];
    auto fetch = [&](int k0) {
#pragma unroll
      for (int u = 0; u < EPT; ++u) {
        const int e = tid + 256 * u;
        const int a = e / SBP_TILE, b = e % SBP_TILE;
        const int r = r0 + a, k = k0 + b;
        pa[u] = (r < n && k < n) ? s2thb[(size_t)r * n + k] : 0.0;
        const int kb = k0 + a, cb = c0 + b;
        // Z: rows 0 and n-1 of XY do not contribute
        pb[u] = (kb > 0 && kb < n - 1 && cb < ld) ? XY[(size_t)kb * ld + cb] : 0.0;
      }
    };
    fetch(0);
    for (int k0 = 0; k0 < n; k0 += SBP_TILE) {
#pragma unroll
      for (int u = 0; u < EPT; ++u) {
        const int e = tid + 256 * u;
        As[e / SBP_TILE][e % SBP_TILE] = pa[u];
        Bs[e / SBP_TILE][e % SBP_TILE] = pb[u];
      }
      __syncthreads();
      if (k0 + SBP_TILE < n) fetch(k0 + SBP_TILE);
#pragma unroll 8
      for (int kk = 0; kk < SBP_TILE; ++kk) {
        const double bv = Bs[kk][tc];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = fma(As[tr * 4 + q][kk], bv, acc[q]);
      }
      __syncthreads();
    }
    const int c = c0 + tc;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + tr * 4 + q;
      if (r < n && c < ld) Mall[(size_t)r * ld + c] = (c >= n) ? -mu_sdt * acc[q] : acc[q];
    }
    return;
  }
  const int idx = (blockIdx.x - ntr * ntc) * 256 + tid;
  if (idx >= n * ld) return;
  const int i = idx / ld, c = idx % ld;
  double v = 0.0;
#pragma unroll
  for (int o = -1; o <= 1; ++o) {
    const int j = i + o;
    if (j < 0 || j >= n) continue;
    const double dij = band3(d1b, i, j, n);
    v += dij * (c < n ? thb2s[(size_t)j * n + c] + XY[(size_t)j * ld + c] : XY[(size_t)j * ld + c]);
  }
  Mall[(size_t)(n + i) * ld + c] = v * ((c < n) ? 0.5 : -0.5 * mu_sdt);
}

// One substep: F = sum_p partial[:, p] (each block redoes this small sum),
// x = [th_src; F], out = Mall x; rows [0, n) -> th_dst, rows [n, 2n) -> g.
#define SBP_APPLY_ROWS 8
__global__ void __launch_bounds__(256)
sbp_apply_kernel(int n, int P, const double* __restrict__ partial, const double* __restrict__ Mall,
                 const double* __restrict__ th_src, double* __restrict__ th_dst,
                 double* __restrict__ g_out) {
  __shared__ double x[2 * SBP_NMAX];
  const int tid = threadIdx.x, ld = 2 * n;
  for (int i = tid; i < n; i += 256) {
    x[i] = th_src[i];
    double sum = 0.0;
    for (int p = 0; p < P; ++p) sum += partial[(size_t)i * P + p];
    x[n + i] = sum;
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  constexpr int RW = SBP_APPLY_ROWS / 4;   // rows per wave
  const int r0 = blockIdx.x * SBP_APPLY_ROWS + wave * RW;
  double t[RW];
#pragma unroll
  for (int q = 0; q < RW; ++q) t[q] = 0.0;
  for (int c = lane; c < ld; c += 64) {
    const double xc = x[c];
#pragma unroll
    for (int q = 0; q < RW; ++q) t[q] = fma(Mall[(size_t)min(r0 + q, ld - 1) * ld + c], xc, t[q]);
  }
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    const double v = wave_sum(t[q]);
    const int r = r0 + q;
    if (lane == 0 && r < ld) {
      if (r < n) th_dst[r] = v;
      else g_out[r - n] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// SBP substep transfer as a level stencil.  With the input's boundary rows
// zeroed (Z), one substep cycle Psbp = T2S S2T (SBP grid -> Galerkin theta
// -> SBP grid) is c2s s2c on the interior: the Dirichlet projection is the
// identity on the CGL interpolant of values that vanish at both ends, and
// the two linear interpolations compose to a band of half-width 2.  The
// host keeps the five diagonals pb[i][o] (o = -2..2, input rows 1..n-2) when
// every other entry of Psbp Z is roundoff (create_impl), and each substep
// but the last is then
//   out[i] = sum_o pb[i][o] x[i+o],  x[j] = Z_j (th[j] - bc g[j] w[j])
// (the split kernel's fused update), with the next substep's horizontal
// mean <w, out> summed on the way: bpart[i * nb + blockIdx.x] over the
// block's columns with the diagonal Parseval weights pwc (per real column);
// sbp_cross_kernel adds the nb block sums of a level and the
// self-conjugate columns' cross terms into partial[i].  One thread per real column, LB levels per block (a window of
// five x and three w in registers); a bandwidth-bound pass instead of a
// 257 x 257 contraction.
// ---------------------------------------------------------------------------
#define SBP_BAND_LB 32
__global__ void __launch_bounds__(256)
sbp_band_kernel(int n, int ld, const double* __restrict__ pb, const double* __restrict__ th,
                const double* __restrict__ w, const double* __restrict__ g, double bc,
                const double* __restrict__ pwc, double* __restrict__ out,
                double* __restrict__ bpart) {
  __shared__ double red[4][SBP_BAND_LB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = blockIdx.x * 256 + tid;   // ld is a multiple of 64: every lane is a column
  const bool on = c < ld;
  const int cc = on ? c : ld - 1;
  const int i0 = blockIdx.y * SBP_BAND_LB, i1 = min(n, i0 + SBP_BAND_LB);
  const double wt = on ? pwc[cc] : 0.0;
  auto ld_x = [&](int j, double& xv, double& wv) {
    if (j < 0 || j >= n) {
      xv = 0.0;
      wv = 0.0;
      return;
    }
    wv = w[(size_t)j * ld + cc];
    const double t = th[(size_t)j * ld + cc];
    xv = (j == 0 || j == n - 1) ? 0.0 : t - bc * g[j] * wv;
  };
  // window: x[i-2 .. i+2], w[i .. i+2]
  double x0, x1, x2, x3, x4, w0, w1, w2, dummy;
  ld_x(i0 - 2, x0, dummy);
  ld_x(i0 - 1, x1, dummy);
  ld_x(i0, x2, w0);
  ld_x(i0 + 1, x3, w1);
#pragma unroll 4
  for (int i = i0; i < i1; ++i) {
    ld_x(i + 2, x4, w2);
    const double* p = pb + 5 * i;
    const double v = p[0] * x0 + p[1] * x1 + p[2] * x2 + p[3] * x3 + p[4] * x4;
    if (on) out[(size_t)i * ld + c] = v;
    const double s = wave_sum(wt * w0 * v);
    if (lane == 0) red[wave][i - i0] = s;
    x0 = x1;
    x1 = x2;
    x2 = x3;
    x3 = x4;
    w0 = w1;
    w1 = w2;
  }
  __syncthreads();
  if (tid < i1 - i0)
    bpart[(size_t)(i0 + tid) * gridDim.x + blockIdx.x] =
        (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
}

// Cross terms of the horizontal mean's Hermitian part on the self-conjugate
// columns: for each pair (m, q) (q = m for a self-partnered mode) with
// weight wt, wt (a_m.r b_q.r + a_q.r b_m.r - a_m.i b_q.i - a_q.i b_m.i);
// together with the diagonal weights of sbp_band_kernel this is exactly
// the sum of parseval_partial_kernel.  One block per level: the level's nb
// block sums of sbp_band_kernel plus the cross terms -> partial[i].
__global__ void __launch_bounds__(256)
sbp_cross_kernel(int n, int ld, int npairs, const int2* __restrict__ pq,
                 const double* __restrict__ pwt, const double* __restrict__ a,
                 const double* __restrict__ b, const double* __restrict__ bpart, int nb,
                 double* __restrict__ partial) {
  __shared__ double red[4];
  const int i = blockIdx.x, tid = threadIdx.x;
  const double* ar = a + (size_t)i * ld;
  const double* br = b + (size_t)i * ld;
  double s = 0.0;
  for (int k = tid; k < nb; k += 256) s += bpart[(size_t)i * nb + k];
  for (int k = tid; k < npairs; k += 256) {
    const int m = pq[k].x, q = pq[k].y;
    const double amr = ar[2 * m], ami = ar[2 * m + 1], aqr = ar[2 * q], aqi = ar[2 * q + 1];
    const double bmr = br[2 * m], bmi = br[2 * m + 1], bqr = br[2 * q], bqi = br[2 * q + 1];
    s += pwt[k] * ((amr * bqr + aqr * bmr) - (ami * bqi + aqi * bmi));
  }
  s = wave_sum(s);
  if ((tid & 63) == 0) red[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) partial[i] = (red[0] + red[1]) + (red[2] + red[3]);
}

hipError_t launch_sbp_band(int n, int ld, const double* pb, const double* th, const double* w,
                           const double* g, double bc, const double* pwc, double* out,
                           double* partial, double* bpart, int npairs, const int2* pq,
                           const double* pwt, hipStream_t s) {
  const int ncb = (ld + 255) / 256;
  if (ld % 64 != 0 || th == out) return hipErrorInvalidValue;
  const dim3 grid(ncb, (n + SBP_BAND_LB - 1) / SBP_BAND_LB);
  hipLaunchKernelGGL(sbp_band_kernel, grid, dim3(256), 0, s, n, ld, pb, th, w, g, bc, pwc, out,
                     bpart);
  hipLaunchKernelGGL(sbp_cross_kernel, dim3(n), dim3(256), 0, s, n, ld, npairs, pq, pwt, w, out,
                     bpart, ncb, partial);
  return hipGetLastError();
}

// sbp buffer (doubles, n = Nz+1): [0,5n) B band, [5n,10n) LU bands,
// [10n,11n) reciprocal pivots.
hipError_t launch_sbp_prepare(int n, int P, const double* partial, const double* d1b,
                              const double* lb, double c1, double c2, double* sbp,
                              hipStream_t s) {
  if (n > SBP_NMAX) return hipErrorInvalidValue;
  double* Bg = sbp;
  double* LUg = sbp + 5 * (size_t)n;
  hipLaunchKernelGGL(sbp_prepare_kernel, dim3(1), dim3(512), 0, s, n, P, partial, d1b, lb, c1, c2,
                     Bg, LUg);
  return hipGetLastError();
}

hipError_t launch_sbp_maps(int n, const double* sbp, const double* thb2s, const double* s2thb,
                           const double* d1b, double mu_sdt, double* XY, double* Mall,
                           hipStream_t s) {
  if (n > SBP_NMAX) return hipErrorInvalidValue;
  const double* Bg = sbp;
  const double* LUg = sbp + 5 * (size_t)n;
  hipLaunchKernelGGL(sbp_rhs_kernel, dim3((2 * n * n + 255) / 256), dim3(256), 0, s, n, Bg, thb2s,
                     d1b, XY);
  hipLaunchKernelGGL(sbp_cols_kernel, dim3((2 * n + 63) / 64), dim3(SBP_CW * 64), 0, s, n, LUg,
                     XY);
  const int ntr = (n + SBP_TILE - 1) / SBP_TILE, ntc = (2 * n + SBP_TILE - 1) / SBP_TILE;
  const int nband = (2 * n * n + 255) / 256;
  hipLaunchKernelGGL(sbp_compose_kernel, dim3(ntr * ntc + nband), dim3(256), 0, s, n, XY, s2thb,
                     thb2s, d1b, mu_sdt, Mall);
  return hipGetLastError();
}

hipError_t launch_sbp_apply(int n, int P, const double* partial, const double* Mall,
                            const double* th_src, double* th_dst, double* g_out, hipStream_t s) {
  if (n > SBP_NMAX || th_src == th_dst) return hipErrorInvalidValue;
  hipLaunchKernelGGL(sbp_apply_kernel, dim3((2 * n + SBP_APPLY_ROWS - 1) / SBP_APPLY_ROWS),
                     dim3(256), 0, s, n, P, partial, Mall, th_src, th_dst, g_out);
  return hipGetLastError();
}

// y = M x for small dense M (n_out x n_in <= SBP_NMAX), SBP_MV_ROWS rows
// per wavefront over many blocks.
__global__ void __launch_bounds__(256)
small_matvec_kernel(int n_out, int n_in, const double* __restrict__ M, const double* __restrict__ x,
                    double* __restrict__ y) {
  block_matvec(n_out, n_in, M, x, y);
}

// y = x0 + c * (x1 - x2) (n entries; one block).
__global__ void __launch_bounds__(256)
small_axpy_kernel(int n, const double* x0, const double* x1, const double* x2, double c, double* y) {
  for (int i = threadIdx.x; i < n; i += 256) y[i] = x0[i] + c * (x1[i] - x2[i]);
}

hipError_t launch_small_matvec(int n_out, int n_in, const double* M, const double* x, double* y,
                               hipStream_t s) {
  if (n_in > SBP_NMAX) return hipErrorInvalidValue;
  const int nb = (n_out + 4 * SBP_MV_ROWS - 1) / (4 * SBP_MV_ROWS);
  hipLaunchKernelGGL(small_matvec_kernel, dim3(nb), dim3(256), 0, s, n_out, n_in, M, x, y);
  return hipGetLastError();
}

hipError_t launch_small_axpy(int n, const double* x0, const double* x1, const double* x2, double c,
                             double* y, hipStream_t s) {
  hipLaunchKernelGGL(small_axpy_kernel, dim3(1), dim3(256), 0, s, n, x0, x1, x2, c, y);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Post-step sanitation (nhqg/solver.py:98-146, 993-999): zero the k=0 mode,
// Hermitian-project the self-conjugate columns pairwise, and (full layout
// only) multiply by the 2/3 mask.  pairs[p] = (m, m') with m <= m'.
// ---------------------------------------------------------------------------
// Rows of up to three fields in one launch: blockIdx.y < nl0 in f0, then nl1
// rows of f1, then the rest of f2 (the same per-row arithmetic as three
// launches).
__global__ void __launch_bounds__(256)
sanitize_kernel(int nl0, int nl1, int ld, int npairs, const int2* __restrict__ pairs, int m0,
                double* __restrict__ f0, double* __restrict__ f1, double* __restrict__ f2) {
  int l = blockIdx.y;
  double* f = f0;
  if (l >= nl0) {
    l -= nl0;
    f = f1;
    if (l >= nl1) {
      l -= nl1;
      f = f2;
    }
  }
  double* r = f + (size_t)l * ld;
  for (int p = blockIdx.x * 256 + threadIdx.x; p < npairs; p += gridDim.x * 256) {
    const int2 pr = pairs[p];
    double ax = r[2 * pr.x], ay = r[2 * pr.x + 1];
    double bx = r[2 * pr.y], by = r[2 * pr.y + 1];
    if (pr.x == m0) { ax = ay = 0.0; }
    if (pr.y == m0) { bx = by = 0.0; }
    if (pr.x == pr.y) {
      r[2 * pr.x] = 0.5 * (ax + ax);
      r[2 * pr.x + 1] = 0.5 * (ay - ay);
    } else {
      r[2 * pr.x] = 0.5 * (ax + bx);
      r[2 * pr.x + 1] = 0.5 * (ay - by);
      r[2 * pr.y] = 0.5 * (bx + ax);
      r[2 * pr.y + 1] = 0.5 * (by - ay);
    }
  }
}

hipError_t launch_sanitize(int nl, int ld, int npairs, const int2* pairs, int m0, double* f,
                           hipStream_t s) {
  const int bx = (npairs + 255) / 256;
  hipLaunchKernelGGL(sanitize_kernel, dim3(bx > 0 ? bx : 1, nl), dim3(256), 0, s, nl, 0, ld,
                     npairs, pairs, m0, f, f, f);
  return hipGetLastError();
}

hipError_t launch_sanitize3(int nl0, int nl12, int ld, int npairs, const int2* pairs, int m0,
                            double* f0, double* f1, double* f2, hipStream_t s) {
  const int bx = (npairs + 255) / 256;
  hipLaunchKernelGGL(sanitize_kernel, dim3(bx > 0 ? bx : 1, nl0 + 2 * nl12), dim3(256), 0, s,
                     nl0, nl12, ld, npairs, pairs, m0, f0, f1, f2);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256)
mask_kernel(size_t total, int ld, double* f, const double* mask) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= total) return;
  f[idx] *= mask[(idx % (size_t)ld) >> 1];
}

hipError_t launch_mask(int nl, int ld, double* f, const double* mask, hipStream_t s) {
  const size_t total = (size_t)nl * ld;
  hipLaunchKernelGGL(mask_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, total,
                     ld, f, mask);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Layout conversion between the reference's (levels, Nx, Nk) complex arrays
// and the band-compact layout.  full_idx[m] = row*Nk + ky for kept modes.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
gather_kernel(int nl, int nm, int nm_pad, int full_n, const int* __restrict__ full_idx,
              const double2* __restrict__ full, double2* __restrict__ comp) {
  const int l = blockIdx.y;
  for (int m = blockIdx.x * 256 + threadIdx.x; m < nm_pad; m += gridDim.x * 256)
    comp[(size_t)l * nm_pad + m] =
        m < nm ? full[(size_t)l * full_n + full_idx[m]] : make_double2(0.0, 0.0);
}

__global__ void __launch_bounds__(256)
zero_kernel(size_t n, double* p) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx < n) p[idx] = 0.0;
}

__global__ void __launch_bounds__(256)
scatter_kernel(int nl, int nm, int nm_pad, int full_n, const int* __restrict__ full_idx,
               const double2* __restrict__ comp, double2* __restrict__ full) {
  const int l = blockIdx.y;
  for (int m = blockIdx.x * 256 + threadIdx.x; m < nm; m += gridDim.x * 256)
    full[(size_t)l * full_n + full_idx[m]] = comp[(size_t)l * nm_pad + m];
}

hipError_t launch_gather(int nl, int nm, int nm_pad, int full_n, const int* full_idx,
                         const double2* full, double2* comp, hipStream_t s) {
  const int bx = (nm_pad + 255) / 256;
  hipLaunchKernelGGL(gather_kernel, dim3(bx, nl), dim3(256), 0, s, nl, nm, nm_pad, full_n,
                     full_idx, full, comp);
  return hipGetLastError();
}

hipError_t launch_scatter(int nl, int nm, int nm_pad, int full_n, const int* full_idx,
                          const double2* comp, double2* full, hipStream_t s) {
  const size_t nz = (size_t)nl * full_n * 2;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s, nz,
                     (double*)full);
  const int bx = (nm + 255) / 256;
  hipLaunchKernelGGL(scatter_kernel, dim3(bx, nl), dim3(256), 0, s, nl, nm, nm_pad, full_n,
                     full_idx, comp, full);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Diagnostics reductions.
// vsum[l][p]: partial sums over modes of pw * w_extra[m] * |Ha|^2 ...
// (reuse parseval with a per-mode extra weight) and a max reduction.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
max_reduce_kernel(size_t n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double red[256];
  double m = 0.0;
  for (size_t i = threadIdx.x; i < n; i += 256) m = fmax(m, x[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

hipError_t launch_max_reduce(size_t n, const double* x, double* out, hipStream_t s) {
  hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(256), 0, s, n, x, out);
  return hipGetLastError();
}

// E_q += i beta kx inv_denom q  (the -i beta kx psi term, psi = -q inv_denom;
// nhqg/solver.py:547).
__global__ void __launch_bounds__(256)
beta_kernel(int nm, int ldc, double beta, const double* __restrict__ kx_mode,
            const double* __restrict__ inv_denom, const double2* __restrict__ q,
            double2* __restrict__ eq) {
  const int l = blockIdx.y;
  for (int m = blockIdx.x * 256 + threadIdx.x; m < nm; m += gridDim.x * 256) {
    const size_t idx = (size_t)l * ldc + m;
    eq[idx] = cadd(eq[idx], cmul_i(q[idx], beta * kx_mode[m] * inv_denom[m]));
  }
}

hipError_t launch_beta(int nl, int ldc, int nm, double beta, const double* kx_mode,
                       const double* inv_denom, const double2* q, double2* eq, hipStream_t s) {
  hipLaunchKernelGGL(beta_kernel, dim3((nm + 255) / 256, nl), dim3(256), 0, s, nm, ldc, beta,
                     kx_mode, inv_denom, q, eq);
  return hipGetLastError();
}

// Plain streaming copy (calibration of the HBM traffic counters and a
// practical bandwidth ceiling): dst[i] = src[i], 16-byte elements.
__global__ void __launch_bounds__(256)
stream_copy_kernel(size_t n, const double2* __restrict__ src, double2* __restrict__ dst) {
  // four 16-byte loads in flight per lane, nontemporal stores (the written
  // lines are not re-read): 5.07 TB/s on a 1 GiB copy against 4.53 for one
  // load per iteration and plain stores
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t base = (size_t)blockIdx.x * 256 + threadIdx.x; base < n; base += 4 * stride) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t i = base + u * stride;
      v[u] = i < n ? src[i] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t i = base + u * stride;
      if (i < n) {
        __builtin_nontemporal_store(v[u].x, &dst[i].x);
        __builtin_nontemporal_store(v[u].y, &dst[i].y);
      }
    }
  }
}

hipError_t launch_stream_copy(size_t n16, const double2* src, double2* dst, hipStream_t s) {
  hipLaunchKernelGGL(stream_copy_kernel, dim3(8192), dim3(256), 0, s, n16, src, dst);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// 1-D and per-level helpers of the mean-temperature equation and the
// diagnostics (evolve_mean with the legacy exchange, the RK4 validator).
// ---------------------------------------------------------------------------

// y = c0 x0 + c1 x1 + c2 x2 + c3 x3 over n entries (null x_k skipped)
__global__ void __launch_bounds__(256)
small_lincomb_kernel(int n, double* __restrict__ y, double c0, const double* x0, double c1,
                     const double* x1, double c2, const double* x2, double c3, const double* x3) {
  for (int i = threadIdx.x; i < n; i += 256) {
    double v = x0 ? c0 * x0[i] : 0.0;
    if (x1) v += c1 * x1[i];
    if (x2) v += c2 * x2[i];
    if (x3) v += c3 * x3[i];
    y[i] = v;
  }
}

hipError_t launch_small_lincomb(int n, double* y, double c0, const double* x0, double c1,
                                const double* x1, double c2, const double* x2, double c3,
                                const double* x3, hipStream_t s) {
  hipLaunchKernelGGL(small_lincomb_kernel, dim3(1), dim3(256), 0, s, n, y, c0, x0, c1, x1, c2, x2,
                     c3, x3);
  return hipGetLastError();
}

// y = M v with v[l] = sum_p partial[l][p] (the per-level horizontal means of
// launch_parseval), one block; both sums in a fixed order
__global__ void __launch_bounds__(256)
partial_matvec_kernel(int n_out, int nl, int P, const double* __restrict__ partial,
                      const double* __restrict__ M, double* __restrict__ y) {
  __shared__ double v[SBP_NMAX];
  for (int l = threadIdx.x; l < nl; l += 256) {
    double t = 0.0;
    for (int p = 0; p < P; ++p) t += partial[(size_t)l * P + p];
    v[l] = t;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_out; i += 256) {
    double t = 0.0;
    for (int l = 0; l < nl; ++l) t += M[(size_t)i * nl + l] * v[l];
    y[i] = t;
  }
}

hipError_t launch_partial_matvec(int n_out, int nl, int P, const double* partial, const double* M,
                                 double* y, hipStream_t s) {
  if (nl > SBP_NMAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(partial_matvec_kernel, dim3(1), dim3(256), 0, s, n_out, nl, P, partial, M, y);
  return hipGetLastError();
}

// y[l][i] = g[l] x[l][i] over the ld doubles of each level
__global__ void __launch_bounds__(256)
level_scale_kernel(int ld, const double* __restrict__ g, const double* __restrict__ x,
                   double* __restrict__ y) {
  const int l = blockIdx.y;
  const double a = g[l];
  for (int i = blockIdx.x * 256 + threadIdx.x; i < ld; i += gridDim.x * 256)
    y[(size_t)l * ld + i] = a * x[(size_t)l * ld + i];
}

hipError_t launch_level_scale(int nl, int ld, const double* g, const double* x, double* y,
                              hipStream_t s) {
  const int nb = (ld + 255) / 256 < 64 ? (ld + 255) / 256 : 64;
  hipLaunchKernelGGL(level_scale_kernel, dim3(nb, nl), dim3(256), 0, s, ld, g, x, y);
  return hipGetLastError();
}

// out[m] = sum_l cc[l] |f[l][m]|^2, the depth-integrated power of each mode
// of a nodal field (diagnostics.py:99-108 before the shell binning)
__global__ void __launch_bounds__(256)
depth_power_kernel(int nl, int ld, int nm, const double* __restrict__ cc,
                   const double* __restrict__ f, double* __restrict__ out) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= nm) return;
  double s = 0.0;
  for (int l = 0; l < nl; ++l) {
    const double re = f[(size_t)l * ld + 2 * m], im = f[(size_t)l * ld + 2 * m + 1];
    s += cc[l] * (re * re + im * im);
  }
  out[m] = s;
}

hipError_t launch_depth_power(int nl, int ld, int nm, const double* cc, const double* f,
                              double* out, hipStream_t s) {
  hipLaunchKernelGGL(depth_power_kernel, dim3((nm + 255) / 256), dim3(256), 0, s, nl, ld, nm, cc, f,
                     out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Horizontal spectral upsampling Nx -> Nd (Nd >= Nx, same parity), the
// method of scripts/upsample_checkpoint_horiz.py:27-59 written as a direct
// mode map: the real field's spectrum (the self-conjugate columns of the
// source projected onto their Hermitian part, as irfft2 does) scaled by
// (Nd/Nx)^2, its Nyquist row kx = -Nx/2 split in half between +-Nx/2 and its
// Nyquist column ky = Nx/2 halved, every other destination mode zero.
// One thread per (level, destination mode) of the destination layout.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
upsample_kernel(int nm_d, int nkx_d, int nd, const int* __restrict__ rows_d, int ld_d,
                double* __restrict__ out, int nx, int nkx_s, int nky_s,
                const int* __restrict__ kidx_s, const int* __restrict__ partner_s, int ld_s,
                const double* __restrict__ src, double scale) {
  const int l = blockIdx.y;
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= nm_d) return;
  const int ky = m / nkx_d;
  const int row = rows_d[m % nkx_d];
  const int kx = row < nd / 2 ? row : row - nd;   // fftfreq order
  double re = 0.0, im = 0.0;
  const int akx = kx < 0 ? -kx : kx;
  if (akx <= nx / 2 && ky <= nx / 2 && ky < nky_s) {
    const int srow = ((kx % nx) + nx) % nx;
    const int si = kidx_s[srow];
    if (si >= 0) {
      const int ms = ky * nkx_s + si;
      const double* f = src + (size_t)l * ld_s;
      re = f[2 * ms];
      im = f[2 * ms + 1];
      const int p = partner_s[ms];
      if (p >= 0) {   // self-conjugate column: its Hermitian part
        re = 0.5 * (re + f[2 * p]);
        im = 0.5 * (im - f[2 * p + 1]);
      }
      double fac = scale;
      if ((nx % 2 == 0) && akx == nx / 2) fac *= 0.5;
      if ((nx % 2 == 0) && ky == nx / 2) fac *= 0.5;
      re *= fac;
      im *= fac;
    }
  }
  out[(size_t)l * ld_d + 2 * m] = re;
  out[(size_t)l * ld_d + 2 * m + 1] = im;
}

hipError_t launch_upsample(int nl, int nm_d, int nkx_d, int nd, const int* rows_d, int ld_d,
                           double* out, int nx, int nkx_s, int nky_s, const int* kidx_s,
                           const int* partner_s, int ld_s, const double* src, double scale,
                           hipStream_t s) {
  hipLaunchKernelGGL(upsample_kernel, dim3((nm_d + 255) / 256, nl), dim3(256), 0, s, nm_d, nkx_d,
                     nd, rows_d, ld_d, out, nx, nkx_s, nky_s, kidx_s, partner_s, ld_s, src, scale);
  return hipGetLastError();
}
