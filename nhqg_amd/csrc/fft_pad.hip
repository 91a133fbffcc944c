// Horizontal spectral products of the NHQG explicit tendency, 3/2 rule.
//
// The reference zero-pads each level's (Nx, Nx/2+1) spectrum to the
// M = 3Nx/2 grid -- rows kx = 0..Nx/2-1 keep their index, rows
// kx = -Nx/2..-1 move to the top of the padded range, the ky columns keep
// theirs (nhqg/spectral.py:18-34) -- forms the products there with irfft2 /
// rfft2 of size M (flux form :187-212, Jacobian form :155-184, sponge
// :62-74) and truncates back to the kept modes with the factor (M/Nx)^2
// (:37-50).  M must be exactly 3Nx/2: a larger grid would drop the alias of
// the products of two kx = +-Nx/2 (or ky = Nx/2) components, which the
// reference keeps.  Here the three passes of the 2/3 rule run on M-point
// transforms, M = 3P with P = Nx/2 a power of two:
//
//   col_inverse_pad : per (level, ky <= Nx/2): the Nx entries of the column
//                     at their padded rows, the others zero, inverse DFT
//                     along x (M points)                           -> X5
//   row_flux_pad /  : per (level, padded row pair): c2r along y of the
//   row_jac_pad       half spectrum (zero beyond ky = Nx/2), pointwise
//                     products, r2c; only ky <= Nx/2 are kept        -> Y6
//   col_forward_pad : per (level, ky): forward DFT along x, the Nx kept
//                     rows, i kx, times (M/Nx)^2                     -> A
//
// An M-point transform is one radix-3 step and three P-point radix-4
// transforms (team_fft), each run by a group of P/4 of the team's 3P/4
// threads:
//   inverse transforms (team_fft3, decimation in frequency): the radix-3
//     stage over the team's LDS buffer, then the sub-transforms; the bins
//     come out in decimated order (bin 3m + p at sub-buffer p, position m);
//   forward transforms (decimation in time): input x[3n + r] goes to
//     sub-buffer r, and each output bin k' + pP = sum_r w3^(rp) w^(rk') Y_r[k']
//     is combined when it is read (dit_bin), only for the rows and ky that
//     are kept (the column pass needs 2P of the 3P rows).
// Building the inverse transforms' radix-3 inputs straight from global memory
// instead (each group loading the inputs of its own sub-transform) needed
// 2-3x the load instructions and more registers: it took the row pass from
// 3.1 to 5.8 ms per step at 256^3.  Every barrier that spans the groups of a
// team is a block barrier, which every thread of a workgroup reaches, idle
// teams included.
#include "fft_team.h"

// Columns per column-pass workgroup (C) and its threads (C teams of 3P/4):
// 192..768 threads, LDS below 124 KB at every size.
template <int P>
struct ColBlock3 {
  static constexpr int C = P <= 8 ? 32 : P <= 32 ? 16 : P <= 128 ? 8 : P == 256 ? 4 : 2;
  static constexpr int v = C * 3 * P / 4;
};
// Threads per row-pass workgroup: 192 (256/P teams) up to P = 256, one team
// of 384 at P = 512.
template <int P>
struct RowBlock3 {
  static constexpr int v = (3 * P / 4 > 192) ? 3 * P / 4 : 192;
};

// Position in a team buffer of bin x of an inverse (decimation-in-frequency)
// transform's output.
template <int P>
NHQG_DEV int dec3(int x) {
  return (x % 3) * P + swz(x / 3);
}

// tw3[e] = exp(-2 pi i e / M), e < 2P (the radix-3 stage); twP the
// stage-major P-point table (every third entry of the M-point one).
template <int P>
NHQG_DEV void load_twiddles3(double2* tw3, double2* twP, const double2* twg) {
  for (int i = threadIdx.x; i < 2 * P; i += blockDim.x) tw3[i] = twg[i];
  load_twiddles<P>(twP, twg, 3);
}

// x0 + w3^p x1 + w3^2p x2 with w3 = exp(-2 pi i / 3), or its conjugate (INV)
template <bool INV>
NHQG_DEV double2 radix3(double2 x0, double2 x1, double2 x2, int p) {
  constexpr double H = 0.86602540378443864676;  // sqrt(3) / 2
  const double2 s = cadd(x1, x2), d = csub(x1, x2);
  const double2 m = make_double2(x0.x - 0.5 * s.x, x0.y - 0.5 * s.y);
  const double2 hd = make_double2(-H * d.y, H * d.x);  // i (sqrt(3)/2) d
  const double2 r = ((p == 1) != INV) ? csub(m, hd) : cadd(m, hd);
  return p == 0 ? cadd(x0, s) : r;
}

// The P-point transform of the caller's group (threads t / (P/4) of a 3P/4
// team) on sub-buffer r of the team buffers a, b (natural order in, natural
// order out); returns the team buffer, a or b, that holds every group's
// result.  Opens with the group's barrier: the group's threads wrote their
// sub-buffer (a caller whose sub-buffers were written across groups takes a
// block barrier first).
template <int P, bool INV>
NHQG_DEV double2* sub_fft3(double2* a, double2* b, int t, const double2* twP) {
  constexpr int G = P / 4;
  const int r = t / G;
  team_sync<G>();
  double2* res = team_fft<P, INV>(a + r * P, b + r * P, t % G, twP);
  return res == a + r * P ? a : b;
}

// Forward-transform bin k from the three sub-transform results in buffer z
// (decimation in time): sum_r w3^(rp) w^(r k') Y_r[k'], k = k' + pP.
template <int P>
NHQG_DEV double2 dit_bin(const double2* z, const double2* tw3, int k) {
  const int kp = k % P, p = k / P;
  const double2 a1 = cmul(z[P + swz(kp)], tw3[kp]);
  const double2 a2 = cmul(z[2 * P + swz(kp)], tw3[2 * kp]);
  return radix3<false>(z[swz(kp)], a1, a2, p);
}

// Unnormalised DFT of M = 3P points by a team of 3P/4 threads (index t):
// input in natural order in a (element x at swz(x)), result in decimated
// order (dec3) in the returned buffer, a or b.  INV selects exp(+i...).
//   X[3m + p] = sum_n w_P^(nm) [w_M^(np) sum_q x[n + qP] w_3^(qp)]
// Ends with a block barrier: the caller may read any team's result.
template <int P, bool INV>
NHQG_DEV double2* team_fft3(double2* a, double2* b, int t, const double2* tw3,
                            const double2* twP) {
  constexpr int T = 3 * P / 4, G = P / 4;
  constexpr double H = 0.86602540378443864676;  // sqrt(3) / 2
  for (int n = t; n < P; n += T) {
    const double2 x0 = a[swz(n)], x1 = a[swz(n + P)], x2 = a[swz(n + 2 * P)];
    const double2 s = cadd(x1, x2), d = csub(x1, x2);
    const double2 m = make_double2(x0.x - 0.5 * s.x, x0.y - 0.5 * s.y);
    const double2 hd = make_double2(-H * d.y, H * d.x);  // i (sqrt(3)/2) d
    // forward: w_3 = -1/2 - i sqrt(3)/2
    const double2 y1 = INV ? cadd(m, hd) : csub(m, hd);
    const double2 y2 = INV ? csub(m, hd) : cadd(m, hd);
    b[swz(n)] = cadd(x0, s);
    b[P + swz(n)] = cmul(y1, twid<INV>(tw3, n));
    b[2 * P + swz(n)] = cmul(y2, twid<INV>(tw3, 2 * n));
  }
  __syncthreads();
  const int r = t / G;
  double2* res = team_fft<P, INV>(b + r * P, a + r * P, t % G, twP);
  __syncthreads();
  return res == b + r * P ? b : a;
}

// Inverse DFT along y of the packed spectrum of (p + i q) from two half
// spectra Pp, Qp (entries ky < nky, zero beyond; irfft's convention that
// only the real parts of ky = 0 and M/2 count).  When iky, Pp's entries
// are multiplied by i ky first.  p[j], q[j] at y = t + j*3P/4, times scale.
template <int P>
NHQG_DEV void inv3_pair(double2* ba, double2* bb, int t, const double2* tw3, const double2* twP,
                        int nky, const double2* Pp, const double2* Qp, const double* ky_col,
                        bool iky, double* p, double* q, double scale) {
  constexpr int M = 3 * P, T = 3 * P / 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = t + j * T;
    const int ky = (k <= M / 2) ? k : M - k;
    double2 v = make_double2(0.0, 0.0);
    if (ky < nky) {
      double2 a = Pp[ky];
      if (iky) a = cmul_i(a, ky_col[ky]);
      const double2 b = Qp ? Qp[ky] : make_double2(0.0, 0.0);
      if (k == 0 || k == M / 2) v = make_double2(a.x, b.x);
      else if (k < M / 2) v = make_double2(a.x - b.y, a.y + b.x);
      else v = make_double2(a.x + b.y, b.x - a.y);
    }
    ba[swz(k)] = v;
  }
  __syncthreads();
  const double2* r = team_fft3<P, true>(ba, bb, t, tw3, twP);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double2 z = r[dec3<P>(t + j * T)];
    p[j] = z.x * scale;
    if (q) q[j] = z.y * scale;
  }
  __syncthreads();
}

// Forward DFT along y of two real rows (a, b) given at y = t + j*3P/4;
// half spectra A, B at ky = t + s*3P/4 (s < 2), zero for ky >= nky.
template <int P>
NHQG_DEV void fwd3_pair(double2* ba, double2* bb, int t, const double2* tw3, const double2* twP,
                        const double* a, const double* b, int nky, double2 (&A)[2],
                        double2 (&B)[2]) {
  constexpr int M = 3 * P, T = 3 * P / 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int y = t + j * T;
    ba[(y % 3) * P + swz(y / 3)] = make_double2(a[j], b ? b[j] : 0.0);
  }
  __syncthreads();
  const double2* z = sub_fft3<P, false>(ba, bb, t, twP);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int ky = t + s * T;
    if (ky < nky) {
      const double2 zk = dit_bin<P>(z, tw3, ky);
      const double2 zm = dit_bin<P>(z, tw3, (M - ky) % M);
      A[s] = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
      B[s] = make_double2(0.5 * (zk.y + zm.y), 0.5 * (zm.x - zk.x));
    } else {
      A[s] = make_double2(0.0, 0.0);
      B[s] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Pass 1: inverse DFT along x of the advected inputs on the padded rows.
// Same fields and X5 slots as the 2/3 rule's col_inverse_flux; X5[f][l][x]
// [ky] for the M padded rows x.  kidx_of_x (M entries): kept row of padded
// row x, or -1 for the zero rows.
// ---------------------------------------------------------------------------
template <int P, bool SPONGE, bool JAC, bool NS = false>
__global__ void __launch_bounds__(ColBlock3<P>::v)
col_inverse_pad(int nkx, int nky, int ldc, int ng, const double2* __restrict__ qn,
                const double2* __restrict__ wn, const double2* __restrict__ thn,
                const double2* __restrict__ qc, const double2* __restrict__ wc,
                const double2* __restrict__ thc, const int* __restrict__ kidx_of_x,
                const double* __restrict__ kx_row, const double* __restrict__ ky_col,
                const double* __restrict__ inv_denom, const double2* __restrict__ twg,
                double2* __restrict__ X5, size_t fstride, int nb) {
  constexpr int M = 3 * P, T = 3 * P / 4;
  constexpr int BS = ColBlock3<P>::v, TEAMS = ColBlock3<P>::C;
  constexpr int NA = JAC ? 8 : 5;
  constexpr int NF = NA + (SPONGE ? (NS ? 1 : 3) : 0);   // NS: q' alone
  __shared__ double2 buf[2][TEAMS][M + FFT_PAD];
  __shared__ double2 tw3[2 * P], twP[P];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  int bx, l;
  if (!xcd_column_block((nky + TEAMS - 1) / TEAMS, nb, bx, l)) return;   // whole workgroup
  const int ky0 = bx * TEAMS;
  const int ky = ky0 + team;
  const bool active = ky < nky;
  load_twiddles3<P>(tw3, twP, twg);
  const size_t lb = (size_t)l * ldc;
  int mi[4];
  double kxv[4], idn[4];
  const double kyv = active ? ky_col[ky] : 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = kidx_of_x[t + j * T];
    mi[j] = (active && i >= 0) ? ky * nkx + i : -1;
    kxv[j] = mi[j] >= 0 ? kx_row[i] : 0.0;
    idn[j] = mi[j] >= 0 ? inv_denom[mi[j]] : 0.0;
  }
  auto src_of = [&](int f) { return f < 2 ? 0 : f < NA ? (JAC ? (f - 2) / 2 : f - 2) : 3 + f - NA; };
  auto valid = [&](int f) { return !(f >= NA + 1 && l >= ng); };
#pragma unroll 1
  for (int f = 0; f < NF; ++f) {
    if (!valid(f)) continue;   // block-uniform
    const int g = src_of(f);
    const double2* Psrc = g == 0 ? qn : g == 1 ? wn : g == 2 ? thn : g == 3 ? qc : g == 4 ? wc : thc;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double2 v = mi[j] >= 0 ? Psrc[lb + mi[j]] : make_double2(0.0, 0.0);
      if (f == 0) v = cmul_i(v, kyv * idn[j]);
      else if (f == 1) v = cmul_i(v, -kxv[j] * idn[j]);
      else if (JAC && f < NA && (f - 2) % 2 == 1) v = cmul_i(v, kxv[j]);
      buf[0][team][swz(t + j * T)] = v;
    }
    __syncthreads();
    const double2* res = team_fft3<P, true>(buf[0][team], buf[1][team], t, tw3, twP);
    const int par = (res == buf[0][team]) ? 0 : 1;
    double2* dst = X5 + f * fstride + (size_t)l * M * nky;
    for (int idx = tid; idx < M * TEAMS; idx += BS) {
      const int cc = idx % TEAMS, x = idx / TEAMS;
      const int kyc = ky0 + cc;
      if (kyc < nky) dst[(size_t)x * nky + kyc] = buf[par][cc][dec3<P>(x)];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Pass 2 (flux form) on the padded grid: Y6 slots as row_flux2 (G_f =
// F_y(u f), H_f = i ky F_y(v f), sponge F_y(sigma f)), rows x, x+1 of one
// level packed in each complex transform.  eta_phys, sponge_phys: M x M.
// NS (with SPONGE): the w and theta sponges act on their nodal padded rows
// and F_y(sigma f) is added into H_f, as in row_flux2; q' keeps its slot.
// ---------------------------------------------------------------------------
template <int P, bool SPONGE, bool TRAP, bool NS = false>
__global__ void __launch_bounds__(RowBlock3<P>::v)
row_flux_pad(int nky, int ng, const double2* __restrict__ X5, size_t fstride,
             double2* __restrict__ Y6, const double* __restrict__ ky_col,
             const double* __restrict__ eta_phys, const double* __restrict__ sponge_phys,
             const double2* __restrict__ twg) {
  constexpr int M = 3 * P, T = 3 * P / 4, TEAMS = RowBlock3<P>::v / T;
  __shared__ double2 buf[2][TEAMS][M + FFT_PAD];
  __shared__ double2 tw3[2 * P], twP[P];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  const int l = blockIdx.y;
  int x0 = 2 * (blockIdx.x * TEAMS + team);
  const bool active = x0 < M;
  if (!active) x0 = M - 2;  // duplicate valid rows; results are not stored
  load_twiddles3<P>(tw3, twP, twg);
  __syncthreads();
  double2* ba = buf[0][team];
  double2* bb = buf[1][team];
  const size_t r0 = ((size_t)l * M + x0) * nky, r1 = r0 + nky;
  const double sc = 1.0 / ((double)M * (double)M);
  auto phys_rows = [&](const double* src, double (&p0)[4], double (&p1)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const size_t y = t + j * T;
      p0[j] = src[(size_t)x0 * M + y];
      p1[j] = src[(size_t)(x0 + 1) * M + y];
    }
  };
  auto store_pair = [&](int slot, const double2 (&A)[2], const double2 (&B)[2], bool iky) {
    if (!active) return;
    double2* d0 = Y6 + slot * fstride + r0;
    double2* d1 = Y6 + slot * fstride + r1;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int ky = t + s * T;
      if (ky < nky) {
        d0[ky] = iky ? cmul_i(A[s], ky_col[ky]) : A[s];
        d1[ky] = iky ? cmul_i(B[s], ky_col[ky]) : B[s];
      }
    }
  };
  double u0[4], u1[4], v0[4], v1[4];
  inv3_pair<P>(ba, bb, t, tw3, twP, nky, X5 + r0, X5 + r1, ky_col, false, u0, u1, sc);
  inv3_pair<P>(ba, bb, t, tw3, twP, nky, X5 + fstride + r0, X5 + fstride + r1, ky_col, false,
               v0, v1, sc);
#pragma unroll 1
  for (int f = 0; f < 3; ++f) {
    double f0[4], f1[4];
    inv3_pair<P>(ba, bb, t, tw3, twP, nky, X5 + (2 + f) * fstride + r0,
                 X5 + (2 + f) * fstride + r1, ky_col, false, f0, f1, sc);
    if (TRAP && f == 0) {
      double e0[4], e1[4];
      phys_rows(eta_phys, e0, e1);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        f0[j] += e0[j];
        f1[j] += e1[j];
      }
    }
    double a0[4], a1[4];
    double2 A[2], B[2];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a0[j] = u0[j] * f0[j];
      a1[j] = u1[j] * f1[j];
    }
    fwd3_pair<P>(ba, bb, t, tw3, twP, a0, a1, nky, A, B);
    store_pair(2 * f, A, B, false);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a0[j] = v0[j] * f0[j];
      a1[j] = v1[j] * f1[j];
    }
    fwd3_pair<P>(ba, bb, t, tw3, twP, a0, a1, nky, A, B);
    if (NS && f > 0) {
      // H_f = i ky F_y(v f) + F_y(sigma f) on the nodal padded rows
      double s0[4], s1[4];
      phys_rows(sponge_phys, s0, s1);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a0[j] = s0[j] * f0[j];
        a1[j] = s1[j] * f1[j];
      }
      double2 A2[2], B2[2];
      fwd3_pair<P>(ba, bb, t, tw3, twP, a0, a1, nky, A2, B2);
      if (active) {
        double2* d0 = Y6 + (2 * f + 1) * fstride + r0;
        double2* d1 = Y6 + (2 * f + 1) * fstride + r1;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int ky = t + s * T;
          if (ky < nky) {
            d0[ky] = cadd(cmul_i(A[s], ky_col[ky]), A2[s]);
            d1[ky] = cadd(cmul_i(B[s], ky_col[ky]), B2[s]);
          }
        }
      }
    } else {
      store_pair(2 * f + 1, A, B, true);
    }
  }
  if (SPONGE) {
    // sigma_pad times each coefficient level of q', w, theta (spectral.py:62-74),
    // of q' alone with NS
    double s0[4], s1[4];
    phys_rows(sponge_phys, s0, s1);
#pragma unroll 1
    for (int f = 0; f < (NS ? 1 : 3); ++f) {
      if (f > 0 && l >= ng) break;  // block-uniform
      double c0[4], c1[4], a0[4], a1[4];
      inv3_pair<P>(ba, bb, t, tw3, twP, nky, X5 + (5 + f) * fstride + r0,
                   X5 + (5 + f) * fstride + r1, ky_col, false, c0, c1, sc);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a0[j] = s0[j] * c0[j];
        a1[j] = s1[j] * c1[j];
      }
      double2 A[2], B[2];
      fwd3_pair<P>(ba, bb, t, tw3, twP, a0, a1, nky, A, B);
      store_pair(6 + f, A, B, false);
    }
  }
}

// ---------------------------------------------------------------------------
// Pass 2, Jacobian form, on the padded grid (nhqg/spectral.py:155-184): per
// (level, padded row x), f_y and f_x as one packed inverse, J_f = u f_x +
// v f_y, G_f = F_y(J_f) into Y[f]; sponge products into Y[3 + f].
// eta_x, eta_y: grad(eta) on the padded grid (M x M).
// ---------------------------------------------------------------------------
template <int P, bool SPONGE, bool TRAP>
__global__ void __launch_bounds__(RowBlock3<P>::v)
row_jac_pad(int nky, int ng, const double2* __restrict__ X, size_t fstride,
            double2* __restrict__ Y, const double* __restrict__ ky_col,
            const double* __restrict__ eta_x, const double* __restrict__ eta_y,
            const double* __restrict__ sponge_phys, const double2* __restrict__ twg) {
  constexpr int M = 3 * P, T = 3 * P / 4, TEAMS = RowBlock3<P>::v / T;
  __shared__ double2 buf[2][TEAMS][M + FFT_PAD];
  __shared__ double2 tw3[2 * P], twP[P];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  const int l = blockIdx.y;
  int x = blockIdx.x * TEAMS + team;
  const bool active = x < M;
  if (!active) x = M - 1;  // duplicate a valid row; results are not stored
  load_twiddles3<P>(tw3, twP, twg);
  __syncthreads();
  double2* ba = buf[0][team];
  double2* bb = buf[1][team];
  const size_t ro = ((size_t)l * M + x) * nky;
  const double sc = 1.0 / ((double)M * (double)M);
  auto store = [&](int slot, const double2 (&G)[2]) {
    if (!active) return;
    double2* d = Y + slot * fstride + ro;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int ky = t + s * T;
      if (ky < nky) d[ky] = G[s];
    }
  };
  double u[4], v[4];
  inv3_pair<P>(ba, bb, t, tw3, twP, nky, X + ro, X + fstride + ro, ky_col, false, u, v, sc);
#pragma unroll 1
  for (int f = 0; f < 3; ++f) {
    double fy[4], fx[4];
    inv3_pair<P>(ba, bb, t, tw3, twP, nky, X + (2 + 2 * f) * fstride + ro,
                 X + (3 + 2 * f) * fstride + ro, ky_col, true, fy, fx, sc);
    if (TRAP && f == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const size_t y = (size_t)x * M + t + j * T;
        fx[j] += eta_x[y];
        fy[j] += eta_y[y];
      }
    }
    double jv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) jv[j] = u[j] * fx[j] + v[j] * fy[j];
    double2 G[2], dummy[2];
    fwd3_pair<P>(ba, bb, t, tw3, twP, jv, nullptr, nky, G, dummy);
    store(f, G);
  }
  if (SPONGE) {
    double sig[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) sig[j] = sponge_phys[(size_t)x * M + t + j * T];
#pragma unroll 1
    for (int f = 0; f < 3; ++f) {
      if (f > 0 && l >= ng) break;  // block-uniform
      double c[4], a[4];
      inv3_pair<P>(ba, bb, t, tw3, twP, nky, X + (8 + f) * fstride + ro, nullptr, ky_col, false,
                   c, nullptr, sc);
#pragma unroll
      for (int j = 0; j < 4; ++j) a[j] = sig[j] * c[j];
      double2 S[2], dummy[2];
      fwd3_pair<P>(ba, bb, t, tw3, twP, a, nullptr, nky, S, dummy);
      store(3 + f, S);
    }
  }
}

// ---------------------------------------------------------------------------
// Pass 3 on the padded grid: forward DFT along x (M points); A_f = (i kx
// F_x(G_f) + F_x(H_f)) (M/Nx)^2 at the Nx kept rows (rows[i]: padded row
// of kept row i); Jacobian form A_f = F_x(G_f) (M/Nx)^2.
// ---------------------------------------------------------------------------
template <int P, bool SPONGE, bool JAC, bool NS = false>
__global__ void __launch_bounds__(ColBlock3<P>::v)
col_forward_pad(int nkx, int nky, int ldc, int ng, const double2* __restrict__ Y6,
                size_t fstride, const int* __restrict__ rows, const double* __restrict__ kx_row,
                double out_scale, double2* __restrict__ out_q, double2* __restrict__ out_w,
                double2* __restrict__ out_th, double2* __restrict__ sp_q,
                double2* __restrict__ sp_w, double2* __restrict__ sp_th,
                const double2* __restrict__ twg, int nb) {
  constexpr int M = 3 * P, T = 3 * P / 4;
  constexpr int BS = ColBlock3<P>::v, TEAMS = ColBlock3<P>::C;
  __shared__ double2 buf[2][TEAMS][M + FFT_PAD];
  __shared__ double2 tw3[2 * P], twP[P];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  int bx, l;
  if (!xcd_column_block((nky + TEAMS - 1) / TEAMS, nb, bx, l)) return;   // whole workgroup
  const int ky0 = bx * TEAMS;
  const int ky = ky0 + team;
  const bool active = ky < nky;
  load_twiddles3<P>(tw3, twP, twg);
  constexpr int NA = JAC ? 1 : 2;
  constexpr int NG = NA + (SPONGE ? 1 : 0);
  constexpr int NQ = 3 * NG;
  constexpr int NL = M * TEAMS / BS;  // = 4 slot values gathered per thread
  int rowv[4];
  double kxv[4];
  bool in[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = t + j * T;
    in[j] = i < nkx && active;
    rowv[j] = i < nkx ? rows[i] : 0;
    kxv[j] = i < nkx ? kx_row[i] : 0.0;
  }
  // w and theta sponge slots: on their ng levels, and none with NS
  auto valid = [&](int q) { return !(q % NG == NA && q / NG > 0 && (NS || l >= ng)); };
  double2 g[4];
#pragma unroll 1
  for (int q = 0; q < NQ; ++q) {
    if (!valid(q)) continue;  // block-uniform
    const int f = q / NG, h = q % NG;
    const int slot = JAC ? (h == 0 ? f : 3 + f) : (h < 2) ? 2 * f + h : 6 + f;
    const double2* src = Y6 + slot * fstride + (size_t)l * M * nky;
#pragma unroll
    for (int it = 0; it < NL; ++it) {
      const int idx = tid + it * BS;
      const int cc = idx % TEAMS, x = idx / TEAMS;
      const int kyc = ky0 + cc;
      // decimation in time: row x to sub-buffer x % 3, position x / 3
      buf[0][cc][(x % 3) * P + swz(x / 3)] =
          (kyc < nky) ? src[(size_t)x * nky + kyc] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const double2* zb = sub_fft3<P, false>(buf[0][team], buf[1][team], t, twP);
    __syncthreads();   // a kept row's bin combines all three groups' results
    double2* const dout = f == 0 ? out_q : f == 1 ? out_w : out_th;
    double2* const dsp = f == 0 ? sp_q : f == 1 ? sp_w : sp_th;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!in[j]) continue;
      const double2 z = dit_bin<P>(zb, tw3, rowv[j]);
      const size_t o = (size_t)l * ldc + ky * nkx + t + j * T;
      if (JAC) {
        (h == 0 ? dout : dsp)[o] = cscale(z, out_scale);
      } else if (h == 0) {
        g[j] = cmul_i(z, kxv[j]);
      } else if (h == 1) {
        dout[o] = cscale(cadd(g[j], z), out_scale);
      } else {
        dsp[o] = cscale(z, out_scale);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

template <int P, bool SP, bool JAC, bool NS = false>
static void flux_pass_pad(const FluxArgs& a, int pass, hipStream_t s) {
  constexpr int M = 3 * P, T = 3 * P / 4;
  constexpr int BC = ColBlock3<P>::v, TC = ColBlock3<P>::C;
  constexpr int BR = RowBlock3<P>::v, TR = BR / T;
  const int nb = (a.nky + TC - 1) / TC * a.nq;
  const dim3 gc = xcd_grid(nb);
  if (pass == 0 || pass == 1)
    hipLaunchKernelGGL((col_inverse_pad<P, SP, JAC, NS>), gc, dim3(BC), 0, s, a.nkx, a.nky,
                       a.ldc, a.ng, a.qn, a.wn, a.thn, a.qc, a.wc, a.thc, a.kidx_of_x, a.kx_row,
                       a.ky_col, a.inv_denom, a.tw, a.X5, a.fstride, nb);
  if ((pass == 0 || pass == 2) && JAC) {
    const dim3 gr((M + TR - 1) / TR, a.nq);
    if (a.eta_x)
      hipLaunchKernelGGL((row_jac_pad<P, SP, true>), gr, dim3(BR), 0, s, a.nky, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_x, a.eta_y, a.sponge_phys, a.tw);
    else
      hipLaunchKernelGGL((row_jac_pad<P, SP, false>), gr, dim3(BR), 0, s, a.nky, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_x, a.eta_y, a.sponge_phys, a.tw);
  } else if (pass == 0 || pass == 2) {
    const dim3 g2((M / 2 + TR - 1) / TR, a.nq);   // two rows per team
    if (a.eta_phys)
      hipLaunchKernelGGL((row_flux_pad<P, SP, true, NS>), g2, dim3(BR), 0, s, a.nky, a.ng,
                         a.X5, a.fstride, a.Y6, a.ky_col, a.eta_phys, a.sponge_phys, a.tw);
    else
      hipLaunchKernelGGL((row_flux_pad<P, SP, false, NS>), g2, dim3(BR), 0, s, a.nky, a.ng,
                         a.X5, a.fstride, a.Y6, a.ky_col, a.eta_phys, a.sponge_phys, a.tw);
  }
  if (pass == 0 || pass == 3)
    hipLaunchKernelGGL((col_forward_pad<P, SP, JAC, NS>), gc, dim3(BC), 0, s, a.nkx, a.nky,
                       a.ldc, a.ng, a.Y6, a.fstride, a.rows, a.kx_row, a.out_scale, a.aq, a.aw,
                       a.ath, a.sq, a.sw, a.sth, a.tw, nb);
}

template <int P>
static hipError_t flux_pad_p(const FluxArgs& a, int pass, hipStream_t s) {
  // the layout must hold every kx row and ky <= Nx/2 (nkx = 2P, nky = P + 1)
  if (a.nkx != 2 * P || a.nky != P + 1) return hipErrorInvalidValue;
  if (a.nsp) {
    // flux form with a sponge only
    if (a.jac || !a.sponge_phys) return hipErrorInvalidValue;
    flux_pass_pad<P, true, false, true>(a, pass, s);
    return hipGetLastError();
  }
  if (a.jac) {
    if (a.sponge_phys) flux_pass_pad<P, true, true>(a, pass, s);
    else flux_pass_pad<P, false, true>(a, pass, s);
  } else {
    if (a.sponge_phys) flux_pass_pad<P, true, false>(a, pass, s);
    else flux_pass_pad<P, false, false>(a, pass, s);
  }
  return hipGetLastError();
}

hipError_t launch_flux_pad(const FluxArgs& a, int pass, hipStream_t s) {
  switch (a.N) {
    case 24: return flux_pad_p<8>(a, pass, s);
    case 48: return flux_pad_p<16>(a, pass, s);
    case 96: return flux_pad_p<32>(a, pass, s);
    case 192: return flux_pad_p<64>(a, pass, s);
    case 384: return flux_pad_p<128>(a, pass, s);
    case 768: return flux_pad_p<256>(a, pass, s);
    case 1536: return flux_pad_p<512>(a, pass, s);
    default: return hipErrorInvalidValue;
  }
}
