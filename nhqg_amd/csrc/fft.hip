// Horizontal spectral products of the NHQG explicit tendency, 2/3 rule.
//
// The reference evaluates, per vertical level, div(u f, v f) with
// u = -psi_y, v = psi_x for f in {q'+eta, w, theta} using irfft2 / rfft2 on
// the Nx grid and masks the result (nhqg/spectral.py:307-328), and adds the
// sponge products rfft2(sigma * irfft2(f)) (spectral.py:53-60).  Here the
// whole per-level chain is three fused passes over LDS-resident Stockham
// FFTs (radix 4, one radix-2 stage when log2 Nx is odd; twiddles from a
// host table, no recursion):
//
//   col_inverse : per (level, ky): inverse DFT along x of
//                 U = i ky psi, V = -i kx psi, Q, W, Theta (psi = -q / |k|^2
//                 is formed on the fly from the nodal q)           -> X5
//   row_flux2   : per (level, row pair x, x+1): c2r along y (the two rows
//                 packed in one complex FFT), pointwise products
//                 u(q+eta), v(q+eta), uw, vw, u theta, v theta, sigma q,
//                 sigma w, sigma theta, r2c along y (row pairs packed);
//                 applies i ky and the
//                 sponge                                           -> Y6
//   col_forward : per (level, ky): forward DFT along x, i kx G + H,
//                 gather of the kept rows (2/3 band)                -> A
//
// A "team" of Nx/4 threads owns one transform.  Row passes run 256-thread
// blocks (1024/Nx teams); the column passes run ColBlock<N> threads, so a
// block holds at least 4 columns and their transposed stores and gathers
// move >= 64-byte runs (4 columns at Nx = 1024, 8 at 512 and 256).  The
// trap background is added to q' in physical space.
// The sponge rides in the same passes but, as in the reference, acts on
// the coefficient levels of q', w and theta (X5 slots 5-7, Y6 slots 6-8):
// acting on nodal values instead is equal in exact arithmetic, but the
// transform back to coefficients then leaves roundoff of the large nodal
// products in the tiny high Chebyshev coefficients, which d/dZ amplifies.
// That matters for q' (its nodal values carry the barotropic part: ~6e-12
// of w per step), not for w and theta (~1e-15): at N = 256 in flux form
// (NS) their sponges act on the nodal rows inside the row pass (fft_r16.hip)
// and only q' keeps a coefficient slot (X5 slot 5, Y6 slot 6).
#include "fft_team.h"

#ifndef NHQG_ROW16
#define NHQG_ROW16 1   // the 16-thread row pass of fft_r16.hip at N = 256
#endif

// Threads per column-pass workgroup: 8 columns (128-byte runs of every
// X5 / Y6 row) at N = 256 and 512, 4 at N = 1024, and 256 threads (1024/N
// columns) below.  At N = 256, 4-column blocks wrote 21% more than the
// algorithmic bytes and made the inverse pass 0.67 instead of 0.55 ms per
// step.
template <int N>
struct ColBlock {
  static constexpr int v = (N >= 512) ? 1024 : (N == 256 ? 512 : 256);
};

// ---------------------------------------------------------------------------
// Pass 1: inverse DFT along x of the advected inputs.
// qn/wn/thn: nodal spectral fields (level-major, row stride ldc double2).
// Flux form: X5[f][l][x][ky], f = U, V, Q, W, Theta (+ sponge q, w, theta
// coefficient levels).  Jacobian form: U, V, then per field f and i kx f
// (Q, iKx Q, W, iKx W, Theta, iKx Theta; i ky is applied along y in pass 2),
// then the sponge inputs.
// ---------------------------------------------------------------------------
template <int N, bool SPONGE, bool JAC, bool NS = false>
__global__ void __launch_bounds__(ColBlock<N>::v)
col_inverse_flux(int nkx, int nky, int ldc, int ng, const double2* __restrict__ qn,
                 const double2* __restrict__ wn, const double2* __restrict__ thn,
                 const double2* __restrict__ qc, const double2* __restrict__ wc,
                 const double2* __restrict__ thc, const int* __restrict__ kidx_of_x,
                 const double* __restrict__ kx_row, const double* __restrict__ ky_col,
                 const double* __restrict__ inv_denom, const double2* __restrict__ twg,
                 double2* __restrict__ X5, size_t fstride, int nb) {
  constexpr int BS = ColBlock<N>::v;
  constexpr int T = N / 4, TEAMS = BS / T;
  constexpr int NA = JAC ? 8 : 5;   // advection inputs
  // with NS only q' has a coefficient-level sponge slot
  constexpr int NF = NA + (SPONGE ? (NS ? 1 : 3) : 0);
  __shared__ double2 buf[2][TEAMS][N + FFT_PAD];
  __shared__ double2 tw[N];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  int bx, l;
  if (!xcd_column_block((nky + TEAMS - 1) / TEAMS, nb, bx, l)) return;   // whole workgroup
  const int ky0 = bx * TEAMS;
  const int ky = ky0 + team;
  const bool active = ky < nky;
  load_twiddles<N>(tw, twg);
  const size_t lb = (size_t)l * ldc;
  // mode index, kx and 1/|k|^2 of this thread's 4 points, the same for every
  // field (-1: x outside the kept rows, or an idle team)
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
  // source of field f: 0..2 = nodal q, w, theta; 3..5 = coefficient q, w, theta
  auto src_of = [&](int f) { return f < 2 ? 0 : f < NA ? (JAC ? (f - 2) / 2 : f - 2) : 3 + f - NA; };
  auto load = [&](int g, double2 (&v)[4]) {
    const double2* P = g == 0 ? qn : g == 1 ? wn : g == 2 ? thn : g == 3 ? qc : g == 4 ? wc : thc;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = mi[j] >= 0 ? P[lb + mi[j]] : make_double2(0.0, 0.0);
  };
  // w and theta coefficient fields have ng < Nz+1 levels (block-uniform)
  auto valid = [&](int f) { return !(f >= NA + 1 && l >= ng); };
  // The next field's column is loaded into registers before this field's
  // transform and stores, so its load latency overlaps them; fields with
  // the same source (U, V and Q from q; the Jacobian form's f and i kx f)
  // share one load.
  double2 craw[4], nraw[4];
  load(src_of(0), nraw);
#pragma unroll 1
  for (int f = 0; f < NF; ++f) {
    if (!valid(f)) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) craw[j] = nraw[j];
    int fn = f + 1;
    while (fn < NF && !valid(fn)) ++fn;
    if (fn < NF && src_of(fn) != src_of(f)) load(src_of(fn), nraw);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double2 v = craw[j];
      if (f == 0) v = cmul_i(v, kyv * idn[j]);
      else if (f == 1) v = cmul_i(v, -kxv[j] * idn[j]);
      else if (JAC && f < NA && (f - 2) % 2 == 1) v = cmul_i(v, kxv[j]);
      buf[0][team][swz(t + j * T)] = v;
    }
    __syncthreads();
    double2* res = team_fft<N, true>(buf[0][team], buf[1][team], t, tw);
    const int par = (res == buf[0][team]) ? 0 : 1;
    __syncthreads();   // the store below reads every team's result
    double2* dst = X5 + f * fstride + (size_t)l * N * nky;
    for (int idx = tid; idx < N * TEAMS; idx += BS) {
      const int cc = idx % TEAMS, x = idx / TEAMS;
      const int kyc = ky0 + cc;
      if (kyc < nky) dst[(size_t)x * nky + kyc] = buf[par][cc][swz(x)];
    }
    __syncthreads();
  }
}

// Build the packed spectrum Z of (p + i q) along y from two half spectra
// (ky < nky) so that its inverse DFT is p(y) + i q(y), with irfft's
// convention that only the real parts of the ky=0 and ky=N/2 entries count.
template <int N>
NHQG_DEV void build_packed(double2* z, int t, int nky, const double2* P, const double2* Q) {
  constexpr int T = N / 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = t + j * T;
    const int ky = (k <= N / 2) ? k : N - k;
    double2 v = make_double2(0.0, 0.0);
    if (ky < nky) {
      const double2 p = P[ky];
      const double2 q = Q ? Q[ky] : make_double2(0.0, 0.0);
      if (k == 0 || k == N / 2) v = make_double2(p.x, q.x);
      else if (k < N / 2) v = make_double2(p.x - q.y, p.y + q.x);
      else v = make_double2(p.x + q.y, q.x - p.y);
    }
    z[swz(k)] = v;
  }
}

// Forward DFT along y of two real rows (a, b) held as 4 values per thread
// at y = t + j*N/4; returns half-spectrum slots ky = t + s*N/4 (s < 3).
template <int N>
NHQG_DEV void fwd_pair(double2* ba, double2* bb, int t, const double2* tw, const double* a,
                       const double* b, int nky, double2* A, double2* B) {
  constexpr int T = N / 4;
  const double2* r;
  if constexpr (Log2<N>::v % 2 == 0 && T <= 64) {
    // The team is one wavefront (or part of one) and the bins come back in
    // registers: bin t + s*T of this lane pairs with bin N - t - s*T =
    // (3-s)*T + (T-t), register 3-s of lane T-t (for t = 0, register
    // (4-s) % 4 of this lane), fetched by cross-lane shuffles, no LDS.
    double2 v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = make_double2(a[j], b ? b[j] : 0.0);
    team_fft_reg<N, false>(v, ba, bb, t, tw);
    const int lane = threadIdx.x & 63;
    const int src = (lane & ~(T - 1)) | ((T - t) & (T - 1));
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      double2 zm = make_double2(__shfl(v[3 - s].x, src, 64), __shfl(v[3 - s].y, src, 64));
      if (t == 0) zm = v[(4 - s) & 3];
      const double2 zk = v[s];
      if (t + s * T < nky) {
        A[s] = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
        B[s] = make_double2(0.5 * (zk.y + zm.y), 0.5 * (zm.x - zk.x));
      } else {
        A[s] = make_double2(0.0, 0.0);
        B[s] = make_double2(0.0, 0.0);
      }
    }
    return;
  } else if constexpr (Log2<N>::v % 2 == 0) {
    // the bins of the thread come back in registers; the pairing with bin
    // N - ky needs every bin of the team, through LDS
    double2 v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = make_double2(a[j], b ? b[j] : 0.0);
    team_fft_reg<N, false>(v, ba, bb, t, tw);
#pragma unroll
    for (int j = 0; j < 4; ++j) ba[swz(t + j * T)] = v[j];
    team_sync<T>();
    r = ba;
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) ba[swz(t + j * T)] = make_double2(a[j], b ? b[j] : 0.0);
    team_sync<T>();
    r = team_fft<N, false>(ba, bb, t, tw);
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int ky = t + s * T;
    if (ky < nky) {
      const double2 zk = r[swz(ky)];
      const double2 zm = r[swz((N - ky) & (N - 1))];
      A[s] = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
      B[s] = make_double2(0.5 * (zk.y + zm.y), 0.5 * (zm.x - zk.x));
    } else {
      A[s] = make_double2(0.0, 0.0);
      B[s] = make_double2(0.0, 0.0);
    }
  }
  team_sync<T>();
}

template <int N>
NHQG_DEV void inv_pair(double2* ba, double2* bb, int t, const double2* tw, int nky,
                       const double2* P, const double2* Q, double* p, double* q, double scale) {
  constexpr int T = N / 4;
  build_packed<N>(ba, t, nky, P, Q);
  team_sync<T>();
  const double2* r = team_fft<N, true>(ba, bb, t, tw);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double2 z = r[swz(t + j * T)];
    p[j] = z.x * scale;
    if (q) q[j] = z.y * scale;
  }
  team_sync<T>();
}

// Half spectra P, Q (entries ky < nky) at the packed slots k = t + j*N/4 a
// thread fills, loaded into registers ahead of use.
template <int N>
struct Half4 {
  double2 p[4], q[4];
};

template <int N>
NHQG_DEV void load_half(Half4<N>& h, int t, int nky, const double2* P, const double2* Q) {
  constexpr int T = N / 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = t + j * T;
    const int ky = (k <= N / 2) ? k : N - k;
    const bool in = ky < nky;
    h.p[j] = in ? P[ky] : make_double2(0.0, 0.0);
    h.q[j] = (in && Q) ? Q[ky] : make_double2(0.0, 0.0);
  }
}

// inverse of the packed spectrum built from registers (irfft convention on
// ky = 0 and ky = N/2): p(y) + i q(y) at y = t + j*N/4
template <int N>
NHQG_DEV void inv_pair_regs(double2* ba, double2* bb, int t, const double2* tw, const Half4<N>& h,
                            double* p, double* q, double scale) {
  constexpr int T = N / 4;
  double2 v[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = t + j * T;
    const double2 a = h.p[j], b = h.q[j];
    if (k == 0 || k == N / 2) v[j] = make_double2(a.x, b.x);
    else if (k < N / 2) v[j] = make_double2(a.x - b.y, a.y + b.x);
    else v[j] = make_double2(a.x + b.y, b.x - a.y);
  }
  if constexpr (Log2<N>::v % 2 == 0) {
    team_fft_reg<N, true>(v, ba, bb, t, tw);
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) ba[swz(t + j * T)] = v[j];
    team_sync<T>();
    const double2* r = team_fft<N, true>(ba, bb, t, tw);
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = r[swz(t + j * T)];
    team_sync<T>();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    p[j] = v[j].x * scale;
    if (q) q[j] = v[j].y * scale;
  }
}

// ---------------------------------------------------------------------------
// Pass 2 (flux form): Y6[g][l][x][ky], g = Gq, Hq, Gw, Hw, Gth, Hth with
// G_f = F_y(u f) and H_f = i ky F_y(v f) + F_y(sigma f).  Two rows per team:
// rows x and x+1 of one level share every packed transform -- u, v, each
// field and each sponge input as (row x) + i (row x+1) inverses, and the
// products as forward pairs.  The two members of a pair are the same field
// at the same level, so they are of equal scale and the unpacking's
// roundoff leak is harmless.  Fields are never packed with each other: the
// leak across fields of very different size (a strong barotropic q next to
// a weak w) is not.  17 complex transforms per two rows instead of the 26 of
// a one-row pass that packs u with v and u f with v f.  Each half-spectrum
// is loaded just before its transform: loading one transform ahead costs
// 44 VGPRs and drops occupancy from 3 to 2 waves per SIMD, which loses more
// than the hidden load latency gains.  The pass is bound by LDS bandwidth;
// its transforms start and end in registers (team_fft_reg) where log2 N is
// even, and the forward transforms pair bins N - k across lanes by shuffles.
// The kernel is held at 3 waves per SIMD (amdgpu_waves_per_eu): 15% faster
// than 2 waves (at N = 1024 it spills a few registers to stay there).
// ---------------------------------------------------------------------------
// With NS (nodal sponge, with SPONGE) sigma acts on the nodal w and theta
// rows the products already hold and F_y(sigma f) is added into H_f; only
// q' keeps its coefficient-level sponge slot (X5 slot 5 -> Y6 slot 6).
template <int N, bool SPONGE, bool TRAP, bool NS = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
row_flux2(int nky, int ng, const double2* __restrict__ X5, size_t fstride, double2* __restrict__ Y6,
          const double* __restrict__ ky_col, const double* __restrict__ eta_phys,
          const double* __restrict__ sponge_phys, const double2* __restrict__ twg) {
  constexpr int T = N / 4, TEAMS = 256 / T;
  __shared__ double2 buf[2][TEAMS][N + FFT_PAD];
  __shared__ double2 tw[N];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  const int l = blockIdx.y;
  int x0 = 2 * (blockIdx.x * TEAMS + team);
  const bool active = x0 < N;
  if (!active) x0 = N - 2;  // duplicate valid rows; results are not stored
  load_twiddles<N>(tw, twg);
  __syncthreads();
  double2* ba = buf[0][team];
  double2* bb = buf[1][team];
  const size_t r0 = ((size_t)l * N + x0) * nky, r1 = r0 + nky;
  const double sc = 1.0 / ((double)N * (double)N);
  // physical-space trap and sponge rows: loaded up front for N <= 256; at
  // N >= 512 just before use, which keeps them out of the register peak
  // (at N = 1024 holding them spilled 52 B per lane)
  constexpr bool LATE = N >= 512;
  double e0[4], e1[4], s0[4], s1[4];
  auto phys_rows = [&](const double* src, double (&p0)[4], double (&p1)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const size_t y = t + j * T;
      p0[j] = src[(size_t)x0 * N + y];
      p1[j] = src[(size_t)(x0 + 1) * N + y];
    }
  };
  if (!LATE) {
    if (TRAP) phys_rows(eta_phys, e0, e1);
    if (SPONGE) phys_rows(sponge_phys, s0, s1);
  }
  auto store_pair = [&](int slot, const double2 (&A)[3], const double2 (&B)[3], bool iky) {
    if (!active) return;
    double2* d0 = Y6 + slot * fstride + r0;
    double2* d1 = Y6 + slot * fstride + r1;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int ky = t + q * T;
      if (ky < nky) {
        d0[ky] = iky ? cmul_i(A[q], ky_col[ky]) : A[q];
        d1[ky] = iky ? cmul_i(B[q], ky_col[ky]) : B[q];
      }
    }
  };
  Half4<N> hc;
  load_half<N>(hc, t, nky, X5 + 0 * fstride + r0, X5 + 0 * fstride + r1);
  double u0[4], u1[4], v0[4], v1[4];
  inv_pair_regs<N>(ba, bb, t, tw, hc, u0, u1, sc);
  load_half<N>(hc, t, nky, X5 + 1 * fstride + r0, X5 + 1 * fstride + r1);
  inv_pair_regs<N>(ba, bb, t, tw, hc, v0, v1, sc);
#pragma unroll 1
  for (int f = 0; f < 3; ++f) {
    load_half<N>(hc, t, nky, X5 + (2 + f) * fstride + r0, X5 + (2 + f) * fstride + r1);
    double f0[4], f1[4];
    inv_pair_regs<N>(ba, bb, t, tw, hc, f0, f1, sc);
    if (TRAP && f == 0) {
      if (LATE) phys_rows(eta_phys, e0, e1);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        f0[j] += e0[j];
        f1[j] += e1[j];
      }
    }
    double a0[4], a1[4];
    double2 A[3], B[3];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a0[j] = u0[j] * f0[j];
      a1[j] = u1[j] * f1[j];
    }
    fwd_pair<N>(ba, bb, t, tw, a0, a1, nky, A, B);
    store_pair(2 * f, A, B, false);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a0[j] = v0[j] * f0[j];
      a1[j] = v1[j] * f1[j];
    }
    fwd_pair<N>(ba, bb, t, tw, a0, a1, nky, A, B);
    if (NS && f > 0) {
      // H_f = i ky F_y(v f) + F_y(sigma f) on the nodal rows
      if (LATE) phys_rows(sponge_phys, s0, s1);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a0[j] = s0[j] * f0[j];
        a1[j] = s1[j] * f1[j];
      }
      double2 A2[3], B2[3];
      fwd_pair<N>(ba, bb, t, tw, a0, a1, nky, A2, B2);
      if (active) {
        double2* d0 = Y6 + (2 * f + 1) * fstride + r0;
        double2* d1 = Y6 + (2 * f + 1) * fstride + r1;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const int ky = t + q * T;
          if (ky < nky) {
            d0[ky] = cadd(cmul_i(A[q], ky_col[ky]), A2[q]);
            d1[ky] = cadd(cmul_i(B[q], ky_col[ky]), B2[q]);
          }
        }
      }
    } else {
      store_pair(2 * f + 1, A, B, true);
    }
  }
  if (SPONGE) {
    // sigma times each coefficient level of q', w, theta (spectral.py:53-60),
    // of q' alone with NS
    if (LATE) phys_rows(sponge_phys, s0, s1);
#pragma unroll 1
    for (int f = 0; f < (NS ? 1 : 3); ++f) {
      if (f > 0 && l >= ng) break;  // block-uniform
      load_half<N>(hc, t, nky, X5 + (5 + f) * fstride + r0, X5 + (5 + f) * fstride + r1);
      double c0[4], c1[4], a0[4], a1[4];
      inv_pair_regs<N>(ba, bb, t, tw, hc, c0, c1, sc);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a0[j] = s0[j] * c0[j];
        a1[j] = s1[j] * c1[j];
      }
      double2 A[3], B[3];
      fwd_pair<N>(ba, bb, t, tw, a0, a1, nky, A, B);
      store_pair(6 + f, A, B, false);
    }
  }
}

// ---------------------------------------------------------------------------
// Pass 2, Jacobian form (nhqg/spectral.py:281-304): per (level, x) row,
// f_y = F_y^-1(i ky Q_f) and f_x = F_y^-1(iKx Q_f) as one packed inverse,
// J_f = u f_x + v f_y (= psi_x f_y - psi_y f_x), and G_f = F_y(J_f) into
// Y[f]; sponge products into Y[3 + f].  The trap adds grad(eta) to q's
// gradient (J(psi, q + eta)).
// ---------------------------------------------------------------------------
template <int N, bool SPONGE, bool TRAP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
row_jac(int nky, int ng, const double2* __restrict__ X, size_t fstride, double2* __restrict__ Y,
        const double* __restrict__ ky_col, const double* __restrict__ eta_x,
        const double* __restrict__ eta_y, const double* __restrict__ sponge_phys,
        const double2* __restrict__ twg) {
  constexpr int T = N / 4, TEAMS = 256 / T;
  __shared__ double2 buf[2][TEAMS][N + FFT_PAD];
  __shared__ double2 tw[N];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  const int l = blockIdx.y;
  int x = blockIdx.x * TEAMS + team;
  const bool active = x < N;
  if (!active) x = N - 1;  // duplicate a valid row; results are not stored
  load_twiddles<N>(tw, twg);
  __syncthreads();
  double2* ba = buf[0][team];
  double2* bb = buf[1][team];
  const size_t ro = ((size_t)l * N + x) * nky;
  const double sc = 1.0 / ((double)N * (double)N);
  double ex[4], ey[4], sig[4], kyv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    ex[j] = TRAP ? eta_x[(size_t)x * N + t + j * T] : 0.0;
    ey[j] = TRAP ? eta_y[(size_t)x * N + t + j * T] : 0.0;
    sig[j] = SPONGE ? sponge_phys[(size_t)x * N + t + j * T] : 0.0;
    const int k = t + j * T;
    const int kk = (k <= N / 2) ? k : N - k;
    kyv[j] = kk < nky ? ky_col[kk] : 0.0;
  }
  Half4<N> hc, hn;
  load_half<N>(hc, t, nky, X + 0 * fstride + ro, X + 1 * fstride + ro);
  load_half<N>(hn, t, nky, X + 2 * fstride + ro, X + 3 * fstride + ro);
  double u[4], v[4];
  inv_pair_regs<N>(ba, bb, t, tw, hc, u, v, sc);
#pragma unroll 1
  for (int f = 0; f < 3; ++f) {
    hc = hn;
    if (f < 2)
      load_half<N>(hn, t, nky, X + (4 + 2 * f) * fstride + ro, X + (5 + 2 * f) * fstride + ro);
    else if (SPONGE)
      load_half<N>(hn, t, nky, X + 8 * fstride + ro, nullptr);
    // i ky on the half spectrum of f (entries stored at |ky|)
#pragma unroll
    for (int j = 0; j < 4; ++j) hc.p[j] = cmul_i(hc.p[j], kyv[j]);
    double fy[4], fx[4];
    inv_pair_regs<N>(ba, bb, t, tw, hc, fy, fx, sc);
    if (TRAP && f == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        fx[j] += ex[j];
        fy[j] += ey[j];
      }
    }
    double jv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) jv[j] = u[j] * fx[j] + v[j] * fy[j];
    double2 G[3], dummy[3];
    fwd_pair<N>(ba, bb, t, tw, jv, nullptr, nky, G, dummy);
    if (active) {
      double2* dg = Y + f * fstride + ro;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int ky = t + s * T;
        if (ky < nky) dg[ky] = G[s];
      }
    }
  }
  if (SPONGE) {
    // sigma times each coefficient level of q', w, theta (spectral.py:53-60)
#pragma unroll 1
    for (int f = 0; f < 3; ++f) {
      if (f > 0 && l >= ng) break;  // block-uniform
      hc = hn;
      if (f < 2 && l < ng) load_half<N>(hn, t, nky, X + (9 + f) * fstride + ro, nullptr);
      double c[4], a[4];
      inv_pair_regs<N>(ba, bb, t, tw, hc, c, nullptr, sc);
#pragma unroll
      for (int j = 0; j < 4; ++j) a[j] = sig[j] * c[j];
      double2 S[3], dummy[3];
      fwd_pair<N>(ba, bb, t, tw, a, nullptr, nky, S, dummy);
      if (active) {
        double2* ds = Y + (3 + f) * fstride + ro;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const int ky = t + s * T;
          if (ky < nky) ds[ky] = S[s];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Pass 3: forward DFT along x; A_f(kx, ky) = i kx F_x(G_f) + F_x(H_f) at the
// kept rows (Jacobian form: A_f = F_x(G_f)), times the output mask.  Writes
// A into the nodal buffers.
// ---------------------------------------------------------------------------
template <int N, bool SPONGE, bool JAC, bool NS = false>
__global__ void __launch_bounds__(ColBlock<N>::v)
col_forward_flux(int nkx, int nky, int ldc, int ng, const double2* __restrict__ Y6,
                 size_t fstride, const int* __restrict__ rows, const double* __restrict__ kx_row,
                 const double* __restrict__ mask, double2* __restrict__ out_q,
                 double2* __restrict__ out_w, double2* __restrict__ out_th,
                 double2* __restrict__ sp_q, double2* __restrict__ sp_w,
                 double2* __restrict__ sp_th, const double2* __restrict__ twg, int nb) {
  constexpr int BS = ColBlock<N>::v;
  constexpr int T = N / 4, TEAMS = BS / T;
  __shared__ double2 buf[2][TEAMS][N + FFT_PAD];
  __shared__ double2 tw[N];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  int bx, l;
  if (!xcd_column_block((nky + TEAMS - 1) / TEAMS, nb, bx, l)) return;   // whole workgroup
  const int ky0 = bx * TEAMS;
  const int ky = ky0 + team;
  const bool active = ky < nky;
  load_twiddles<N>(tw, twg);
  double2* outs[3] = {out_q, out_w, out_th};
  double2* sps[3] = {sp_q, sp_w, sp_th};
  constexpr int NA = JAC ? 1 : 2;     // G (and H) per field
  constexpr int NG = NA + (SPONGE ? 1 : 0);
  constexpr int NQ = 3 * NG;
  constexpr int NL = N * TEAMS / BS;  // slot values gathered per thread
  // output rows, kx and mask of this thread's modes, the same for every field
  int rowv[4];
  double kxv[4], mkv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = t + j * T;
    const bool in = i < nkx;
    rowv[j] = in ? rows[i] : 0;
    kxv[j] = in ? kx_row[i] : 0.0;
    mkv[j] = (in && active) ? mask[ky * nkx + i] : 0.0;
  }
  // the sponge slot of w and theta exists only on their ng levels (and not
  // at all with NS, where their sponge is part of the H slot)
  auto valid = [&](int q) { return !(q % NG == NA && q / NG > 0 && (NS || l >= ng)); };
  auto gather = [&](int q, double2 (&v)[NL]) {
    const int f = q / NG, h = q % NG;
    const int slot = JAC ? (h == 0 ? f : 3 + f) : (h < 2) ? 2 * f + h : 6 + f;
    const double2* src = Y6 + slot * fstride + (size_t)l * N * nky;
#pragma unroll
    for (int it = 0; it < NL; ++it) {
      const int idx = tid + it * BS;
      const int cc = idx % TEAMS, x = idx / TEAMS;
      const int kyc = ky0 + cc;
      v[it] = (kyc < nky) ? src[(size_t)x * nky + kyc] : make_double2(0.0, 0.0);
    }
  };
  double2 cur[NL], nxt[NL], g[4];
  gather(0, nxt);
#pragma unroll 1
  for (int q = 0; q < NQ; ++q) {
    if (!valid(q)) continue;  // block-uniform
    const int f = q / NG, h = q % NG;
#pragma unroll
    for (int it = 0; it < NL; ++it) cur[it] = nxt[it];
    int qn = q + 1;
    while (qn < NQ && !valid(qn)) ++qn;
    if (qn < NQ) gather(qn, nxt);
#pragma unroll
    for (int it = 0; it < NL; ++it) {
      const int idx = tid + it * BS;
      buf[0][idx % TEAMS][swz(idx / TEAMS)] = cur[it];
    }
    __syncthreads();
    const double2* r = team_fft<N, false>(buf[0][team], buf[1][team], t, tw);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = t + j * T;
      if (i < nkx) {
        const double2 z = r[swz(rowv[j])];
        const int m = ky * nkx + i;
        if (JAC) {
          if (active) (h == 0 ? outs : sps)[f][(size_t)l * ldc + m] = cscale(z, mkv[j]);
        } else if (h == 0) {
          g[j] = cmul_i(z, kxv[j]);
        } else if (h == 1) {
          if (active) outs[f][(size_t)l * ldc + m] = cscale(cadd(g[j], z), mkv[j]);
        } else if (active) {
          sps[f][(size_t)l * ldc + m] = cscale(z, mkv[j]);
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Velocity magnitude for the max_speed diagnostic: per (level, x) row,
// c2r of U = i ky psi, V = -i kx psi (from X5 slots 0, 1) and a row max of
// sqrt(u^2 + v^2).  rowmax[l*N + x].
// ---------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(256)
row_speed(int nky, const double2* __restrict__ X5, size_t fstride, double* __restrict__ rowmax,
          const double2* __restrict__ twg) {
  constexpr int T = N / 4, TEAMS = 256 / T;
  __shared__ double2 buf[2][TEAMS][N + FFT_PAD];
  __shared__ double2 tw[N];
  __shared__ double red[256];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  const int l = blockIdx.y;
  int x = blockIdx.x * TEAMS + team;
  const bool active = x < N;
  if (!active) x = N - 1;
  load_twiddles<N>(tw, twg);
  __syncthreads();
  const size_t ro = ((size_t)l * N + x) * nky;
  const double sc = 1.0 / ((double)N * (double)N);
  double u[4], v[4];
  inv_pair<N>(buf[0][team], buf[1][team], t, tw, nky, X5 + ro, X5 + fstride + ro, u, v, sc);
  double mx = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) mx = fmax(mx, sqrt(u[j] * u[j] + v[j] * v[j]));
  red[tid] = mx;
  __syncthreads();
  if (t == 0 && active) {
    double r = red[tid];
    for (int i = 1; i < T; ++i) r = fmax(r, red[tid + i]);
    rowmax[(size_t)l * N + x] = r;
  }
}

// max |w|, |theta| and |w theta| over each grid row x of one level, from the
// column transforms of the nodal w and theta that the speed pass leaves in
// X5 fields 3 and 4 (compute_diagnostics' max_w, max_theta, max_tw,
// nhqg/diagnostics.py:640-643); rowmax[k * rstride + l * N + x], k = w, theta,
// w theta.  A NaN anywhere in a row makes that row's maxima NaN.
template <int N>
__global__ void __launch_bounds__(256)
row_wth(int nky, const double2* __restrict__ X5, size_t fstride, double* __restrict__ rowmax,
        size_t rstride, const double2* __restrict__ twg) {
  constexpr int T = N / 4, TEAMS = 256 / T;
  __shared__ double2 buf[2][TEAMS][N + FFT_PAD];
  __shared__ double2 tw[N];
  __shared__ double red[3][256];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  const int l = blockIdx.y;
  int x = blockIdx.x * TEAMS + team;
  const bool active = x < N;
  if (!active) x = N - 1;
  load_twiddles<N>(tw, twg);
  __syncthreads();
  const size_t ro = ((size_t)l * N + x) * nky;
  const double sc = 1.0 / ((double)N * (double)N);
  double w[4], th[4];
  inv_pair<N>(buf[0][team], buf[1][team], t, tw, nky, X5 + 3 * fstride + ro,
              X5 + 4 * fstride + ro, w, th, sc);
  auto nmax = [](double a, double b) { return (a != a || b != b) ? a + b : (a > b ? a : b); };
  double mw = 0.0, mt = 0.0, mwt = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    mw = nmax(mw, fabs(w[j]));
    mt = nmax(mt, fabs(th[j]));
    mwt = nmax(mwt, fabs(w[j] * th[j]));
  }
  red[0][tid] = mw;
  red[1][tid] = mt;
  red[2][tid] = mwt;
  __syncthreads();
  if (t == 0 && active) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double r = red[k][tid];
      for (int i = 1; i < T; ++i) r = nmax(r, red[k][tid + i]);
      rowmax[k * rstride + (size_t)l * N + x] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------



template <int N, bool SP, bool JAC, bool NS = false>
static void flux_pass(const FluxArgs& a, int pass, hipStream_t s) {
  constexpr int TEAMS = 1024 / N;                       // row passes
  constexpr int BC = ColBlock<N>::v, TC = 4 * BC / N;   // column passes
  const int nb = (a.nky + TC - 1) / TC * a.nq;
  const dim3 gc = xcd_grid(nb);
  dim3 gr((N + TEAMS - 1) / TEAMS, a.nq);
  if (pass == 0 || pass == 1)
    hipLaunchKernelGGL((col_inverse_flux<N, SP, JAC, NS>), gc, dim3(BC), 0, s, a.nkx, a.nky,
                       a.ldc, a.ng, a.qn, a.wn, a.thn, a.qc, a.wc, a.thc, a.kidx_of_x, a.kx_row,
                       a.ky_col, a.inv_denom, a.tw, a.X5, a.fstride, nb);
  if ((pass == 0 || pass == 2) && JAC) {
    if (a.eta_x)
      hipLaunchKernelGGL((row_jac<N, SP, true>), gr, dim3(256), 0, s, a.nky, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_x, a.eta_y, a.sponge_phys, a.tw);
    else
      hipLaunchKernelGGL((row_jac<N, SP, false>), gr, dim3(256), 0, s, a.nky, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_x, a.eta_y, a.sponge_phys, a.tw);
  } else if ((pass == 0 || pass == 2) && !(NHQG_ROW16 && N == 256 && launch_row_flux16(a, s))) {
    dim3 g2((N / 2 + TEAMS - 1) / TEAMS, a.nq);   // two rows per team
    if (a.eta_phys)
      hipLaunchKernelGGL((row_flux2<N, SP, true, NS>), g2, dim3(256), 0, s, a.nky, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_phys, a.sponge_phys, a.tw);
    else
      hipLaunchKernelGGL((row_flux2<N, SP, false, NS>), g2, dim3(256), 0, s, a.nky, a.ng,
                         a.X5, a.fstride, a.Y6, a.ky_col, a.eta_phys, a.sponge_phys, a.tw);
  }
  if (pass == 0 || pass == 3)
    hipLaunchKernelGGL((col_forward_flux<N, SP, JAC, NS>), gc, dim3(BC), 0, s, a.nkx, a.nky,
                       a.ldc, a.ng, a.Y6, a.fstride, a.rows, a.kx_row, a.mask, a.aq, a.aw, a.ath,
                       a.sq, a.sw, a.sth, a.tw, nb);
}

template <int N>
static hipError_t flux_n(const FluxArgs& a, int pass, hipStream_t s) {
  if (a.nsp) {
    // flux form with a sponge only
    if (a.jac || !a.sponge_phys) return hipErrorInvalidValue;
    flux_pass<N, true, false, true>(a, pass, s);
    return hipGetLastError();
  }
  if (a.jac) {
    if (a.sponge_phys) flux_pass<N, true, true>(a, pass, s);
    else flux_pass<N, false, true>(a, pass, s);
  } else {
    if (a.sponge_phys) flux_pass<N, true, false>(a, pass, s);
    else flux_pass<N, false, false>(a, pass, s);
  }
  return hipGetLastError();
}

hipError_t launch_flux(const FluxArgs& a, int pass, hipStream_t s) {
  switch (a.N) {
    case 16: return flux_n<16>(a, pass, s);
    case 32: return flux_n<32>(a, pass, s);
    case 64: return flux_n<64>(a, pass, s);
    case 128: return flux_n<128>(a, pass, s);
    case 256: return flux_n<256>(a, pass, s);
    case 512: return flux_n<512>(a, pass, s);
    case 1024: return flux_n<1024>(a, pass, s);
    default: return a.N % 3 == 0 ? launch_flux_pad(a, pass, s) : hipErrorInvalidValue;
  }
}

template <int N>
static hipError_t speed_n(const FluxArgs& a, double* rowmax, hipStream_t s) {
  constexpr int TEAMS = 1024 / N;
  constexpr int BC = ColBlock<N>::v, TC = 4 * BC / N;
  const int nb = (a.nky + TC - 1) / TC * a.nq;
  // X5 fields 0, 1: u, v from psi; 3, 4: the nodal w and theta when given
  const double2* wn = a.wn ? a.wn : a.qn;
  const double2* thn = a.thn ? a.thn : a.qn;
  hipLaunchKernelGGL((col_inverse_flux<N, false, false>), xcd_grid(nb), dim3(BC), 0, s, a.nkx,
                     a.nky, a.ldc, 0, a.qn, wn, thn, a.qn, a.qn, a.qn, a.kidx_of_x, a.kx_row,
                     a.ky_col, a.inv_denom, a.tw, a.X5, a.fstride, nb);
  dim3 gr((N + TEAMS - 1) / TEAMS, a.nq);
  hipLaunchKernelGGL(row_speed<N>, gr, dim3(256), 0, s, a.nky, a.X5, a.fstride, rowmax, a.tw);
  if (a.wn && a.thn)
    hipLaunchKernelGGL(row_wth<N>, gr, dim3(256), 0, s, a.nky, a.X5, a.fstride,
                       rowmax + (size_t)a.nq * N, (size_t)a.nq * N, a.tw);
  return hipGetLastError();
}

hipError_t launch_speed(const FluxArgs& a, double* rowmax, hipStream_t s) {
  switch (a.N) {
    case 16: return speed_n<16>(a, rowmax, s);
    case 32: return speed_n<32>(a, rowmax, s);
    case 64: return speed_n<64>(a, rowmax, s);
    case 128: return speed_n<128>(a, rowmax, s);
    case 256: return speed_n<256>(a, rowmax, s);
    case 512: return speed_n<512>(a, rowmax, s);
    case 1024: return speed_n<1024>(a, rowmax, s);
    default: return hipErrorInvalidValue;
  }
}
