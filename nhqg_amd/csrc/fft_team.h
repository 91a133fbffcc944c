// Team-level FFT building blocks shared by the spectral-product passes.
//
// A "team" of threads owns one transform held in LDS; the building blocks
// here are the radix-4 (one radix-2 stage for odd log2 N) Stockham
// transform of N = 2^k points, its register-ended variant, the LDS
// swizzle, the stage-major twiddle table and the XCD-aware mapping of
// column workgroups.
#pragma once

#include "common.h"

template <int N> struct Log2;
template <> struct Log2<8> { static constexpr int v = 3; };
template <> struct Log2<16> { static constexpr int v = 4; };
template <> struct Log2<32> { static constexpr int v = 5; };
template <> struct Log2<64> { static constexpr int v = 6; };
template <> struct Log2<128> { static constexpr int v = 7; };
template <> struct Log2<256> { static constexpr int v = 8; };
template <> struct Log2<512> { static constexpr int v = 9; };
template <> struct Log2<1024> { static constexpr int v = 10; };

// LDS layout of a team's transform buffer.  Element i of a team's N-point
// buffer sits at swz(i): bits 2..5 are XORed into bits 0..3 (a bijection on
// every aligned block of 64, the identity above bit 5).  The radix-4
// stages write at strides 4 and in runs of 4 spaced 16 apart, and the
// column passes stash and gather across teams; with the plain layout those
// accesses put 4 lanes of one 16-byte phase on the same banks.  With swz and
// team buffers FFT_PAD elements longer than N (so that neighbouring teams
// start on different banks) every such access is conflict-free or within
// 1.1x of it, for 8 or 16 lanes per 16-byte LDS phase, N = 16..1024.
constexpr int FFT_PAD = 2;
NHQG_DEV int swz(int i) { return i ^ ((i >> 2) & 15); }

template <bool INV>
NHQG_DEV double2 twid(const double2* tw, int e) {
  const double2 w = tw[e];
  return INV ? make_double2(w.x, -w.y) : w;
}

// Barrier between the stages of one team's FFT.  A team of T <= 64 threads
// lies inside one wavefront, whose LDS operations complete in issue order:
// a wave-scope fence (a compiler barrier for the memory operations) is
// enough, and the other teams of the block run on unhindered.  Larger
// teams span several waves and take the block barrier.
template <int T>
NHQG_DEV void team_sync() {
  if constexpr (T <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

// Unnormalised DFT of N points held in LDS buffer a (b is scratch; both in
// the swz layout), by a team of N/4 threads with index t, twiddles from the
// stage-major table of load_twiddles.  INV selects exp(+i...).  Every thread
// of the block must call it (team barriers inside; a caller that reads
// another team's result after it needs a block barrier).  Returns the
// buffer that holds the result in natural order.
template <int N, bool INV>
NHQG_DEV double2* team_fft(double2* a, double2* b, int t, const double2* tw) {
  constexpr int T = N / 4;
  constexpr int L2 = Log2<N>::v;
  double2* in = a;
  double2* out = b;
  int Ns = 1;
#pragma unroll
  for (int s = 0; s < L2 / 2; ++s) {
    const int j = t;
    const int k = j & (Ns - 1);
    double2 v0 = in[swz(j)], v1 = in[swz(j + T)], v2 = in[swz(j + 2 * T)],
            v3 = in[swz(j + 3 * T)];
    if (s > 0) {
      // stage-major table: w^(p k N / (4 Ns)) at tw[4^s - 4 + (p - 1) Ns + k]
      const double2* ts = tw + (Ns - 4);
      v1 = cmul(v1, twid<INV>(ts, k));
      v2 = cmul(v2, twid<INV>(ts, Ns + k));
      v3 = cmul(v3, twid<INV>(ts, 2 * Ns + k));
    }
    const double2 a0 = cadd(v0, v2), a1 = csub(v0, v2);
    const double2 a2 = cadd(v1, v3), a3 = csub(v1, v3);
    const double2 ia3 = make_double2(-a3.y, a3.x);
    const int d = (j - k) * 4 + k;
    out[swz(d)] = cadd(a0, a2);
    out[swz(d + Ns)] = INV ? cadd(a1, ia3) : csub(a1, ia3);
    out[swz(d + 2 * Ns)] = csub(a0, a2);
    out[swz(d + 3 * Ns)] = INV ? csub(a1, ia3) : cadd(a1, ia3);
    team_sync<T>();
    double2* tmp = in;
    in = out;
    out = tmp;
    Ns *= 4;
  }
  if (L2 & 1) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = t + h * T;
      const int k = j & (Ns - 1);
      const double2 v0 = in[swz(j)];
      // radix-2 table: w^(k N / (2 Ns)) at tw[Ns - 4 + k]
      const double2 v1 = cmul(in[swz(j + N / 2)], twid<INV>(tw + (Ns - 4), k));
      const int d = (j - k) * 2 + k;
      out[swz(d)] = cadd(v0, v1);
      out[swz(d + Ns)] = csub(v0, v1);
    }
    team_sync<T>();
    double2* tmp = in;
    in = out;
    out = tmp;
  }
  return in;
}

// Stage-major twiddle table from the natural one (twg[e] = exp(-2 pi i e/N)):
// radix-4 stage s >= 1 (Ns = 4^s) holds w^(p k N / (4 Ns)) for p = 1, 2, 3
// and k < Ns at 4^s - 4 + (p - 1) Ns + k; a final radix-2 stage (odd log2 N)
// holds w^(k N / (2 Ns)) at Ns - 4 + k.  A stage's lanes read consecutive
// (or broadcast) entries, where strided reads of the natural table put up to
// 4 lanes on one bank.  At most N - 4 entries in all.
// The same transform for even log2 N with the data in registers at both
// ends: v[m] holds point t + m*N/4 on entry and bin t + m*N/4 on return.
// Those are exactly the points the first radix-4 stage reads and the bins
// the last one writes in the thread itself, so the LDS write before the
// first stage and the LDS read after the last are skipped (16 of the 32
// element accesses per thread at N = 256).  a and b are scratch; ends with a
// team barrier, so the buffers may be reused at once.
template <int N, bool INV>
NHQG_DEV void team_fft_reg(double2 (&v)[4], double2* a, double2* b, int t, const double2* tw) {
  constexpr int T = N / 4;
  constexpr int S = Log2<N>::v / 2;
  static_assert(Log2<N>::v % 2 == 0, "register-ended transform needs even log2 N");
  double2* in = a;
  double2* out = b;
  int Ns = 1;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int j = t;
    const int k = j & (Ns - 1);
    double2 v0, v1, v2, v3;
    if (s == 0) {
      v0 = v[0]; v1 = v[1]; v2 = v[2]; v3 = v[3];
    } else {
      v0 = in[swz(j)]; v1 = in[swz(j + T)]; v2 = in[swz(j + 2 * T)]; v3 = in[swz(j + 3 * T)];
      const double2* ts = tw + (Ns - 4);
      v1 = cmul(v1, twid<INV>(ts, k));
      v2 = cmul(v2, twid<INV>(ts, Ns + k));
      v3 = cmul(v3, twid<INV>(ts, 2 * Ns + k));
    }
    const double2 a0 = cadd(v0, v2), a1 = csub(v0, v2);
    const double2 a2 = cadd(v1, v3), a3 = csub(v1, v3);
    const double2 ia3 = make_double2(-a3.y, a3.x);
    const double2 r0 = cadd(a0, a2), r2 = csub(a0, a2);
    const double2 r1 = INV ? cadd(a1, ia3) : csub(a1, ia3);
    const double2 r3 = INV ? csub(a1, ia3) : cadd(a1, ia3);
    if (s == S - 1) {
      // last stage: Ns = T, k = j, so its bins are d + m*Ns = t + m*T
      v[0] = r0; v[1] = r1; v[2] = r2; v[3] = r3;
    } else {
      const int d = (j - k) * 4 + k;
      out[swz(d)] = r0;
      out[swz(d + Ns)] = r1;
      out[swz(d + 2 * Ns)] = r2;
      out[swz(d + 3 * Ns)] = r3;
      team_sync<T>();
      double2* tmp = in;
      in = out;
      out = tmp;
    }
    Ns *= 4;
  }
  team_sync<T>();
}

// Workgroups are handed to the 8 XCDs round-robin by linear id, each XCD
// with its own L2.  The column passes run on a 1-D grid of 8 * per
// workgroups (xcd_grid) remapped here so that XCD x takes the contiguous
// logical range [x * per, (x + 1) * per), logical = column block + gx *
// level: neighbouring column blocks, whose pieces of every X5 / Y6 row share
// cache lines, then meet in one L2 instead of each XCD fetching the lines
// from HBM.  Returns false for the padding workgroups (logical >= nb).
NHQG_DEV bool xcd_column_block(int gx, int nb, int& bx, int& l) {
  const int per = gridDim.x >> 3;
  const int b = blockIdx.x;
  const int logical = (b & 7) * per + (b >> 3);
  if (logical >= nb) return false;
  bx = logical % gx;
  l = logical / gx;
  return true;
}

// twg[stride * e] = exp(-2 pi i e / N): the natural table of N points, or
// every stride-th entry of a longer one
template <int N>
NHQG_DEV void load_twiddles(double2* tw, const double2* twg, int stride = 1) {
  constexpr int L2 = Log2<N>::v;
  int Ns = 4;
#pragma unroll
  for (int s = 1; s < L2 / 2; ++s) {
    for (int i = threadIdx.x; i < 3 * Ns; i += blockDim.x) {
      const int p = i / Ns + 1, k = i % Ns;
      tw[Ns - 4 + i] = twg[stride * (p * k * (N / (4 * Ns)))];
    }
    Ns *= 4;
  }
  if (L2 & 1)
    for (int i = threadIdx.x; i < Ns; i += blockDim.x)
      tw[Ns - 4 + i] = twg[stride * (i * (N / (2 * Ns)))];
}

// 1-D grid for xcd_column_block: nb logical workgroups padded to 8 per XCD
static inline dim3 xcd_grid(int nb) { return dim3(8 * ((nb + 7) / 8)); }
