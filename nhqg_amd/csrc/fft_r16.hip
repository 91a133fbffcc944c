// Row pass of the flux-form spectral products at Nx = 256 with 16 threads
// per transform (the product pass of fft.hip's three-pass chain; the
// reference's per-level products are nhqg/spectral.py:307-328 and the
// sponge products spectral.py:53-60).
//
// A 256-point DFT is 16 x 16: with k = k1 + 16 k2 and y = b + 16 a,
//   z[b + 16 a] = sum_k1 w16^(k1 a) [ w256^(k1 b) sum_k2 Z[k1 + 16 k2] w16^(k2 b) ],
// so thread k1 of a 16-thread team runs a 16-point DFT over its registers,
// multiplies by w256^(k1 b), and one LDS transpose (pitch 17, conflict-free
// both ways) hands thread b the 16 values it transforms again.  Thread t
// holds point t + 16 r in register r on entry and bin t + 16 r on return,
// for either direction.  Against the radix-4 team of fft.hip (64 threads x
// 4 values, three LDS exchanges per transform) this moves a third of the
// LDS bytes and has one dependent LDS round trip per transform instead of
// three.
//
// Two rows x, x+1 of one level per team, packed as (row x) + i (row x+1) in
// every transform, as in fft.hip.  Holding u and v of both rows in
// registers beside a field and the product would not fit two waves per
// SIMD, so the pass runs in two sweeps: u with the three advected fields,
// then v with the same fields transformed again (their X5 rows come back
// from L2).  20 transforms per row pair instead of 17; the bins pair with
// their mirror N - k across lanes by shuffles.
#include "fft_team.h"

namespace {

constexpr int R16_P = 17;       // LDS pitch of the 16 x 16 exchange (complex entries)
constexpr int R16_TEAMS = 16;   // teams of 16 threads per 256-thread block

constexpr double C8 = 0.92387953251128675613;   // cos(pi/8)
constexpr double S8 = 0.38268343236508977173;   // sin(pi/8)
constexpr double RH = 0.70710678118654752440;   // sqrt(1/2)

template <bool INV>
NHQG_DEV void dft4(double2& x0, double2& x1, double2& x2, double2& x3) {
  const double2 a0 = cadd(x0, x2), a1 = csub(x0, x2), a2 = cadd(x1, x3), a3 = csub(x1, x3);
  const double2 ia3 = make_double2(-a3.y, a3.x);
  x0 = cadd(a0, a2);
  x2 = csub(a0, a2);
  x1 = INV ? cadd(a1, ia3) : csub(a1, ia3);
  x3 = INV ? csub(a1, ia3) : cadd(a1, ia3);
}

// z * exp(+-2 pi i j / 16), j in {1, 2, 3, 4, 6, 9}
template <bool INV, int J>
NHQG_DEV double2 w16(double2 z) {
  constexpr double s = INV ? 1.0 : -1.0;
  if constexpr (J == 4) return make_double2(-s * z.y, s * z.x);
  constexpr double c = J == 1 ? C8 : J == 2 ? RH : J == 3 ? S8 : J == 6 ? -RH : -C8;
  constexpr double n = s * (J == 1 ? S8 : J == 2 ? RH : J == 3 ? C8 : J == 6 ? RH : -S8);
  return make_double2(z.x * c - z.y * n, z.x * n + z.y * c);
}

// 16-point DFT of the registers, natural order in and out (k = 4m + n,
// bin p + 4q: 4-point DFTs over m, twiddles w16^(np), 4-point DFTs over n)
template <bool INV>
NHQG_DEV void dft16(double2 (&v)[16]) {
#pragma unroll
  for (int n = 0; n < 4; ++n) dft4<INV>(v[n], v[n + 4], v[n + 8], v[n + 12]);
  // v[n + 4p] = T[n][p]
  v[5] = w16<INV, 1>(v[5]);
  v[6] = w16<INV, 2>(v[6]);
  v[7] = w16<INV, 3>(v[7]);
  v[9] = w16<INV, 2>(v[9]);
  v[10] = w16<INV, 4>(v[10]);
  v[11] = w16<INV, 6>(v[11]);
  v[13] = w16<INV, 3>(v[13]);
  v[14] = w16<INV, 6>(v[14]);
  v[15] = w16<INV, 9>(v[15]);
  double2 o[16];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    double2 x0 = v[4 * p], x1 = v[4 * p + 1], x2 = v[4 * p + 2], x3 = v[4 * p + 3];
    dft4<INV>(x0, x1, x2, x3);
    o[p] = x0;
    o[p + 4] = x1;
    o[p + 8] = x2;
    o[p + 12] = x3;
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = o[i];
}

// The thread's twiddles exp(-2 pi i t r / 256), r = 1..15, come from an
// LDS table, entry [r - 1][t] (lanes read consecutive entries): fifteen LDS
// reads per transform and no registers held between transforms.

template <bool INV>
NHQG_DEV double2 conj_if(double2 z) {
  return INV ? make_double2(z.x, -z.y) : z;
}

// Unnormalised 256-point DFT by a team of 16 threads (index t, LDS tile xs
// of 16 x R16_P entries, twiddle table tws): point t + 16 r in v[r] on
// entry, bin t + 16 r on return.
template <bool INV>
NHQG_DEV void fft256(double2 (&v)[16], double2* xs, int t, const double2* tws) {
  dft16<INV>(v);
#pragma unroll
  for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], conj_if<INV>(tws[16 * (r - 1) + t]));
#pragma unroll
  for (int r = 0; r < 16; ++r) xs[t * R16_P + r] = v[r];
  team_sync<16>();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = xs[r * R16_P + t];
  team_sync<16>();
  dft16<INV>(v);
}

}  // namespace

// Y6[g][l][x][ky] as row_flux2 of fft.hip writes it: G_f = F_y(u f),
// H_f = i ky F_y(v f), and the sponge products F_y(sigma c_f) of the
// coefficient levels (slots 6..8).  Grid: (N / 32, levels), 256 threads.
//
// The eleven inverse transforms of a row pair (u, q, w, theta, v, q, w,
// theta, then the sponge inputs) run as one software pipeline: the X5 rows
// of the next are loaded while the current one is transformed.  Thread t
// loads only ky = t + 16 j (j < 6, the kept rows ky < nky <= 96) of both
// rows, 12 values; the packed spectrum's bins N - ky, which the Hermitian
// packing forms from the same values, belong to lane 16 - t and go there by
// shuffles.  That halves the loads and leaves the registers for the
// prefetch at two waves per SIMD.
//
// NS (nodal sponge, with SPONGE): sigma acts on the nodal w and theta rows
// that the products already hold, and F_y(sigma f) is added to the H_f slot
// (H_f = i ky F_y(v f) + F_y(sigma f)); only q' keeps a coefficient-level
// sponge slot (X5 slot 5 -> Y6 slot 6).  The sponge is linear and acts on
// each level alike, so it commutes with the vertical maps: Pw S(VSw w) =
// S(w) in exact arithmetic, and the roundoff the nodal route leaves in the
// back-transformed w and theta tendencies stays ~1e-15 of the field per
// step.  q' keeps the coefficient route: its nodal values carry the large
// barotropic part, and their roundoff in the small high Chebyshev
// coefficients, amplified by d/dZ in the implicit solve, reached ~6e-12 of
// w per step.  Then v is transformed a third time (sweeps u | v with q',
// w | v with theta) so that v is dead when a field's two products are
// transformed: 19 transforms per row pair, 6 X5 and 7 Y6 slots.
template <int nky, bool SPONGE, bool TRAP, bool NS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
row_flux16(int ng, const double2* __restrict__ X5, size_t fstride,
           double2* __restrict__ Y6, const double* __restrict__ ky_col,
           const double2* __restrict__ eta_rows, const double* __restrict__ sponge_phys,
           const double2* __restrict__ twg) {
  constexpr int N = 256;
  constexpr int J = (nky + 15) / 16;   // loaded ky = t + 16 j, j < J
  // bins 16 J .. N - 16 J hold no kept row; the mirrors N - ky fill
  // registers 16 - J .. 15
  static_assert(nky > 16 * (J - 1) && J == 6, "sized for the 2/3 band at N = 256");
  __shared__ double2 xb[R16_TEAMS][16 * R16_P];
  __shared__ double2 w[15 * 16];     // exp(-2 pi i t r / N) at [r - 1][t]
  __shared__ double kys[16 * J];     // ky of bin k, zero beyond the kept rows
  const int tid = threadIdx.x, team = tid >> 4, t = tid & 15;
  const int l = blockIdx.y;
  const int x0 = 2 * (blockIdx.x * R16_TEAMS + team);
  double2* xs = xb[team];
  if (tid < 15 * 16) w[tid] = twg[(tid & 15) * ((tid >> 4) + 1)];
  if (tid < 16 * J) kys[tid] = tid < nky ? ky_col[tid] : 0.0;
  __syncthreads();
  // uniform bases (level l) and 32-bit lane offsets: the loads and stores
  // then take a scalar base, one offset register and immediate steps
  const size_t lb = (size_t)l * N * nky;
  const int orow = x0 * nky + t;       // element ky = t of row x0 within the level
  const int ophys = x0 * N + t;        // point (x0, t) of a physical row
  const int lane = threadIdx.x & 63;
  const int mirror = (lane & ~15) | ((16 - t) & 15);
  const double sc = 1.0 / ((double)N * (double)N);
  // (ky comes from LDS in the store phase: a global load there would wait
  // for the prefetch issued before it, vector memory counts completing in
  // order)

  // half spectra of rows x0 (rp) and x0 + 1 (rq) at ky = t + 16 j: the
  // prefetched input of the next inverse
  double2 rp[J], rq[J], ep[J], eq[J];
  auto load_raw = [&](int slot) __attribute__((always_inline)) {
    const double2* P = X5 + slot * fstride + lb;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const bool in = t + 16 * j < nky;
      rp[j] = in ? P[orow + 16 * j] : make_double2(0.0, 0.0);
      rq[j] = in ? P[orow + 16 * j + nky] : make_double2(0.0, 0.0);
    }
  };
  // q' + eta: the trap's rows added in spectral form, loaded with q's
  auto load_trap = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const bool in = t + 16 * j < nky;
      ep[j] = in ? eta_rows[orow + 16 * j] : make_double2(0.0, 0.0);
      eq[j] = in ? eta_rows[orow + 16 * j + nky] : make_double2(0.0, 0.0);
    }
  };
  // the packed spectrum Z of (row x0) + i (row x0 + 1) (irfft's convention
  // on ky = 0; ky = N/2 is not kept): bin t + 16 r in v[r]
  auto build = [&](double2 (&v)[16], bool trap) __attribute__((always_inline)) {
    // pinned here: computed ahead across a loop edge, the packing's
    // intermediates would be carried (and spilled) instead of the loads
#pragma unroll
    for (int j = 0; j < J; ++j)
      __asm__ volatile("" : "+v"(rp[j].x), "+v"(rp[j].y), "+v"(rq[j].x), "+v"(rq[j].y));
    if (trap) {
#pragma unroll
      for (int j = 0; j < J; ++j) {
        rp[j] = cadd(rp[j], ep[j]);
        rq[j] = cadd(rq[j], eq[j]);
      }
    }
    double2 m[J];   // Z[N - ky] for this lane's ky, wanted by lane 16 - t
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const double2 p = rp[j], q = rq[j];
      v[j] = (t == 0 && j == 0) ? make_double2(p.x, q.x) : make_double2(p.x - q.y, p.y + q.x);
      m[j] = make_double2(p.x + q.y, q.x - p.y);
    }
#pragma unroll
    for (int j = J; j < 16 - J; ++j) v[j] = make_double2(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < J; ++j) {
      // register 15 - j holds bin N - (16 - t) - 16 j: lane 16 - t's ky
      // (16 - t) + 16 j; for t = 0, bin 16 (16 - j) - 16 = N - 16 (j + 1)
      // is this lane's own ky = 16 (j + 1)
      const double2 o = make_double2(__shfl(m[j].x, mirror, 64), __shfl(m[j].y, mirror, 64));
      const double2 own = j + 1 < J ? m[j + 1] : make_double2(0.0, 0.0);
      v[15 - j] = make_double2(t == 0 ? own.x : o.x, t == 0 ? own.y : o.y);   // selects
    }
  };
  // forward DFT of (row x0) + i (row x0 + 1), unpacked into the two half
  // spectra and stored to Y6 slot `slot` (times i ky when iky)
  auto forward_store = [&](double2 (&v)[16], int slot, bool iky, double half)
      __attribute__((always_inline)) {
    fft256<false>(v, xs, t, w);
    double2* d = Y6 + slot * fstride + lb;
#pragma unroll
    for (int c = 0; c < J; ++c) {
      // bin N - k of k = t + 16 c: register 15 - c of lane 16 - t, or for
      // t = 0 register (16 - c) % 16 of this lane
      const double2 zo = make_double2(__shfl(v[15 - c].x, mirror, 64),
                                      __shfl(v[15 - c].y, mirror, 64));
      const double2 zs = v[(16 - c) & 15];
      const double2 zm = make_double2(t == 0 ? zs.x : zo.x, t == 0 ? zs.y : zo.y);
      const double2 zk = v[c];
      const int k = t + 16 * c;
      if (k < nky) {
        // times (re, im) = (1, 0), or (0, ky) for i ky
        const double re = iky ? 0.0 : 1.0, im = iky ? kys[k] : 0.0;
        const double2 A = make_double2(half * (zk.x + zm.x), half * (zk.y - zm.y));
        const double2 B = make_double2(half * (zk.y + zm.y), half * (zm.x - zk.x));
        d[orow + 16 * c] = make_double2(A.x * re - A.y * im, A.x * im + A.y * re);
        d[orow + 16 * c + nky] = make_double2(B.x * re - B.y * im, B.x * im + B.y * re);
      }
    }
  };

  // inverse sequence: X5 slots u, q, w, theta, v, q, w, theta, then the
  // sponge inputs 5, 6, 7
  // the physical rows x0, x0 + 1 of eta or sigma at y = t + 16 a, loaded
  // before the next prefetch is issued, so that waiting for them does not
  // wait for it
  auto load_phys = [&](const double* src, double (&e)[32]) __attribute__((always_inline)) {
    int op = ophys;
    __asm__ volatile("" : "+v"(op));   // not hoisted out of the loop
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      e[a] = src[op + 16 * a];
      e[16 + a] = src[op + 16 * a + N];
    }
  };
  double2 vel[16], z[16];
  load_raw(0);
  // the inverses are left unscaled: the forward transform's unpacking
  // applies sc (sponge products) or sc^2 (velocity times field), powers of
  // two, so every value is what scaling first would give
  auto inverse_t = [&]() __attribute__((always_inline)) { fft256<true>(z, xs, t, w); };
  auto product_store = [&](int fi, int sweep) __attribute__((always_inline)) {
#pragma unroll
    for (int a = 0; a < 16; ++a) z[a] = make_double2(vel[a].x * z[a].x, vel[a].y * z[a].y);
    forward_store(z, 2 * fi + sweep, sweep == 1, 0.5 * sc * sc);
  };
  if constexpr (NS) {
    // forward DFT of one row's packed products v = sc (v f) + i (sigma f)
    // (sc a power of two), unpacked into sc^2 F_y(v f) (A) and sc
    // F_y(sigma f) (B) at the kept ky of this lane: H = i ky A + B -> dst
    auto row_flux_sponge = [&](double2 (&v)[16], double2* dst) __attribute__((always_inline)) {
      fft256<false>(v, xs, t, w);
      const double half = 0.5 * sc;
#pragma unroll
      for (int c = 0; c < J; ++c) {
        const double2 zo = make_double2(__shfl(v[15 - c].x, mirror, 64),
                                        __shfl(v[15 - c].y, mirror, 64));
        const double2 zs = v[(16 - c) & 15];
        const double2 zm = make_double2(t == 0 ? zs.x : zo.x, t == 0 ? zs.y : zo.y);
        const double2 zk = v[c];
        const int k = t + 16 * c;
        const double ky = kys[k];
        const double2 A = make_double2(half * (zk.x + zm.x), half * (zk.y - zm.y));
        const double2 B = make_double2(half * (zk.y + zm.y), half * (zm.x - zk.x));
        if (k < nky) dst[16 * c] = make_double2(B.x - ky * A.y, B.y + ky * A.x);
      }
    };
    // z: the field's physical rows x0 (re) and x0 + 1 (im), vel: v's; e:
    // sigma on the two rows.  H_f of both rows -> Y6 slot `slot`
    auto flux_sponge_store = [&](int slot, const double (&e)[32]) __attribute__((always_inline)) {
      double2 c1[16];
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        const double2 f = z[a];
        z[a] = make_double2(sc * (vel[a].x * f.x), e[a] * f.x);
        c1[a] = make_double2(sc * (vel[a].y * f.y), e[16 + a] * f.y);
      }
      double2* d = Y6 + slot * fstride + lb + orow;
      row_flux_sponge(z, d);
      row_flux_sponge(c1, d + nky);
    };
    // sweep 0: u with q' + eta, w, theta (G slots), as without NS
    build(z, false);
    load_raw(2);
    if (TRAP) load_trap();
    inverse_t();
#pragma unroll
    for (int a = 0; a < 16; ++a) vel[a] = z[a];
    build(z, TRAP);
    load_raw(3);
    inverse_t();
    product_store(0, 0);
#pragma unroll 1
    for (int fi = 1; fi < 3; ++fi) {
      build(z, false);
      load_raw(fi == 1 ? 4 : 1);
      inverse_t();
      product_store(fi, 0);
    }
    // sweep 1: v with q' + eta (H slot), then w with its sponge (H slot)
    build(z, false);
    load_raw(2);
    if (TRAP) load_trap();
    inverse_t();
#pragma unroll
    for (int a = 0; a < 16; ++a) vel[a] = z[a];
    build(z, TRAP);
    load_raw(3);
    inverse_t();
    product_store(0, 1);
    // sigma's rows and the next prefetch after the transform: held across
    // it they would not fit beside the two products
    double e[32];
    build(z, false);
    inverse_t();
    load_phys(sponge_phys, e);
    flux_sponge_store(3, e);
    load_raw(1);
    // sweep 2: v again, theta with its sponge (H slot)
    build(z, false);
    load_raw(4);
    inverse_t();
#pragma unroll
    for (int a = 0; a < 16; ++a) vel[a] = z[a];
    build(z, false);
    inverse_t();
    load_phys(sponge_phys, e);
    flux_sponge_store(5, e);
    load_raw(5);
    // sigma times the coefficient levels of q'
    build(z, false);
    load_phys(sponge_phys, e);
    fft256<true>(z, xs, t, w);
#pragma unroll
    for (int a = 0; a < 16; ++a) z[a] = make_double2(z[a].x * e[a], z[a].y * e[16 + a]);
    forward_store(z, 6, false, 0.5 * sc);
    return;
  }
#pragma unroll 1
  for (int sweep = 0; sweep < 2; ++sweep) {
    // u, then v
    build(z, false);
    load_raw(2);
    if (TRAP) load_trap();
    inverse_t();
#pragma unroll
    for (int a = 0; a < 16; ++a) vel[a] = z[a];
    // q' + eta
    build(z, TRAP);
    load_raw(3);
    inverse_t();
    product_store(0, sweep);
    // w, theta
#pragma unroll 1
    for (int fi = 1; fi < 3; ++fi) {
      build(z, false);
      const int next = fi == 1 ? 4 : (sweep == 0 ? 1 : 5);
      if (fi == 1 || sweep == 0 || SPONGE) load_raw(next);
      inverse_t();
      product_store(fi, sweep);
    }
  }
  if (SPONGE) {
    // sigma times each coefficient level of q', w, theta (levels >= ng
    // hold no w or theta)
    const int nsp = l < ng ? 3 : 1;
#pragma unroll 1
    for (int fi = 0; fi < nsp; ++fi) {
      build(z, false);
      double e[32];
      load_phys(sponge_phys, e);
      fft256<true>(z, xs, t, w);
      if (fi + 1 < nsp) load_raw(6 + fi);
#pragma unroll
      for (int a = 0; a < 16; ++a)
        z[a] = make_double2(z[a].x * e[a], z[a].y * e[16 + a]);
      forward_store(z, 6 + fi, false, 0.5 * sc);
    }
  }
}

// The trap's rows in the row pass's input form: eta_rows[x][ky] = N *
// sum_y eta(x, y) exp(-2 pi i ky y / N) for ky < nky, the x-inverse
// transform of eta's spectrum (the form X5 holds q' in), so that the row
// pass adds eta to q' before its inverse transform.  A direct DFT per row
// with the exact table twiddles, once per context.
__global__ void __launch_bounds__(128)
eta_rows_kernel(int N, int nky, const double* __restrict__ eta_phys,
                const double2* __restrict__ twg, double2* __restrict__ out) {
  __shared__ double row[1024];
  const int x = blockIdx.x;
  for (int y = threadIdx.x; y < N; y += blockDim.x) row[y] = eta_phys[(size_t)x * N + y];
  __syncthreads();
  for (int ky = threadIdx.x; ky < nky; ky += blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int y = 0; y < N; ++y) {
      const double2 w = twg[(ky * y) & (N - 1)];
      re = fma(row[y], w.x, re);
      im = fma(row[y], w.y, im);
    }
    out[(size_t)x * nky + ky] = make_double2(N * re, N * im);
  }
}

hipError_t launch_eta_rows(int N, int nky, const double* eta_phys, const double2* twg,
                           double2* out, hipStream_t s) {
  if (N > 1024 || (N & (N - 1))) return hipErrorInvalidValue;
  hipLaunchKernelGGL(eta_rows_kernel, dim3(N), dim3(128), 0, s, N, nky, eta_phys, twg, out);
  return hipGetLastError();
}

bool row_flux16_fits(const FluxArgs& a) {
  // the band-compact layout's 86 kept ky rows (ky <= 256/3)
  return a.N == 256 && !a.jac && a.nky == 86 && !(a.eta_phys && !a.eta_rows);
}

bool launch_row_flux16(const FluxArgs& a, hipStream_t s) {
  constexpr int NKY = 86;
  if (!row_flux16_fits(a) || (a.nsp && !a.sponge_phys)) return false;
  const dim3 g(256 / (2 * R16_TEAMS), a.nq);
  if (a.nsp) {
    if (a.eta_rows)
      hipLaunchKernelGGL((row_flux16<NKY, true, true, true>), g, dim3(256), 0, s, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_rows, a.sponge_phys, a.tw);
    else
      hipLaunchKernelGGL((row_flux16<NKY, true, false, true>), g, dim3(256), 0, s, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_rows, a.sponge_phys, a.tw);
  } else if (a.sponge_phys) {
    if (a.eta_rows)
      hipLaunchKernelGGL((row_flux16<NKY, true, true, false>), g, dim3(256), 0, s, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_rows, a.sponge_phys, a.tw);
    else
      hipLaunchKernelGGL((row_flux16<NKY, true, false, false>), g, dim3(256), 0, s, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_rows, a.sponge_phys, a.tw);
  } else {
    if (a.eta_rows)
      hipLaunchKernelGGL((row_flux16<NKY, false, true, false>), g, dim3(256), 0, s, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_rows, a.sponge_phys, a.tw);
    else
      hipLaunchKernelGGL((row_flux16<NKY, false, false, false>), g, dim3(256), 0, s, a.ng, a.X5,
                         a.fstride, a.Y6, a.ky_col, a.eta_rows, a.sponge_phys, a.tw);
  }
  return true;
}
