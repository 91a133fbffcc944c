// Shared definitions for the NHQG gfx950 kernels.
//
// Field layout in HBM (band-compact, level-major): a field with nl vertical
// levels is a row-major real matrix of nl rows and ld = 2*nm_pad columns;
// column 2m / 2m+1 hold Re / Im of mode m.  Modes are ordered ky-major,
// m = ky_i * nkx + kx_i, over the kept rows (|kx| <= Nx/3 for the 2/3 band,
// or all rows) and kept ky columns; nm_pad rounds the mode count up to a
// multiple of 64 so every contraction workgroup (128 columns) is full.  Padding modes carry
// zeros and zero per-mode factors.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

#define NHQG_DEV __device__ __forceinline__

typedef double v4d __attribute__((ext_vector_type(4)));

NHQG_DEV double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
NHQG_DEV double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
NHQG_DEV double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
NHQG_DEV double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
NHQG_DEV double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// multiply by i*s
NHQG_DEV double2 cmul_i(double2 a, double s) { return make_double2(-a.y * s, a.x * s); }

// Linear epilogue shared by the contraction and the elementwise kernels:
//   out = a_acc * s_acc[mode] * acc + sum_t aX[t] * sX[t][mode] * X[t][idx]
//   out *= s_out[mode]                       (when s_out is set)
//   out *= d_tab[row * d_ld + mode]          (when d_tab is set)
// A null per-mode pointer means a factor of one; X terms with null
// pointers are skipped.  idx is the element index in the output field,
// row its level.  The last factor is the diagonal of a shell solve in the
// eigenbasis of B0 (A_k = alpha_k I - c_k B0).
#define NHQG_EPI_TERMS 8
struct Epi {
  double a_acc;
  const double* s_acc;
  int nX;
  const double* X[NHQG_EPI_TERMS];
  double aX[NHQG_EPI_TERMS];
  const double* sX[NHQG_EPI_TERMS];
  const double* s_out;
  const double* d_tab;
  int d_ld;
};

NHQG_DEV double epi_apply(const Epi& e, double acc, size_t idx, int row, int mode) {
  double o = e.a_acc * acc;
  if (e.s_acc) o *= e.s_acc[mode];
  for (int t = 0; t < e.nX; ++t) {
    double v = e.aX[t] * e.X[t][idx];
    if (e.sX[t]) v *= e.sX[t][mode];
    o += v;
  }
  if (e.s_out) o *= e.s_out[mode];
  if (e.d_tab) o *= e.d_tab[(size_t)row * e.d_ld + mode];
  return o;
}

// Epilogue of up to NV results of one lane at arbitrary output rows: the
// arithmetic of epi_apply, with each extra field read for all NV rows at
// once (one memory latency per term).  Rows with ok false are neither read
// (row 0 stands in) nor stored.
template <int NV>
NHQG_DEV void epi_rows(const Epi& e, double* C, const double (&val)[NV], const int (&row)[NV],
                       const bool (&ok)[NV], int ld, int col, int mode) {
  double o[NV];
  // the per-mode scales load with the first term, not after the last
  const double sa = e.s_acc ? e.s_acc[mode] : 1.0;
  const double so = e.s_out ? e.s_out[mode] : 1.0;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    o[i] = e.a_acc * val[i];
    if (e.s_acc) o[i] *= sa;
  }
  // byte offsets of the NV elements, formed once for every term and the
  // stores (fields stay below 4 GiB: Nz + 1 <= 257 rows of ld doubles)
  unsigned bo[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) bo[i] = ((unsigned)(ok[i] ? row[i] : 0) * (unsigned)ld + col) * 8u;
  for (int x = 0; x < e.nX; ++x) {
    const char* X = reinterpret_cast<const char*>(e.X[x]);
    const double ax = e.aX[x];
    const bool hs = e.sX[x] != nullptr;
    const double sx = hs ? e.sX[x][mode] : 1.0;
    double v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = *reinterpret_cast<const double*>(X + bo[i]);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double w = ax * v[i];
      if (hs) w *= sx;
      o[i] += w;
    }
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    if (ok[i]) {
      double w = o[i];
      if (e.s_out) w *= so;
      if (e.d_tab) w *= e.d_tab[(size_t)row[i] * e.d_ld + mode];
      // nontemporal: the outputs stream through (+1% on the 256^3 step in
      // alternating runs; the same hint on the FFT intermediates lost)
      __builtin_nontemporal_store(w, reinterpret_cast<double*>(reinterpret_cast<char*>(C) + bo[i]));
    }
  }
}

static inline Epi epi_make(double a_acc) {
  Epi e;
  e.a_acc = a_acc;
  e.s_acc = nullptr;
  e.nX = 0;
  for (int t = 0; t < NHQG_EPI_TERMS; ++t) {
    e.X[t] = nullptr;
    e.aX[t] = 0.0;
    e.sX[t] = nullptr;
  }
  e.s_out = nullptr;
  e.d_tab = nullptr;
  e.d_ld = 0;
  return e;
}

static inline void epi_add(Epi& e, double a, const double* X, const double* s = nullptr) {
  e.X[e.nX] = X;
  e.aX[e.nX] = a;
  e.sX[e.nX] = s;
  e.nX++;
}

// ---- launch wrappers (defined in the .hip files) ----
// N: the transform length, nx under the 2/3 rule, 3nx/2 (the zero-padded
// grid) under the 3/2 rule; the tables kidx_of_x (N), rows, tw (N) and the
// physical fields (N x N) are those of that grid.  out_scale multiplies the
// outputs of the 3/2 rule ((N/nx)^2); the 2/3 rule multiplies by mask.
struct FluxArgs {
  int N, nq, nkx, nky, ldc, ng;
  double out_scale;
  const double2 *qn, *wn, *thn;   // nodal fields (advection)
  const double2 *qc, *wc, *thc;   // coefficient fields (sponge)
  double2 *aq, *aw, *ath;          // advection outputs (nodal spectral)
  double2 *sq, *sw, *sth;          // sponge outputs (coefficient levels)
  const int *kidx_of_x, *rows;
  const double *kx_row, *ky_col, *inv_denom, *mask, *eta_phys, *sponge_phys;
  int jac;                        // 0 flux form, 1 Jacobian form
  const double *eta_x, *eta_y;    // grad(eta) on the grid (Jacobian form with a trap)
  const double2* tw;
  double2 *X5, *Y6;
  size_t fstride;
  // the trap's rows in X5's form (launch_eta_rows), or null: the Nx = 256
  // flux-form row pass adds eta to q' in spectral form when it is given
  const double2* eta_rows = nullptr;
  // nodal sponge of w and theta (flux form at N = 256 with a sponge, see
  // row_flux16): no sponge outputs sw / sth; the sponge is part of aw / ath
  int nsp = 0;
};
// pass: 0 all three passes, 1 col_inverse, 2 row products, 3 col_forward
hipError_t launch_flux(const FluxArgs& a, int pass, hipStream_t s);
// single-level 2-D real transforms and init steps (init.hip): forward
// spec = rfft2(f), else f = irfft2(spec) with spec overwritten; spec is
// N x (N/2+1) in rfft2 layout
hipError_t launch_init_fft2(int N, bool forward, double* f, double2* spec, const double2* tw,
                            hipStream_t s);
hipError_t launch_init_spec_filter(int N, double2* spec, double k0, double k_peak,
                                   double k_width, int annulus, hipStream_t s);
hipError_t launch_init_taper(int N, double* f, double L, double R, double width, hipStream_t s);
hipError_t launch_init_normalize(int N, double* f, double L, double R, double amplitude,
                                 hipStream_t s);
hipError_t launch_init_cyclones(int N, double* f, double L, const double* centers, int count,
                                double amplitude, double radius, hipStream_t s);
hipError_t launch_init_put_q0(int nm, const int* full_idx, const double2* spec, double2* q0,
                              int add, hipStream_t s);
// the flux-form row pass at N = 256 with 16 threads per transform
// (fft_r16.hip); false when a does not fit it
bool launch_row_flux16(const FluxArgs& a, hipStream_t s);
// whether a fits the 16-thread row pass (flux form, N = 256, 86 kept ky)
bool row_flux16_fits(const FluxArgs& a);
// out[x][ky] = N * sum_y eta_phys[x][y] exp(-2 pi i ky y / N), ky < nky
hipError_t launch_eta_rows(int N, int nky, const double* eta_phys, const double2* twg,
                           double2* out, hipStream_t s);
// the 3/2 rule's passes (a.N = 3P, P = 8..512; fft_pad.hip)
hipError_t launch_flux_pad(const FluxArgs& a, int pass, hipStream_t s);
// rowmax: Nq*N row maxima of the speed from psi (a.qn); with a.wn and a.thn
// set, three more Nq*N sections: max |w|, max |theta|, max |w theta|
hipError_t launch_speed(const FluxArgs& a, double* rowmax, hipStream_t s);
// Shell matrices for launch_shell_solve: order np = 64*ceil(Ng/64), shell s
// at s*np*np, inv_s[j][i] at j*np + i (the transposed inverse), zero padded.
int shell_pad_order(int Ng);
// packs shells [s0, s1) of invT (n_shells x Ng x Ng, transposed) into out
void shell_pack_inv(const double* invT, int n_shells, int Ng, int s0, int s1, double* out);
// tab[r][m] = 1 / (alpha[m] - c[m] lam[r]) for r < nl, m < nmp (row pitch nmp)
hipError_t launch_eigen_diag(int nl, int nmp, const double* alpha, const double* c,
                             const double* lam, double* tab, hipStream_t s);
// builds and inverts every shell factor on the device into invP (layout
// above); *status is set nonzero on a zero pivot
hipError_t launch_shell_build(int n_shells, int Ng, const double* B0T, const double* KT,
                              const double* coef, double* invP, int* status, hipStream_t s);
hipError_t launch_shell_solve(int n_shells, int Ng, int ld, const int* shell_ptr,
                              const int* shell_modes, const double* invP, const double* rhs,
                              double* w_out, const double* Rth, double* th_out,
                              const double* iath, double alpha, hipStream_t s);
hipError_t launch_parseval(int nl, int nm, int ld, int P, const double* a, const double* b,
                           const double* pw, const int* partner, double* partial,
                           hipStream_t s);
hipError_t launch_sbp_prepare(int n, int P, const double* partial, const double* d1b,
                              const double* lb, double c1, double c2, double* sbp,
                              hipStream_t s);
// X, Y and the 2n x 2n substep map of the corrector call (columns.hip)
hipError_t launch_sbp_maps(int n, const double* sbp, const double* thb2s, const double* s2thb,
                           const double* d1b, double mu_sdt, double* XY, double* Mall,
                           hipStream_t s);
// one substep: th_dst, g from th_src and the Parseval partials
hipError_t launch_sbp_apply(int n, int P, const double* partial, const double* Mall,
                            const double* th_src, double* th_dst, double* g_out, hipStream_t s);
// one SBP substep transfer as a five-diagonal level stencil with the fused
// update and the next substep's horizontal mean, one value per level in
// partial (bpart: n * ceil(ld / 256) block sums of scratch; columns.hip)
hipError_t launch_sbp_band(int n, int ld, const double* pb, const double* th, const double* w,
                           const double* g, double bc, const double* pwc, double* out,
                           double* partial, double* bpart, int npairs, const int2* pq,
                           const double* pwt, hipStream_t s);
hipError_t launch_small_matvec(int n_out, int n_in, const double* M, const double* x, double* y,
                               hipStream_t s);
hipError_t launch_small_axpy(int n, const double* x0, const double* x1, const double* x2, double c,
                             double* y, hipStream_t s);
hipError_t launch_sanitize(int nl, int ld, int npairs, const int2* pairs, int m0, double* f,
                           hipStream_t s);
// the same on three fields in one launch: nl0 rows of f0, nl12 of f1 and f2
hipError_t launch_sanitize3(int nl0, int nl12, int ld, int npairs, const int2* pairs, int m0,
                            double* f0, double* f1, double* f2, hipStream_t s);
hipError_t launch_mask(int nl, int ld, double* f, const double* mask, hipStream_t s);
hipError_t launch_gather(int nl, int nm, int nm_pad, int full_n, const int* full_idx,
                         const double2* full, double2* comp, hipStream_t s);
hipError_t launch_scatter(int nl, int nm, int nm_pad, int full_n, const int* full_idx,
                          const double2* comp, double2* full, hipStream_t s);
hipError_t launch_max_reduce(size_t n, const double* x, double* out, hipStream_t s);
hipError_t launch_small_lincomb(int n, double* y, double c0, const double* x0, double c1,
                                const double* x1, double c2, const double* x2, double c3,
                                const double* x3, hipStream_t s);
hipError_t launch_partial_matvec(int n_out, int nl, int P, const double* partial, const double* M,
                                 double* y, hipStream_t s);
hipError_t launch_level_scale(int nl, int ld, const double* g, const double* x, double* y,
                              hipStream_t s);
hipError_t launch_depth_power(int nl, int ld, int nm, const double* cc, const double* f,
                              double* out, hipStream_t s);
hipError_t launch_upsample(int nl, int nm_d, int nkx_d, int nd, const int* rows_d, int ld_d,
                           double* out, int nx, int nkx_s, int nky_s, const int* kidx_s,
                           const int* partner_s, int ld_s, const double* src, double scale,
                           hipStream_t s);
hipError_t launch_stream_copy(size_t n16, const double2* src, double2* dst, hipStream_t s);
hipError_t launch_beta(int nl, int ldc, int nm, double beta, const double* kx_mode,
                       const double* inv_denom, const double2* q, double2* eq, hipStream_t s);

// Contraction operator on the device: A (M x K) stored k-major, At[k * lda
// + r] = A[r][k], zero padded to kp columns (a multiple of the k-chunk) and
// lda = 16*mt + 8 rows; row 16*mt holds A's 257th row when M = 16*mt + 1.
struct GemmOp {
  const double* At;
  int M, K, mt, lda, kp;
  // int8 slices for the mixed-precision kernel (gemm_i8.hip), or null:
  // A8 in tile order, sa[r] = 2^ea_r, afirst[r] = A[r][0] (k = 0 runs in
  // FP64); nti row tiles of 16, nk k-steps of 64 over k = 1 .. K-1
  const void* A8 = nullptr;
  const double* sa = nullptr;
  const double* afirst = nullptr;
  int slices = 0, nti = 0, nk = 0;
  // Reflection split (gemm_split.hip), used when sp_At is set: two products
  // of half size, E (half 0) and O (half 1), each at most 16*sp_mth rows.
  // Input: ib = 0 reads rows istr*k + ioff[h] of B for half h; ib = 1 feeds
  // E with B[k] + B[N-k] and O with B[k] - B[N-k], N = sp_nbi (B[N/2] alone
  // to E at k = N/2).  Output: ob = 0 writes half h's row r to C row
  // ostr*r + ooff[h] for r < nr[h]; ob = 1 writes E_r + O_r to row r and
  // E_r - O_r to row N - r, N = sp_nbo, r <= N/2.  sp_At holds the halves
  // k-major, sp_kp rows of 2*16*sp_mth + 16 (E rows, O rows, then E's row
  // 16*sp_mth, computed on the VALU, when sp_xrow).
  const double* sp_At = nullptr;
  int sp_mth = 0, sp_kp = 0, sp_ib = 0, sp_ob = 0, sp_xrow = 0;
  int sp_istr = 0, sp_ioff[2] = {0, 0}, sp_ostr = 0, sp_ooff[2] = {0, 0}, sp_nr[2] = {0, 0};
  int sp_nk[2] = {0, 0};
  int sp_nbi = 0, sp_nbo = 0;
  int sp_pref = 0;   // the prefer_cols the split was planned with
  int sp_kc = 0;     // the k-chunk sp_kp is padded to (gemm_split_kc)
  // int8 slices of the two halves (gemm_i8.hip, split kernel): tile t, half
  // h, slice p, k-step kk at (((t*2 + h)*S + p)*snkh + kk)*64 + lane (v4i);
  // ssa / safirst at h*snti*16 + r; snti row tiles per half, snkh k-steps
  // of 64 per half (1 or 2)
  const void* sA8 = nullptr;
  const double* ssa = nullptr;
  const double* safirst = nullptr;
  int snti = 0, snkh = 0;
};
// One contraction C = epilogue(A * B') of a batch.  B' = B, or with W set
// the SBP substep update B - bc * W * g[row] with rows 0 and K-1 zeroed.
struct GemmJob {
  GemmOp op;
  const double* B;
  double* C;
  const double* W;
  const double* g;
  double bc;
  Epi e;
  double* C2;   // optional second output, epilogue e2 of the same product
  Epi e2;
};
#define GEMM_MAX_JOBS 3
struct GemmBatch {
  GemmJob job[GEMM_MAX_JOBS];
  int ld;       // row stride (doubles) of every B, C, W and epilogue field
  int slices;   // 0: FP64 MFMA kernel; 4..7: int8 slice kernel (every op split)
};
int gemm_row_tiles(int M);
int gemm_op_lda(int mt);
int gemm_op_kp(int K);
void gemm_pack_op(const double* A, int M, int K, int mt, double* out);
// Runs b.job[0..nj) in one launch; every job has the same mt and all or
// none use the update.  ld must be a multiple of 128.
hipError_t launch_gemm_batch(const GemmBatch& b, int nj, hipStream_t s);
// Reflection split of a dense M x K operator (row-major): fills the split
// fields of op (not sp_At) and the packed halves, or returns false when the
// operator has none of the structures within tol * max|A|.  prefer_cols
// tries the column butterfly before the row butterfly; the halves get at
// least min_mth row tiles (operators that share a batch share one count)
// and are padded to a multiple of kc rows of depth.
bool gemm_split_plan(const double* A, int M, int K, bool prefer_cols, double tol, int min_mth,
                     int kc, GemmOp& op, std::vector<double>& packed);
// k-chunk of the split kernel with (upd) or without the fused substep update
int gemm_split_kc(bool upd);
// Every job of the batch split with the same sp_mth, sp_ib and sp_ob (and
// all with or all without the substep update).
bool gemm_split_batch_ok(const GemmBatch& b, int nj);
hipError_t launch_gemm_split(const GemmBatch& b, int nj, hipStream_t s);
// int8 slices of the split halves of an operator planned by gemm_split_plan
// (packed = its FP64 halves); fills the sA8 layout, ssa, safirst, snti, snkh
void gemm_i8_split_halves(const std::vector<double>& packed, const GemmOp& op, int S,
                          std::vector<int8_t>& a8, std::vector<double>& sa,
                          std::vector<double>& afirst, int& snti, int& snkh);
// every job carries split slices with the same sp_ib / sp_ob, no substep update
bool gemm_i8_split_batch_ok(const GemmBatch& b, int nj);
hipError_t launch_gemm_i8_split(const GemmBatch& b, int nj, hipStream_t s);
// Mixed precision (gemm_i8.hip): the same contract from int8 slice products,
// b.slices slices per operand; no fused SBP update.
bool gemm_i8_supported(int S);
hipError_t launch_gemm_i8(const GemmBatch& b, int nj, hipStream_t s);
void gemm_i8_split(const double* A, int M, int K, int S, std::vector<int8_t>& a8,
                   std::vector<double>& sa, std::vector<double>& afirst);
// Learned closure (closure.hip): E_q[n0 + j][c] += out_scale * MLP(in_scale *
// q[n0 .. n0+15][c]) for every column c < nm2 of the level-major fields.
hipError_t launch_closure(int nm2, int ld, int n0, const void* wb, const double* q, double* Eq,
                          float in_scale, float out_scale, hipStream_t s);
struct nhqg_mlp;
int closure_export_bf16(const nhqg_mlp* m, void** dst, hipStream_t s, std::string& err);
// out[l][c] = epilogue(0) over nl rows of ld doubles.
hipError_t launch_combine(int nl, int ld, double* out, const Epi& e, hipStream_t s);
