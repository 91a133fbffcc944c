// Context, step orchestration and the C ABI (include/nhqg.h).
//
// One context = one HIP stream, device-resident state in the band-compact
// layout, and the whole ARS(2,2,2) balanced_sbp2_pc step
// (nhqg/solver.py:1167-1225, dispatched by imex_step :1228-1242) captured
// once into a hipGraph and replayed per step.
#include <string>
#include <limits>
#include <vector>
#include <cstring>
#include <cstdlib>
#include <cstdio>
#include <cmath>
#include <algorithm>
#include <initializer_list>

#include "common.h"
#include "../../include/nhqg.h"

namespace {

std::string g_create_error;

struct Ops {
  // contraction operators, packed k-major (GemmOp); Psbp = T2S * S2T is one
  // full SBP substep cycle (SBP grid -> Galerkin theta -> SBP grid)
  GemmOp V, VSw, VSd, Vinv, Pw, Pd, Qw, GSw, W2S, T2S, S2T, Psbp, Mtw, Mwt, Mth2w, Sw, Sd, row_cc;
  // Dirichlet-top shell solves in the eigenbasis of B0 (eigQ.At set)
  GemmOp eigQ, eigQinv;
  // vertical_cutoff_n: w and theta maps of _apply_vertical_cutoff (Cw.At set)
  GemmOp Cw, Cth;
  double *lam, *mal, *mc, *dtab;
  double *thb2s, *s2thb, *proj_d;
  double *d1b, *lb, *cc, *cc_sbp;
  int *rows, *kidx_of_x, *full_idx, *partner, *shell_ptr, *shell_modes;
  int2* pairs;
  double *kx_row, *ky_col, *kx_mode, *ksq, *inv_denom, *iaq, *iath, *idq, *diss_q, *diss_w,
      *diss_th, *mask, *pw, *ones, *pw_ke, *pw_cc;
  double *shell_inv, *eta_phys, *sponge_phys, *eta_x, *eta_y;
  double2* eta_rows;   // the trap in the Nx = 256 row pass's input form
  double2* tw;
  // mean temperature (evolve_mean): V G_Z, tau solve, diffusion, flux map
  double *vgz, *msolve, *mdiff, *mflux;
  // horizontal-mean weights and partners of the padded 3/2 grid
  double* pw_pad;
  int* partner_pad;
  // the SBP corrector's horizontal means: the Nx grid's under the 2/3 rule,
  // the padded grid's under the 3/2 rule (solver.py:873-906)
  double* pw_corr;
  int* partner_corr;
  // tables of the grid the spectral products run on (the Nx grid, or the
  // padded npad grid under the 3/2 rule): row map, kept rows, twiddles
  int *fkidx, *frows;
  double2* ftw;
  // SBP substep transfer as a five-diagonal level stencil (when Psbp Z is
  // banded): pb (Nq x 5), the corrector mean's diagonal weights per real
  // column and its cross-term pairs on the self-conjugate columns
  double* psbp_band = nullptr;
  double* pwc = nullptr;
  int2* xpairs = nullptr;
  double* xpw = nullptr;
  int nxpairs = 0;
};

}  // namespace

struct nhqg_ctx {
  nhqg_params P;
  int Nx = 0, Nk = 0, Nq = 0, Ng = 0, nkx = 0, nky = 0, nm = 0, nm_pad = 0, ld = 0, ldc = 0;
  int Npad = 0;   // transform size of the spectral products (nx, or 3nx/2)
  int Pred = 0;
  size_t bytes = 0;
  hipStream_t st = nullptr;
  // side stream of the step (default priority, like st): the SBP corrector's
  // preamble runs there, forked from st by ev_fork and joined by ev_join
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool have_state = false;
  double last_ms = 0.0;
  std::vector<void*> allocs;
  std::string err;
  Ops o;
  // state
  double *yq, *yw, *yth, *ythb;
  // workspace
  double *qn, *wn, *thn;
  double *E1q, *E1w, *E1th, *E2q, *E2w, *E2th;
  double *Rq, *Rw, *Rth, *Rthb;
  double *s1q, *s1w, *s1th, *s1thb, *c1th, *c1thb;
  double *rhs_w, *tmpg;
  double *wsbp, *thsbp, *thsbp2;
  double *spq, *spw, *spth;
  double *partial, *sbp, *g;
  double *sbp_xy, *sbp_map, *thbt;   // corrector substep map (X|Y, [M1 M2; G1 G2]) and scratch
  double *rk4_dq = nullptr, *rk4_dw = nullptr, *rk4_dth = nullptr;   // exp(-rate dt) per mode
  double2 *X5, *Y6;
  size_t fstride;
  double2* staging;
  double *diag, *rowmax;
  // mean-temperature vectors (Nq): dTheta/dZ at the CGL levels, E_th_bar of
  // the two stages, I_th_bar, and the RK4 validator's th_bar stage values
  double *mgrad = nullptr, *etb1 = nullptr, *etb2 = nullptr, *itb = nullptr, *ktb = nullptr,
         *acctb = nullptr, *stb = nullptr;
  double* mpow = nullptr;            // nhqg_mode_power: 3 x nm
  std::vector<double> h_thb2s, h_d1b;  // host copies for the exchange monitor
  // per-launch timing of one eager step (nhqg_profile_step)
  struct PRec {
    int cls;
    hipEvent_t a, b;
    double work;
  };
  bool prof = false;
  std::vector<PRec> recs;
  // learned closure (nhqg_set_closure): BF16 weights, top-level offset, scales
  void* cl_w = nullptr;
  bool cl_on = false;
  int cl_n0 = 0;
  float cl_in = 1.0f, cl_out = 1.0f;
  // mixed precision: int8 slices per operand (0 = FP64); mixed_now is set
  // while the step's contractions are enqueued (diagnostics stay FP64)
  int mixed_slices = 0;
  bool mixed_now = false;
  // w and theta sponges on the nodal rows of the flux products where the
  // row pass supports it (row_flux16); false: all three on coefficient levels
  bool nodal_sponge = true;
  // stage 2's V q1 and VSw w1 beside the stage-1 corrector preamble
  bool qw_early = true;
  // stage 1 stores y + dt delta E1 in place of E1 where E1 is final
  bool y1_tail = true;
  // SBP substeps by the banded stencil (o.psbp_band); Pband: its partial
  // buffer in levels' worth of doubles (one mean + the block sums)
  bool sbp_band = true;
  int Pband = 0;
};

namespace {

#define CK(expr)                                                                      \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      c->err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
      return NHQG_ERR_HIP;                                                            \
    }                                                                                 \
  } while (0)

template <typename T>
int dalloc(nhqg_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  const size_t b = n * sizeof(T);
  CK(hipMalloc(&q, b > 0 ? b : 16));
  c->allocs.push_back(q);
  c->bytes += b;
  *p = (T*)q;
  return NHQG_OK;
}

template <typename T>
int upload(nhqg_ctx* c, T** p, const T* host, size_t n) {
  if (!host) {
    *p = nullptr;
    return NHQG_OK;
  }
  int r = dalloc(c, p, n);
  if (r) return r;
  CK(hipMemcpy(*p, host, n * sizeof(T), hipMemcpyHostToDevice));
  return NHQG_OK;
}

// per-mode table padded with zeros to nm_pad
int upload_modes(nhqg_ctx* c, double** p, const double* host) {
  std::vector<double> v((size_t)c->nm_pad, 0.0);
  if (host) std::memcpy(v.data(), host, sizeof(double) * c->nm);
  return upload(c, p, v.data(), v.size());
}

// contraction operator (M x K row-major on the host) packed for
// launch_gemm_batch with mt row tiles
int upload_op(nhqg_ctx* c, GemmOp* op, const double* host, int M, int K, int mt) {
  *op = GemmOp{nullptr, M, K, mt, gemm_op_lda(mt), gemm_op_kp(K)};
  if (!host) return NHQG_OK;
  std::vector<double> v((size_t)op->lda * op->kp);
  gemm_pack_op(host, M, K, mt, v.data());
  double* d = nullptr;
  int r = upload(c, &d, v.data(), v.size());
  op->At = d;
  return r;
}

// Reflection-split halves of an uploaded operator (gemm_split.hip), unless
// dense contractions are asked for or the operator lacks the structure;
// prefer_cols picks the column butterfly for operators that have both (the
// CGL transform), so that it shares a batch with its neighbours; with_update
// marks the operators applied with the fused SBP substep update, whose
// kernel runs shallower k-chunks.
int upload_op_split(nhqg_ctx* c, GemmOp* op, const double* host, bool prefer_cols,
                    bool with_update = false) {
  if (!host || c->P.dense_contractions) return NHQG_OK;
  // every operator of a context gets the tile count of the largest half
  // (its Nz/2 + 1 rows: the last one on the VALU when Nz/2 is a multiple of 16)
  const int min_mth = ((c->Nq - 1) / 2 + 15) / 16;
  std::vector<double> packed;
  if (!gemm_split_plan(host, op->M, op->K, prefer_cols, 1e-11, min_mth,
                       gemm_split_kc(with_update), *op, packed))
    return NHQG_OK;
  op->sp_pref = prefer_cols;
  double* d = nullptr;
  int r = upload(c, &d, packed.data(), packed.size());
  op->sp_At = d;
  return r;
}

// int8 slices of an uploaded operator for the mixed-precision kernel
int upload_op_i8(nhqg_ctx* c, GemmOp* op, const double* host, int S) {
  if (!host || S == 0 || op->K < 2) return NHQG_OK;
  if (op->sp_At) {
    // slices of the split halves (the packed FP64 halves planned again)
    GemmOp plan = *op;
    std::vector<double> packed;
    const int min_mth = op->sp_mth;
    if (gemm_split_plan(host, op->M, op->K, op->sp_pref != 0, 1e-11, min_mth, op->sp_kc, plan,
                        packed) &&
        plan.sp_mth == op->sp_mth) {
      std::vector<int8_t> a8;
      std::vector<double> sa, al;
      int snti = 0, snkh = 0;
      gemm_i8_split_halves(packed, plan, S, a8, sa, al, snti, snkh);
      int8_t* d8 = nullptr;
      double *dsa = nullptr, *dal = nullptr;
      int r = upload(c, &d8, a8.data(), a8.size());
      if (!r) r = upload(c, &dsa, sa.data(), sa.size());
      if (!r) r = upload(c, &dal, al.data(), al.size());
      if (r) return r;
      op->sA8 = d8;
      op->ssa = dsa;
      op->safirst = dal;
      op->snti = snti;
      op->snkh = snkh;
    }
  }
  std::vector<int8_t> a8;
  std::vector<double> sa, al;
  gemm_i8_split(host, op->M, op->K, S, a8, sa, al);
  int8_t* d8 = nullptr;
  double *dsa = nullptr, *dal = nullptr;
  int r = upload(c, &d8, a8.data(), a8.size());
  if (!r) r = upload(c, &dsa, sa.data(), sa.size());
  if (!r) r = upload(c, &dal, al.data(), al.size());
  if (r) return r;
  op->A8 = d8;
  op->sa = dsa;
  op->afirst = dal;
  op->slices = S;
  op->nti = (op->M + 15) / 16;
  op->nk = (op->K - 1 + 63) / 64;
  return NHQG_OK;
}

int field(nhqg_ctx* c, double** p, int nl) {
  int r = dalloc(c, p, (size_t)nl * c->ld);
  if (r) return r;
  CK(hipMemset(*p, 0, (size_t)nl * c->ld * sizeof(double)));
  return NHQG_OK;
}

// ---------------------------------------------------------------------------
// step pieces
// ---------------------------------------------------------------------------
#define LK(expr)                                                                      \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      c->err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
      return NHQG_ERR_HIP;                                                            \
    }                                                                                 \
  } while (0)

// Kernel classes timed by nhqg_profile_step; work is the class's unit:
// flops for contractions (4*M*K per complex mode), bytes moved otherwise.
enum {
  CL_GEMM = 0,
  CL_FLUX_COLINV,
  CL_FLUX_ROW,
  CL_FLUX_COLFWD,
  CL_SHELL,
  CL_PARSEVAL,
  CL_SBP,
  CL_ELEMENTWISE,
  CL_SMALL,
  CL_GEMM_I8,
  CL_SBP_BAND,
  CL_COUNT
};
const char* const kClassNames[CL_COUNT] = {
    "gemm_f64_mfma", "flux_col_inverse", "flux_row_products", "flux_col_forward",
    "shell_solve", "parseval_means", "sbp_serial", "elementwise", "small", "gemm_i8_mixed",
    "sbp_band"};

void prof_begin(nhqg_ctx* c, hipEvent_t* a) {
  (void)hipEventCreate(a);
  (void)hipEventRecord(*a, c->st);
}

void prof_end(nhqg_ctx* c, int cls, hipEvent_t a, double work) {
  hipEvent_t b = nullptr;
  (void)hipEventCreate(&b);
  (void)hipEventRecord(b, c->st);
  c->recs.push_back({cls, a, b, work});
}

#define PL(cls, work, expr)                         \
  do {                                              \
    hipEvent_t pa_ = nullptr;                       \
    if (c->prof) prof_begin(c, &pa_);               \
    LK(expr);                                       \
    if (c->prof) prof_end(c, (cls), pa_, (work));   \
  } while (0)

// One contraction of a batch (launch_gemm_batch); W/g/bc select the fused
// SBP substep update of the input.
struct Job {
  const GemmOp* op;
  const double* B;
  double* C;
  Epi e;
  const double* W = nullptr;
  const double* g = nullptr;
  double bc = 0.0;
  double* C2 = nullptr;
  Epi e2 = epi_make(0.0);
};

int gemm_jobs(nhqg_ctx* c, std::initializer_list<Job> jobs) {
  GemmBatch b;
  std::memset(&b, 0, sizeof(b));
  b.ld = c->ld;
  int n = 0;
  double work = 0.0;
  for (const Job& j : jobs) {
    if (n == GEMM_MAX_JOBS) {
      c->err = "too many contractions in one batch";
      return NHQG_ERR_ARG;
    }
    GemmJob& g = b.job[n++];
    g.op = *j.op;
    g.B = j.B;
    g.C = j.C;
    g.W = j.W;
    g.g = j.g;
    g.bc = j.bc;
    g.e = j.e;
    g.C2 = j.C2;
    g.e2 = j.e2;
    work += 4.0 * j.op->M * j.op->K * (double)c->nm;
  }
  // int8 slice kernel inside the step for operators that carry slices
  // (not the SBP transfers, not the fused substep update)
  bool i8 = c->mixed_now && c->mixed_slices > 0;
  for (int k = 0; k < n && i8; ++k) i8 = b.job[k].op.A8 != nullptr && b.job[k].W == nullptr;
  if (i8) b.slices = c->mixed_slices;
  if (i8 && gemm_i8_split_batch_ok(b, n)) {
    // the split halves as int8 slice products
    double sw = 0.0;
    for (int k = 0; k < n; ++k) {
      const GemmOp& o = b.job[k].op;
      sw += 4.0 * ((double)o.sp_nr[0] * o.sp_nk[0] + (double)o.sp_nr[1] * o.sp_nk[1]) * c->nm;
    }
    PL(CL_GEMM_I8, sw, launch_gemm_i8_split(b, n, c->st));
  } else if (i8) {
    b.slices = c->mixed_slices;
    PL(CL_GEMM_I8, work, launch_gemm_i8(b, n, c->st));
  } else if (gemm_split_batch_ok(b, n)) {
    // the split products' own flops: 4 * (rows * depth of each half) per mode
    double sw = 0.0;
    for (int k = 0; k < n; ++k) {
      const GemmOp& o = b.job[k].op;
      sw += 4.0 * ((double)o.sp_nr[0] * o.sp_nk[0] + (double)o.sp_nr[1] * o.sp_nk[1]) * c->nm;
    }
    PL(CL_GEMM, sw, launch_gemm_split(b, n, c->st));
  } else {
    PL(CL_GEMM, work, launch_gemm_batch(b, n, c->st));
  }
  return NHQG_OK;
}

int gemm(nhqg_ctx* c, const GemmOp& op, const double* B, double* C, const Epi& e) {
  return gemm_jobs(c, {Job{&op, B, C, e}});
}

int combine(nhqg_ctx* c, int nl, double* out, const Epi& e) {
  PL(CL_ELEMENTWISE, 16.0 * (e.nX + 1) * nl * (double)c->nm,
     launch_combine(nl, c->ld, out, e, c->st));
  return NHQG_OK;
}

#define RUN(expr)             \
  do {                        \
    int r_ = (expr);          \
    if (r_) return r_;        \
  } while (0)

// explicit_rhs on the base grid (nhqg/solver.py:510-599; E_th_bar = 0).
// With R set, the back-transform contractions also write R_f = y_f + al*E_f
// (the stage-1 predictor right-hand sides) from the same accumulators, for
// each field whose E is final there (E_q only without beta and closure);
// the return value flags the fields done that way (bit f).
//
// Under evolve_mean with the legacy exchange (thb and Etb given) it also
// forms E_th_bar = -eps^2 G_Z V_inv <w theta> from the padded-grid means of
// the nodal w, theta, and adds dTheta/dZ * w (nodal) to the advection of
// theta before its back transform: E_th = -P_d V_inv (A_th + dTheta/dZ w)
// (nhqg/solver.py:575-590).
int explicit_rhs(nhqg_ctx* c, const double* q, const double* w, const double* th, double* Eq,
                 double* Ew, double* Eth, double* const* R = nullptr, double al = 0.0,
                 int* fused = nullptr, const Epi* th_tail = nullptr, double th_scale = 0.0,
                 const double* thb = nullptr, double* Etb = nullptr, bool qw_nodal = false,
                 double ytail = 0.0, int* ytailed = nullptr) {
  const Ops& o = c->o;
  // qw_nodal: qn = V q and wn = VSw w are already on the stream (the step
  // issues them beside the stage-1 corrector's preamble)
  if (qw_nodal)
    RUN(gemm_jobs(c, {Job{&o.VSd, th, c->thn, epi_make(1.0)}}));
  else
    RUN(gemm_jobs(c, {Job{&o.V, q, c->qn, epi_make(1.0)}, Job{&o.VSw, w, c->wn, epi_make(1.0)},
                      Job{&o.VSd, th, c->thn, epi_make(1.0)}}));
  const bool legacy = thb && Etb;
  if (legacy) {
    const int n = c->Nq;
    PL(CL_PARSEVAL, 32.0 * n * (double)c->nm,
       launch_parseval(n, c->nm, c->ld, c->Pred, c->wn, c->thn, o.pw_pad, o.partner_pad,
                       c->partial, c->st));
    PL(CL_SMALL, 0.0, launch_partial_matvec(n, n, c->Pred, c->partial, o.mflux, Etb, c->st));
    PL(CL_SMALL, 0.0, launch_small_matvec(n, n, o.vgz, thb, c->mgrad, c->st));
    // dTheta/dZ * w at the CGL levels, held until the advection is done
    PL(CL_ELEMENTWISE, 32.0 * n * (double)c->nm,
       launch_level_scale(n, c->ld, c->mgrad, c->wn, c->thsbp2, c->st));
  }
  FluxArgs a;
  a.N = c->Npad;
  a.out_scale = (double)c->Npad * c->Npad / ((double)c->Nx * c->Nx);
  a.nq = c->Nq;
  a.nkx = c->nkx;
  a.nky = c->nky;
  a.ldc = c->ldc;
  a.qn = (const double2*)c->qn;
  a.wn = (const double2*)c->wn;
  a.thn = (const double2*)c->thn;
  a.ng = c->Ng;
  a.qc = (const double2*)q;
  a.wc = (const double2*)w;
  a.thc = (const double2*)th;
  a.aq = (double2*)c->qn;
  a.aw = (double2*)c->wn;
  a.ath = (double2*)c->thn;
  a.sq = (double2*)c->spq;
  a.sw = (double2*)c->spw;
  a.sth = (double2*)c->spth;
  a.kidx_of_x = o.fkidx;
  a.rows = o.frows;
  a.kx_row = o.kx_row;
  a.ky_col = o.ky_col;
  a.inv_denom = o.inv_denom;
  a.mask = o.mask;
  a.eta_phys = o.eta_phys;
  a.eta_rows = o.eta_rows;
  a.sponge_phys = o.sponge_phys;
  a.jac = c->P.advection;
  a.eta_x = o.eta_x;
  a.eta_y = o.eta_y;
  a.tw = o.ftw;
  a.X5 = c->X5;
  a.Y6 = c->Y6;
  a.fstride = c->fstride;
  // flux form: the sponge of w and theta on their nodal rows inside the
  // row pass, only q' on its coefficient levels
  a.nsp = (c->nodal_sponge && a.sponge_phys && !a.jac) ? 1 : 0;
  const double F = 16.0 * c->Nq * (double)c->Npad * c->nky;   // one intermediate field
  const double Fc = 16.0 * c->Nq * (double)c->nm;            // one compact field
  const int sp = a.sponge_phys ? (a.nsp ? 1 : 3) : 0;
  // intermediates per pass: flux form 5 in / 6 out, Jacobian form 8 in / 3 out
  const int nxin = a.jac ? 8 : 5, nyout = a.jac ? 3 : 6;
  PL(CL_FLUX_COLINV, (3 + sp) * Fc + (nxin + sp) * F, launch_flux(a, 1, c->st));
  PL(CL_FLUX_ROW, (nxin + nyout + 2 * sp) * F, launch_flux(a, 2, c->st));
  PL(CL_FLUX_COLFWD, (nyout + sp) * F + (3 + sp) * Fc, launch_flux(a, 3, c->st));
  if (legacy) {
    Epi e = epi_make(0.0);
    epi_add(e, 1.0, c->thn);
    epi_add(e, 1.0, c->thsbp2);
    RUN(combine(c, c->Nq, c->thn, e));
  }
  // E = -(back transform of the advection) - sponge (solver.py:541-549, 595-597)
  Epi eq = epi_make(-1.0), ew = epi_make(-1.0), et = epi_make(-1.0);
  if (a.sponge_phys) {
    epi_add(eq, -1.0, c->spq);
    if (!a.nsp) {
      epi_add(ew, -1.0, c->spw);
      epi_add(et, -1.0, c->spth);
    }
  }
  if (th_tail) {
    // Eth receives th_scale * E_th + (the tail's terms) instead of E_th
    if (et.nX + th_tail->nX > NHQG_EPI_TERMS) {
      c->err = "too many epilogue terms";
      return NHQG_ERR_ARG;
    }
    et.a_acc *= th_scale;
    for (int t = 0; t < et.nX; ++t) et.aX[t] *= th_scale;
    for (int t = 0; t < th_tail->nX; ++t)
      epi_add(et, th_tail->aX[t], th_tail->X[t], th_tail->sX[t]);
  }
  // ytail: store y + ytail * E instead of E for every field whose E is final
  // here (q' not when the beta term or the closure add to it below); the
  // fields done are flagged in *ytailed
  if (ytailed) *ytailed = 0;
  if (ytail != 0.0) {
    Epi* es[3] = {&eq, &ew, &et};
    const double* ys[3] = {q, w, th};
    const bool q_ok = c->P.beta == 0.0 && !c->cl_on && !th_tail;
    for (int f = 0; f < 3; ++f) {
      if ((f == 0 && !q_ok) || (f == 2 && th_tail) || es[f]->nX + 1 > NHQG_EPI_TERMS) continue;
      Epi& e = *es[f];
      e.a_acc *= ytail;
      for (int t = 0; t < e.nX; ++t) e.aX[t] *= ytail;
      epi_add(e, 1.0, ys[f]);
      if (ytailed) *ytailed |= 1 << f;
    }
  }
  Job jq{&o.Vinv, c->qn, Eq, eq}, jw{&o.Pw, c->wn, Ew, ew}, jt{&o.Pd, c->thn, Eth, et};
  if (fused) *fused = 0;
  if (R) {
    // R = y + al * E = y - al * acc - al * sponge
    Job* js[3] = {&jq, &jw, &jt};
    const double* ys[3] = {q, w, th};
    double* sps[3] = {c->spq, c->spw, c->spth};
    const bool q_final = c->P.beta == 0.0 && !c->cl_on;
    for (int f = 0; f < 3; ++f) {
      if (f == 0 && !q_final) continue;
      Epi e2 = epi_make(-al);
      epi_add(e2, 1.0, ys[f]);
      if (a.sponge_phys && (f == 0 || !a.nsp)) epi_add(e2, -al, sps[f]);
      js[f]->C2 = R[f];
      js[f]->e2 = e2;
      if (fused) *fused |= 1 << f;
    }
  }
  RUN(gemm_jobs(c, {jq, jw, jt}));
  if (c->P.beta != 0.0)
    LK(launch_beta(c->Nq, c->ldc, c->nm, c->P.beta, o.kx_mode, o.inv_denom, (const double2*)q,
                   (double2*)Eq, c->st));
  if (c->cl_on)
    PL(CL_SMALL, 0.0,
       launch_closure(2 * c->nm, c->ld, c->cl_n0, c->cl_w, q, Eq, c->cl_in, c->cl_out, c->st));
  return NHQG_OK;
}

// matrix stream + rhs, R_th in + w, theta out
double shell_bytes(const nhqg_ctx* c) {
  const double np = shell_pad_order(c->Ng);
  return 8.0 * c->P.n_shells * np * np + 4.0 * 16.0 * c->Ng * (double)c->nm;
}

// imex_implicit_solve (nhqg/solver.py:691-738), q_boundary='none', in two
// parts: w and theta (implicit_solve_wth), then q from w (implicit_solve_q).
int implicit_solve_wth(nhqg_ctx* c, const double* Rq, const double* Rw, const double* Rth,
                       double* ow, double* oth) {
  const Ops& o = c->o;
  const double al = c->P.gamma_imex * c->P.dt;
  const double ars = al * c->P.ra_sigma;
  if (!c->P.w_neumann) {
    Epi e = epi_make(al);
    e.s_acc = o.idq;
    epi_add(e, 1.0, Rw);
    epi_add(e, ars, Rth, o.iath);
    RUN(gemm(c, o.Qw, Rq, c->rhs_w, e));
    if (o.eigQ.At) {
      // w = Q diag(1 / (alpha_m - c_m lam)) Q^-1 rhs_w, theta = (R_th + alpha w) iath
      Epi d = epi_make(1.0);
      d.d_tab = o.dtab;
      d.d_ld = c->nm_pad;
      RUN(gemm(c, o.eigQinv, c->rhs_w, c->tmpg, d));
      Job j{&o.eigQ, c->tmpg, ow, epi_make(1.0)};
      j.C2 = oth;
      j.e2 = epi_make(al);
      epi_add(j.e2, 1.0, Rth);
      j.e2.s_out = o.iath;
      RUN(gemm_jobs(c, {j}));
    } else {
      PL(CL_SHELL, shell_bytes(c),
         launch_shell_solve(c->P.n_shells, c->Ng, c->ld, o.shell_ptr, o.shell_modes,
                            o.shell_inv, c->rhs_w, ow, Rth, oth, o.iath, al, c->st));
    }
  } else {
    RUN(gemm(c, o.Mth2w, Rth, c->tmpg, epi_make(1.0)));
    Epi e = epi_make(al);
    e.s_acc = o.idq;
    epi_add(e, 1.0, Rw);
    epi_add(e, ars, c->tmpg, o.iath);
    RUN(gemm(c, o.Qw, Rq, c->rhs_w, e));
    // the fused theta output is overwritten below with the mapped form
    PL(CL_SHELL, shell_bytes(c),
       launch_shell_solve(c->P.n_shells, c->Ng, c->ld, o.shell_ptr, o.shell_modes, o.shell_inv,
                          c->rhs_w, ow, Rth, c->tmpg, o.iath, al, c->st));
    Epi et = epi_make(al);
    epi_add(et, 1.0, Rth);
    et.s_out = o.iath;
    RUN(gemm(c, o.Mwt, ow, oth, et));
  }
  return NHQG_OK;
}

int implicit_solve_q(nhqg_ctx* c, const double* Rq, const double* ow, double* oq) {
  const Ops& o = c->o;
  const double al = c->P.gamma_imex * c->P.dt;
  Epi eq = epi_make(al);
  epi_add(eq, 1.0, Rq);
  eq.s_out = o.iaq;
  RUN(gemm(c, o.GSw, ow, oq, eq));
  return NHQG_OK;
}

int implicit_solve(nhqg_ctx* c, const double* Rq, const double* Rw, const double* Rth,
                   double* oq, double* ow, double* oth) {
  RUN(implicit_solve_wth(c, Rq, Rw, Rth, ow, oth));
  return implicit_solve_q(c, Rq, ow, oq);
}

// Launches in the scope of a SideStream go to the side stream st2 (the
// helpers all launch on c->st); the destructor restores st on every path.
struct SideStream {
  nhqg_ctx* c;
  explicit SideStream(nhqg_ctx* x) : c(x) { std::swap(c->st, c->st2); }
  ~SideStream() { std::swap(c->st, c->st2); }
};

// SBP2 thermal corrector with hoisted invariants (nhqg/solver.py:909-950).
// th_out may equal th_in; thb_out is written (thb_in copied first).
// The substeps stay on the SBP grid: each one maps its updated field back
// and forth with the single operator Psbp = T2S * S2T (the reference applies
// S2T then T2S), and the update th - sdt * w_sbp * g with zeroed boundary
// rows is applied as the contraction reads its input; only the last substep
// returns to Galerkin coefficients (S2T).
//
// The preamble (w_sbp = W2S w, <w_sbp^2>, the factorisation and the substep
// maps) depends on w alone.  corrector_preamble_fork enqueues it on the side
// stream as soon as w is final, so that it runs beside the launches that
// precede the corrector; corrector(..., forked = true) then joins it before
// the first substep.
int corrector_preamble(nhqg_ctx* c, const double* w_in, double total_dt, bool with_t2s,
                       const double* th_in) {
  const Ops& o = c->o;
  const int n = c->Nq;
  const int ns = c->P.substeps > 0 ? c->P.substeps : 1;
  const double sdt = total_dt / ns;
  const double mu = c->P.eps_sq, kappa = 1.0 / c->P.sigma;
  if (with_t2s)
    RUN(gemm_jobs(c, {Job{&o.W2S, w_in, c->wsbp, epi_make(1.0)},
                      Job{&o.T2S, th_in, c->thsbp, epi_make(1.0)}}));
  else
    RUN(gemm_jobs(c, {Job{&o.W2S, w_in, c->wsbp, epi_make(1.0)}}));
  PL(CL_PARSEVAL, 16.0 * n * (double)c->nm,
     launch_parseval(n, c->nm, c->ld, c->Pred, c->wsbp, c->wsbp, o.pw_corr, o.partner_corr, c->partial,
                     c->st));
  PL(CL_SBP, 0.0, launch_sbp_prepare(n, c->Pred, c->partial, o.d1b, o.lb, 0.5 * mu * kappa * sdt,
                        0.25 * mu * sdt * sdt, c->sbp, c->st));
  PL(CL_SBP, 0.0, launch_sbp_maps(n, c->sbp, o.thb2s, o.s2thb, o.d1b, mu * sdt, c->sbp_xy,
                                  c->sbp_map, c->st));
  return NHQG_OK;
}

// The side-stream preamble (not in a profiled step, whose classes are timed
// on one stream).  Returns whether it forked.
int corrector_preamble_fork(nhqg_ctx* c, const double* w_in, double total_dt, bool* forked) {
  *forked = false;
  if (c->prof || !c->st2) return NHQG_OK;
  LK(hipEventRecord(c->ev_fork, c->st));
  LK(hipStreamWaitEvent(c->st2, c->ev_fork, 0));
  {
    SideStream side(c);
    RUN(corrector_preamble(c, w_in, total_dt, false, nullptr));
    LK(hipEventRecord(c->ev_join, c->st));
  }
  *forked = true;
  return NHQG_OK;
}

int corrector(nhqg_ctx* c, const double* w_in, const double* th_in, const double* thb_in,
              double* th_out, double* thb_out, double total_dt, bool forked = false) {
  const Ops& o = c->o;
  const int n = c->Nq;
  const int ns = c->P.substeps > 0 ? c->P.substeps : 1;
  const double sdt = total_dt / ns;
  double* cur = c->thsbp;
  double* nxt = c->thsbp2;
  if (forked) {
    RUN(gemm(c, o.T2S, th_in, cur, epi_make(1.0)));
    LK(hipStreamWaitEvent(c->st, c->ev_join, 0));
  } else {
    RUN(corrector_preamble(c, w_in, total_dt, true, th_in));
  }
  // th_bar ping-pongs between thb_out and a scratch level vector; an even
  // number of substeps ends in thb_out (an odd one gets a final copy)
  double* tb_a = c->thbt;
  double* tb_b = thb_out;
  if (ns % 2 == 0) std::swap(tb_a, tb_b);
  PL(CL_SMALL, 0.0, launch_small_axpy(n, thb_in, thb_in, thb_in, 0.0, tb_a, c->st));
  // with the banded transfer each substep but the last also sums the next
  // substep's <w_sbp theta_sbp> (one value per level)
  const bool band = o.psbp_band != nullptr;
  int P = c->Pred;
  for (int s = 0; s < ns; ++s) {
    if (s == 0 || !band) {
      PL(CL_PARSEVAL, 32.0 * n * (double)c->nm,
         launch_parseval(n, c->nm, c->ld, c->Pred, c->wsbp, cur, o.pw_corr, o.partner_corr,
                         c->partial, c->st));
      P = c->Pred;
    }
    PL(CL_SBP, 0.0, launch_sbp_apply(n, P, c->partial, c->sbp_map, tb_a, tb_b, c->g, c->st));
    std::swap(tb_a, tb_b);
    if (band && s + 1 < ns) {
      PL(CL_SBP_BAND, 48.0 * n * (double)c->nm,
         launch_sbp_band(n, c->ld, o.psbp_band, cur, c->wsbp, c->g, sdt, o.pwc, nxt, c->partial,
                         c->partial + n, o.nxpairs, o.xpairs, o.xpw, c->st));
      P = 1;
    } else {
      Job j{s + 1 < ns ? &o.Psbp : &o.S2T, cur, s + 1 < ns ? nxt : th_out, epi_make(1.0)};
      j.W = c->wsbp;
      j.g = c->g;
      j.bc = sdt;
      RUN(gemm_jobs(c, {j}));
    }
    std::swap(cur, nxt);
  }
  // after the loop tb_a holds the last result
  if (tb_a != thb_out)
    PL(CL_SMALL, 0.0, launch_small_axpy(n, tb_a, tb_a, tb_a, 0.0, thb_out, c->st));
  return NHQG_OK;
}

// _apply_vertical_cutoff (nhqg/solver.py:177-191): w <- Cw w and theta <-
// Cth theta, Cw = w_pinv proj_w diag(n <= cutoff) w_stencil (Cth likewise
// with the Dirichlet maps); the identity without a cutoff
int vertical_cutoff(nhqg_ctx* c, double* w, double* th) {
  const Ops& o = c->o;
  if (!o.Cw.At) return NHQG_OK;
  RUN(gemm_jobs(c, {Job{&o.Cw, w, c->tmpg, epi_make(1.0)},
                    Job{&o.Cth, th, c->rhs_w, epi_make(1.0)}}));
  const size_t bytes = sizeof(double) * c->Ng * c->ld;
  CK(hipMemcpyAsync(w, c->tmpg, bytes, hipMemcpyDeviceToDevice, c->st));
  CK(hipMemcpyAsync(th, c->rhs_w, bytes, hipMemcpyDeviceToDevice, c->st));
  return NHQG_OK;
}

int sanitize(nhqg_ctx* c) {
  const Ops& o = c->o;
  PL(CL_SMALL, 0.0, launch_sanitize3(c->Nq, c->Ng, c->ld, c->P.n_pairs, o.pairs, 0, c->yq,
                                     c->yw, c->yth, c->st));
  if (c->P.apply_mask) {
    LK(launch_mask(c->Nq, c->ld, c->yq, o.mask, c->st));
    LK(launch_mask(c->Ng, c->ld, c->yw, o.mask, c->st));
    LK(launch_mask(c->Ng, c->ld, c->yth, o.mask, c->st));
  }
  return NHQG_OK;
}

// One full step y <- imex_step(y).
int enqueue_step_body(nhqg_ctx* c) {
  const Ops& o = c->o;
  const double dt = c->P.dt;
  const double g = c->P.gamma_imex;
  const double al = g * dt;
  const double om = dt * (1.0 - g);
  const double delta = -0.70710678118654752440;  // -sqrt(2)/2
  const int Nq = c->Nq, Ng = c->Ng;
  // evolve_mean: balanced_sbp2_pc (predictors on the base grid, an SBP
  // corrector after each stage, solver.py:1167-1225) or the legacy exchange
  // (imex_step_ars222 with the exchange in the explicit RHS and the mean's
  // tau solve, solver.py:1012-1053)
  const bool legacy = c->P.evolve_mean != 0 && c->P.exchange == 1;
  const bool em = c->P.evolve_mean != 0 && !legacy;

  // stage 1 predictor; R = y + al*E1 comes out of the back transforms
  // where E1 is final there
  double* const R1[3] = {c->Rq, c->Rw, c->Rth};
  int fused = 0;
  // stage 2 uses E1 only as y + dt delta E1: where E1 is final in its back
  // transform, that sum is stored in its place (E1 then holds Y1), which
  // saves stage 2 the read of y
  int yt = 0;
  RUN(explicit_rhs(c, c->yq, c->yw, c->yth, c->E1q, c->E1w, c->E1th, R1, al, &fused, nullptr,
                   0.0, legacy ? c->ythb : nullptr, legacy ? c->etb1 : nullptr, false,
                   c->y1_tail ? dt * delta : 0.0, &yt));
  if (yt & ~fused) {
    c->err = "stage-1 tail on a field without its fused right-hand side";
    return NHQG_ERR_ARG;
  }
  {
    const double* ys[3] = {c->yq, c->yw, c->yth};
    const double* es[3] = {c->E1q, c->E1w, c->E1th};
    const int nl[3] = {Nq, Ng, Ng};
    for (int f = 0; f < 3; ++f) {
      if (fused & (1 << f)) continue;
      Epi e = epi_make(0.0);
      epi_add(e, 1.0, ys[f]);
      epi_add(e, al, es[f]);
      RUN(combine(c, nl[f], R1[f], e));
    }
  }
  // the corrector's preamble forks as soon as w1 is final
  bool fk = false;
  RUN(implicit_solve_wth(c, c->Rq, c->Rw, c->Rth, c->s1w, c->s1th));
  if (em) RUN(corrector_preamble_fork(c, c->s1w, al, &fk));
  RUN(implicit_solve_q(c, c->Rq, c->s1w, c->s1q));
  if (legacy) {
    // th_bar1 = mean_solve (th_bar + al E1_th_bar)
    PL(CL_SMALL, 0.0, launch_small_lincomb(Nq, c->Rthb, 1.0, c->ythb, al, c->etb1, 0.0, nullptr,
                                           0.0, nullptr, c->st));
    PL(CL_SMALL, 0.0, launch_small_matvec(Nq, Nq, o.msolve, c->Rthb, c->s1thb, c->st));
  } else {
    PL(CL_SMALL, 0.0, launch_small_matvec(Nq, Nq, o.proj_d, c->ythb, c->s1thb, c->st));
  }
  // stage 2's nodal q1 and w1 do not depend on the corrector (it changes
  // theta alone): with the preamble forked, the main stream computes them
  // while the side stream runs the preamble's one-block kernels
  const bool qw_early = em && fk && c->qw_early;
  if (qw_early)
    RUN(gemm_jobs(c, {Job{&o.V, c->s1q, c->qn, epi_make(1.0)},
                      Job{&o.VSw, c->s1w, c->wn, epi_make(1.0)}}));
  // stage 1 corrector
  const double* th1 = c->s1th;
  const double* thb1 = c->s1thb;
  if (em) {
    RUN(corrector(c, c->s1w, c->s1th, c->s1thb, c->c1th, c->c1thb, al, fk));
    th1 = c->c1th;
    thb1 = c->c1thb;
  }
  // stage 2.  With a Dirichlet w (I_th = w1), R_th2 below comes out of the
  // theta back transform: dt(1-delta) E_th2 plus the other terms
  const bool rth_fused = !c->P.w_neumann;
  Epi htail = epi_make(0.0);
  if (rth_fused) {
    if (yt & 4) {
      epi_add(htail, 1.0, c->E1th);
    } else {
      epi_add(htail, 1.0, c->yth);
      epi_add(htail, dt * delta, c->E1th);
    }
    epi_add(htail, om, c->s1w);
    epi_add(htail, -om, th1, o.diss_th);
    if (em) {
      epi_add(htail, om / al, th1);
      epi_add(htail, -om / al, c->s1th);
    }
  }
  RUN(explicit_rhs(c, c->s1q, c->s1w, th1, c->E2q, c->E2w, rth_fused ? c->Rth : c->E2th, nullptr,
                   0.0, nullptr, rth_fused ? &htail : nullptr, dt * (1.0 - delta),
                   legacy ? c->s1thb : nullptr, legacy ? c->etb2 : nullptr, qw_early));
  {
    // R_q2 = y + dt(delta E1 + (1-delta) E2) + om*(G_Z S_w w1) - om*diss_q*q1
    Epi e = epi_make(om);
    if (yt & 1) {
      epi_add(e, 1.0, c->E1q);
    } else {
      epi_add(e, 1.0, c->yq);
      epi_add(e, dt * delta, c->E1q);
    }
    epi_add(e, dt * (1.0 - delta), c->E2q);
    epi_add(e, -om, c->s1q, o.diss_q);
    // R_w2 = ... + om*(inv_denom Qw q1 + Ra_s Mtw th1) - om diss_w w1
    const double* thw = th1;
    if (c->P.w_neumann) {
      RUN(gemm(c, o.Mtw, th1, c->tmpg, epi_make(1.0)));
      thw = c->tmpg;
    }
    Epi f = epi_make(om);
    f.s_acc = o.inv_denom;
    if (yt & 2) {
      epi_add(f, 1.0, c->E1w);
    } else {
      epi_add(f, 1.0, c->yw);
      epi_add(f, dt * delta, c->E1w);
    }
    epi_add(f, dt * (1.0 - delta), c->E2w);
    epi_add(f, om * c->P.ra_sigma, thw);
    epi_add(f, -om, c->s1w, o.diss_w);
    RUN(gemm_jobs(c, {Job{&o.GSw, c->s1w, c->Rq, e}, Job{&o.Qw, c->s1q, c->Rw, f}}));
    // R_th2 = ... + om*I_th - om diss_th th1 + om*(th1 - th1p)/alpha
    if (!rth_fused) {
      const double* ith = c->s1w;
      if (c->P.w_neumann) {
        RUN(gemm(c, o.Mwt, c->s1w, c->tmpg, epi_make(1.0)));
        ith = c->tmpg;
      }
      Epi h = epi_make(0.0);
      if (yt & 4) {
        epi_add(h, 1.0, c->E1th);
      } else {
        epi_add(h, 1.0, c->yth);
        epi_add(h, dt * delta, c->E1th);
      }
      epi_add(h, dt * (1.0 - delta), c->E2th);
      epi_add(h, om, ith);
      epi_add(h, -om, th1, o.diss_th);
      if (em) {
        epi_add(h, om / al, th1);
        epi_add(h, -om / al, c->s1th);
      }
      RUN(combine(c, Ng, c->Rth, h));
    }
    if (em) {
      PL(CL_SMALL, 0.0, launch_small_axpy(Nq, c->ythb, thb1, c->s1thb, om / al, c->Rthb, c->st));
    } else if (legacy) {
      // R_th_bar2 = th_bar + dt (delta E1 + (1-delta) E2) + om eps^2/sigma G_Z2 th_bar1
      PL(CL_SMALL, 0.0, launch_small_matvec(Nq, Nq, o.mdiff, c->s1thb, c->itb, c->st));
      PL(CL_SMALL, 0.0, launch_small_lincomb(Nq, c->Rthb, 1.0, c->ythb, dt * delta, c->etb1,
                                             dt * (1.0 - delta), c->etb2, om, c->itb, c->st));
    } else {
      PL(CL_SMALL, 0.0, launch_small_axpy(Nq, c->ythb, c->ythb, c->ythb, 0.0, c->Rthb, c->st));
    }
  }
  RUN(implicit_solve_wth(c, c->Rq, c->Rw, c->Rth, c->yw, c->yth));
  if (em) RUN(corrector_preamble_fork(c, c->yw, al, &fk));
  RUN(implicit_solve_q(c, c->Rq, c->yw, c->yq));
  PL(CL_SMALL, 0.0, launch_small_matvec(Nq, Nq, legacy ? o.msolve : o.proj_d, c->Rthb, c->ythb,
                                        c->st));
  if (em) RUN(corrector(c, c->yw, c->yth, c->ythb, c->yth, c->ythb, al, fk));
  // _finalize_state (solver.py:993-999): zero_mean (in sanitize, which
  // commutes with the per-mode cutoff) and the optional vertical cutoff
  RUN(vertical_cutoff(c, c->yw, c->yth));
  RUN(sanitize(c));
  return NHQG_OK;
}

// ---------------------------------------------------------------------------
// RK4 validator (nhqg/solver.py:1249-1336): the full RHS is the explicit RHS
// plus the implicit tendency.  Under evolve_mean the th_bar tendency is
// I_th_bar = eps^2/sigma G_Z2 th_bar, plus E_th_bar and theta's
// -dTheta/dZ w term with the legacy exchange (split exchanges leave both to
// their substep, solver.py:552-556).  The stage states' _apply_bcs is the
// identity on Galerkin-basis fields (pinv * proj * S = I) and is skipped for
// them; th_bar is projected at every stage and at the end, as in the
// reference.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
exp_scale_kernel(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = exp(a * x[i]);
}

// K <- explicit_rhs(s) + implicit_tendency(s) (solver.py:1249-1258)
int full_rhs(nhqg_ctx* c, const double* q, const double* w, const double* th, const double* thb,
             double* Kq, double* Kw, double* Kth, double* Ktb) {
  const Ops& o = c->o;
  const bool em = c->P.evolve_mean != 0;
  const bool legacy = em && c->P.exchange == 1;
  RUN(explicit_rhs(c, q, w, th, Kq, Kw, Kth, nullptr, 0.0, nullptr, nullptr, 0.0,
                   legacy ? thb : nullptr, legacy ? Ktb : nullptr));
  // th_bar: (E_th_bar) + eps^2/sigma G_Z2 th_bar, or zero
  if (em) {
    PL(CL_SMALL, 0.0, launch_small_matvec(c->Nq, c->Nq, o.mdiff, thb, c->itb, c->st));
    PL(CL_SMALL, 0.0, launch_small_lincomb(c->Nq, Ktb, legacy ? 1.0 : 0.0, legacy ? Ktb : nullptr,
                                           1.0, c->itb, 0.0, nullptr, 0.0, nullptr, c->st));
  } else {
    PL(CL_SMALL, 0.0, launch_small_lincomb(c->Nq, Ktb, 0.0, nullptr, 0.0, nullptr, 0.0, nullptr,
                                           0.0, nullptr, c->st));
  }
  // I_q = G_Z S_w w, added to E_q in the contraction's epilogue (in place)
  Epi e = epi_make(1.0);
  epi_add(e, 1.0, Kq);
  RUN(gemm(c, o.GSw, w, Kq, e));
  // I_w = inv_denom Qw q + Ra_s (Mtw) th
  const double* thw = th;
  if (c->P.w_neumann) {
    RUN(gemm(c, o.Mtw, th, c->tmpg, epi_make(1.0)));
    thw = c->tmpg;
  }
  Epi f = epi_make(1.0);
  f.s_acc = o.inv_denom;
  epi_add(f, c->P.ra_sigma, thw);
  epi_add(f, 1.0, Kw);
  RUN(gemm(c, o.Qw, q, Kw, f));
  // I_th = (Mwt) w
  if (c->P.w_neumann) {
    Epi h = epi_make(1.0);
    epi_add(h, 1.0, Kth);
    RUN(gemm(c, o.Mwt, w, Kth, h));
  } else {
    Epi h = epi_make(0.0);
    epi_add(h, 1.0, Kth);
    epi_add(h, 1.0, w);
    RUN(combine(c, c->Ng, Kth, h));
  }
  return NHQG_OK;
}

int rk4_step(nhqg_ctx* c) {
  const Ops& o = c->o;
  const double dt = c->P.dt;
  const int Nq = c->Nq, Ng = c->Ng;
  if (c->P.evolve_mean && (!o.mdiff || (c->P.exchange == 1 && (!o.mflux || !o.vgz)))) {
    c->err = "rk4_step under evolve_mean needs the mean_diff (and, for the legacy exchange, "
             "mean_flux and VGZ) operators";
    return NHQG_ERR_ARG;
  }
  if (!c->rk4_dq) {
    RUN(dalloc(c, &c->rk4_dq, (size_t)c->nm_pad));
    RUN(dalloc(c, &c->rk4_dw, (size_t)c->nm_pad));
    RUN(dalloc(c, &c->rk4_dth, (size_t)c->nm_pad));
    const unsigned b = (unsigned)((c->nm_pad + 255) / 256);
    hipLaunchKernelGGL(exp_scale_kernel, dim3(b), dim3(256), 0, c->st, c->nm_pad, -dt, o.diss_q,
                       c->rk4_dq);
    hipLaunchKernelGGL(exp_scale_kernel, dim3(b), dim3(256), 0, c->st, c->nm_pad, -dt, o.diss_w,
                       c->rk4_dw);
    hipLaunchKernelGGL(exp_scale_kernel, dim3(b), dim3(256), 0, c->st, c->nm_pad, -dt, o.diss_th,
                       c->rk4_dth);
    CK(hipGetLastError());
  }
  double* y[3] = {c->yq, c->yw, c->yth};
  double* sv[3] = {c->s1q, c->s1w, c->s1th};
  double* k[3] = {c->E1q, c->E1w, c->E1th};
  double* acc[3] = {c->Rq, c->Rw, c->Rth};
  const int nl[3] = {Nq, Ng, Ng};
  // stage state s = y + a k (th_bar projected); with_acc: acc += a k
  auto axpy = [&](double* const* out, double a, bool with_acc) -> int {
    for (int f = 0; f < 3; ++f) {
      Epi e = epi_make(0.0);
      epi_add(e, 1.0, with_acc ? acc[f] : y[f]);
      epi_add(e, a, k[f]);
      RUN(combine(c, nl[f], out[f], e));
    }
    if (with_acc) {
      PL(CL_SMALL, 0.0, launch_small_lincomb(Nq, c->acctb, 1.0, c->acctb, a, c->ktb, 0.0, nullptr,
                                             0.0, nullptr, c->st));
    } else {
      PL(CL_SMALL, 0.0, launch_small_lincomb(Nq, c->Rthb, 1.0, c->ythb, a, c->ktb, 0.0, nullptr,
                                             0.0, nullptr, c->st));
      PL(CL_SMALL, 0.0, launch_small_matvec(Nq, Nq, o.proj_d, c->Rthb, c->stb, c->st));
      // _apply_bcs ends with the vertical cutoff (solver.py:1261-1273)
      RUN(vertical_cutoff(c, out[1], out[2]));
    }
    return NHQG_OK;
  };
  RUN(full_rhs(c, y[0], y[1], y[2], c->ythb, k[0], k[1], k[2], c->ktb));   // k1
  {
    // acc = y + dt/6 k1
    for (int f = 0; f < 3; ++f) {
      Epi e = epi_make(0.0);
      epi_add(e, 1.0, y[f]);
      epi_add(e, dt / 6.0, k[f]);
      RUN(combine(c, nl[f], acc[f], e));
    }
    PL(CL_SMALL, 0.0, launch_small_lincomb(Nq, c->acctb, 1.0, c->ythb, dt / 6.0, c->ktb, 0.0,
                                           nullptr, 0.0, nullptr, c->st));
  }
  RUN(axpy(sv, 0.5 * dt, false));                                             // s2
  RUN(full_rhs(c, sv[0], sv[1], sv[2], c->stb, k[0], k[1], k[2], c->ktb));   // k2
  RUN(axpy(acc, dt / 3.0, true));
  RUN(axpy(sv, 0.5 * dt, false));                                             // s3
  RUN(full_rhs(c, sv[0], sv[1], sv[2], c->stb, k[0], k[1], k[2], c->ktb));   // k3
  RUN(axpy(acc, dt / 3.0, true));
  RUN(axpy(sv, dt, false));                                                   // s4
  RUN(full_rhs(c, sv[0], sv[1], sv[2], c->stb, k[0], k[1], k[2], c->ktb));   // k4
  // y = (acc + dt/6 k4) * exp(-rate dt), the RK4-only post-step dissipation
  const double* dz[3] = {c->rk4_dq, c->rk4_dw, c->rk4_dth};
  for (int f = 0; f < 3; ++f) {
    Epi e = epi_make(0.0);
    epi_add(e, 1.0, acc[f]);
    epi_add(e, dt / 6.0, k[f]);
    e.s_out = dz[f];
    RUN(combine(c, nl[f], y[f], e));
  }
  // th_bar = proj_d (acc + dt/6 k4)
  PL(CL_SMALL, 0.0, launch_small_lincomb(Nq, c->Rthb, 1.0, c->acctb, dt / 6.0, c->ktb, 0.0,
                                         nullptr, 0.0, nullptr, c->st));
  PL(CL_SMALL, 0.0, launch_small_matvec(Nq, Nq, o.proj_d, c->Rthb, c->ythb, c->st));
  RUN(vertical_cutoff(c, c->yw, c->yth));
  RUN(sanitize(c));
  return NHQG_OK;
}

// the step's contractions take the int8 slice kernel when mixed precision
// is on; everything else (diagnostics, component calls) stays FP64 unless
// it sets mixed_now itself
int enqueue_step(nhqg_ctx* c) {
  c->mixed_now = c->mixed_slices > 0;
  const int r = enqueue_step_body(c);
  c->mixed_now = false;
  return r;
}

int build_graph(nhqg_ctx* c) {
  if (c->gexec) return NHQG_OK;
  CK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  int r = enqueue_step(c);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(c->st, &gr);
  if (r) {
    if (gr) (void)hipGraphDestroy(gr);
    return r;
  }
  if (e != hipSuccess) {
    c->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
    return NHQG_ERR_HIP;
  }
  c->graph = gr;
  CK(hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0));
  return NHQG_OK;
}

int check_params(const nhqg_params* p, std::string& err) {
  const int nx = p->nx;
  if (nx < 16 || nx > 1024 || (nx & (nx - 1))) {
    err = "nx must be a power of two in [16, 1024]";
    return NHQG_ERR_ARG;
  }
  if (p->nz < 3 || p->nz > 256) {
    err = "nz must be in [3, 256]";
    return NHQG_ERR_ARG;
  }
  if (p->nkx < 1 || p->nkx > nx || p->nky < 1 || p->nky > nx / 2 + 1) {
    err = "mode layout out of range";
    return NHQG_ERR_ARG;
  }
  if (p->advection != 0 && p->advection != 1) {
    err = "advection must be 0 (flux form) or 1 (Jacobian form)";
    return NHQG_ERR_UNSUPPORTED;
  }
  if (p->mixed_slices != 0 && !gemm_i8_supported(p->mixed_slices)) {
    err = "mixed_slices must be 0 (FP64) or in [4, 7]";
    return NHQG_ERR_ARG;
  }
  if (p->dense_contractions != 0 && p->dense_contractions != 1) {
    err = "dense_contractions must be 0 or 1";
    return NHQG_ERR_ARG;
  }
  if (p->n_shells < 1) {
    err = "n_shells must be >= 1";
    return NHQG_ERR_ARG;
  }
  if (p->npad != 0 && p->npad != nx) {
    if (p->npad != 3 * nx / 2) {
      err = "npad must be nx (2/3 rule) or 3nx/2 (3/2 rule)";
      return NHQG_ERR_ARG;
    }
    if (p->apply_mask || p->nkx != nx || p->nky != nx / 2 + 1) {
      err = "the 3/2 rule needs the full mode layout (nkx = nx, nky = nx/2+1) and apply_mask 0";
      return NHQG_ERR_ARG;
    }
  }
  return NHQG_OK;
}

int create_impl(nhqg_ctx* c, const nhqg_params* p, const nhqg_operators* h) {
  c->P = *p;
  c->Nx = p->nx;
  c->Nk = p->nx / 2 + 1;
  c->Nq = p->nz + 1;
  c->Ng = p->nz - 1;
  c->nkx = p->nkx;
  c->nky = p->nky;
  c->nm = p->nkx * p->nky;
  c->nm_pad = (c->nm + 63) / 64 * 64;
  c->Npad = p->npad ? p->npad : p->nx;
  c->ld = 2 * c->nm_pad;
  c->ldc = c->nm_pad;
  c->Pred = 16;
  CK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  CK(hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking));
  CK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  CK(hipEventCreate(&c->ev0));
  CK(hipEventCreate(&c->ev1));
  const int Nq = c->Nq, Ng = c->Ng, nm = c->nm;
  Ops& o = c->o;
  const int mt = gemm_row_tiles(Nq);
  RUN(upload_op(c, &o.V, h->V, Nq, Nq, mt));
  RUN(upload_op(c, &o.VSw, h->VSw, Nq, Ng, mt));
  RUN(upload_op(c, &o.VSd, h->VSd, Nq, Ng, mt));
  RUN(upload_op(c, &o.Vinv, h->Vinv, Nq, Nq, mt));
  RUN(upload_op(c, &o.Pw, h->Pw, Ng, Nq, mt));
  RUN(upload_op(c, &o.Pd, h->Pd, Ng, Nq, mt));
  RUN(upload_op(c, &o.Qw, h->Qw, Ng, Nq, mt));
  RUN(upload_op(c, &o.GSw, h->GSw, Nq, Ng, mt));
  RUN(upload_op(c, &o.W2S, h->W2S, Nq, Ng, mt));
  RUN(upload_op(c, &o.T2S, h->T2S, Nq, Ng, mt));
  RUN(upload_op(c, &o.S2T, h->S2T, Ng, Nq, mt));
  // coefficient -> value maps split on their output rows (batched V, VSw,
  // VSd; W2S, T2S), value -> coefficient maps on their input rows (batched
  // Vinv, Pw, Pd; S2T)
  RUN(upload_op_split(c, &o.V, h->V, false));
  RUN(upload_op_split(c, &o.VSw, h->VSw, false));
  RUN(upload_op_split(c, &o.VSd, h->VSd, false));
  RUN(upload_op_split(c, &o.Vinv, h->Vinv, true));
  RUN(upload_op_split(c, &o.Pw, h->Pw, true));
  RUN(upload_op_split(c, &o.Pd, h->Pd, true));
  RUN(upload_op_split(c, &o.Qw, h->Qw, false));
  RUN(upload_op_split(c, &o.GSw, h->GSw, false));
  RUN(upload_op_split(c, &o.W2S, h->W2S, false));
  RUN(upload_op_split(c, &o.T2S, h->T2S, false));
  RUN(upload_op_split(c, &o.S2T, h->S2T, true, true));
  std::vector<double> Psbp_h;
  {
    // one SBP substep cycle: Psbp = T2S * S2T (Nq x Nq)
    std::vector<double>& P = Psbp_h;
    P.assign((size_t)Nq * Nq, 0.0);
    for (int i = 0; i < Nq; ++i)
      for (int k = 0; k < Ng; ++k) {
        const double a = h->T2S[(size_t)i * Ng + k];
        for (int j = 0; j < Nq; ++j) P[(size_t)i * Nq + j] += a * h->S2T[(size_t)k * Nq + j];
      }
    RUN(upload_op(c, &o.Psbp, P.data(), Nq, Nq, mt));
    RUN(upload_op_split(c, &o.Psbp, P.data(), false, true));
  }
  RUN(upload(c, &o.thb2s, h->thb2s, (size_t)Nq * Nq));
  RUN(upload(c, &o.s2thb, h->s2thb, (size_t)Nq * Nq));
  RUN(upload(c, &o.proj_d, h->proj_d, (size_t)Nq * Nq));
  RUN(upload(c, &o.d1b, h->d1_band, (size_t)Nq * 3));
  RUN(upload(c, &o.lb, h->lap_band, (size_t)Nq * 3));
  RUN(upload(c, &o.cc, h->cc, (size_t)Nq));
  RUN(upload_op(c, &o.row_cc, h->row_cc, 1, Nq, 1));
  RUN(upload(c, &o.cc_sbp, h->cc_sbp, (size_t)Nq));
  RUN(upload_op(c, &o.Sw, h->Sw, Nq, Ng, mt));
  RUN(upload_op(c, &o.Sd, h->Sd, Nq, Ng, mt));
  if ((h->cutoff_w == nullptr) != (h->cutoff_th == nullptr)) {
    c->err = "the vertical cutoff needs both cutoff_w and cutoff_th";
    return NHQG_ERR_ARG;
  }
  if (h->cutoff_w) {
    RUN(upload_op(c, &o.Cw, h->cutoff_w, Ng, Ng, mt));
    RUN(upload_op(c, &o.Cth, h->cutoff_th, Ng, Ng, mt));
    RUN(upload_op_split(c, &o.Cw, h->cutoff_w, false));
    RUN(upload_op_split(c, &o.Cth, h->cutoff_th, false));
  }
  if (p->w_neumann) {
    if (!h->Mtw || !h->Mwt || !h->Mth2w) {
      c->err = "w_neumann requires Mtw, Mwt and Mth2w";
      return NHQG_ERR_ARG;
    }
    RUN(upload_op(c, &o.Mtw, h->Mtw, Ng, Ng, mt));
    RUN(upload_op(c, &o.Mwt, h->Mwt, Ng, Ng, mt));
    RUN(upload_op(c, &o.Mth2w, h->Mth2w, Ng, Ng, mt));
  }
  // int8 slices of the Chebyshev transforms and the implicit-solve maps
  // (the SBP transfers, Psbp and the diagnostics operators stay FP64)
  c->mixed_slices = p->mixed_slices;
  if (const char* e = std::getenv("NHQG_NODAL_SPONGE")) c->nodal_sponge = std::atoi(e) != 0;
  if (const char* e = std::getenv("NHQG_QW_EARLY")) c->qw_early = std::atoi(e) != 0;
  if (const char* e = std::getenv("NHQG_Y1_TAIL")) c->y1_tail = std::atoi(e) != 0;
  if (const char* e = std::getenv("NHQG_SBP_BAND")) c->sbp_band = std::atoi(e) != 0;
  {
    const int S = c->mixed_slices;
    RUN(upload_op_i8(c, &o.V, h->V, S));
    RUN(upload_op_i8(c, &o.VSw, h->VSw, S));
    RUN(upload_op_i8(c, &o.VSd, h->VSd, S));
    RUN(upload_op_i8(c, &o.Vinv, h->Vinv, S));
    RUN(upload_op_i8(c, &o.Pw, h->Pw, S));
    RUN(upload_op_i8(c, &o.Pd, h->Pd, S));
    RUN(upload_op_i8(c, &o.Qw, h->Qw, S));
    RUN(upload_op_i8(c, &o.GSw, h->GSw, S));
    if (p->w_neumann) {
      RUN(upload_op_i8(c, &o.Mtw, h->Mtw, S));
      RUN(upload_op_i8(c, &o.Mwt, h->Mwt, S));
      RUN(upload_op_i8(c, &o.Mth2w, h->Mth2w, S));
    }
  }
  RUN(upload(c, &o.rows, h->rows, (size_t)c->nkx));
  RUN(upload(c, &o.kidx_of_x, h->kidx_of_x, (size_t)c->Nx));
  RUN(upload(c, &o.full_idx, h->full_idx, (size_t)nm));
  RUN(upload(c, &o.kx_row, h->kx_row, (size_t)c->nkx));
  RUN(upload(c, &o.ky_col, h->ky_col, (size_t)c->nky));
  RUN(upload_modes(c, &o.kx_mode, h->kx_mode));
  RUN(upload_modes(c, &o.ksq, h->ksq));
  RUN(upload_modes(c, &o.inv_denom, h->inv_denom));
  RUN(upload_modes(c, &o.iaq, h->inv_alpha_q));
  RUN(upload_modes(c, &o.iath, h->inv_alpha_th));
  RUN(upload_modes(c, &o.diss_q, h->diss_q));
  RUN(upload_modes(c, &o.diss_w, h->diss_w));
  RUN(upload_modes(c, &o.diss_th, h->diss_th));
  RUN(upload_modes(c, &o.mask, h->mask));
  RUN(upload_modes(c, &o.pw, h->pweight));
  {
    std::vector<double> idq((size_t)nm), ones((size_t)nm, 1.0), pke((size_t)nm);
    for (int m = 0; m < nm; ++m) {
      idq[m] = h->inv_denom[m] * h->inv_alpha_q[m];
      pke[m] = h->pweight[m] * h->ksq[m] * h->inv_denom[m] * h->inv_denom[m];
    }
    RUN(upload_modes(c, &o.idq, idq.data()));
    RUN(upload_modes(c, &o.ones, ones.data()));
    RUN(upload_modes(c, &o.pw_ke, pke.data()));
  }
  {
    std::vector<int> part((size_t)c->nm_pad, -1);
    std::memcpy(part.data(), h->partner, sizeof(int) * nm);
    RUN(upload(c, &o.partner, part.data(), part.size()));
  }
  RUN(upload(c, (int**)&o.pairs, h->herm_pairs, (size_t)2 * p->n_pairs));
  RUN(upload(c, &o.shell_ptr, h->shell_ptr, (size_t)p->n_shells + 1));
  RUN(upload(c, &o.shell_modes, h->shell_modes, (size_t)nm));
  const bool eigen = h->eig_Q && h->eig_Qinv && h->eig_lam && h->mode_alpha && h->mode_c &&
                     !p->w_neumann;
  if (eigen) {
    RUN(upload_op(c, &o.eigQ, h->eig_Q, Ng, Ng, mt));
    RUN(upload_op(c, &o.eigQinv, h->eig_Qinv, Ng, Ng, mt));
    RUN(upload_op_split(c, &o.eigQ, h->eig_Q, false));
    RUN(upload_op_split(c, &o.eigQinv, h->eig_Qinv, false));
    RUN(upload_op_i8(c, &o.eigQ, h->eig_Q, c->mixed_slices));
    RUN(upload_op_i8(c, &o.eigQinv, h->eig_Qinv, c->mixed_slices));
    RUN(upload(c, &o.lam, h->eig_lam, (size_t)Ng));
    // padding modes: alpha 1, c 0 (their columns are zero)
    std::vector<double> a((size_t)c->nm_pad, 1.0), cm((size_t)c->nm_pad, 0.0);
    std::memcpy(a.data(), h->mode_alpha, sizeof(double) * nm);
    std::memcpy(cm.data(), h->mode_c, sizeof(double) * nm);
    RUN(upload(c, &o.mal, a.data(), a.size()));
    RUN(upload(c, &o.mc, cm.data(), cm.size()));
    RUN(dalloc(c, &o.dtab, (size_t)Ng * c->nm_pad));
    LK(launch_eigen_diag(Ng, c->nm_pad, o.mal, o.mc, o.lam, o.dtab, c->st));
    CK(hipStreamSynchronize(c->st));
    o.shell_inv = nullptr;
  } else {
    // shell inverses, padded for the MFMA solve: uploaded (packed in slabs
    // of shells) or built and inverted on the device
    o.eigQ = GemmOp{nullptr, 0, 0, 0, 0, 0};
    o.eigQinv = o.eigQ;
    o.lam = o.mal = o.mc = o.dtab = nullptr;
    const int np = shell_pad_order(Ng);
    const size_t per = (size_t)np * np;
    RUN(dalloc(c, &o.shell_inv, (size_t)p->n_shells * per));
    if (h->shell_inv) {
      const int slab = 128;
      std::vector<double> buf((size_t)slab * per);
      for (int s0 = 0; s0 < p->n_shells; s0 += slab) {
        const int s1 = std::min(p->n_shells, s0 + slab);
        shell_pack_inv(h->shell_inv, p->n_shells, Ng, s0, s1, buf.data());
        CK(hipMemcpy(o.shell_inv + (size_t)s0 * per, buf.data(), (size_t)(s1 - s0) * per * 8,
                     hipMemcpyHostToDevice));
      }
    } else {
      if (!h->shell_B0 || !h->shell_coef) {
        c->err = "shell factors need shell_inv or shell_B0 + shell_coef";
        return NHQG_ERR_ARG;
      }
      std::vector<double> t((size_t)Ng * Ng);
      double *B0T = nullptr, *KT = nullptr, *coef = nullptr;
      int* status = nullptr;
      for (int i = 0; i < Ng; ++i)
        for (int j = 0; j < Ng; ++j) t[(size_t)j * Ng + i] = h->shell_B0[(size_t)i * Ng + j];
      RUN(upload(c, &B0T, t.data(), t.size()));
      if (h->shell_K) {
        for (int i = 0; i < Ng; ++i)
          for (int j = 0; j < Ng; ++j) t[(size_t)j * Ng + i] = h->shell_K[(size_t)i * Ng + j];
        RUN(upload(c, &KT, t.data(), t.size()));
      }
      RUN(upload(c, &coef, h->shell_coef, (size_t)3 * p->n_shells));
      RUN(dalloc(c, &status, 1));
      CK(hipMemset(status, 0, sizeof(int)));
      LK(launch_shell_build(p->n_shells, Ng, B0T, KT, coef, o.shell_inv, status, c->st));
      int st = 0;
      CK(hipStreamSynchronize(c->st));
      CK(hipMemcpy(&st, status, sizeof(int), hipMemcpyDeviceToHost));
      if (st) {
        c->err = "singular shell factor (zero pivot)";
        return NHQG_ERR_ARG;
      }
    }
  }
  // trap and sponge on the grid of the spectral products: the Nx grid, or
  // under the 3/2 rule the padded grid (the *_pad fields; the Nx-grid ones
  // are not read)
  const bool pad = c->Npad != c->Nx;
  const size_t nphys = (size_t)c->Npad * c->Npad;
  const double* eta = pad ? h->eta_phys_pad : h->eta_phys;
  const double* eta_x = pad ? h->eta_x_phys_pad : h->eta_x_phys;
  const double* eta_y = pad ? h->eta_y_phys_pad : h->eta_y_phys;
  RUN(upload(c, &o.eta_phys, eta, nphys));
  if (p->advection == 1 && eta) {
    if (!eta_x || !eta_y) {
      c->err = pad ? "the Jacobian form with a trap needs eta_x_phys_pad and eta_y_phys_pad"
                   : "the Jacobian form with a trap needs eta_x_phys and eta_y_phys";
      return NHQG_ERR_ARG;
    }
    RUN(upload(c, &o.eta_x, eta_x, nphys));
    RUN(upload(c, &o.eta_y, eta_y, nphys));
  } else {
    o.eta_x = o.eta_y = nullptr;
  }
  RUN(upload(c, &o.sponge_phys, pad ? h->sponge_phys_pad : h->sponge_phys, nphys));
  RUN(upload(c, (double**)&o.tw, h->twiddle, (size_t)2 * c->Nx));
  if (pad) {
    if (!h->kidx_of_x_pad || !h->rows_pad || !h->twiddle_pad) {
      c->err = "the 3/2 rule needs kidx_of_x_pad, rows_pad and twiddle_pad";
      return NHQG_ERR_ARG;
    }
    RUN(upload(c, &o.fkidx, h->kidx_of_x_pad, (size_t)c->Npad));
    RUN(upload(c, &o.frows, h->rows_pad, (size_t)c->nkx));
    RUN(upload(c, (double**)&o.ftw, h->twiddle_pad, (size_t)2 * c->Npad));
  } else {
    o.fkidx = o.kidx_of_x;
    o.frows = o.rows;
    o.ftw = o.tw;
  }
  o.eta_rows = nullptr;
  if (!pad && c->Nx == 256 && o.eta_phys && p->advection == 0) {
    RUN(dalloc(c, &o.eta_rows, (size_t)c->Nx * c->nky));
    CK(launch_eta_rows(c->Nx, c->nky, o.eta_phys, o.tw, o.eta_rows, c->st));
  }
  // mean temperature: the legacy exchange needs all four maps, RK4 under
  // evolve_mean the diffusion
  if (p->evolve_mean && p->exchange == 1 &&
      (!h->VGZ || !h->mean_solve || !h->mean_diff || !h->mean_flux)) {
    c->err = "evolve_mean with the legacy exchange needs VGZ, mean_solve, mean_diff, mean_flux";
    return NHQG_ERR_ARG;
  }
  RUN(upload(c, &o.vgz, h->VGZ, (size_t)Nq * Nq));
  RUN(upload(c, &o.msolve, h->mean_solve, (size_t)Nq * Nq));
  RUN(upload(c, &o.mdiff, h->mean_diff, (size_t)Nq * Nq));
  RUN(upload(c, &o.mflux, h->mean_flux, (size_t)Nq * Nq));
  // padded-grid horizontal means: the Nx-grid weights when not given
  if (h->pweight_pad) {
    RUN(upload_modes(c, &o.pw_pad, h->pweight_pad));
  } else {
    o.pw_pad = o.pw;
  }
  if (h->partner_pad) {
    std::vector<int> part((size_t)c->nm_pad, -1);
    std::memcpy(part.data(), h->partner_pad, sizeof(int) * nm);
    RUN(upload(c, &o.partner_pad, part.data(), part.size()));
  } else {
    o.partner_pad = o.partner;
  }
  o.pw_corr = pad ? o.pw_pad : o.pw;
  o.partner_corr = pad ? o.partner_pad : o.partner;
  if (!Psbp_h.empty() && c->sbp_band) {
    // Psbp Z (input rows 0 and Nq-1 zeroed by the substep update) as five
    // diagonals when every other entry is construction roundoff
    const std::vector<double>& P = Psbp_h;
    double mx = 0.0;
    for (double v : P) mx = std::max(mx, std::fabs(v));
    bool band = mx > 0.0 && Nq >= 5;
    for (int i = 0; i < Nq && band; ++i)
      for (int j = 1; j < Nq - 1; ++j)
        if (std::abs(i - j) > 2 && std::fabs(P[(size_t)i * Nq + j]) > 1e-12 * mx) {
          band = false;
          break;
        }
    if (band) {
      std::vector<double> pb((size_t)Nq * 5, 0.0);
      for (int i = 0; i < Nq; ++i)
        for (int o2 = -2; o2 <= 2; ++o2) {
          const int j = i + o2;
          if (j >= 1 && j < Nq - 1) pb[(size_t)i * 5 + o2 + 2] = P[(size_t)i * Nq + j];
        }
      RUN(upload(c, &o.psbp_band, pb.data(), pb.size()));
      // the corrector mean (parseval_partial_kernel with pw_corr and
      // partner_corr) split into diagonal weights per real column and
      // cross terms of the partnered (self-conjugate) columns
      const double* hpw = (pad && h->pweight_pad) ? h->pweight_pad : h->pweight;
      const int32_t* hpt = (pad && h->partner_pad) ? h->partner_pad : h->partner;
      std::vector<double> pwc((size_t)c->ld, 0.0);
      std::vector<int2> xp;
      std::vector<double> xw;
      for (int m = 0; m < nm; ++m) {
        const int q = hpt ? hpt[m] : -1;
        double d = hpw[m];
        if (q == m) {
          d = 0.5 * hpw[m];
          xp.push_back(make_int2(m, m));
          xw.push_back(0.25 * hpw[m]);
        } else if (q >= 0) {
          d = 0.25 * (hpw[m] + hpw[q]);
          if (m < q) {
            xp.push_back(make_int2(m, q));
            xw.push_back(0.25 * (hpw[m] + hpw[q]));
          }
        }
        pwc[2 * (size_t)m] = d;
        pwc[2 * (size_t)m + 1] = d;
      }
      RUN(upload(c, &o.pwc, pwc.data(), pwc.size()));
      o.nxpairs = (int)xp.size();
      if (o.nxpairs) {
        RUN(upload(c, &o.xpairs, xp.data(), xp.size()));
        RUN(upload(c, &o.xpw, xw.data(), xw.size()));
      }
    }
  }
  c->h_thb2s.assign(h->thb2s, h->thb2s + (size_t)Nq * Nq);
  c->h_d1b.assign(h->d1_band, h->d1_band + (size_t)Nq * 3);

  // state and workspace
  RUN(field(c, &c->yq, Nq));
  RUN(field(c, &c->yw, Ng));
  RUN(field(c, &c->yth, Ng));
  RUN(dalloc(c, &c->ythb, (size_t)Nq));
  RUN(field(c, &c->qn, Nq));
  RUN(field(c, &c->wn, Nq));
  RUN(field(c, &c->thn, Nq));
  RUN(field(c, &c->E1q, Nq));
  RUN(field(c, &c->E1w, Ng));
  RUN(field(c, &c->E1th, Ng));
  RUN(field(c, &c->E2q, Nq));
  RUN(field(c, &c->E2w, Ng));
  RUN(field(c, &c->E2th, Ng));
  RUN(field(c, &c->Rq, Nq));
  RUN(field(c, &c->Rw, Ng));
  RUN(field(c, &c->Rth, Ng));
  RUN(dalloc(c, &c->Rthb, (size_t)Nq));
  RUN(field(c, &c->s1q, Nq));
  RUN(field(c, &c->s1w, Ng));
  RUN(field(c, &c->s1th, Ng));
  RUN(dalloc(c, &c->s1thb, (size_t)Nq));
  RUN(field(c, &c->c1th, Ng));
  RUN(dalloc(c, &c->c1thb, (size_t)Nq));
  RUN(field(c, &c->rhs_w, Ng));
  RUN(field(c, &c->tmpg, Ng));
  RUN(field(c, &c->wsbp, Nq));
  RUN(field(c, &c->thsbp, Nq));
  RUN(field(c, &c->thsbp2, Nq));
  RUN(field(c, &c->spq, Nq));
  RUN(field(c, &c->spw, Ng));
  RUN(field(c, &c->spth, Ng));
  c->Pband = (c->ld + 255) / 256 + 1;
  RUN(dalloc(c, &c->partial, (size_t)Nq * std::max(c->Pred, c->Pband) + 64));
  RUN(dalloc(c, &c->sbp, (size_t)Nq * 16));
  RUN(dalloc(c, &c->sbp_xy, (size_t)Nq * 2 * Nq));
  RUN(dalloc(c, &c->sbp_map, (size_t)4 * Nq * Nq));
  RUN(dalloc(c, &c->thbt, (size_t)Nq));
  RUN(dalloc(c, &c->g, (size_t)Nq));
  // FFT intermediates, one block of 17 fields: flux form 8 in + 9 out,
  // Jacobian form 11 in + 6 out
  c->fstride = (size_t)Nq * c->Npad * c->nky;
  RUN(dalloc(c, &c->X5, c->fstride * 17));
  c->Y6 = c->X5 + c->fstride * (p->advection ? 11 : 8);
  RUN(dalloc(c, &c->staging, (size_t)Nq * c->Nx * c->Nk));
  RUN(dalloc(c, &c->diag, (size_t)64 + 8 * Nq));
  RUN(dalloc(c, &c->rowmax, (size_t)4 * Nq * c->Nx));
  double** vecs[] = {&c->mgrad, &c->etb1, &c->etb2, &c->itb, &c->ktb, &c->acctb, &c->stb};
  for (double** v : vecs) RUN(dalloc(c, v, (size_t)Nq));
  CK(hipMemset(c->ythb, 0, sizeof(double) * Nq));
  CK(hipDeviceSynchronize());
  return NHQG_OK;
}

int copy_full_in(nhqg_ctx* c, const double* host, int nl, double* comp) {
  const size_t full_n = (size_t)c->Nx * c->Nk;
  CK(hipMemcpyAsync(c->staging, host, nl * full_n * sizeof(double2), hipMemcpyHostToDevice,
                    c->st));
  LK(launch_gather(nl, c->nm, c->nm_pad, (int)full_n, c->o.full_idx, c->staging,
                   (double2*)comp, c->st));
  CK(hipStreamSynchronize(c->st));
  return NHQG_OK;
}

int copy_full_out(nhqg_ctx* c, double* host, int nl, const double* comp) {
  const size_t full_n = (size_t)c->Nx * c->Nk;
  LK(launch_scatter(nl, c->nm, c->nm_pad, (int)full_n, c->o.full_idx, (const double2*)comp,
                    c->staging, c->st));
  CK(hipMemcpyAsync(host, c->staging, nl * full_n * sizeof(double2), hipMemcpyDeviceToHost,
                    c->st));
  CK(hipStreamSynchronize(c->st));
  return NHQG_OK;
}

// Sum of partial[l][p] weighted by wl[l] (host side, Pred partials per level).
double reduce_partials(const std::vector<double>& part, int nl, int P, const double* wl) {
  double s = 0.0;
  for (int l = 0; l < nl; ++l) {
    double t = 0.0;
    for (int p = 0; p < P; ++p) t += part[(size_t)l * P + p];
    s += (wl ? wl[l] : 1.0) * t;
  }
  return s;
}

int diagnostics_impl(nhqg_ctx* c, double* out, int n_out) {
  const Ops& o = c->o;
  const int Nq = c->Nq, Ng = c->Ng, P = c->Pred;
  if (n_out < NHQG_DIAG_SCALARS + 3 * Nq) {
    c->err = "diagnostics output too short";
    return NHQG_ERR_ARG;
  }
  std::vector<double> cc(Nq), part((size_t)Nq * P);
  CK(hipMemcpy(cc.data(), o.cc, sizeof(double) * Nq, hipMemcpyDeviceToHost));
  auto fetch = [&](int nl) -> int {
    CK(hipStreamSynchronize(c->st));
    CK(hipMemcpy(part.data(), c->partial, sizeof(double) * nl * P, hipMemcpyDeviceToHost));
    return NHQG_OK;
  };
  // nodal q (psi = -q inv_denom, applied through the per-mode weights)
  RUN(gemm(c, o.V, c->yq, c->qn, epi_make(1.0)));
  LK(launch_parseval(Nq, c->nm, c->ld, P, c->qn, c->qn, o.pw_ke, nullptr, c->partial, c->st));
  RUN(fetch(Nq));
  const double ke_tot = 0.5 * reduce_partials(part, Nq, P, cc.data());
  LK(launch_parseval(Nq, c->nm, c->ld, P, c->qn, c->qn, o.pw, nullptr, c->partial, c->st));
  RUN(fetch(Nq));
  const double q2 = reduce_partials(part, Nq, P, cc.data());
  // barotropic: depth average of nodal q, then the same weights
  RUN(gemm(c, o.row_cc, c->qn, c->E2q, epi_make(1.0)));
  LK(launch_parseval(1, c->nm, c->ld, P, c->E2q, c->E2q, o.pw_ke, nullptr, c->partial, c->st));
  RUN(fetch(1));
  const double ke_bt = 0.5 * reduce_partials(part, 1, P, nullptr);
  // w, theta at the CGL levels
  RUN(gemm_jobs(c, {Job{&o.VSw, c->yw, c->wn, epi_make(1.0)},
                    Job{&o.VSd, c->yth, c->thn, epi_make(1.0)}}));
  // max speed from the nodal psi; max |w|, |theta|, |w theta| from the nodal
  // w and theta, transformed in the same column pass
  double maxes[4] = {0.0, 0.0, 0.0, 0.0};
  {
    FluxArgs a;
    std::memset(&a, 0, sizeof(a));
    a.N = c->Nx;
    a.nq = Nq;
    a.nkx = c->nkx;
    a.nky = c->nky;
    a.ldc = c->ldc;
    a.qn = (const double2*)c->qn;
    a.wn = (const double2*)c->wn;
    a.thn = (const double2*)c->thn;
    a.out_scale = 1.0;
    a.kidx_of_x = o.kidx_of_x;
    a.rows = o.rows;
    a.kx_row = o.kx_row;
    a.ky_col = o.ky_col;
    a.inv_denom = o.inv_denom;
    a.tw = o.tw;
    a.X5 = c->X5;
    a.Y6 = c->Y6;
    a.fstride = c->fstride;
    LK(launch_speed(a, c->rowmax, c->st));
    const size_t sec = (size_t)Nq * c->Nx;
    for (int k = 0; k < 4; ++k)
      LK(launch_max_reduce(sec, c->rowmax + k * sec, c->diag + k, c->st));
  }
  CK(hipStreamSynchronize(c->st));
  CK(hipMemcpy(maxes, c->diag, 4 * sizeof(double), hipMemcpyDeviceToHost));
  const double max_speed = maxes[0];
  LK(launch_parseval(Nq, c->nm, c->ld, P, c->wn, c->thn, o.pw, nullptr, c->partial, c->st));
  RUN(fetch(Nq));
  const double vol_tw = reduce_partials(part, Nq, P, cc.data());
  LK(launch_parseval(Nq, c->nm, c->ld, P, c->wn, c->wn, o.pw, nullptr, c->partial, c->st));
  RUN(fetch(Nq));
  const double w2 = reduce_partials(part, Nq, P, cc.data());
  LK(launch_parseval(Nq, c->nm, c->ld, P, c->thn, c->thn, o.pw, nullptr, c->partial, c->st));
  RUN(fetch(Nq));
  const double t2 = reduce_partials(part, Nq, P, cc.data());
  std::vector<double> tb(Nq);
  CK(hipMemcpy(tb.data(), c->ythb, sizeof(double) * Nq, hipMemcpyDeviceToHost));
  // solver-consistent flux profile -> Nu_d (diagnostics.py:206-212, 655-675),
  // horizontal means on the padded 3/2 grid as there: on the SBP grid for the
  // balanced_sbp2_pc exchange (vol = sum_j cc_j prof_j with prof = sbp_to_cgl
  // <w_sbp th_sbp>, so the host folds cc^T sbp_to_cgl into cc_sbp), at the
  // CGL levels otherwise
  double vol_d = 0.0;
  double mon[5];
  for (double& v : mon) v = std::numeric_limits<double>::quiet_NaN();
  {
    const double* wl = cc.data();
    std::vector<double> ccs(Nq);
    if (c->P.exchange == 0) {
      RUN(gemm_jobs(c, {Job{&o.W2S, c->yw, c->wsbp, epi_make(1.0)},
                        Job{&o.T2S, c->yth, c->thsbp, epi_make(1.0)}}));
      LK(launch_parseval(Nq, c->nm, c->ld, P, c->wsbp, c->thsbp, o.pw_pad, o.partner_pad,
                         c->partial, c->st));
      CK(hipMemcpy(ccs.data(), o.cc_sbp, sizeof(double) * Nq, hipMemcpyDeviceToHost));
      wl = ccs.data();
    } else {
      LK(launch_parseval(Nq, c->nm, c->ld, P, c->wn, c->thn, o.pw_pad, o.partner_pad, c->partial,
                         c->st));
    }
    RUN(fetch(Nq));
    vol_d = reduce_partials(part, Nq, P, wl);
    if (c->P.exchange == 0 && c->P.evolve_mean) {
      // SBP-grid exchange monitor (diagnostics.py:334-369): the theta-side
      // term -(D1 th_bar_sbp)^T H F and the mean-side term
      // -th_bar_sbp^T H D1 F cancel up to the boundary term
      std::vector<double> F(Nq), ts(Nq, 0.0);
      for (int l = 0; l < Nq; ++l) {
        double t = 0.0;
        for (int q = 0; q < P; ++q) t += part[(size_t)l * P + q];
        F[l] = t;
      }
      for (int i = 0; i < Nq; ++i)
        for (int j = 0; j < Nq; ++j) ts[i] += c->h_thb2s[(size_t)i * Nq + j] * tb[j];
      auto d1 = [&](const std::vector<double>& v, int i) {
        double r = 0.0;
        for (int k = -1; k <= 1; ++k)
          if (i + k >= 0 && i + k < Nq) r += c->h_d1b[(size_t)i * 3 + k + 1] * v[i + k];
        return r;
      };
      const double dz = 1.0 / (Nq - 1);
      double fb = 0.0, ex = 0.0;
      for (int i = 0; i < Nq; ++i) {
        const double h = (i == 0 || i == Nq - 1) ? 0.5 * dz : dz;
        fb -= d1(ts, i) * h * F[i];
        ex -= ts[i] * h * d1(F, i);
      }
      mon[0] = fb;
      mon[1] = ex;
      mon[2] = ts[Nq - 1] * F[Nq - 1] - ts[0] * F[0];
      mon[3] = fb + ex;
      mon[4] = mon[3] / std::max(std::fabs(fb) + std::fabs(ex), 1e-300);
    }
  }
  // vertical spectra: raw sums of |f_n|^2 of q, w_cheb, th_cheb
  // (diagnostics.py:583-585)
  std::vector<double> spec(3 * (size_t)Nq, 0.0);
  RUN(gemm_jobs(c, {Job{&o.Sw, c->yw, c->wn, epi_make(1.0)},
                    Job{&o.Sd, c->yth, c->thn, epi_make(1.0)}}));
  const double* srcs[3] = {c->yq, c->wn, c->thn};
  for (int f = 0; f < 3; ++f) {
    LK(launch_parseval(Nq, c->nm, c->ld, P, srcs[f], srcs[f], o.ones, nullptr, c->partial, c->st));
    RUN(fetch(Nq));
    for (int l = 0; l < Nq; ++l) {
      double t = 0.0;
      for (int q = 0; q < P; ++q) t += part[(size_t)l * P + q];
      spec[(size_t)f * Nq + l] = t;
    }
  }
  double thb_max = 0.0;
  for (double v : tb) thb_max = std::max(thb_max, std::fabs(v));
  // energy fraction of the top quarter of each vertical spectrum
  // (high_mode_fraction, diagnostics.py:590-596)
  double hf[3];
  for (int f = 0; f < 3; ++f) {
    const int n_tail = std::max(1, Nq / 4);
    double tot = 0.0, tail = 0.0;
    for (int l = 0; l < Nq; ++l) tot += spec[(size_t)f * Nq + l];
    for (int l = Nq - n_tail; l < Nq; ++l) tail += spec[(size_t)f * Nq + l];
    hf[f] = tot > 0.0 ? tail / tot : 0.0;
  }
  for (int i = 0; i < n_out; ++i) out[i] = 0.0;
  out[NHQG_DIAG_KE_TOT] = ke_tot;
  out[NHQG_DIAG_KE_BT] = ke_bt;
  out[NHQG_DIAG_KE_BC] = ke_tot - ke_bt;
  out[NHQG_DIAG_MAX_SPEED] = max_speed;
  out[NHQG_DIAG_ENSTROPHY] = 0.5 * q2;
  out[NHQG_DIAG_NUSSELT] = 1.0 + vol_tw;
  out[NHQG_DIAG_NUSSELT_D] = 1.0 + vol_d;
  out[NHQG_DIAG_VOLTW_D] = vol_d;
  for (int i = 0; i < 3 * Nq; ++i) out[NHQG_DIAG_SCALARS + i] = spec[i];
  out[NHQG_DIAG_THBAR_MAX] = thb_max;
  out[NHQG_DIAG_Q_RMS] = std::sqrt(q2);
  out[NHQG_DIAG_W_RMS] = std::sqrt(w2);
  out[NHQG_DIAG_TH_RMS] = std::sqrt(t2);
  out[NHQG_DIAG_MAX_W] = maxes[1];
  out[NHQG_DIAG_MAX_THETA] = maxes[2];
  out[NHQG_DIAG_MAX_TW] = maxes[3];
  out[NHQG_DIAG_VOLTW] = vol_tw;
  out[NHQG_DIAG_HEAT_FLUX_MISMATCH] = vol_d - vol_tw;
  out[NHQG_DIAG_Q_HIGH_FRAC] = hf[0];
  out[NHQG_DIAG_W_HIGH_FRAC] = hf[1];
  out[NHQG_DIAG_TH_HIGH_FRAC] = hf[2];
  for (int k = 0; k < 5; ++k) out[NHQG_DIAG_MON_FEEDBACK + k] = mon[k];
  const bool mon_on = c->P.exchange == 0 && c->P.evolve_mean;
  for (int i = 0; i < NHQG_DIAG_SCALARS + 3 * Nq; ++i) {
    if (!mon_on && i >= NHQG_DIAG_MON_FEEDBACK && i <= NHQG_DIAG_MON_RESIDUAL_REL) continue;
    if (!std::isfinite(out[i])) return NHQG_NONFINITE;
  }
  return NHQG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int nhqg_create(const nhqg_params* p, const nhqg_operators* ops, nhqg_ctx** out) {
  if (!p || !ops || !out) {
    g_create_error = "null argument";
    return NHQG_ERR_ARG;
  }
  *out = nullptr;
  std::string err;
  int r = check_params(p, err);
  if (r) {
    g_create_error = err;
    return r;
  }
  nhqg_ctx* c = new nhqg_ctx();
  r = create_impl(c, p, ops);
  if (r) {
    g_create_error = c->err;
    nhqg_destroy(c);
    return r;
  }
  *out = c;
  return NHQG_OK;
}

int nhqg_destroy(nhqg_ctx* c) {
  if (!c) return NHQG_OK;
  if (c->st) (void)hipStreamSynchronize(c->st);
  if (c->st2) (void)hipStreamSynchronize(c->st2);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->cl_w) (void)hipFree(c->cl_w);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return NHQG_OK;
}

int nhqg_set_state(nhqg_ctx* c, const double* q, const double* w, const double* th,
                   const double* th_bar) {
  if (!c || !q || !w || !th || !th_bar) return NHQG_ERR_ARG;
  RUN(copy_full_in(c, q, c->Nq, c->yq));
  RUN(copy_full_in(c, w, c->Ng, c->yw));
  RUN(copy_full_in(c, th, c->Ng, c->yth));
  CK(hipMemcpy(c->ythb, th_bar, sizeof(double) * c->Nq, hipMemcpyHostToDevice));
  c->have_state = true;
  return NHQG_OK;
}

int nhqg_get_state(nhqg_ctx* c, double* q, double* w, double* th, double* th_bar) {
  if (!c || !q || !w || !th || !th_bar) return NHQG_ERR_ARG;
  CK(hipStreamSynchronize(c->st));
  RUN(copy_full_out(c, q, c->Nq, c->yq));
  RUN(copy_full_out(c, w, c->Ng, c->yw));
  RUN(copy_full_out(c, th, c->Ng, c->yth));
  CK(hipMemcpy(th_bar, c->ythb, sizeof(double) * c->Nq, hipMemcpyDeviceToHost));
  return NHQG_OK;
}

int nhqg_step_async(nhqg_ctx* c, int32_t n) {
  if (!c) return NHQG_ERR_ARG;
  if (!c->have_state) {
    c->err = "nhqg_step before nhqg_set_state";
    return NHQG_ERR_STATE;
  }
  if (n <= 0) return NHQG_OK;
  RUN(build_graph(c));
  CK(hipEventRecord(c->ev0, c->st));
  for (int i = 0; i < n; ++i) CK(hipGraphLaunch(c->gexec, c->st));
  CK(hipEventRecord(c->ev1, c->st));
  return NHQG_OK;
}

int nhqg_sync(nhqg_ctx* c) {
  if (!c) return NHQG_ERR_ARG;
  CK(hipStreamSynchronize(c->st));
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) c->last_ms = ms;
  return NHQG_OK;
}

int nhqg_step(nhqg_ctx* c, int32_t n) {
  RUN(nhqg_step_async(c, n));
  return nhqg_sync(c);
}

int nhqg_last_step_ms(nhqg_ctx* c, double* ms) {
  if (!c || !ms) return NHQG_ERR_ARG;
  *ms = c->last_ms;
  return NHQG_OK;
}

int nhqg_diagnostics(nhqg_ctx* c, double* out, int32_t n_out) {
  if (!c || !out) return NHQG_ERR_ARG;
  if (!c->have_state) return NHQG_ERR_STATE;
  return diagnostics_impl(c, out, n_out);
}

int nhqg_upsample(nhqg_ctx* c, nhqg_ctx* src) {
  if (!c || !src) return NHQG_ERR_ARG;
  if (!src->have_state) {
    c->err = "nhqg_upsample: the source context has no state";
    return NHQG_ERR_STATE;
  }
  if (src->Nq != c->Nq || c->Nx < src->Nx || (c->Nx - src->Nx) % 2) {
    c->err = "nhqg_upsample: needs the same nz and a destination nx >= source nx of the same "
             "parity";
    return NHQG_ERR_ARG;
  }
  CK(hipStreamSynchronize(src->st));
  const double scale = ((double)c->Nx / src->Nx) * ((double)c->Nx / src->Nx);
  const double* sf[3] = {src->yq, src->yw, src->yth};
  double* df[3] = {c->yq, c->yw, c->yth};
  const int nl[3] = {c->Nq, c->Ng, c->Ng};
  for (int f = 0; f < 3; ++f)
    LK(launch_upsample(nl[f], c->nm, c->nkx, c->Nx, c->o.rows, c->ld, df[f], src->Nx, src->nkx,
                       src->nky, src->o.kidx_of_x, src->o.partner, src->ld, sf[f], scale, c->st));
  CK(hipMemcpyAsync(c->ythb, src->ythb, sizeof(double) * c->Nq, hipMemcpyDeviceToDevice, c->st));
  RUN(sanitize(c));
  CK(hipStreamSynchronize(c->st));
  c->have_state = true;
  return NHQG_OK;
}

// Work arrays of the init entry points in the FFT scratch: the level f
// (nx*nx doubles), its rfft2 spec (nx*(nx/2+1)) and up to `extra` doubles.
static int init_scratch(nhqg_ctx* c, size_t extra, double** f, double2** spec, double** tail) {
  const size_t N = c->Nx, nf = (N * N + 1) / 2 + 32, ns = N * (N / 2 + 1) + 32;
  if (nf + ns + (extra + 1) / 2 > c->fstride * 17) {
    c->err = "init: the FFT scratch is too small for this call";
    return NHQG_ERR_ARG;
  }
  *f = (double*)c->X5;
  *spec = c->X5 + nf;
  if (tail) *tail = (double*)(c->X5 + nf + ns);
  return NHQG_OK;
}

int nhqg_init_barotropic(nhqg_ctx* c, const double* noise, double L, double k_peak,
                         double k_width, double amplitude, double confine_radius) {
  if (!c || !noise || !(L > 0.0) || !(k_peak > 0.0) || !(k_width > 0.0) || !(amplitude >= 0.0))
    return NHQG_ERR_ARG;
  const int N = c->Nx;
  double* f;
  double2* spec;
  RUN(init_scratch(c, 0, &f, &spec, nullptr));
  const double k0 = 2.0 * M_PI / L;
  CK(hipMemcpyAsync(f, noise, sizeof(double) * N * N, hipMemcpyHostToDevice, c->st));
  // zeta = irfft2(rfft2(noise) * annulus * band)
  LK(launch_init_fft2(N, true, f, spec, c->o.tw, c->st));
  LK(launch_init_spec_filter(N, spec, k0, k_peak, k_width, 1, c->st));
  LK(launch_init_fft2(N, false, f, spec, c->o.tw, c->st));
  const bool confine = confine_radius > 0.0;
  if (confine) {
    // polar-cap taper two initial wavelengths wide, then band-limited again
    LK(launch_init_taper(N, f, L, confine_radius, 2.0 * 2.0 * M_PI / k_peak, c->st));
    LK(launch_init_fft2(N, true, f, spec, c->o.tw, c->st));
    LK(launch_init_spec_filter(N, spec, k0, k_peak, k_width, 0, c->st));
    LK(launch_init_fft2(N, false, f, spec, c->o.tw, c->st));
  }
  LK(launch_init_normalize(N, f, L, confine ? confine_radius : 0.0, amplitude, c->st));
  LK(launch_init_fft2(N, true, f, spec, c->o.tw, c->st));
  CK(hipMemsetAsync(c->yq, 0, sizeof(double) * c->Nq * c->ld, c->st));
  CK(hipMemsetAsync(c->yw, 0, sizeof(double) * c->Ng * c->ld, c->st));
  CK(hipMemsetAsync(c->yth, 0, sizeof(double) * c->Ng * c->ld, c->st));
  CK(hipMemsetAsync(c->ythb, 0, sizeof(double) * c->Nq, c->st));
  LK(launch_init_put_q0(c->nm, c->o.full_idx, spec, (double2*)c->yq, 0, c->st));
  CK(hipStreamSynchronize(c->st));
  c->have_state = true;
  return NHQG_OK;
}

int nhqg_inject_cyclones(nhqg_ctx* c, const double* centers, int32_t count, double L,
                         double amplitude, double radius) {
  if (!c || count < 0 || (count > 0 && !centers) || !(L > 0.0) || !(radius > 0.0))
    return NHQG_ERR_ARG;
  if (!c->have_state) {
    c->err = "nhqg_inject_cyclones: the context has no state";
    return NHQG_ERR_STATE;
  }
  if (count == 0) return NHQG_OK;
  const int N = c->Nx;
  double *f, *cen;
  double2* spec;
  RUN(init_scratch(c, (size_t)2 * count, &f, &spec, &cen));
  CK(hipStreamSynchronize(c->st));
  CK(hipMemcpyAsync(cen, centers, sizeof(double) * 2 * count, hipMemcpyHostToDevice, c->st));
  LK(launch_init_cyclones(N, f, L, cen, count, amplitude, radius, c->st));
  LK(launch_init_fft2(N, true, f, spec, c->o.tw, c->st));
  LK(launch_init_spec_filter(N, spec, 0.0, 0.0, 1.0, 0, c->st));
  LK(launch_init_put_q0(c->nm, c->o.full_idx, spec, (double2*)c->yq, 1, c->st));
  CK(hipStreamSynchronize(c->st));
  return NHQG_OK;
}

int nhqg_mode_power(nhqg_ctx* c, double* out, int32_t n_out) {
  if (!c || !out) return NHQG_ERR_ARG;
  if (!c->have_state) return NHQG_ERR_STATE;
  if (n_out < 3 * c->nm) {
    c->err = "mode power output shorter than 3 * nm";
    return NHQG_ERR_ARG;
  }
  const Ops& o = c->o;
  if (!c->mpow) RUN(dalloc(c, &c->mpow, (size_t)3 * c->nm_pad));
  RUN(gemm_jobs(c, {Job{&o.V, c->yq, c->qn, epi_make(1.0)},
                    Job{&o.VSw, c->yw, c->wn, epi_make(1.0)},
                    Job{&o.VSd, c->yth, c->thn, epi_make(1.0)}}));
  const double* f[3] = {c->qn, c->wn, c->thn};
  for (int k = 0; k < 3; ++k)
    LK(launch_depth_power(c->Nq, c->ld, c->nm, o.cc, f[k], c->mpow + (size_t)k * c->nm_pad,
                          c->st));
  CK(hipStreamSynchronize(c->st));
  for (int k = 0; k < 3; ++k)
    CK(hipMemcpy(out + (size_t)k * c->nm, c->mpow + (size_t)k * c->nm_pad,
                 sizeof(double) * c->nm, hipMemcpyDeviceToHost));
  return NHQG_OK;
}

static int apply_impl(nhqg_ctx* c, int32_t op, double arg);

int nhqg_apply(nhqg_ctx* c, int32_t op, double arg) {
  if (!c) return NHQG_ERR_ARG;
  if (!c->have_state) return NHQG_ERR_STATE;
  c->mixed_now = c->mixed_slices > 0;
  const int r = apply_impl(c, op, arg);
  c->mixed_now = false;
  return r;
}

static int apply_impl(nhqg_ctx* c, int32_t op, double arg) {
  switch (op) {
    case NHQG_OP_EXPLICIT_RHS:
      RUN(explicit_rhs(c, c->yq, c->yw, c->yth, c->E1q, c->E1w, c->E1th));
      break;
    case NHQG_OP_IMPLICIT_TEND: {
      // I_q = G_Z S_w w; I_w = inv_denom Qw q + Ra_s (Mtw) th; I_th = (Mwt) w
      const Ops& o = c->o;
      RUN(gemm(c, o.GSw, c->yw, c->E1q, epi_make(1.0)));
      const double* thw = c->yth;
      if (c->P.w_neumann) {
        RUN(gemm(c, o.Mtw, c->yth, c->tmpg, epi_make(1.0)));
        thw = c->tmpg;
      }
      Epi f = epi_make(1.0);
      f.s_acc = o.inv_denom;
      epi_add(f, c->P.ra_sigma, thw);
      RUN(gemm(c, o.Qw, c->yq, c->E1w, f));
      if (c->P.w_neumann) {
        RUN(gemm(c, o.Mwt, c->yw, c->E1th, epi_make(1.0)));
      } else {
        Epi h = epi_make(0.0);
        epi_add(h, 1.0, c->yw);
        RUN(combine(c, c->Ng, c->E1th, h));
      }
      break;
    }
    case NHQG_OP_IMPLICIT_SOLVE: {
      Epi e = epi_make(0.0);
      epi_add(e, 1.0, c->yq);
      RUN(combine(c, c->Nq, c->Rq, e));
      Epi f = epi_make(0.0);
      epi_add(f, 1.0, c->yw);
      RUN(combine(c, c->Ng, c->Rw, f));
      Epi h = epi_make(0.0);
      epi_add(h, 1.0, c->yth);
      RUN(combine(c, c->Ng, c->Rth, h));
      RUN(implicit_solve(c, c->Rq, c->Rw, c->Rth, c->yq, c->yw, c->yth));
      LK(launch_small_axpy(c->Nq, c->ythb, c->ythb, c->ythb, 0.0, c->Rthb, c->st));
      LK(launch_small_matvec(c->Nq, c->Nq, c->o.proj_d, c->Rthb, c->ythb, c->st));
      break;
    }
    case NHQG_OP_CORRECTOR:
      RUN(corrector(c, c->yw, c->yth, c->ythb, c->yth, c->ythb, arg));
      break;
    case NHQG_OP_SANITIZE:
      RUN(sanitize(c));
      break;
    case NHQG_OP_RK4_STEP:
      RUN(rk4_step(c));
      break;
    default:
      c->err = "unknown op";
      return NHQG_ERR_ARG;
  }
  CK(hipStreamSynchronize(c->st));
  return NHQG_OK;
}

int nhqg_get_aux(nhqg_ctx* c, double* q, double* w, double* th) {
  if (!c || !q || !w || !th) return NHQG_ERR_ARG;
  RUN(copy_full_out(c, q, c->Nq, c->E1q));
  RUN(copy_full_out(c, w, c->Ng, c->E1w));
  RUN(copy_full_out(c, th, c->Ng, c->E1th));
  return NHQG_OK;
}

int nhqg_profile_step(nhqg_ctx* c, double* out, int32_t n_out) {
  if (!c || !out) return NHQG_ERR_ARG;
  if (!c->have_state) return NHQG_ERR_STATE;
  if (n_out < 3 * CL_COUNT) {
    c->err = "profile output needs 3 entries per kernel class";
    return NHQG_ERR_ARG;
  }
  CK(hipStreamSynchronize(c->st));
  c->recs.clear();
  c->prof = true;
  const int r = enqueue_step(c);
  c->prof = false;
  const hipError_t e = hipStreamSynchronize(c->st);
  for (int i = 0; i < 3 * CL_COUNT; ++i) out[i] = 0.0;
  for (auto& p : c->recs) {
    float ms = 0.0f;
    if (r == NHQG_OK && e == hipSuccess) (void)hipEventElapsedTime(&ms, p.a, p.b);
    out[3 * p.cls] += ms;
    out[3 * p.cls + 1] += 1.0;
    out[3 * p.cls + 2] += p.work;
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  c->recs.clear();
  if (r) return r;
  CK(e);
  return NHQG_OK;
}

const char* nhqg_timing_name(int32_t i) {
  return (i >= 0 && i < CL_COUNT) ? kClassNames[i] : "";
}

int nhqg_stream_copy(nhqg_ctx* c, uint64_t bytes, int32_t reps, double* gbps) {
  if (!c || !gbps || reps < 1 || bytes < 16) return NHQG_ERR_ARG;
  const size_t n16 = (size_t)(bytes / 16);
  void *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, n16 * 16));
  hipError_t e = hipMalloc(&b, n16 * 16);
  if (e != hipSuccess) {
    (void)hipFree(a);
    c->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return NHQG_ERR_HIP;
  }
  int r = NHQG_OK;
  float ms = 0.0f;
  e = hipMemsetAsync(a, 0, n16 * 16, c->st);
  if (e == hipSuccess) e = launch_stream_copy(n16, (const double2*)a, (double2*)b, c->st);
  if (e == hipSuccess) e = hipEventRecord(c->ev0, c->st);
  for (int i = 0; e == hipSuccess && i < reps; ++i)
    e = launch_stream_copy(n16, (const double2*)a, (double2*)b, c->st);
  if (e == hipSuccess) e = hipEventRecord(c->ev1, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
  if (e == hipSuccess) gbps[0] = 2.0 * n16 * 16.0 * reps / (ms * 1e-3) / 1e9;
  // the runtime's own device-to-device copy over the same buffers
  if (e == hipSuccess) e = hipEventRecord(c->ev0, c->st);
  for (int i = 0; e == hipSuccess && i < reps; ++i)
    e = hipMemcpyAsync(b, a, n16 * 16, hipMemcpyDeviceToDevice, c->st);
  if (e == hipSuccess) e = hipEventRecord(c->ev1, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
  if (e != hipSuccess) {
    c->err = std::string("stream copy: ") + hipGetErrorString(e);
    r = NHQG_ERR_HIP;
  } else {
    gbps[1] = 2.0 * n16 * 16.0 * reps / (ms * 1e-3) / 1e9;
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return r;
}

int nhqg_get_shell_inverse(nhqg_ctx* c, int32_t s, double* out) {
  if (!c || !out || s < 0 || s >= c->P.n_shells) return NHQG_ERR_ARG;
  if (!c->o.shell_inv) {
    c->err = "shell solves run in the eigenbasis of B0: no per-shell matrices";
    return NHQG_ERR_UNSUPPORTED;
  }
  const int Ng = c->Ng, np = shell_pad_order(Ng);
  std::vector<double> t((size_t)np * np);
  CK(hipStreamSynchronize(c->st));
  CK(hipMemcpy(t.data(), c->o.shell_inv + (size_t)s * np * np, t.size() * sizeof(double),
               hipMemcpyDeviceToHost));
  // the slot holds inv_s[j][i] at j*np + i
  for (int i = 0; i < Ng; ++i)
    for (int j = 0; j < Ng; ++j) out[(size_t)i * Ng + j] = t[(size_t)j * np + i];
  return NHQG_OK;
}

int nhqg_set_closure(nhqg_ctx* c, const nhqg_mlp* mlp, int32_t n0, double in_scale,
                     double out_scale) {
  if (!c) return NHQG_ERR_ARG;
  if (!mlp) {
    c->cl_on = false;
  } else {
    if (n0 < 0 || n0 + 16 > c->Nq) {
      c->err = "closure levels n0 .. n0+15 must lie within the Nz+1 q levels";
      return NHQG_ERR_ARG;
    }
    int r = closure_export_bf16(mlp, &c->cl_w, c->st, c->err);
    if (r) return r;
    c->cl_on = true;
    c->cl_n0 = n0;
    c->cl_in = (float)in_scale;
    c->cl_out = (float)out_scale;
  }
  // the captured step graph no longer matches
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  return NHQG_OK;
}

int nhqg_device_bytes(nhqg_ctx* c, uint64_t* bytes) {
  if (!c || !bytes) return NHQG_ERR_ARG;
  *bytes = c->bytes;
  return NHQG_OK;
}

const char* nhqg_last_error(const nhqg_ctx* c) {
  return c ? c->err.c_str() : g_create_error.c_str();
}

}  // extern "C"
