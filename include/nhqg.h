/*
 * nhqg.h -- C ABI of the MI355X (gfx950) NHQG time-step library.
 *
 * The reference exposes the hot path as two Python calls,
 *     grid = make_grid(cfg)             nhqg/grid.py:439
 *     new  = imex_step(state, grid)     nhqg/solver.py:1228
 * over State(q_hat, w_hat, th_hat, th_bar)  (nhqg/solver.py:33-38), and its
 * drivers run imex_step in a jit+scan loop (scripts/run_polar.py:338-343,
 * nhqg/solver.py:1385-1416) followed by compute_diagnostics
 * (nhqg/diagnostics.py:597-710).  The reference has no native FFI; this
 * header is the boundary below the Python mirror of that interface
 * (nhqg_amd/solver.py), and a maintainer binding it from elsewhere needs
 * only these symbols.
 *
 * Conventions: plain pointers and sizes; all arrays row-major; complex
 * arrays are interleaved float64 pairs (numpy complex128).  Every call
 * returns an NHQG_* status; nothing aborts.  A context owns one HIP stream
 * (the context stream) on the current device, plus a side stream that the
 * step forks from it and joins back within each call, and is not thread
 * safe.
 */
#ifndef NHQG_H
#define NHQG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NHQG_OK 0
#define NHQG_ERR_ARG 1         /* bad argument or unsupported shape */
#define NHQG_ERR_HIP 2         /* HIP runtime error (see nhqg_last_error) */
#define NHQG_ERR_UNSUPPORTED 3 /* configuration branch not built */
#define NHQG_ERR_STATE 4       /* call out of order (e.g. step before set_state) */
#define NHQG_NONFINITE 5        /* nhqg_diagnostics: output written, some value non-finite */

typedef struct nhqg_ctx nhqg_ctx;

/* Scalar parameters of the step (nhqg/config.py:17-73 after make_grid's
 * derivations, nhqg/grid.py:669-687). */
typedef struct {
  int32_t nx;          /* horizontal grid points, power of two 16..1024 */
  int32_t nz;          /* Chebyshev intervals; Nz+1 CGL levels, Nz <= 256 */
  int32_t nkx, nky;    /* kept kx rows / ky columns of the mode layout */
  int32_t n_shells;    /* number of |k|^2 shells in shell_inv */
  int32_t n_pairs;     /* entries of herm_pairs */
  int32_t substeps;    /* SBP corrector substeps (sbp_corrector_substeps) */
  int32_t evolve_mean; /* 1: thermal_closure = evolve_mean (th_bar evolves; see exchange) */
  int32_t advection;   /* 0: flux form (the production branch), 1: Jacobian form
                        * (nonlinear_advection = "jacobian", nhqg/spectral.py:281-304) */
  int32_t w_neumann;   /* 1: w in the mixed bottom-Dirichlet/top-Neumann basis */
  int32_t apply_mask;  /* 1: layout keeps out-of-band modes; mask outputs */
  int32_t mixed_slices; /* 0: plain FP64 step.  4..7: mixed-precision step (no reference
                         * counterpart): the Chebyshev transforms and implicit-solve maps
                         * run as int8 slice products with exact int32 accumulation and
                         * FP64 reconstruction, this many 7-bit slices per operand; the
                         * SBP exchange, the k = 0 equation, FFT products and
                         * diagnostics stay FP64 */
  int32_t dense_contractions; /* 0: a vertical operator with the reflection symmetry of the
                               * Chebyshev grid (z -> -z: parity blocks, or a butterfly
                               * on CGL-value rows or columns) is applied as two products
                               * of half size, half the flops of the dense product, same
                               * result to roundoff; 1: every operator dense (no reference
                               * counterpart: the reference applies dense matrices) */
  int32_t exchange;    /* mean_exchange_discretization: 0 balanced_sbp2_pc (the SBP corrector
                        * under evolve_mean; the diagnostics' heat-flux profile on the SBP
                        * grid, diagnostics.py:206-208), 1 legacy (under evolve_mean the
                        * -dTheta/dZ w and -eps^2 d<w theta>/dZ terms of explicit_rhs,
                        * solver.py:575-590, and the mean's tau solve, solver.py:746-758) */
  int32_t npad;        /* horizontal_dealiasing: 0 or nx for the 2/3 rule (products on the
                        * nx grid, outputs times mask, spectral.py:307-358); 3nx/2 for the
                        * 3/2 rule (products on the zero-padded npad grid, truncated with the
                        * factor (npad/nx)^2, spectral.py:18-50, 62-74, 155-253; the SBP
                        * corrector's horizontal means on that grid).  The 3/2 rule needs the
                        * full layout, apply_mask 0 and the *_pad operator fields */
  double dt, gamma_imex, ra_sigma, sigma, eps_sq, beta;
} nhqg_params;

/* Host operator tables, all float64 row-major unless noted.  Nq = nz+1,
 * Ng = nz-1, nm = nkx*nky modes (ky-major: m = ky_i*nkx + kx_i). */
typedef struct {
  /* fused vertical operators (products of the reference's grid matrices) */
  const double *V;      /* Nq x Nq  coefficients -> CGL values */
  const double *VSw;    /* Nq x Ng  w Galerkin -> CGL values */
  const double *VSd;    /* Nq x Ng  theta Galerkin -> CGL values */
  const double *Vinv;   /* Nq x Nq */
  const double *Pw;     /* Ng x Nq  w_pinv proj_w V_inv */
  const double *Pd;     /* Ng x Nq  d_pinv proj_d V_inv */
  const double *Qw;     /* Ng x Nq  w_pinv proj_w G_Z */
  const double *GSw;    /* Nq x Ng  G_Z w_stencil */
  const double *W2S;    /* Nq x Ng  cgl_to_sbp V w_stencil */
  const double *T2S;    /* Nq x Ng  cgl_to_sbp V d_stencil */
  const double *S2T;    /* Ng x Nq  d_pinv proj_d V_inv sbp_to_cgl */
  const double *thb2s;  /* Nq x Nq  cgl_to_sbp V */
  const double *s2thb;  /* Nq x Nq  proj_d V_inv sbp_to_cgl */
  const double *proj_d; /* Nq x Nq */
  const double *d1_band;  /* Nq x 3, SBP D1 tridiagonal band */
  const double *lap_band; /* Nq x 3, SBP Laplacian tridiagonal band */
  const double *Mtw;    /* Ng x Ng  w_pinv proj_w d_stencil (w_neumann only) */
  const double *Mwt;    /* Ng x Ng  map_w_to_th (w_neumann only) */
  const double *Mth2w;  /* Ng x Ng  map_th_to_w (w_neumann only) */
  const double *cc;     /* Nq Clenshaw-Curtis weights */
  const double *row_cc; /* 1 x Nq (cc as a row, for depth averages) */
  const double *cc_sbp; /* Nq: cc^T sbp_to_cgl (Nusselt on the SBP grid) */
  const double *Sw;     /* Nq x Ng  w Galerkin stencil */
  const double *Sd;     /* Nq x Ng  theta Galerkin stencil */
  /* mode layout */
  const int32_t *rows;      /* nkx: storage row of each kept kx */
  const int32_t *kidx_of_x; /* nx: kept index of each row or -1 */
  const int32_t *full_idx;  /* nm: row*(nx/2+1) + ky of each mode */
  const double *kx_row;     /* nkx */
  const double *ky_col;     /* nky */
  /* per-mode tables (nm) */
  const double *kx_mode, *ksq, *inv_denom, *inv_alpha_q, *inv_alpha_th;
  const double *diss_q, *diss_w, *diss_th, *mask, *pweight;
  const int32_t *partner;   /* nm: (-kx) partner on self-conjugate columns or -1 */
  const int32_t *herm_pairs; /* 2*n_pairs: (m, m') with m <= m' */
  /* |k|^2 shells */
  const int32_t *shell_ptr;   /* n_shells+1 offsets into shell_modes */
  const int32_t *shell_modes; /* modes grouped by shell */
  const double *shell_inv;    /* n_shells x Ng x Ng, each TRANSPOSED */
  /* horizontal physical fields, nx*nx, or NULL */
  const double *eta_phys;     /* trap background PV on the grid */
  const double *sponge_phys;  /* band-limited sponge rate */
  const double *twiddle;      /* 2*nx: exp(-2 pi i k / nx) */
  /* Shell factors built on the device (used when shell_inv is NULL):
   * A_s = coef[s][0] I - coef[s][1] K - coef[s][2] B0, inverted in FP64 by
   * Gauss-Jordan with partial pivoting (replaces the host inverse loop of
   * _build_imex_inv, nhqg/grid.py:349-432, for q_boundary='none'). */
  const double *shell_B0;     /* Ng x Ng: w_pinv G_Z G_Z w_stencil */
  const double *shell_K;      /* Ng x Ng buoyancy map (Neumann top) or NULL (K = I) */
  const double *shell_coef;   /* n_shells x 3: alpha, beta, c */
  /* Dirichlet-top w: B0 = Q diag(lam) Q^-1 (real spectrum), so every shell
   * solve is w = Q diag(1 / (alpha_m - c_m lam)) Q^-1 rhs.  When eig_Q is
   * set (and nz w is Dirichlet) no shell matrices are built or streamed. */
  const double *eig_Q;        /* Ng x Ng eigenvectors of B0 */
  const double *eig_Qinv;     /* Ng x Ng */
  const double *eig_lam;      /* Ng eigenvalues */
  const double *mode_alpha;   /* per mode: alpha of its shell */
  const double *mode_c;       /* per mode: c of its shell */
  /* Jacobian form with a trap: d(eta)/dx, d(eta)/dy on the grid (Nx x Nx,
   * row x, column y), the gradient J(psi, q + eta) adds to q's */
  const double *eta_x_phys, *eta_y_phys;
  /* mean temperature, Nq x Nq: VGZ = V G_Z (dTheta/dZ at the CGL levels);
   * mean_solve = inverse of (I - gamma dt eps^2/sigma G_Z2) with its last two
   * rows replaced by the Dirichlet tau rows, applied to a right-hand side
   * whose last two entries are dropped (imex_mean_temp_solve,
   * solver.py:746-758); mean_diff = eps^2/sigma G_Z2 (I_th_bar, :638);
   * mean_flux = -eps^2 G_Z V_inv (E_th_bar from the CGL flux profile,
   * :585-590, 324-329) */
  const double *VGZ, *mean_solve, *mean_diff, *mean_flux;
  /* per-mode weights and (-kx) partners of the horizontal mean on the
   * zero-padded 3/2 grid (horizontal_mean_from_nodal_spectral with Npad,
   * solver.py:231-245): equal to pweight/partner on band-limited layouts */
  const double *pweight_pad;
  const int32_t *partner_pad;
  /* the padded grid of the 3/2 rule (params.npad = 3nx/2; NULL otherwise):
   * kidx_of_x_pad (npad): kept row of each padded row x (rows kx = 0..nx/2-1
   * at x = kx, kx = -nx/2..-1 at x = npad + kx, spectral.py:18-34) or -1;
   * rows_pad (nkx): padded row of each kept row; twiddle_pad (2*npad):
   * exp(-2 pi i k / npad); eta_phys_pad, eta_x_phys_pad, eta_y_phys_pad,
   * sponge_phys_pad (npad x npad): irfft2 of the zero-padded spectra on the
   * padded grid, without rescaling (grid.py:663-667).  Under the 3/2 rule
   * these replace eta_phys, eta_x_phys, eta_y_phys and sponge_phys. */
  const int32_t *kidx_of_x_pad, *rows_pad;
  const double *twiddle_pad;
  const double *eta_phys_pad, *eta_x_phys_pad, *eta_y_phys_pad, *sponge_phys_pad;
  /* vertical_cutoff_n (nhqg/solver.py:177-191), Ng x Ng each, or NULL for
   * none: cutoff_w = w_pinv proj_w diag(n <= cutoff) w_stencil and cutoff_th
   * = d_pinv proj_d diag(n <= cutoff) d_stencil, applied to w and theta at
   * the end of every step (_finalize_state) and to every RK4 stage state
   * (_apply_bcs). */
  const double *cutoff_w, *cutoff_th;
} nhqg_operators;

/* Diagnostics vector layout returned by nhqg_diagnostics (length
 * NHQG_DIAG_SCALARS + 3*(nz+1)). */
#define NHQG_DIAG_KE_TOT 0
#define NHQG_DIAG_KE_BT 1
#define NHQG_DIAG_KE_BC 2
#define NHQG_DIAG_MAX_SPEED 3
#define NHQG_DIAG_ENSTROPHY 4
#define NHQG_DIAG_NUSSELT 5
#define NHQG_DIAG_NUSSELT_D 6
#define NHQG_DIAG_VOLTW_D 7
#define NHQG_DIAG_THBAR_MAX 8
#define NHQG_DIAG_Q_RMS 9
#define NHQG_DIAG_W_RMS 10
#define NHQG_DIAG_TH_RMS 11
#define NHQG_DIAG_MAX_W 12           /* max |w| on the grid at the CGL levels */
#define NHQG_DIAG_MAX_THETA 13       /* max |theta| */
#define NHQG_DIAG_MAX_TW 14          /* max |w theta| */
#define NHQG_DIAG_VOLTW 15           /* vol_avg_tw (CGL, Parseval) */
#define NHQG_DIAG_HEAT_FLUX_MISMATCH 16 /* vol_avg_tw_dealiased - vol_avg_tw */
#define NHQG_DIAG_Q_HIGH_FRAC 17     /* energy fraction of the top quarter of */
#define NHQG_DIAG_W_HIGH_FRAC 18     /* the vertical spectrum (diagnostics.py:590-596) */
#define NHQG_DIAG_TH_HIGH_FRAC 19
/* SBP-grid exchange monitor (diagnostics.py:334-369; NaN unless evolve_mean
 * with the balanced_sbp2_pc exchange, and then not counted as non-finite) */
#define NHQG_DIAG_MON_FEEDBACK 20    /* th_mean_feedback_sum_sbp */
#define NHQG_DIAG_MON_EXCHANGE 21    /* mean_flux_exchange_tendency_sbp */
#define NHQG_DIAG_MON_BOUNDARY 22    /* mean_theta_exchange_boundary_sbp */
#define NHQG_DIAG_MON_RESIDUAL 23    /* mean_theta_exchange_residual_sbp */
#define NHQG_DIAG_MON_RESIDUAL_REL 24 /* mean_theta_exchange_residual_sbp_rel */
#define NHQG_DIAG_SCALARS 25

/* Build a context: uploads operators, allocates the device state and
 * workspace, captures the step graph lazily.  Replaces make_grid's device
 * side (nhqg/grid.py:439-789). */
int nhqg_create(const nhqg_params *p, const nhqg_operators *ops, nhqg_ctx **out);
int nhqg_destroy(nhqg_ctx *ctx);

/* State in the reference layout (nhqg/solver.py:33-38): q (nz+1, nx, nx/2+1)
 * complex128, w and th (nz-1, nx, nx/2+1) complex128, th_bar (nz+1,) float64.
 * Modes outside the layout are ignored on set and returned as zero. */
int nhqg_set_state(nhqg_ctx *ctx, const double *q, const double *w, const double *th,
                   const double *th_bar);
int nhqg_get_state(nhqg_ctx *ctx, double *q, double *w, double *th, double *th_bar);

/* Advance n_steps of imex_step (nhqg/solver.py:1228-1242) on the device
 * state by replaying the captured step graph; synchronous. */
int nhqg_step(nhqg_ctx *ctx, int32_t n_steps);
/* Same, asynchronous on the context stream (call nhqg_sync to wait). */
int nhqg_step_async(nhqg_ctx *ctx, int32_t n_steps);
int nhqg_sync(nhqg_ctx *ctx);

/* Judged diagnostics of the current state (nhqg/diagnostics.py:597-675):
 * out[NHQG_DIAG_*] scalars then q, w, theta vertical spectra.  Returns
 * NHQG_NONFINITE (output still written) when any value is NaN or infinite:
 * the condition on which the reference driver saves a checkpoint and stops
 * (scripts/run_polar.py:364-367). */
int nhqg_diagnostics(nhqg_ctx *ctx, double *out, int32_t n_out);

/* Depth-integrated power of each stored mode of the nodal q, w, theta,
 * sum_j cc_j |f_j(k)|^2: out[f * nm + m] for f = q, w, theta and the mode
 * layout's modes m (3 * nm values).  Binned over |k| shells with the rfft
 * weights it gives compute_diagnostics' q/w/th_horiz_spec
 * (nhqg/diagnostics.py:99-108, 683-685). */
int nhqg_mode_power(nhqg_ctx *ctx, double *out, int32_t n_out);

/* dst's state <- src's state spectrally upsampled from src's nx to dst's
 * (same nz, dst nx >= src nx, same parity), then sanitized for dst's layout:
 * the method of scripts/upsample_checkpoint_horiz.py:27-59 (the old Nyquist
 * row and column split in half, factor (nx_dst/nx_src)^2), computed on the
 * device from the resident state with no host round trip.  Both contexts
 * on the current device; th_bar is copied. */
int nhqg_upsample(nhqg_ctx *dst, nhqg_ctx *src);

/* The driver's annular barotropic vorticity initial state
 * (scripts/run_polar.py:115-158) as the context state.  noise: nx*nx
 * row-major, the caller's standard-normal draw (the driver's
 * numpy.random.default_rng(seed).standard_normal((Nx, Nx))).  On the device:
 * rfft2, the Gaussian annulus exp(-((|k| - k_peak)/k_width)^2 / 2) with
 * k = 0 removed and the 2/3 band, irfft2; when confine_radius > 0 a polar-cap
 * taper 0.5 (1 - tanh((r - confine_radius) / (4 pi / k_peak))) around the
 * box centre and the 2/3 band again; scaled to population standard deviation
 * `amplitude` (over r < confine_radius when confined); its rfft2 is q's T_0
 * coefficient, everything else zero.  L: the box side.  Uses the context's
 * FFT scratch; allocates nothing. */
int nhqg_init_barotropic(nhqg_ctx *ctx, const double *noise, double L, double k_peak,
                         double k_width, double amplitude, double confine_radius);

/* Gaussian cyclones added to q's T_0 coefficient (scripts/run_polar.py:
 * 91-112): centers holds count (x0, y0) pairs, the caller's uniform draws on
 * [0, L); the patch sum_c amplitude exp(-(dx^2 + dy^2) / (2 radius^2)) with
 * periodic separations is rfft2'd, 2/3-masked and added. */
int nhqg_inject_cyclones(nhqg_ctx *ctx, const double *centers, int32_t count, double L,
                         double amplitude, double radius);

/* Component entry points on the device state, for parity testing:
 *   NHQG_OP_EXPLICIT_RHS    aux <- explicit_rhs(state)    solver.py:510-599
 *   NHQG_OP_IMPLICIT_TEND   aux <- implicit_tendency(state) solver.py:615-641
 *   NHQG_OP_IMPLICIT_SOLVE  state <- imex_implicit_solve(state as R)
 *                           and th_bar <- proj_d th_bar     solver.py:691-744
 *   NHQG_OP_CORRECTOR       state <- SBP corrector(state, total_dt=arg)
 *                                                         solver.py:909-950
 *   NHQG_OP_SANITIZE        state <- finalize + sanitize  solver.py:128-146
 *   NHQG_OP_RK4_STEP        state <- rk4_step(state), the explicit RK4
 *                           validator, including th_bar's tendency under
 *                           evolve_mean (either exchange)  solver.py:1276-1336
 * nhqg_get_aux copies the aux fields (same layout as the state, th_bar
 * part returned as zeros). */
#define NHQG_OP_EXPLICIT_RHS 1
#define NHQG_OP_IMPLICIT_TEND 2
#define NHQG_OP_IMPLICIT_SOLVE 3
#define NHQG_OP_CORRECTOR 4
#define NHQG_OP_SANITIZE 5
#define NHQG_OP_RK4_STEP 6
int nhqg_apply(nhqg_ctx *ctx, int32_t op, double arg);
int nhqg_get_aux(nhqg_ctx *ctx, double *q, double *w, double *th);

/* Device time of the last nhqg_step call's replays in milliseconds, from
 * HIP events on the context stream. */
int nhqg_last_step_ms(nhqg_ctx *ctx, double *ms);

/* Per-kernel-class timing of one eagerly launched step (HIP events around
 * each group on the context stream): out[0..n) in ms, names in
 * nhqg_timing_name(i). */
int nhqg_profile_step(nhqg_ctx *ctx, double *out, int32_t n_out);
const char *nhqg_timing_name(int32_t i);

/* Measurement utility with no counterpart in the reference: a plain
 * streaming copy of `bytes` (read + write), repeated `reps` times on the
 * context stream.  gbps[0]: the library's copy kernel (stream_copy_kernel,
 * the known-size dispatch that calibrates HBM traffic counters); gbps[1]:
 * hipMemcpyAsync device-to-device over the same buffers.  Both count read +
 * write bytes; the larger is the practical bandwidth ceiling. */
int nhqg_stream_copy(nhqg_ctx *ctx, uint64_t bytes, int32_t reps, double gbps[2]);

/* Copy the inverse of the s-th shell factor held by the context
 * (imex_inv[s] of nhqg/grid.py:430, Ng x Ng row-major) to out. */
int nhqg_get_shell_inverse(nhqg_ctx *ctx, int32_t s, double *out);

/* ---- Learned vertical-tail closure (no counterpart in the reference) ----
 * A per-column MLP x (16) -> tanh (128) -> tanh (128) -> 16 acting on the
 * top 16 Chebyshev coefficients of q; its fallback and judge is the plain
 * solver at higher Nz.  Parameters: one flat float32 array W1 (128 x 16),
 * b1, W2 (128 x 128), b2, W3 (16 x 128), b3.  Training on the device:
 * FP32 forward/backward, mean-squared loss, Adam.
 * NHQG_MLP_ODD holds every bias at zero (set_params clears them, training
 * leaves them): the network is then odd, MLP(-x) = -MLP(x), so the step
 * closure adds nothing to zero columns and keeps the k / -k columns of
 * ky = 0 Hermitian. */
#define NHQG_MLP_ODD 1
typedef struct nhqg_mlp nhqg_mlp;
int nhqg_mlp_create(int32_t n_in, int32_t width, int32_t n_out, int32_t max_batch,
                    int32_t flags, nhqg_mlp **out);
int nhqg_mlp_destroy(nhqg_mlp *m);
int nhqg_mlp_n_params(const nhqg_mlp *m, int64_t *n);
int nhqg_mlp_set_params(nhqg_mlp *m, const float *w);  /* also resets Adam */
int nhqg_mlp_get_params(nhqg_mlp *m, float *w);
/* y (batch x n_out) = MLP(x (batch x n_in)) */
int nhqg_mlp_forward(nhqg_mlp *m, const float *x, int32_t batch, float *y);

/* Allocation-free forward pass.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory during the call (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * result as nhqg_mlp_forward, but x (batch x n_in) and y (batch x n_out) are
 * DEVICE pointers owned by the caller, the three layer kernels are enqueued
 * on `stream` (a hipStream_t; NULL = the null stream) and the call returns
 * without waiting.  The hidden activations use a fixed device array of
 * 2 x NHQG_MLP_NOALLOC_MAX_BATCH x NHQG_MLP_NOALLOC_MAX_WIDTH floats (4 MiB per
 * GPU, kept until the process ends); batch above the maximum, or a network
 * wider than the maximum, returns NHQG_ERR_ARG with the reason in
 * nhqg_mlp_last_error.  Every call on one GPU shares that array: calls that
 * overlap in time on one GPU race.  Nothing in it needs clearing: each call
 * writes every activation it reads. */
#define NHQG_MLP_NOALLOC_MAX_BATCH 4096
#define NHQG_MLP_NOALLOC_MAX_WIDTH 128
int nhqg_mlp_forward_noalloc(nhqg_mlp *m, const float *x, int32_t batch, float *y,
                             void *stream);
/* gradient of mean((MLP(x) - t)^2) w.r.t. the parameters, and the loss */
int nhqg_mlp_gradient(nhqg_mlp *m, const float *x, const float *t, int32_t batch, float *grad,
                      float *loss);
/* training set resident on the device (n samples) */
int nhqg_mlp_set_data(nhqg_mlp *m, const float *x, const float *t, int32_t n);
/* n_steps Adam steps over consecutive minibatches starting at sample start;
 * loss[i] (may be NULL) = minibatch loss before step i */
int nhqg_mlp_train(nhqg_mlp *m, int32_t n_steps, int32_t batch, int32_t start, float lr,
                   float *loss);
const char *nhqg_mlp_last_error(const nhqg_mlp *m);
/* Attach the network (BF16-rounded weights, FP32 accumulation) to every
 * explicit RHS of the step: E_q[n0 + j] += out_scale * MLP(in_scale *
 * q[n0 .. n0+15])_j for each mode and each of Re, Im.  The network must be
 * 16 -> 128 -> 128 -> 16 and created with NHQG_MLP_ODD (NHQG_ERR_ARG
 * otherwise); mlp = NULL detaches.  The context keeps a BF16 snapshot of
 * the weights taken by this call: training the network afterwards changes
 * the step only after it is attached again. */
int nhqg_set_closure(nhqg_ctx *ctx, const nhqg_mlp *mlp, int32_t n0, double in_scale,
                     double out_scale);

/* Bytes of device memory held by the context. */
int nhqg_device_bytes(nhqg_ctx *ctx, uint64_t *bytes);

/* Last error text of a context (or of the last failed nhqg_create when
 * ctx is NULL). */
const char *nhqg_last_error(const nhqg_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* NHQG_H */
