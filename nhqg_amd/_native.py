"""ctypes binding of the HIP library (include/nhqg.h).

The library ``libnhqg.so`` lives next to this file and is built by
``nhqg_amd/build.py``.  There is no CPU fallback: if the library or a GPU is
missing, creating a context raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnhqg.so")
_lib = None

NHQG_OK = 0
NHQG_NONFINITE = 5
OP_EXPLICIT_RHS = 1
OP_IMPLICIT_TEND = 2
OP_IMPLICIT_SOLVE = 3
OP_CORRECTOR = 4
OP_SANITIZE = 5
OP_RK4_STEP = 6
MLP_ODD = 1
DIAG_NAMES = ("KE_tot", "KE_bt", "KE_bc", "max_speed", "enstrophy", "Nusselt",
              "Nusselt_dealiased", "vol_avg_tw_dealiased", "th_bar_max", "q_rms", "w_rms",
              "th_rms", "max_w", "max_theta", "max_tw", "vol_avg_tw", "heat_flux_mismatch",
              "q_high_frac", "w_high_frac", "th_high_frac", "th_mean_feedback_sum_sbp",
              "mean_flux_exchange_tendency_sbp", "mean_theta_exchange_boundary_sbp",
              "mean_theta_exchange_residual_sbp", "mean_theta_exchange_residual_sbp_rel")
N_DIAG_SCALARS = len(DIAG_NAMES)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Params(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "nx", "nz", "nkx", "nky", "n_shells", "n_pairs", "substeps", "evolve_mean",
        "advection", "w_neumann", "apply_mask", "mixed_slices", "dense_contractions",
        "exchange", "npad")] + [
        (n, C.c_double) for n in ("dt", "gamma_imex", "ra_sigma", "sigma", "eps_sq", "beta")]


_OPS_FIELDS = [
    ("V", _dp), ("VSw", _dp), ("VSd", _dp), ("Vinv", _dp), ("Pw", _dp), ("Pd", _dp),
    ("Qw", _dp), ("GSw", _dp), ("W2S", _dp), ("T2S", _dp), ("S2T", _dp), ("thb2s", _dp),
    ("s2thb", _dp), ("proj_d", _dp), ("d1_band", _dp), ("lap_band", _dp), ("Mtw", _dp),
    ("Mwt", _dp), ("Mth2w", _dp), ("cc", _dp), ("row_cc", _dp), ("cc_sbp", _dp), ("Sw", _dp),
    ("Sd", _dp), ("rows", _ip), ("kidx_of_x", _ip), ("full_idx", _ip), ("kx_row", _dp),
    ("ky_col", _dp), ("kx_mode", _dp), ("ksq", _dp), ("inv_denom", _dp),
    ("inv_alpha_q", _dp), ("inv_alpha_th", _dp), ("diss_q", _dp), ("diss_w", _dp),
    ("diss_th", _dp), ("mask", _dp), ("pweight", _dp), ("partner", _ip),
    ("herm_pairs", _ip), ("shell_ptr", _ip), ("shell_modes", _ip), ("shell_inv", _dp),
    ("eta_phys", _dp), ("sponge_phys", _dp), ("twiddle", _dp), ("shell_B0", _dp),
    ("shell_K", _dp), ("shell_coef", _dp), ("eig_Q", _dp), ("eig_Qinv", _dp), ("eig_lam", _dp),
    ("mode_alpha", _dp), ("mode_c", _dp), ("eta_x_phys", _dp), ("eta_y_phys", _dp),
    ("VGZ", _dp), ("mean_solve", _dp), ("mean_diff", _dp), ("mean_flux", _dp),
    ("pweight_pad", _dp), ("partner_pad", _ip),
    ("kidx_of_x_pad", _ip), ("rows_pad", _ip), ("twiddle_pad", _dp), ("eta_phys_pad", _dp),
    ("eta_x_phys_pad", _dp), ("eta_y_phys_pad", _dp), ("sponge_phys_pad", _dp),
    ("cutoff_w", _dp), ("cutoff_th", _dp),
]


class Operators(C.Structure):
    _fields_ = _OPS_FIELDS


EXPORTS = (
    "nhqg_create", "nhqg_destroy", "nhqg_set_state", "nhqg_get_state", "nhqg_step",
    "nhqg_step_async", "nhqg_sync", "nhqg_diagnostics", "nhqg_apply", "nhqg_get_aux",
    "nhqg_last_step_ms", "nhqg_profile_step", "nhqg_timing_name", "nhqg_device_bytes",
    "nhqg_get_shell_inverse",
    "nhqg_last_error", "nhqg_stream_copy",
    "nhqg_mlp_create", "nhqg_mlp_destroy", "nhqg_mlp_n_params", "nhqg_mlp_set_params",
    "nhqg_mlp_get_params", "nhqg_mlp_forward", "nhqg_mlp_gradient", "nhqg_mlp_set_data",
    "nhqg_mlp_train", "nhqg_mlp_last_error", "nhqg_set_closure", "nhqg_mlp_forward_noalloc",
    "nhqg_mode_power", "nhqg_upsample", "nhqg_init_barotropic", "nhqg_inject_cyclones",
)


class NativeError(RuntimeError):
    pass


def lib_path() -> str:
    return _LIB_PATH


def load():
    """Load libnhqg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise NativeError(f"{_LIB_PATH} is missing: run `python nhqg_amd/build.py` "
                          "(or __graft_entry__.build()) first")
    lib = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    lib.nhqg_create.argtypes = [C.POINTER(Params), C.POINTER(Operators), C.POINTER(vp)]
    lib.nhqg_destroy.argtypes = [vp]
    lib.nhqg_set_state.argtypes = [vp, _dp, _dp, _dp, _dp]
    lib.nhqg_get_state.argtypes = [vp, _dp, _dp, _dp, _dp]
    lib.nhqg_step.argtypes = [vp, C.c_int32]
    lib.nhqg_step_async.argtypes = [vp, C.c_int32]
    lib.nhqg_sync.argtypes = [vp]
    lib.nhqg_diagnostics.argtypes = [vp, _dp, C.c_int32]
    lib.nhqg_mode_power.argtypes = [vp, _dp, C.c_int32]
    lib.nhqg_upsample.argtypes = [vp, vp]
    d = C.c_double
    lib.nhqg_init_barotropic.argtypes = [vp, _dp, d, d, d, d, d]
    lib.nhqg_inject_cyclones.argtypes = [vp, _dp, C.c_int32, d, d, d]
    lib.nhqg_apply.argtypes = [vp, C.c_int32, C.c_double]
    lib.nhqg_get_aux.argtypes = [vp, _dp, _dp, _dp]
    lib.nhqg_last_step_ms.argtypes = [vp, _dp]
    lib.nhqg_profile_step.argtypes = [vp, _dp, C.c_int32]
    lib.nhqg_timing_name.argtypes = [C.c_int32]
    lib.nhqg_timing_name.restype = C.c_char_p
    lib.nhqg_device_bytes.argtypes = [vp, C.POINTER(C.c_uint64)]
    lib.nhqg_get_shell_inverse.argtypes = [vp, C.c_int32, _dp]
    _fp = C.POINTER(C.c_float)
    lib.nhqg_mlp_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    C.POINTER(vp)]
    lib.nhqg_mlp_destroy.argtypes = [vp]
    lib.nhqg_mlp_n_params.argtypes = [vp, C.POINTER(C.c_int64)]
    lib.nhqg_mlp_set_params.argtypes = [vp, _fp]
    lib.nhqg_mlp_get_params.argtypes = [vp, _fp]
    lib.nhqg_mlp_forward.argtypes = [vp, _fp, C.c_int32, _fp]
    lib.nhqg_mlp_forward_noalloc.argtypes = [vp, vp, C.c_int32, vp, vp]
    lib.nhqg_mlp_gradient.argtypes = [vp, _fp, _fp, C.c_int32, _fp, _fp]
    lib.nhqg_mlp_set_data.argtypes = [vp, _fp, _fp, C.c_int32]
    lib.nhqg_mlp_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_float, _fp]
    lib.nhqg_mlp_last_error.argtypes = [vp]
    lib.nhqg_mlp_last_error.restype = C.c_char_p
    lib.nhqg_set_closure.argtypes = [vp, vp, C.c_int32, C.c_double, C.c_double]
    lib.nhqg_stream_copy.argtypes = [vp, C.c_uint64, C.c_int32, _dp]
    lib.nhqg_last_error.argtypes = [vp]
    lib.nhqg_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def _ptr(a, kind=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(kind)


class Context:
    """Device context of one Grid and mode layout."""

    def __init__(self, grid, layout):
        from nhqg_amd.grid import device_operators

        self.lib = load()
        self.grid = grid
        self.layout = layout
        t = layout.tables
        ops = device_operators(grid, layout)
        keep = {}

        def arr(a, dt=np.float64):
            if a is None:
                return None
            b = np.ascontiguousarray(a, dtype=dt)
            keep[id(b)] = b
            return b

        o = Operators()
        for name, typ in _OPS_FIELDS:
            src = t[name] if name in t else ops.get(name)
            if typ is _ip:
                setattr(o, name, _ptr(arr(src, np.int32), _ip))
            else:
                setattr(o, name, _ptr(arr(src, np.float64), _dp))
        p = Params()
        p.nx, p.nz = grid.Nx, grid.Nz
        p.nkx, p.nky = layout.nkx, layout.nky
        p.n_shells = int(t["shell_ids"].size)
        p.n_pairs = int(t["herm_pairs"].size // 2)
        p.substeps = max(1, int(grid.sbp_corrector_substeps))
        p.evolve_mean = 1 if grid.evolve_mean else 0
        p.advection = 1 if grid.nonlinear_advection == "jacobian" else 0
        p.w_neumann = 1 if grid.w_bc_top == "neumann" else 0
        rule32 = grid.horizontal_dealiasing == "32_rule"
        # the 2/3 rule masks the outputs of a full layout; the 3/2 rule has no band
        p.apply_mask = 1 if layout.full and not rule32 else 0
        p.npad = grid.Npad if rule32 else grid.Nx
        p.mixed_slices = int(getattr(grid, "mixed_slices", 0))
        p.dense_contractions = int(getattr(grid, "dense_contractions", 0))
        p.exchange = 0 if grid.mean_exchange_discretization == "balanced_sbp2_pc" else 1
        p.dt, p.gamma_imex = grid.dt, grid.gamma_imex
        p.ra_sigma, p.sigma = grid.Ra_sigma, grid.sigma
        p.eps_sq, p.beta = grid.mean_temp_eps_sq, grid.beta
        h = C.c_void_p()
        rc = self.lib.nhqg_create(C.byref(p), C.byref(o), C.byref(h))
        if rc != NHQG_OK:
            raise NativeError(f"nhqg_create failed ({rc}): "
                              f"{self.lib.nhqg_last_error(None).decode()}")
        self.h = h
        self.params = p

    def _check(self, rc, what):
        if rc != NHQG_OK:
            msg = self.lib.nhqg_last_error(self.h).decode() if self.h else ""
            raise NativeError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.nhqg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shapes(self):
        g = self.grid
        return ((g.Nz + 1, g.Nx, g.Nk), (g.Nz - 1, g.Nx, g.Nk), (g.Nz + 1,))

    def set_state(self, q, w, th, thb):
        sq, sg, sb = self.shapes()
        q = np.ascontiguousarray(q, dtype=np.complex128)
        w = np.ascontiguousarray(w, dtype=np.complex128)
        th = np.ascontiguousarray(th, dtype=np.complex128)
        thb = np.ascontiguousarray(thb, dtype=np.float64)
        if q.shape != sq or w.shape != sg or th.shape != sg or thb.shape != sb:
            raise ValueError("state shapes do not match the grid")
        self._check(self.lib.nhqg_set_state(self.h, _ptr(q.view(np.float64)),
                                            _ptr(w.view(np.float64)),
                                            _ptr(th.view(np.float64)), _ptr(thb)),
                    "nhqg_set_state")

    def get_state(self):
        sq, sg, sb = self.shapes()
        q = np.empty(sq, np.complex128)
        w = np.empty(sg, np.complex128)
        th = np.empty(sg, np.complex128)
        thb = np.empty(sb, np.float64)
        self._check(self.lib.nhqg_get_state(self.h, _ptr(q.view(np.float64)),
                                            _ptr(w.view(np.float64)),
                                            _ptr(th.view(np.float64)), _ptr(thb)),
                    "nhqg_get_state")
        return q, w, th, thb

    def step(self, n=1):
        self._check(self.lib.nhqg_step(self.h, int(n)), "nhqg_step")

    def step_async(self, n=1):
        self._check(self.lib.nhqg_step_async(self.h, int(n)), "nhqg_step_async")

    def sync(self):
        self._check(self.lib.nhqg_sync(self.h), "nhqg_sync")

    def last_step_ms(self):
        v = C.c_double()
        self._check(self.lib.nhqg_last_step_ms(self.h, C.byref(v)), "nhqg_last_step_ms")
        return v.value

    def apply(self, op, arg=0.0):
        self._check(self.lib.nhqg_apply(self.h, int(op), float(arg)), "nhqg_apply")

    def get_aux(self):
        sq, sg, _ = self.shapes()
        q = np.empty(sq, np.complex128)
        w = np.empty(sg, np.complex128)
        th = np.empty(sg, np.complex128)
        self._check(self.lib.nhqg_get_aux(self.h, _ptr(q.view(np.float64)),
                                          _ptr(w.view(np.float64)),
                                          _ptr(th.view(np.float64))), "nhqg_get_aux")
        return q, w, th

    def diagnostics(self):
        n = N_DIAG_SCALARS + 3 * (self.grid.Nz + 1)
        out = np.zeros(n, np.float64)
        rc = self.lib.nhqg_diagnostics(self.h, _ptr(out), n)
        if rc != NHQG_NONFINITE:
            self._check(rc, "nhqg_diagnostics")
        return out

    def upsample_from(self, src: "Context") -> None:
        """This context's state <- src's state spectrally upsampled (device)."""
        self._check(self.lib.nhqg_upsample(self.h, src.h), "nhqg_upsample")

    def init_barotropic(self, noise, L, k_peak, k_width, amplitude, confine_radius=None):
        """The driver's annular barotropic vorticity state built on the device
        from the caller's (nx, nx) standard-normal draw."""
        n = self.grid.Nx
        noise = np.ascontiguousarray(noise, dtype=np.float64)
        if noise.shape != (n, n):
            raise ValueError(f"noise must be ({n}, {n})")
        self._check(self.lib.nhqg_init_barotropic(
            self.h, _ptr(noise), float(L), float(k_peak), float(k_width), float(amplitude),
            0.0 if confine_radius is None else float(confine_radius)), "nhqg_init_barotropic")

    def inject_cyclones(self, centers, L, amplitude, radius):
        """Gaussian cyclones at the given (x0, y0) centres added to q's T_0
        coefficient on the device."""
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 2)
        self._check(self.lib.nhqg_inject_cyclones(
            self.h, _ptr(c), int(c.shape[0]), float(L), float(amplitude), float(radius)),
            "nhqg_inject_cyclones")

    def mode_power(self):
        """(3, nm): sum_j cc_j |f_j|^2 of the nodal q, w, theta per mode."""
        nm = self.layout.nm
        out = np.zeros(3 * nm, np.float64)
        self._check(self.lib.nhqg_mode_power(self.h, _ptr(out), 3 * nm), "nhqg_mode_power")
        return out.reshape(3, nm)

    def profile_step(self):
        """Advance one step eagerly with per-kernel-class HIP event timing.
        Returns {class: (ms, launches, work)}."""
        n = 3 * 16
        out = np.zeros(n, np.float64)
        self._check(self.lib.nhqg_profile_step(self.h, _ptr(out), n), "nhqg_profile_step")
        res = {}
        for i in range(16):
            name = self.lib.nhqg_timing_name(i).decode()
            if name:
                res[name] = (float(out[3 * i]), int(out[3 * i + 1]), float(out[3 * i + 2]))
        return res

    def stream_copy(self, nbytes, reps=10):
        """(kernel, hipMemcpyAsync) GB/s, read + write bytes, of a plain
        device-to-device copy of nbytes."""
        v = (C.c_double * 2)()
        self._check(self.lib.nhqg_stream_copy(self.h, int(nbytes), int(reps), v),
                    "nhqg_stream_copy")
        return float(v[0]), float(v[1])

    def shell_inverse(self, s: int) -> np.ndarray:
        """inv(A'(k)) of the s-th touched shell as held on the device."""
        n = self.grid.Nz - 1
        out = np.empty((n, n))
        self._check(self.lib.nhqg_get_shell_inverse(self.h, int(s), _ptr(out)),
                    "nhqg_get_shell_inverse")
        return out

    def set_closure(self, mlp, n0: int, in_scale: float, out_scale: float) -> None:
        """Attach a trained MLP (BF16 weights) to every explicit RHS of the
        step; mlp=None detaches."""
        self._check(self.lib.nhqg_set_closure(self.h, mlp.h if mlp is not None else None,
                                              int(n0), float(in_scale), float(out_scale)),
                    "nhqg_set_closure")

    def device_bytes(self):
        v = C.c_uint64()
        self._check(self.lib.nhqg_device_bytes(self.h, C.byref(v)), "nhqg_device_bytes")
        return int(v.value)


class MLP:
    """Device MLP of the learned closure (nhqg_mlp_*): FP32 parameters,
    forward / gradient / Adam training on the GPU."""

    def __init__(self, n_in: int = 16, width: int = 128, n_out: int = 16, max_batch: int = 4096,
                 odd: bool = False):
        self.lib = load()
        self.n_in, self.width, self.n_out, self.max_batch = n_in, width, n_out, max_batch
        self.odd = bool(odd)
        h = C.c_void_p()
        rc = self.lib.nhqg_mlp_create(n_in, width, n_out, max_batch, MLP_ODD if odd else 0,
                                      C.byref(h))
        if rc != NHQG_OK:
            raise NativeError(f"nhqg_mlp_create failed ({rc})")
        self.h = h
        n = C.c_int64()
        self._check(self.lib.nhqg_mlp_n_params(self.h, C.byref(n)), "nhqg_mlp_n_params")
        self.n_params = int(n.value)

    def _check(self, rc, what):
        if rc != NHQG_OK:
            raise NativeError(f"{what} failed ({rc}): {self.lib.nhqg_mlp_last_error(self.h).decode()}")

    @staticmethod
    def _f(a):
        return np.ascontiguousarray(a, dtype=np.float32)

    def set_params(self, w):
        w = self._f(w)
        assert w.size == self.n_params
        self._check(self.lib.nhqg_mlp_set_params(self.h, _ptr(w, C.POINTER(C.c_float))),
                    "nhqg_mlp_set_params")

    def get_params(self):
        w = np.empty(self.n_params, np.float32)
        self._check(self.lib.nhqg_mlp_get_params(self.h, _ptr(w, C.POINTER(C.c_float))),
                    "nhqg_mlp_get_params")
        return w

    def forward(self, x):
        x = self._f(x)
        y = np.empty((x.shape[0], self.n_out), np.float32)
        fp = C.POINTER(C.c_float)
        self._check(self.lib.nhqg_mlp_forward(self.h, _ptr(x, fp), x.shape[0], _ptr(y, fp)),
                    "nhqg_mlp_forward")
        return y

    def forward_noalloc(self, x_dev: int, batch: int, y_dev: int, stream: int = 0) -> None:
        """Allocation-free forward pass on device pointers (x: batch x n_in,
        y: batch x n_out float32), enqueued on `stream` without waiting."""
        self._check(self.lib.nhqg_mlp_forward_noalloc(self.h, C.c_void_p(x_dev), int(batch),
                                                      C.c_void_p(y_dev), C.c_void_p(stream)),
                    "nhqg_mlp_forward_noalloc")

    def gradient(self, x, t):
        x, t = self._f(x), self._f(t)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        fp = C.POINTER(C.c_float)
        self._check(self.lib.nhqg_mlp_gradient(self.h, _ptr(x, fp), _ptr(t, fp), x.shape[0],
                                               _ptr(g, fp), C.byref(loss)), "nhqg_mlp_gradient")
        return g, float(loss.value)

    def set_data(self, x, t):
        x, t = self._f(x), self._f(t)
        fp = C.POINTER(C.c_float)
        self._check(self.lib.nhqg_mlp_set_data(self.h, _ptr(x, fp), _ptr(t, fp), x.shape[0]),
                    "nhqg_mlp_set_data")
        self.n_data = x.shape[0]

    def train(self, n_steps: int, batch: int, lr: float, start: int = 0):
        loss = np.empty(n_steps, np.float32)
        fp = C.POINTER(C.c_float)
        self._check(self.lib.nhqg_mlp_train(self.h, n_steps, batch, start, lr, _ptr(loss, fp)),
                    "nhqg_mlp_train")
        return loss

    def close(self):
        if getattr(self, "h", None):
            self.lib.nhqg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
