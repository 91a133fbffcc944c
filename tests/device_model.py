"""numpy model of the device algorithm (test helper, not product code).

Executes the same sequence as nhqg_amd/csrc/api.hip on the same tables the
host passes to nhqg_create (band-compact layout, fused vertical operators,
packed real FFTs along y, Parseval horizontal means, banded LU of the SBP
matrix).  Tests compare it with the oracle on the CPU, which validates the
host tables and every algebraic restructuring before the GPU runs them.
"""

from __future__ import annotations

import numpy as np

from nhqg_amd.grid import device_operators


class DeviceModel:
    def __init__(self, grid, layout):
        self.g = grid
        self.lay = layout
        self.t = layout.tables
        self.o = device_operators(grid, layout)
        self.nm = layout.nm
        if self.o.get("eig_Q") is not None:
            # eigenbasis solve: one inverse per shell from its modes' alpha, c
            Q, Qi, lam = self.o["eig_Q"], self.o["eig_Qinv"], self.o["eig_lam"]
            ptr, modes = self.t["shell_ptr"], self.t["shell_modes"]
            inv = []
            for s in range(ptr.size - 1):
                m = modes[ptr[s]]
                d = 1.0 / (self.o["mode_alpha"][m] - self.o["mode_c"][m] * lam)
                inv.append((Q @ np.diag(d) @ Qi).T)
            self.o["shell_inv"] = np.array(inv)
        elif self.o["shell_inv"] is None:
            # device-built factors: A_s = alpha I - beta K - c B0 (transposed
            # inverses, as the device holds them)
            B0, K, coef = self.o["shell_B0"], self.o["shell_K"], self.o["shell_coef"]
            I = np.eye(B0.shape[0])
            inv = []
            for al, be, c in coef:
                A = al * I - c * B0 if K is None else al * I - be * K - c * B0
                inv.append(np.linalg.inv(A).T)
            self.o["shell_inv"] = np.array(inv)

    # ---- layout ----
    def to_compact(self, f):
        fl = np.asarray(f).reshape(f.shape[0], -1)
        return fl[:, self.t["full_idx"]].copy()

    def to_full(self, c, nl):
        g = self.g
        out = np.zeros((nl, g.Nx * g.Nk), complex)
        out[:, self.t["full_idx"]] = c
        return out.reshape(nl, g.Nx, g.Nk)

    # ---- FFT passes ----
    def flux(self, qn, wn, thn):
        g, t, o = self.g, self.t, self.o
        N, nkx, nky = g.Nx, self.lay.nkx, self.lay.nky
        rows = t["rows"]
        nl = qn.shape[0]

        def cols(c):  # (nl, nm) -> (nl, nky, N) full kx columns
            a = np.zeros((nl, nky, N), complex)
            a[:, :, rows] = c.reshape(nl, nky, nkx)
            return a

        idn = t["inv_denom"].reshape(nky, nkx)
        kyc = t["ky_col"][:, None]
        kxr = t["kx_row"][None, :]
        Uc = np.zeros((nl, nky, N), complex)
        Vc = np.zeros((nl, nky, N), complex)
        Uc[:, :, rows] = (1j * kyc * idn)[None] * qn.reshape(nl, nky, nkx)
        Vc[:, :, rows] = (-1j * kxr * idn)[None] * qn.reshape(nl, nky, nkx)
        X = [np.fft.ifft(a, axis=2) * N for a in (Uc, Vc, cols(qn), cols(wn), cols(thn))]
        # X[f]: (nl, nky, x) -> rows along y: (nl, x, ky)
        X = [np.transpose(a, (0, 2, 1)) for a in X]

        def inv_pair(P, Q):
            Z = np.zeros((nl, N, N), complex)
            k = np.arange(N)
            ky = np.where(k <= N // 2, k, N - k)
            ok = ky < nky
            Pk = np.zeros((nl, N, N), complex)
            Qk = np.zeros((nl, N, N), complex)
            Pk[:, :, ok] = P[:, :, ky[ok]]
            if Q is not None:
                Qk[:, :, ok] = Q[:, :, ky[ok]]
            lo = (k > 0) & (k < N // 2)
            hi = k > N // 2
            sp = (k == 0) | (k == N // 2)
            Z[:, :, lo] = Pk[:, :, lo] + 1j * Qk[:, :, lo]
            Z[:, :, hi] = np.conj(Pk[:, :, hi]) + 1j * np.conj(Qk[:, :, hi])
            Z[:, :, sp] = Pk[:, :, sp].real + 1j * Qk[:, :, sp].real
            z = np.fft.ifft(Z, axis=2) * N / (N * N)
            return z.real, z.imag

        def fwd_pair(a, b):
            Z = np.fft.fft(a + 1j * b, axis=2)
            k = np.arange(nky)
            zk = Z[:, :, k]
            zm = Z[:, :, (N - k) % N]
            A = 0.5 * (zk + np.conj(zm))
            B = (zk - np.conj(zm)) / 2j
            return A, B

        u, v = inv_pair(X[0], X[1])
        q, _ = inv_pair(X[2], None)
        w, _ = inv_pair(X[3], None)
        th, _ = inv_pair(X[4], None)
        sg = o["sponge_phys"]
        eta = o["eta_phys"]
        S = [np.zeros((nl, N, nky), complex) for _ in range(3)]
        qa = q + (eta if eta is not None else 0.0)
        out = []
        for f, Sf in ((qa, S[0]), (w, S[1]), (th, S[2])):
            G, Hv = fwd_pair(u * f, v * f)
            H = 1j * t["ky_col"][None, None, :] * Hv + Sf
            # columns: (nl, x, ky) -> FFT along x
            Gx = np.fft.fft(np.transpose(G, (0, 2, 1)), axis=2)[:, :, rows]
            Hx = np.fft.fft(np.transpose(H, (0, 2, 1)), axis=2)[:, :, rows]
            A = 1j * t["kx_row"][None, None, :] * Gx + Hx
            out.append((A.reshape(nl, -1)) * t["mask"][None])
        return out

    # ---- step pieces ----
    def sponge(self, c):
        g = self.g
        f = np.fft.irfft2(self.to_full(c, c.shape[0]), s=(g.Nx, g.Nx))
        return self.to_compact(np.fft.rfft2(self.o["sponge_phys"][None] * f)) * self.t["mask"]

    def explicit_rhs(self, q, w, th):
        o = self.o
        qn, wn, thn = o["V"] @ q, o["VSw"] @ w, o["VSd"] @ th
        Aq, Aw, Ath = self.flux(qn, wn, thn)
        Eq = -(o["Vinv"] @ Aq)
        if self.g.beta != 0.0:
            Eq = Eq + 1j * self.g.beta * self.t["kx_mode"] * self.t["inv_denom"] * q
        Ew, Eth = -(o["Pw"] @ Aw), -(o["Pd"] @ Ath)
        if o["sponge_phys"] is not None:
            Eq, Ew, Eth = Eq - self.sponge(q), Ew - self.sponge(w), Eth - self.sponge(th)
        return Eq, Ew, Eth

    def shell_apply(self, rhs):
        t, o = self.t, self.o
        out = np.zeros_like(rhs)
        ptr, modes = t["shell_ptr"], t["shell_modes"]
        for s in range(ptr.size - 1):
            ms = modes[ptr[s]:ptr[s + 1]]
            out[:, ms] = o["shell_inv"][s].T @ rhs[:, ms]
        return out

    def implicit_solve(self, Rq, Rw, Rth):
        g, t, o = self.g, self.t, self.o
        al = g.gamma_imex * g.dt
        idq = t["inv_denom"] * t["inv_alpha_q"]
        iath = t["inv_alpha_th"]
        rth = Rth if g.w_bc_top == "dirichlet" else o["Mth2w"] @ Rth
        rhs = al * idq * (o["Qw"] @ Rq) + Rw + al * g.Ra_sigma * iath * rth
        w = self.shell_apply(rhs)
        wth = w if g.w_bc_top == "dirichlet" else o["Mwt"] @ w
        th = (Rth + al * wth) * iath
        q = (Rq + al * (o["GSw"] @ w)) * t["inv_alpha_q"]
        return q, w, th

    def parseval(self, a, b):
        t = self.t
        p = t["partner"]
        sc = p >= 0
        A, B = a.copy(), b.copy()
        A[:, sc] = 0.5 * (a[:, sc] + np.conj(a[:, p[sc]]))
        B[:, sc] = 0.5 * (b[:, sc] + np.conj(b[:, p[sc]]))
        return np.sum(t["pweight"] * (A.real * B.real + A.imag * B.imag), axis=1)

    def corrector(self, w, th, thb, total):
        g, o = self.g, self.o
        n = g.Nz + 1
        ns = max(1, g.sbp_corrector_substeps)
        sdt = total / ns
        mu, kap = g.mean_temp_eps_sq, 1.0 / g.sigma
        d1 = o["d1_band"]
        lb = o["lap_band"]

        def dense3(b):
            M = np.zeros((n, n))
            for off in (-1, 0, 1):
                i = np.arange(max(0, -off), min(n, n - off))
                M[i, i + off] = b[i, off + 1]
            return M

        D1, Lp = dense3(d1), dense3(lb)
        wsbp = o["W2S"] @ w
        w2 = self.parseval(wsbp, wsbp)
        half = 0.5 * mu * kap * sdt * Lp + 0.25 * mu * sdt * sdt * (D1 @ np.diag(w2) @ D1)
        A = np.eye(n) - half
        B = np.eye(n) + half
        A[0, :] = 0.0
        A[-1, :] = 0.0
        A[0, 0] = A[-1, -1] = 1.0
        # banded LU without pivoting
        LU = A.copy()
        for k in range(n - 1):
            for i in range(k + 1, min(k + 3, n)):
                f = LU[i, k] / LU[k, k]
                LU[i, k] = f
                for j in range(k + 1, min(k + 3, n)):
                    LU[i, j] -= f * LU[k, j]
        # the substeps stay on the SBP grid: one cycle operator T2S @ S2T
        # per substep, S2T only after the last one
        P = o["T2S"] @ o["S2T"]
        ths = o["T2S"] @ th
        for s in range(ns):
            fl = self.parseval(wsbp, ths)
            tb = o["thb2s"] @ thb
            rhs = B @ tb - mu * sdt * (D1 @ fl)
            rhs[0] = rhs[-1] = 0.0
            x = rhs.copy()
            for i in range(1, n):
                for k in range(max(0, i - 2), i):
                    x[i] -= LU[i, k] * x[k]
            for i in range(n - 1, -1, -1):
                v = x[i]
                for j in range(i + 1, min(i + 3, n)):
                    v -= LU[i, j] * x[j]
                x[i] = v / LU[i, i]
            gh = 0.5 * (D1 @ (tb + x))
            x[0] = x[-1] = 0.0
            thb = o["s2thb"] @ x
            ths = ths - sdt * wsbp * gh[:, None]
            ths[0] = 0.0
            ths[-1] = 0.0
            ths = (P if s + 1 < ns else o["S2T"]) @ ths
        return ths, thb

    def sanitize(self, f):
        t = self.t
        f = f.copy()
        f[:, 0] = 0.0
        pr = t["herm_pairs"].reshape(-1, 2)
        a = f[:, pr[:, 0]].copy()
        b = f[:, pr[:, 1]].copy()
        f[:, pr[:, 0]] = 0.5 * (a + np.conj(b))
        f[:, pr[:, 1]] = 0.5 * (b + np.conj(a))
        if self.lay.full:
            f = f * t["mask"][None]
        return f

    def step(self, q, w, th, thb):
        g, t, o = self.g, self.t, self.o
        dt, gm = g.dt, g.gamma_imex
        al, om = gm * dt, dt * (1 - gm)
        de = -np.sqrt(2.0) / 2.0
        em = g.evolve_mean
        E1 = self.explicit_rhs(q, w, th)
        s1q, s1w, s1th = self.implicit_solve(q + al * E1[0], w + al * E1[1], th + al * E1[2])
        s1thb = o["proj_d"] @ thb
        th1, thb1 = s1th, s1thb
        if em:
            th1, thb1 = self.corrector(s1w, s1th, s1thb, al)
        E2 = self.explicit_rhs(s1q, s1w, th1)
        Rq = q + dt * (de * E1[0] + (1 - de) * E2[0]) + om * (o["GSw"] @ s1w) - om * t["diss_q"] * s1q
        thw = th1 if g.w_bc_top == "dirichlet" else o["Mtw"] @ th1
        Iw = t["inv_denom"] * (o["Qw"] @ s1q) + g.Ra_sigma * thw
        Rw = w + dt * (de * E1[1] + (1 - de) * E2[1]) + om * Iw - om * t["diss_w"] * s1w
        Ith = s1w if g.w_bc_top == "dirichlet" else o["Mwt"] @ s1w
        Rth = th + dt * (de * E1[2] + (1 - de) * E2[2]) + om * Ith - om * t["diss_th"] * th1
        Rthb = thb.copy()
        if em:
            Rth = Rth + om * (th1 - s1th) / al
            Rthb = thb + om * (thb1 - s1thb) / al
        q2, w2, th2 = self.implicit_solve(Rq, Rw, Rth)
        thb2 = o["proj_d"] @ Rthb
        if em:
            th2, thb2 = self.corrector(w2, th2, thb2, al)
        return self.sanitize(q2), self.sanitize(w2), self.sanitize(th2), thb2

    def run_full(self, state, n=1):
        g = self.g
        q, w, th = (self.to_compact(f) for f in state[:3])
        thb = np.array(state[3], dtype=float)
        for _ in range(n):
            q, w, th, thb = self.step(q, w, th, thb)
        return (self.to_full(q, g.Nz + 1), self.to_full(w, g.Nz - 1), self.to_full(th, g.Nz - 1),
                thb)
