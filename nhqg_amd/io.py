"""Checkpoints and the diagnostics-cadence driver loop.

The checkpoint format is the reference's (nhqg/io.py:107-158): one ``.npz``
per step named ``checkpoint_<step:08d>.npz`` with real/imag parts of the
three spectral fields, ``th_bar``, ``step`` and ``t``; loading applies the
Hermitian projection of the self-conjugate columns (historical files carry
an anti-Hermitian ghost there).  Files are read with numpy's default
``allow_pickle=False``.

``run_loop`` follows the reference driver (scripts/run_polar.py:332-378):
advance in chunks of ``diag_steps`` on the device, compute the scalar
diagnostics after each chunk, append a CSV row, and on any non-finite value
save a checkpoint and stop with status 2.  Like the reference driver
(scripts/run_polar.py:233-268) it first records the invocation and the
resolved config in ``run_config.json`` (``run_config.<n>.json`` when an
earlier leg's record exists); ``read_run_config`` rebuilds the config for a
restart.
"""

from __future__ import annotations

import csv
import dataclasses
import datetime
import json
import math
import os
import platform
import shlex
import socket
import sys
from pathlib import Path

import numpy as np

from nhqg_amd.solver import State, Stepper

SCALAR_KEYS = ("KE_tot", "KE_bt", "KE_bc", "max_speed", "enstrophy", "Nusselt",
               "Nusselt_dealiased", "vol_avg_tw_dealiased", "th_bar_max", "q_rms", "w_rms",
               "th_rms")


def save_checkpoint(state: State, step: int, cfg, output_dir: str | os.PathLike) -> str:
    """Reference checkpoint layout (nhqg/io.py:107-125)."""
    Path(output_dir).mkdir(parents=True, exist_ok=True)
    fname = os.path.join(str(output_dir), f"checkpoint_{step:08d}.npz")
    q, w, th, thb = (np.asarray(x) for x in state)
    np.savez(fname, q_hat_real=q.real, q_hat_imag=q.imag, w_hat_real=w.real,
             w_hat_imag=w.imag, th_hat_real=th.real, th_hat_imag=th.imag,
             th_bar=np.asarray(thb, dtype=np.float64), step=step, t=step * cfg.dt)
    return fname


def _hermitian_project(f: np.ndarray) -> np.ndarray:
    Nx = f.shape[1]
    neg = (-np.arange(Nx)) % Nx
    f = np.array(f, dtype=np.complex128)
    f[:, :, 0] = 0.5 * (f[:, :, 0] + np.conj(f[:, neg, 0]))
    if Nx % 2 == 0:
        f[:, :, -1] = 0.5 * (f[:, :, -1] + np.conj(f[:, neg, -1]))
    return f


def load_checkpoint(path: str | os.PathLike) -> tuple[State, int, float]:
    """(state, step, t) from a reference-format checkpoint
    (nhqg/io.py:128-158)."""
    with np.load(path) as d:
        q = _hermitian_project(d["q_hat_real"] + 1j * d["q_hat_imag"])
        w = _hermitian_project(d["w_hat_real"] + 1j * d["w_hat_imag"])
        th = _hermitian_project(d["th_hat_real"] + 1j * d["th_hat_imag"])
        thb = (np.array(d["th_bar"], dtype=np.float64) if "th_bar" in d.files
               else np.zeros(q.shape[0]))
        return State(q, w, th, thb), int(d["step"]), float(d["t"])


def write_run_config(output_dir: str | os.PathLike, cfg, args: dict | None = None) -> Path:
    """Record argv, the resolved config and the host beside the outputs,
    never overwriting an earlier leg's record (the reference's
    run_config.json layout, scripts/run_polar.py:233-268)."""
    out = Path(output_dir)
    out.mkdir(parents=True, exist_ok=True)

    def plain(v):
        return v if isinstance(v, (int, float, str, bool, type(None))) else str(v)

    record = {
        "argv": sys.argv,
        "command": " ".join(shlex.quote(a) for a in [sys.executable, *sys.argv]),
        "args": {k: plain(v) for k, v in (args or {}).items()},
        "config": {f.name: plain(getattr(cfg, f.name)) for f in dataclasses.fields(cfg)},
        "host": socket.gethostname(),
        "python": platform.python_version(),
        "backend": "nhqg_amd (MI355X, gfx950)",
        "started_utc": datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"),
    }
    path = out / "run_config.json"
    if path.exists():
        n = 1
        while (out / f"run_config.{n}.json").exists():
            n += 1
        path = out / f"run_config.{n}.json"
    path.write_text(json.dumps(record, indent=2, default=str) + "\n")
    return path


def read_run_config(path: str | os.PathLike):
    """NHQGConfig from a run_config.json record (fields the config does not
    know are ignored)."""
    from nhqg_amd.config import NHQGConfig

    with open(path) as f:
        rec = json.load(f)
    names = {f.name for f in dataclasses.fields(NHQGConfig)}
    return NHQGConfig(**{k: v for k, v in rec["config"].items() if k in names})


def diagnostics_finite(diag: dict) -> bool:
    return all(math.isfinite(float(diag[k])) for k in SCALAR_KEYS if k in diag)


def _write_row(path: Path, step: int, t: float, diag: dict) -> None:
    keys = sorted(k for k in SCALAR_KEYS if k in diag)
    new = not path.exists()
    with path.open("a", newline="") as f:
        w = csv.writer(f)
        if new:
            w.writerow(["step", "t", *keys])
        w.writerow([step, t, *(float(diag[k]) for k in keys)])


def run_loop(grid, state: State, n_total: int, diag_steps: int, output_dir: str | os.PathLike,
             step0: int = 0, t0: float = 0.0, checkpoint_steps: int = 0):
    """Driver loop with diagnostics cadence and non-finite abort.

    Returns (status, final_state, last_diag): status 0 when all n_total
    steps ran, 2 when the diagnostics turned non-finite (a checkpoint of
    that state is written first, as the reference driver does)."""
    out = Path(output_dir)
    out.mkdir(parents=True, exist_ok=True)
    write_run_config(out, grid.cfg, {"n_total": n_total, "diag_steps": diag_steps,
                                     "step0": step0, "t0": t0,
                                     "checkpoint_steps": checkpoint_steps})
    csv_path = out / "diagnostics.csv"
    sp = Stepper(grid)
    sp.set_state(state)
    step, last = step0, None
    while step < step0 + n_total:
        n = min(diag_steps, step0 + n_total - step)
        sp.run(n)
        step += n
        t = t0 + (step - step0) * float(grid.dt)
        last = sp.diagnostics()
        _write_row(csv_path, step, t, last)
        if not diagnostics_finite(last):
            save_checkpoint(sp.state(), step, grid.cfg, out)
            return 2, sp.state(), last
        if checkpoint_steps > 0 and step % checkpoint_steps == 0:
            save_checkpoint(sp.state(), step, grid.cfg, out)
    final = sp.state()
    save_checkpoint(final, step, grid.cfg, out)
    return 0, final, last
