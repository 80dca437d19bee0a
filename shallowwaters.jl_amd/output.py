"""Run folders, NetCDF output and restarts, and the progress feedback of a run.

Mirrors ``src/output.jl`` (``NcFiles``, ``nc_create``, ``output_nc!``, ``gap``,
``get_run_id_path``), ``src/feedback.jl`` (``Feedback``, ``readable_secs``,
``duration_estimate``, ``feedback_init``, ``feedback!``, ``feedback_end!``, ``progress!``)
and the ``initial_cond="ncfile"`` branch of ``src/initial_conditions.jl:39-120,145-151``.

The files are NetCDF *classic* (64-bit offset) written with ``scipy.io``: the reference writes
NETCDF4 through NetCDF.jl, whose reader opens both formats, so a run written here restarts a
Julia run and the other way round needs only the classic format (NETCDF4 files are HDF5, which
``scipy`` cannot read; reading one raises an error that says so).  Each file holds one
variable, Float32 with dimensions (x, y, t) in the reference's column-major order, stored
(t, y, x) in the file's row-major order as NetCDF.jl does; x and y carry the grid's
coordinate vectors, t (Int32, unlimited) the model time in seconds.
"""
from __future__ import annotations

import datetime
import math
import os
import sys
import time

import numpy as np

# (output_vars entry, file and variable name, grid vectors, units, long name)  output.jl:22-30
_VARS = {
    "u": ("u", "u", "m/s", "zonal velocity"),
    "v": ("v", "v", "m/s", "meridional velocity"),
    "η": ("eta", "T", "m", "sea surface height"),
    "sst": ("sst", "T", "1", "sea surface temperature"),
    "du": ("du", "u", "m^2/s^2", "zonal velocity tendency"),
    "dv": ("dv", "v", "m^2/s^2", "meridional velocity tendency"),
    "dη": ("deta", "T", "m^2/s", "sea surface height tendency"),
}
# q and ζ are intermediates of the last Runge-Kutta stage (Diag.Vorticity), which the GPU
# step keeps in registers: they are not available as output
_NOT_AVAILABLE = {"q": "potential vorticity", "ζ": "relative vorticity"}


def readable_secs(secs: float) -> str:
    """``feedback.jl:16-36``."""
    days = int(math.floor(secs / 3600 / 24))
    hours = int(math.floor((secs / 3600) % 24))
    minutes = int(math.floor((secs / 60) % 60))
    seconds = int(math.floor(secs % 3600 % 60))
    if days > 0:
        return f"{days}d, {hours}h"
    if hours > 0:
        return f"{hours}h, {minutes}min"
    if minutes > 0:
        return f"{minutes}min, {seconds}s"
    if seconds > 10:
        return f"{secs % 3600 % 60:.1f}s"
    return f"{secs % 3600 % 60:.2f}s"


def gap(a) -> int:
    """The smallest run id missing between the smallest and largest existing one, else the
    largest plus one (``output.jl:159-165``)."""
    missing = [i for i in range(min(a), max(a) + 1) if i not in a]
    return min(missing) if missing else max(a) + 1


def get_run_id_path(S):
    """``output.jl:168-207``: create the run folder ``run%04d`` under ``outpath`` and return
    (run_id, runpath)."""
    P = S.parameters
    if not P.output:
        return 0, "no runpath"
    existing = sorted(int(d[3:]) for d in os.listdir(P.outpath) if d.startswith("run") and d[3:].isdigit())
    if not existing:
        runpath = os.path.join(P.outpath, "run0000")
        os.mkdir(runpath)
        return 0, runpath
    if P.get_id_mode == "fill":
        run_id = gap(existing)
    elif P.get_id_mode == "specific":
        run_id = P.run_id
    elif P.get_id_mode == "continue":
        run_id = max(existing) + 1
    else:
        raise ValueError(f"Order '{P.get_id_mode}' is not valid for get_run_id_path(), choose continue or fill.")
    runpath = os.path.join(P.outpath, f"run{run_id:04d}")
    if P.get_id_mode == "specific" and os.path.isdir(runpath):
        import shutil
        shutil.rmtree(runpath)                      # the reference removes and recreates it
    os.mkdir(runpath)
    return run_id, runpath


class Feedback:
    """``feedback.jl:1-12``: timing, NaN state, progress file and run folder of a run."""

    def __init__(self, progress_txt, output, nt, nout, run_id, runpath):
        self.t0 = self.t1 = time.time()
        self.nans_detected = False
        self.progress_txt = progress_txt
        self.output = output
        self.i = 0
        self.nt, self.nout = nt, nout
        self.run_id, self.runpath = run_id, runpath

    def write(self, s: str):
        if self.output and self.progress_txt is not None:
            self.progress_txt.write(s)
            self.progress_txt.flush()


def _now() -> str:
    return datetime.datetime.now(datetime.timezone.utc).strftime("%a, %d %b %Y %H:%M:%S")


def feedback_init(S) -> Feedback:
    """``feedback.jl:75-117``: with output, create the run folder, progress.txt and
    parameter.txt."""
    P, G = S.parameters, S.grid
    if P.output:
        run_id, runpath = get_run_id_path(S)
        txt = open(os.path.join(runpath, "progress.txt"), "w")
        s = f"Starting ShallowWaters run {run_id} on {_now()}"
        print(s)
        txt.write(s + "\n")
        txt.write(f"Model will integrate {P.Ndays}days at a resolution of {G.nx}x{G.ny} with Δ={G.Δ / 1e3}km\n")
        txt.write("Initial conditions are ")
        if P.initial_cond == "rest":
            txt.write("rest.\n")
        elif P.init_starti > 0:
            txt.write(f"time step {P.init_starti} of run {P.init_run_id}.\n")
        else:
            txt.write(f"last time step of run {P.init_run_id}.\n")
        txt.write(f"Boundary conditions are {P.bc} with lbc={P.α}.\n")
        txt.write(f"Number format is {P.T}.\n")
        txt.write(f"\nAll data will be stored in {runpath}\n")
        with open(os.path.join(runpath, "parameter.txt"), "w") as ptxt:
            ptxt.write(repr(P))
    else:
        print(f"Starting ShallowWaters on {_now()} without output.")
        txt, run_id, runpath = None, -1, ""
    return Feedback(txt, P.output, G.nt, G.nout, run_id, runpath)


def duration_estimate(fb: Feedback, S):
    """``feedback.jl:38-59``."""
    nadvstep = S.grid.nadvstep
    time_per_step = (time.time() - fb.t1) / max(fb.i - nadvstep, 1)
    time_total = int(round(time_per_step * fb.nt))
    time_to_go = int(round(time_per_step * (fb.nt - fb.i)))
    done = datetime.datetime.now(datetime.timezone.utc) + datetime.timedelta(seconds=time_to_go)
    s1 = f"Model integration will take approximately {readable_secs(time_total)},"
    s2 = f"and is hopefully done on {done.strftime('%a, %d %b %Y %H:%M:%S')}"
    print("\r\x1b[K", end="")
    print(s1)
    print(s2)
    fb.write("\n" + s1 + "\n" + s2 + "\n")


def progress(fb: Feedback):
    """``feedback.jl:168-181``: print the percentage every 1 %, write it every 5 %."""
    i, nt = fb.i, fb.nt
    if ((i + 1) / nt * 100 % 1) < (i / nt * 100 % 1):
        percent = int(round((i + 1) / nt * 100))
        print(f"\r\x1b[K{percent}%", end="")
        sys.stdout.flush()
        if percent % 5 == 0:
            fb.write(f"\n{percent}%")


def feedback(fb: Feedback, S, nan_step: int = 0):
    """``feedback.jl:120-153`` at step ``fb.i``: the duration estimate after the tracer ran
    twice, the NaN report (the step loop itself checks every ``nout`` steps) and the
    progress."""
    nadvstep = S.grid.nadvstep
    if fb.output:
        if fb.i == nadvstep:
            fb.t1 = time.time()
        elif fb.i == 2 * nadvstep:
            duration_estimate(fb, S)
    if nan_step and not fb.nans_detected:
        fb.nans_detected = True
        print(f" NaNs detected at time step {nan_step}")
        fb.write(f" NaNs detected at time step {nan_step}")
    if fb.i > 100:
        progress(fb)


def feedback_end(fb: Feedback):
    """``feedback.jl:156-165``."""
    s = f" Integration done in {readable_secs(time.time() - fb.t0)}."
    print(s)
    fb.write("\n" + s[1:] + "\n")
    if fb.progress_txt is not None:
        fb.progress_txt.close()
        fb.progress_txt = None


def nc_create(x, y, name: str, path: str, unit: str, long_name: str):
    """``output.jl:57-77`` in the classic format: one Float32 variable (x, y, t) plus the
    coordinate variables and an Int32 time."""
    from scipy.io import netcdf_file
    nc = netcdf_file(os.path.join(path, name + ".nc"), "w", version=2)
    nc.createDimension("t", None)                   # the classic format defines the record dimension first
    nc.createDimension("x", len(x))
    nc.createDimension("y", len(y))
    vx = nc.createVariable("x", "f8", ("x",))
    vx[:] = np.asarray(x, dtype=np.float64)
    vy = nc.createVariable("y", "f8", ("y",))
    vy[:] = np.asarray(y, dtype=np.float64)
    vt = nc.createVariable("t", "i4", ("t",))
    var = nc.createVariable(name, "f4", ("t", "y", "x"))
    var.units, var.long_name, var.missing_value = unit, long_name, np.float32(-999999.0)
    vt.units, vt.long_name = "s", "time"
    vx.units, vx.long_name = "m", "zonal coordinate"
    vy.units, vy.long_name = "m", "meridional coordinate"
    return nc


class NcFiles:
    """``output.jl:1-55``: the NetCDF files of the requested ``output_vars``."""

    def __init__(self, fb: Feedback, S):
        P, G = S.parameters, S.grid
        self.files = {}
        self.iout = 0
        if not P.output:
            return
        for v in P.output_vars:
            if v in _NOT_AVAILABLE:
                raise NotImplementedError(f"output of {v} ({_NOT_AVAILABLE[v]}): it is an intermediate of the last "
                                          "Runge-Kutta stage, which the GPU step does not store")
            if v not in _VARS:
                raise ValueError(f"unknown output variable {v!r}")
            name, grid, unit, long_name = _VARS[v]
            x = {"u": G.x_u, "v": G.x_v, "T": G.x_T}[grid]
            y = {"u": G.y_u, "v": G.y_v, "T": G.y_T}[grid]
            self.files[v] = nc_create(x, y, name, fb.runpath, unit, long_name)

    @property
    def wants_tendencies(self) -> bool:
        return any(v in self.files for v in ("du", "dv", "dη"))

    def output(self, i: int, new, old, S):
        """``output_nc!`` (``output.jl:79-144``) at step ``i``: ``new`` are u, v, η, sst after
        the step (the reference's u0, v0, η0 and Prog.sst), ``old`` u, v, η before it (Prog
        before the copy-back), both without halos; the tendency is their difference in
        Float32."""
        if not self.files:
            return
        self.iout += 1
        k = self.iout - 1
        f32 = lambda a: np.asarray(a, dtype=np.float32)
        vals = {"u": f32(new[0]), "v": f32(new[1]), "η": f32(new[2]), "sst": f32(new[3])}
        if self.wants_tendencies:
            vals.update({"du": vals["u"] - f32(old[0]), "dv": vals["v"] - f32(old[1]), "dη": vals["η"] - f32(old[2])})
        for v, nc in self.files.items():
            name = _VARS[v][0]
            nc.variables[name][k] = vals[v].T
            nc.variables["t"][k] = np.int32(i * S.grid.dtint)
            nc.flush()

    def close(self):
        for nc in self.files.values():
            nc.close()
        self.files = {}


# ---------------------------------------------------------------- restart

def _read_var(path: str, name: str, starti: int):
    from scipy.io import netcdf_file
    with open(path, "rb") as fh:
        magic = fh.read(4)
    if magic[:3] != b"CDF":
        raise ValueError(f"{path} is not a NetCDF classic file (NETCDF4 files are HDF5 and need a NetCDF4 reader)")
    with netcdf_file(path, "r", mmap=False) as nc:
        var = nc.variables[name]
        nt = var.shape[0]
        k = nt if starti == -1 else starti                     # -1: the last record
        if not 1 <= k <= nt:
            raise IndexError(f"{path}: time index {k} outside 1..{nt}")
        return np.array(var[k - 1], dtype=np.float64).T, k     # (x, y) in the reference's order


def _linear_flat(a: np.ndarray, xo: np.ndarray, yo: np.ndarray, xn: np.ndarray, yn: np.ndarray) -> np.ndarray:
    """Bilinear interpolation of ``a`` given at the grid (xo, yo) onto (xn, yn), constant
    beyond the edges (Interpolations.jl ``Gridded(Linear())`` with ``Flat()`` extrapolation)."""
    def weights(old, new):
        newc = np.clip(new, old[0], old[-1])
        j = np.clip(np.searchsorted(old, newc, side="right") - 1, 0, max(len(old) - 2, 0))
        if len(old) == 1:
            return np.zeros_like(j), np.zeros_like(j), np.zeros_like(newc)
        w = (newc - old[j]) / (old[j + 1] - old[j])
        return j, j + 1, w
    if a.shape != (len(xo), len(yo)):
        raise ValueError(f"restart field of shape {a.shape} on a grid of {len(xo)}x{len(yo)} points: the "
                         "interpolation places u on periodic points, as the reference's does")
    i0, i1, wx = weights(xo, xn)
    j0, j1, wy = weights(yo, yn)
    ax0 = a[i0, :] * (1 - wx)[:, None] + a[i1, :] * wx[:, None]
    return ax0[:, j0] * (1 - wy)[None, :] + ax0[:, j1] * wy[None, :]


def read_initial_conditions(S):
    """``initial_conditions.jl:39-120,145-151``: u, v, η (and sst when ``sst_initial ==
    "restart"``) at time index ``init_starti`` of run ``init_run_id`` under ``initpath``, in
    float64 without halos; interpolated (bilinear, flat beyond the edges) when the grid size
    differs and ``init_interpolation`` is set."""
    P, G = S.parameters, S.grid
    runpath = os.path.join(P.initpath, f"run{P.init_run_id:04d}")
    u, starti = _read_var(os.path.join(runpath, "u.nc"), "u", P.init_starti)
    v, _ = _read_var(os.path.join(runpath, "v.nc"), "v", starti)
    η, _ = _read_var(os.path.join(runpath, "eta.nc"), "eta", starti)
    nx_old, ny_old = η.shape
    if (nx_old, ny_old) != (G.nx, G.ny):
        if not P.init_interpolation:
            raise ValueError(f"Grid size {(G.nx, G.ny)} doesn't match initial conditions on a {η.shape} grid.")
        x_T = np.arange(0.5, nx_old, 1.0)
        y_T = np.arange(0.5, ny_old, 1.0)
        x_u, y_u = np.arange(0.0, nx_old, 1.0), y_T                 # as the reference: periodic u points
        x_v, y_v = x_T, np.arange(1.0, ny_old, 1.0)
        dx, dy = nx_old / G.nx, ny_old / G.ny
        x_T_new = dx / 2 + dx * np.arange(G.nx)                    # Δx/2:Δx:nx_old-Δx/2
        y_T_new = dy / 2 + dy * np.arange(G.ny)
        x_u_new, y_u_new = dx * np.arange(G.nx), y_T_new            # 0:Δx:nx_old-Δx
        x_v_new, y_v_new = x_T_new, dy * np.arange(1, G.ny)         # Δy:Δy:ny_old-Δy
        u = _linear_flat(u, x_u, y_u, x_u_new, y_u_new)
        v = _linear_flat(v, x_v, y_v, x_v_new, y_v_new)
        η = _linear_flat(η, x_T, y_T, x_T_new, y_T_new)
    sst = None
    if P.sst_initial == "restart":
        sst, _ = _read_var(os.path.join(runpath, "sst.nc"), "sst", starti)
    return u, v, η, sst
