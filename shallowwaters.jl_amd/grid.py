"""``Grid{T,Tprog}``: derived sizes, halos, time steps and Coriolis fields.

Mirrors ``src/grid.jl:1-159``.  Every quantity is computed in float64 in the same
operation order as the reference and then rounded to ``T``/``Tprog`` where the reference
declares that type.  Arrays follow the reference's column-major layout: a Julia
``Array{T,2}`` of size ``(m, n)`` is a numpy array of shape ``(m, n)`` in Fortran order,
so element ``[i-1, j-1]`` here is element ``[i, j]`` there.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .numtypes import round_scalar, to_storage


_EXACT = {0.0: (0.0, 1.0), 90.0: (1.0, 0.0), 180.0: (0.0, -1.0), 270.0: (-1.0, 0.0)}


def sind(x: float) -> float:
    """Sine of ``x`` degrees: exact at multiples of 90 degrees as Julia's ``sind``, else
    reduced in extended precision (``sin(pi/180*x)`` in float64 is not correctly rounded
    at 30 or 45 degrees; Julia's ``sind`` is)."""
    r = math.fmod(x, 360.0) % 360.0
    if r in _EXACT:
        s = _EXACT[r][0]
        return math.copysign(0.0, x) if s == 0.0 else s
    return float(np.sin(np.longdouble(r) * _PI_L / np.longdouble(180)))


def cosd(x: float) -> float:
    r = math.fmod(x, 360.0) % 360.0
    if r in _EXACT:
        return _EXACT[r][1]
    return float(np.cos(np.longdouble(r) * _PI_L / np.longdouble(180)))


# pi in x87 extended precision (64-bit mantissa), from its decimal expansion
_PI_L = np.longdouble("3.14159265358979323846264338327950288")


def coriolis_at_lat(ω: float, ϕ: float) -> float:
    """``src/grid.jl:130-132``"""
    return 2 * ω * sind(ϕ)


def β_at_lat(ω: float, R: float, ϕ: float) -> float:
    """``src/grid.jl:136-138``"""
    return 2 * ω / R * cosd(ϕ)


def meshgrid(x, y):
    """``src/grid.jl:143-159``: ``xx[i,j]=x[i]``, ``yy[i,j]=y[j]`` of shape (len(x), len(y))."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xx = np.asfortranarray(np.repeat(x[:, None], len(y), axis=1))
    yy = np.asfortranarray(np.repeat(y[None, :], len(x), axis=0))
    return xx, yy


def _rng(a: int, b: int) -> np.ndarray:
    """Julia's ``Array(a:b)`` as float64 (exact integers)."""
    return np.arange(a, b + 1, dtype=np.float64)


@dataclass
class Grid:
    T: str
    Tprog: str
    nx: int
    Lx: float
    L_ratio: float
    bc: str
    g: float
    H: float
    cfl: float
    Ndays: float
    nstep_diff: int
    nstep_advcor: int
    Uadv: float
    output_dt: float
    ω: float
    ϕ: float
    R: float

    def __post_init__(self):
        T, Tprog = self.T, self.Tprog
        nx, Lx, bc = self.nx, float(self.Lx), self.bc
        periodic = bc == "periodic"

        # DOMAIN SIZES (grid.jl:21-23)
        self.Δ = Δ = Lx / nx
        self.ny = ny = int(round(Lx / self.L_ratio / Δ))
        self.Ly = Ly = ny * Δ

        # NUMBER OF GRID POINTS (grid.jl:26-37)
        self.nux = nx if periodic else nx - 1
        self.nuy = ny
        self.nvx = nx
        self.nvy = ny - 1
        self.nqx = nx if periodic else nx + 1
        self.nqy = ny + 1
        self.nT = nx * ny
        self.nu = self.nux * self.nuy
        self.nv = self.nvx * self.nvy
        self.nq = self.nqx * self.nqy

        # GRID VECTORS (grid.jl:40-47)
        self.x_T = Δ * _rng(1, nx) - Δ / 2
        self.y_T = Δ * _rng(1, ny) - Δ / 2
        self.x_u = Δ * _rng(0, nx - 1) if periodic else Δ * _rng(1, nx - 1)
        self.y_u = self.y_T
        self.x_v = self.x_T
        self.y_v = Δ * _rng(1, ny - 1)
        self.x_q = self.x_u if periodic else Δ * _rng(1, nx + 1) - Δ
        self.y_q = Δ * _rng(1, ny + 1) - Δ

        # HALO SIZES (grid.jl:50-53)
        self.halo = 2
        self.haloη = 1
        self.halosstx = 1
        self.halossty = 0

        # EDGE POINT (grid.jl:56)
        self.ep = 1 if periodic else 0

        # GRID VECTORS WITH HALO (grid.jl:59-66)
        self.x_T_halo = Δ * _rng(0, nx + 1) - Δ / 2
        self.y_T_halo = Δ * _rng(0, ny + 1) - Δ / 2
        self.x_u_halo = Δ * _rng(-2, nx + 1) if periodic else Δ * _rng(-1, nx + 1)
        self.y_u_halo = Δ * _rng(-1, ny + 2) - Δ / 2
        self.x_v_halo = Δ * _rng(-1, nx + 2) - Δ / 2
        self.y_v_halo = Δ * _rng(-1, ny + 1)
        self.x_q_halo = self.x_u_halo if periodic else Δ * _rng(-1, nx + 3) - Δ
        self.y_q_halo = Δ * _rng(-1, ny + 3) - Δ

        # TIME STEPS (grid.jl:69-74)
        self.c = math.sqrt(self.g * self.H)
        self.dtint = dtint = int(math.floor(self.cfl * Δ / self.c))
        if dtint < 1:
            raise ValueError(f"time step floor(cfl*Δ/c) = {self.cfl * Δ / self.c} s rounds to 0")
        self.nt = int(math.ceil(self.Ndays * 3600 * 24 / dtint))
        self.dt = round_scalar(float(dtint), T)
        self.Δt = round_scalar(dtint / Δ, T)
        self.Δt_diff = round_scalar(self.nstep_diff * dtint / Δ, Tprog)

        # TIME STEPS FOR ADVECTION (grid.jl:77-83)
        self.nadvstep = nadvstep = max(1, int(math.floor(Δ / self.Uadv / dtint)))
        self.nadvstep_half = nadvstep // 2
        self.dtadvint = dtadvint = nadvstep * dtint
        self.dtadvu = round_scalar(dtadvint * nx / Lx, T)
        self.dtadvv = round_scalar(dtadvint * ny / Ly, T)
        self.half_dtadvu = round_scalar(dtadvint * nx / Lx / 2, T)
        self.half_dtadvv = round_scalar(dtadvint * ny / Ly / 2, T)

        # N TIME STEPS FOR OUTPUT (grid.jl:86-88)
        self.nout = int(math.floor(self.output_dt * 3600 / dtint))
        self.nout_total = (self.nt // self.nout) + 1 if self.nout > 0 else 1
        self.t_vec = np.arange(self.nout_total, dtype=np.int64) * dtint

        # CORIOLIS (grid.jl:91-95)
        self.f0 = f0 = coriolis_at_lat(self.ω, self.ϕ)
        self.β = β = β_at_lat(self.ω, self.R, self.ϕ)
        _, yy_u = meshgrid(self.x_u, self.y_u)
        _, yy_v = meshgrid(self.x_v, self.y_v)
        self.f_u = to_storage(Δ * (f0 + β * (yy_u - Ly / 2)), T)
        self.f_v = to_storage(Δ * (f0 + β * (yy_v - Ly / 2)), T)
        self.f_q = to_storage(Δ * (f0 + β * (yy_q(bc, self.x_q_halo, self.y_q_halo) - Ly / 2)), T)


def yy_q(bc: str, x_q_halo, y_q_halo):
    """``src/grid.jl:99-108`` (1-based ``[3:end-1]`` is python ``[2:-1]``)."""
    if bc == "periodic":
        _, yy = meshgrid(x_q_halo[2:-1], y_q_halo[2:-2])
    else:
        _, yy = meshgrid(x_q_halo[2:-2], y_q_halo[2:-2])
    return yy


def make_grid(P) -> Grid:
    """``Grid{T,Tprog}(P)``, ``src/grid.jl:112-122``."""
    return Grid(T=P.T, Tprog=P.Tprog, nx=P.nx, Lx=P.Lx, L_ratio=P.L_ratio, bc=P.bc, g=P.g, H=P.H,
                cfl=P.cfl, Ndays=P.Ndays, nstep_diff=P.nstep_diff, nstep_advcor=P.nstep_advcor,
                Uadv=P.Uadv, output_dt=P.output_dt, ω=P.ω, ϕ=P.ϕ, R=P.R)
