"""``Forcing{T}``: static wind, topography and continuity-forcing fields.

Mirrors ``src/forcing.jl:1-217``.  Fields are evaluated in float64 with numpy's libm in
the reference's operation order, then rounded to ``T`` (host storage, Julia layout).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .grid import meshgrid
from .numtypes import round_scalar, to_storage, storage_dtype

π = math.pi


@dataclass
class Forcing:
    Fx: np.ndarray
    Fy: np.ndarray
    H: np.ndarray
    η_ref: np.ndarray
    Fη: np.ndarray


def ChannelWind(T, P, G):
    """``src/forcing.jl:57-68``"""
    Δ, Lx, Ly = G.Δ, G.Lx, G.Ly
    xx_u, yy_u = meshgrid(G.x_u, G.y_u)
    Fx = (Δ * P.Fx0 / P.ρ / P.H) * np.cos(π * (yy_u / Ly - 1 / 2)) ** 2
    Fy = (Δ * P.Fy0 / P.ρ / P.H) * np.cos(π * (xx_u / Lx - 1 / 2)) ** 2
    return to_storage(Fx, T), to_storage(Fy, T)


def ShearWind(T, P, G):
    """``src/forcing.jl:72-83``"""
    Δ, Lx, Ly = G.Δ, G.Lx, G.Ly
    xx_u, yy_u = meshgrid(G.x_u, G.y_u)
    Fx = (Δ * P.Fx0 / P.ρ / P.H) * np.tanh(2 * π * (yy_u / Ly - 1 / 2))
    Fy = (Δ * P.Fy0 / P.ρ / P.H) * np.tanh(2 * π * (xx_u / Lx - 1 / 2))
    return to_storage(Fx, T), to_storage(Fy, T)


def DoubleGyreWind(T, P, G):
    """``src/forcing.jl:88-98``"""
    Δ, Lx, Ly = G.Δ, G.Lx, G.Ly
    xx_u, yy_u = meshgrid(G.x_u, G.y_u)
    Fx = (Δ * P.Fx0 / P.ρ / P.H) * (np.cos(2 * π * (yy_u / Ly - 1 / 2)) + 2 * np.sin(π * (yy_u / Ly - 1 / 2)))
    Fy = (Δ * P.Fy0 / P.ρ / P.H) * (np.cos(2 * π * (xx_u / Lx - 1 / 2)) + 2 * np.sin(π * (xx_u / Lx - 1 / 2)))
    return to_storage(Fx, T), to_storage(Fy, T)


def ConstantWind(T, P, G):
    """``src/forcing.jl:101-111``: ``T(Δ*F0/ρ/H)`` times ``ones(T)``, exact in ``T``."""
    Δ = G.Δ
    fx = round_scalar(Δ * P.Fx0 / P.ρ / P.H, T)
    fy = round_scalar(Δ * P.Fy0 / P.ρ / P.H, T)
    Fx = to_storage(np.full((G.nux, G.nuy), fx), T)
    Fy = to_storage(np.full((G.nvx, G.nvy), fy), T)
    return Fx, Fy


def NoWind(T, P, G):
    """``src/forcing.jl:114-123``"""
    dt = storage_dtype(T)
    return np.zeros((G.nux, G.nuy), dtype=dt, order="F"), np.zeros((G.nvx, G.nvy), dtype=dt, order="F")


def InterfaceRelaxation(T, P, G):
    """``src/forcing.jl:127-136``"""
    Ly = G.Ly
    _, yy_T = meshgrid(G.x_T, G.y_T)
    η_ref = -(P.η_refh / 2) * np.tanh(2 * π * (Ly / (4 * P.η_refw)) * (yy_T / Ly - 1 / 2))
    return to_storage(η_ref, T)


def Seamount(T, P, G):
    """``src/forcing.jl:141-151``"""
    Lx, Ly, w = G.Lx, G.Ly, P.topo_width
    xx_T, yy_T = meshgrid(G.x_T_halo, G.y_T_halo)
    bumpx = np.exp(-((xx_T - Lx / 2) ** 2) / (2 * w ** 2))
    bumpy = np.exp(-((yy_T - Ly / 2) ** 2) / (2 * w ** 2))
    return to_storage(P.H - P.topo_height * bumpx * bumpy, T)


def Ridge(T, P, G):
    """``src/forcing.jl:156-168``"""
    Lx, Ly, w = G.Lx, G.Ly, P.topo_width
    xx_T, yy_T = meshgrid(G.x_T_halo, G.y_T_halo)
    bumpx = np.exp(-((xx_T - Lx / 2) ** 2) / (2 * w ** 2))
    bumpy = np.exp(-((yy_T - Ly / 2) ** 2) / (2 * w ** 2))
    Hx = P.H - P.topo_height * bumpx
    Hy = P.H - P.topo_height * bumpy
    return to_storage(Hx, T), to_storage(Hy, T)


def Ridges(T, P, G):
    """``src/forcing.jl:171-188``"""
    Lx, w = G.Lx, P.topo_width
    xx_T, _ = meshgrid(G.x_T_halo, G.y_T_halo)
    b0x = np.exp(-(xx_T ** 2) / (2 * w ** 2))
    b1x = np.exp(-((xx_T - 0.99 * Lx / 4) ** 2) / (2 * w ** 2))
    b2x = np.exp(-((xx_T - 1.01 * Lx / 2) ** 2) / (2 * w ** 2))
    b3x = np.exp(-((xx_T - 0.99 * 3 * Lx / 4) ** 2) / (2 * w ** 2))
    b4x = np.exp(-((xx_T - Lx) ** 2) / (2 * w ** 2))
    th = P.topo_height
    return to_storage(P.H - th * b0x - th * b1x - th * b2x - th * b3x - th * b4x, T)


def FlatBottom(T, P, G):
    """``src/forcing.jl:191-195``"""
    return to_storage(np.full((G.nx + 2 * G.haloη, G.ny + 2 * G.haloη), round_scalar(P.H, T)), T)


def KelvinPump(T, P, G):
    """``src/forcing.jl:198-212``"""
    Ly, Δ = G.Ly, G.Δ
    _, yy_T = meshgrid(G.x_T, G.y_T)
    mϕ = 2 * π * G.R / 360.
    y0 = Ly / 2 - (G.ϕ - P.ϕk) * mϕ
    Fη = P.A * Δ * np.exp(-(yy_T - y0) ** 2 / (2 * P.wk ** 2))
    return to_storage(Fη, T)


def Ftime(T, t: int, ω: float) -> float:
    """``src/forcing.jl:215-217``: ``T(sin(ω*t))``."""
    return round_scalar(math.sin(ω * t), T)


def make_forcing(P, G) -> Forcing:
    """``Forcing{T}(P,G)``, ``src/forcing.jl:11-53``."""
    T = P.T
    winds = {"channel": ChannelWind, "shear": ShearWind, "double_gyre": DoubleGyreWind,
             "constant": ConstantWind, "none": NoWind}
    Fx, _ = winds[P.wind_forcing_x](T, P, G)
    _, Fy = winds[P.wind_forcing_y](T, P, G)
    if P.topography == "ridge":
        H, _ = Ridge(T, P, G)
    elif P.topography == "ridges":
        H = Ridges(T, P, G)
    elif P.topography == "seamount":
        H = Seamount(T, P, G)
    else:
        H = FlatBottom(T, P, G)
    return Forcing(Fx=Fx, Fy=Fy, H=H, η_ref=InterfaceRelaxation(T, P, G), Fη=KelvinPump(T, P, G))
