"""``PrognosticVars``, ``initial_conditions``, ``add_halo``/``remove_halo`` and the host ghost points.

Mirrors ``src/initial_conditions.jl:6-164`` (rest start and the analytic sst fields) and
``src/ghost_points.jl:2-242``.  The ghost-point functions here run once per run on the host
(``add_halo``); inside the time loop the same copies are made on the GPU by the kernels
that produce each field.

The restart path ``initial_cond="ncfile"`` reads a run folder's NetCDF files (``output.py``).
``seeded_initial_conditions`` is an
addition of this package: the seeded, smoothed and geostrophically balanced start used by
the channel, ensemble and mixed-precision workloads, injected through
``time_integration(Prog, Diag, S)`` as the reference allows for any ``Prog``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .grid import meshgrid
from .numtypes import storage_dtype, to_storage, to_float64, canonical_type


@dataclass
class PrognosticVars:
    """``src/initial_conditions.jl:6-11``: u, v, η, sst in Julia layout (Fortran order)."""
    u: np.ndarray
    v: np.ndarray
    η: np.ndarray
    sst: np.ndarray
    T: str = "Float32"

    def as_float64(self):
        return tuple(to_float64(a, self.T) for a in (self.u, self.v, self.η, self.sst))


def zero_prognostic(T, G) -> PrognosticVars:
    """``PrognosticVars{T}(G)``, ``src/initial_conditions.jl:14-24``."""
    dt = storage_dtype(T)
    h, hη = G.halo, G.haloη
    return PrognosticVars(
        u=np.zeros((G.nux + 2 * h, G.nuy + 2 * h), dtype=dt, order="F"),
        v=np.zeros((G.nvx + 2 * h, G.nvy + 2 * h), dtype=dt, order="F"),
        η=np.zeros((G.nx + 2 * hη, G.ny + 2 * hη), dtype=dt, order="F"),
        sst=np.zeros((G.nx + 2 * G.halosstx, G.ny + 2 * G.halossty), dtype=dt, order="F"),
        T=canonical_type(T))


# ---------------------------------------------------------------- ghost points (host)
# Work on float64 views of the values; Tcomm rounding is applied where the reference applies it.

def _comm(x, Tcomm, T):
    """``T(Tcomm(x))`` for float64 values that are exactly representable in ``T``."""
    if Tcomm == T:
        return x
    return to_float64(to_storage(to_float64(to_storage(x, Tcomm), Tcomm), T), T)


def ghost_points_u_nonperiodic(u, one_minus_α):
    """``src/ghost_points.jl:47-57``"""
    u[:, 0] = one_minus_α * u[:, 3]
    u[:, 1] = one_minus_α * u[:, 2]
    u[:, -2] = one_minus_α * u[:, -3]
    u[:, -1] = one_minus_α * u[:, -4]


def ghost_points_u_periodic(u, Tcomm, T):
    """``src/ghost_points.jl:60-73`` (the tangential condition stays commented out there)."""
    u[0, :] = _comm(u[-4, :], Tcomm, T)
    u[1, :] = _comm(u[-3, :], Tcomm, T)
    u[-2, :] = _comm(u[2, :], Tcomm, T)
    u[-1, :] = _comm(u[3, :], Tcomm, T)


def ghost_points_v_nonperiodic(v, one_minus_α):
    """``src/ghost_points.jl:76-86``"""
    v[0, :] = one_minus_α * v[3, :]
    v[1, :] = one_minus_α * v[2, :]
    v[-2, :] = one_minus_α * v[-3, :]
    v[-1, :] = one_minus_α * v[-4, :]


def ghost_points_v_periodic(v, Tcomm, T):
    """``src/ghost_points.jl:89-99``"""
    v[0, :] = _comm(v[-4, :], Tcomm, T)
    v[1, :] = _comm(v[-3, :], Tcomm, T)
    v[-2, :] = _comm(v[2, :], Tcomm, T)
    v[-1, :] = _comm(v[3, :], Tcomm, T)


def ghost_points_η_nonperiodic(η):
    """``src/ghost_points.jl:102-116``"""
    η[0, :] = η[1, :]
    η[-1, :] = η[-2, :]
    η[:, 0] = η[:, 1]
    η[:, -1] = η[:, -2]


def ghost_points_η_periodic(η, Tcomm, T):
    """``src/ghost_points.jl:119-133``"""
    η[0, :] = _comm(η[-2, :], Tcomm, T)
    η[-1, :] = _comm(η[1, :], Tcomm, T)
    η[:, 0] = η[:, 1]
    η[:, -1] = η[:, -2]


def ghost_points_sst_nonperiodic(sst, hx, hy):
    """``src/ghost_points.jl:136-156``"""
    for i in range(1, hx + 1):
        sst[i - 1, :] = sst[hx, :]
        sst[-i, :] = sst[-hx - 1, :]
    for j in range(1, hy + 1):
        sst[:, j - 1] = sst[:, hy]
        sst[:, -j] = sst[:, -hy - 1]


def ghost_points_sst_periodic(sst, hx, hy, Tcomm, T):
    """``src/ghost_points.jl:159-179``"""
    n = sst.shape[0]
    for i in range(1, hx + 1):
        sst[i - 1, :] = _comm(sst[n - 2 * hx + i - 1, :], Tcomm, T)
        sst[n - hx + i - 1, :] = _comm(sst[hx + i - 1, :], Tcomm, T)
    for j in range(1, hy + 1):
        sst[:, j - 1] = sst[:, hy]
        sst[:, -j] = sst[:, -hy - 1]


def ghost_points(u, v, η, S):
    """``src/ghost_points.jl:182-199`` on float64 value arrays of format ``Tprog``."""
    P = S.parameters
    if P.bc == "periodic":
        ghost_points_u_periodic(u, P.Tcomm, P.Tprog)
        ghost_points_v_periodic(v, P.Tcomm, P.Tprog)
        ghost_points_η_periodic(η, P.Tcomm, P.Tprog)
    else:
        oma = S.constants.one_minus_α
        ghost_points_u_nonperiodic(u, oma)
        ghost_points_v_nonperiodic(v, oma)
        ghost_points_η_nonperiodic(η)


def ghost_points_sst(sst, S):
    """``src/ghost_points.jl:232-242``"""
    P, G = S.parameters, S.grid
    if P.bc == "periodic":
        ghost_points_sst_periodic(sst, G.halosstx, G.halossty, P.Tcomm, P.Tprog)
    else:
        ghost_points_sst_nonperiodic(sst, G.halosstx, G.halossty)


def add_halo(u, v, η, sst, S, T=None):
    """``src/ghost_points.jl:2-27``: pad with zeros and apply the ghost points.

    ``u, v, η, sst`` are interior arrays of format ``T`` (default ``Tprog``) in host
    storage; the padded arrays are returned in the same storage.
    """
    G = S.grid
    T = canonical_type(T or S.parameters.Tprog)
    h, hη, hx, hy = G.halo, G.haloη, G.halosstx, G.halossty

    def pad(a, px, py):
        a = to_float64(a, T)
        out = np.zeros((a.shape[0] + 2 * px, a.shape[1] + 2 * py), order="F")
        out[px:px + a.shape[0], py:py + a.shape[1]] = a
        return out

    u, v, η, sst = pad(u, h, h), pad(v, h, h), pad(η, hη, hη), pad(sst, hx, hy)
    ghost_points(u, v, η, S)
    ghost_points_sst(sst, S)
    return tuple(to_storage(a, T) for a in (u, v, η, sst))


def remove_halo(u, v, η, sst, S):
    """``src/ghost_points.jl:30-44`` (copies, like Julia's slicing)."""
    G = S.grid
    h, hη, hx, hy = G.halo, G.haloη, G.halosstx, G.halossty
    return (np.asfortranarray(u[h:u.shape[0] - h, h:u.shape[1] - h]),
            np.asfortranarray(v[h:v.shape[0] - h, h:v.shape[1] - h]),
            np.asfortranarray(η[hη:η.shape[0] - hη, hη:η.shape[1] - hη]),
            np.asfortranarray(sst[hx:sst.shape[0] - hx, hy:sst.shape[1] - hy]))


# ---------------------------------------------------------------- initial conditions

def sst_initial_field(S) -> np.ndarray:
    """The analytic sst start of ``src/initial_conditions.jl:122-141`` in float64."""
    P, G = S.parameters, S.grid
    xx_T, yy_T = meshgrid(G.x_T, G.y_T)
    Lx, Ly = G.Lx, G.Ly
    mid, amp = (P.SSTmin + P.SSTmax) / 2, (P.SSTmin - P.SSTmax)
    if P.sst_initial == "south":
        return mid + np.tanh(2 * math.pi * (Ly / (4 * P.SSTw)) * (yy_T / Ly - P.SSTϕ)) * amp / 2
    if P.sst_initial == "west":
        return mid + np.tanh(2 * math.pi * (Lx / (4 * P.SSTw)) * (xx_T / Lx - P.SSTϕ)) * amp / 2
    if P.sst_initial == "flat":
        return np.full(xx_T.shape, float(P.SSTmin), order="F")
    if P.sst_initial == "rect":
        x0, x1, y0, y1 = P.sst_rect_coords
        sst = np.full(xx_T.shape, float(P.SSTmin), order="F")
        inside = (xx_T / Lx > x0) & (xx_T / Lx < x1) & (yy_T / Ly > y0) & (yy_T / Ly < y1)
        sst[inside] = P.SSTmax
        return sst
    raise ValueError(f"sst_initial '{P.sst_initial}' unsupported")


def initial_conditions(Tprog, S) -> PrognosticVars:
    """``initial_conditions(T,S)``, ``src/initial_conditions.jl:26-164``: rest, or a restart
    from a run folder's NetCDF files (``initial_cond="ncfile"``)."""
    P, G = S.parameters, S.grid
    Tprog = canonical_type(Tprog)
    sst = None
    if P.initial_cond == "ncfile":                         # :39-120, the run folder's NetCDF files
        from .output import read_initial_conditions
        u, v, η, sst = read_initial_conditions(S)
        for name, a, shape in (("u", u, (G.nux, G.nuy)), ("v", v, (G.nvx, G.nvy)), ("η", η, (G.nx, G.ny))):
            if a.shape != shape:
                raise ValueError(f"restart {name} of shape {a.shape}, the grid needs {shape}")
        u, v, η = (np.asfortranarray(a) for a in (u, v, η))
    else:
        u = np.zeros((G.nux, G.nuy), order="F")
        v = np.zeros((G.nvx, G.nvy), order="F")
        η = np.zeros((G.nx, G.ny), order="F")
    if sst is None:
        if P.sst_initial == "restart":
            raise ValueError("sst_initial='restart' needs initial_cond='ncfile'")
        sst = sst_initial_field(S)
    elif sst.shape != (G.nx, G.ny):
        raise ValueError(f"restart sst of shape {sst.shape}, the grid needs {(G.nx, G.ny)}")
    sst = np.asfortranarray(sst)
    u, v, η, sst = (to_storage(a, Tprog) for a in (u, v, η, sst))
    u, v, η, sst = add_halo(u, v, η, sst, S, T=Tprog)
    return PrognosticVars(u=u, v=v, η=η, sst=sst, T=Tprog)


# ---------------------------------------------------------------- seeded starts (new)

_MASK64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """``n`` outputs of SplitMix64 (Steele, Lea, Flood 2014) from ``seed``, as uint64."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = (np.uint64(seed & _MASK64) + idx * np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    """Uniform doubles in [0,1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def gaussian(seed: int, n: int) -> np.ndarray:
    """Standard normals by Box-Muller on SplitMix64 uniforms."""
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m)
    u1 = 1.0 - u[0::2]                      # (0,1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
    return z[:n]


def smoothed_noise(seed: int, nx: int, ny: int, σ: float, passes: int = 8) -> np.ndarray:
    """Gaussian noise of std ``σ`` on the T-grid, smoothed by ``passes`` 5-point filters
    (weights 1/2 centre, 1/8 each neighbour; zero-gradient edges), rescaled to std ``σ``."""
    a = gaussian(seed, nx * ny).reshape((nx, ny), order="F")
    for _ in range(passes):
        p = np.pad(a, 1, mode="edge")
        a = 0.5 * p[1:-1, 1:-1] + 0.125 * (p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:])
    s = a.std()
    return np.asfortranarray(a * (σ / s) if s > 0 else a)


def geostrophic_uv(η: np.ndarray, S):
    """u, v (interior, float64) in geostrophic balance with the T-grid field ``η``:
    ``f u = -g ∂η/∂y``, ``f v = g ∂η/∂x`` with 4-point centred differences; zero-gradient
    at the edges, periodic in x when ``bc="periodic"``."""
    P, G = S.parameters, S.grid
    nx, ny, Δ, g = G.nx, G.ny, G.Δ, P.g
    periodic = P.bc == "periodic"
    ex = np.concatenate([η[-1:, :], η, η[:1, :]], axis=0) if periodic else np.pad(η, ((1, 1), (0, 0)), mode="edge")
    e = np.pad(ex, ((0, 0), (1, 1)), mode="edge")          # e[i+1, j+1] = η[i, j]
    f = lambda y: G.f0 + G.β * (y - G.Ly / 2)
    # u at the face between 1-based T-cells x and x+1 (x = 1-ep .. nx-1): columns x, x+1 of e
    ia = np.arange(0 if periodic else 1, nx)
    dηdy = (e[ia, 2:] - e[ia, :-2] + e[ia + 1, 2:] - e[ia + 1, :-2]) / (4 * Δ)
    u = -g * dηdy / f(G.y_u)[None, :]
    # v at the top face of T-row b (b = 1..ny-1), columns 1..nx
    dηdx = (e[2:, 1:ny] - e[:-2, 1:ny] + e[2:, 2:ny + 1] - e[:-2, 2:ny + 1]) / (4 * Δ)
    v = g * dηdx / f(G.y_v)[None, :]
    return np.asfortranarray(u), np.asfortranarray(v)


def seeded_initial_conditions(S, seed: int, σ: float = 0.1, passes: int = 8,
                              sst_noise: bool = True, Tprog=None) -> PrognosticVars:
    """Smoothed Gaussian η (std ``σ`` m), geostrophic u, v, and an sst of a seeded
    sinusoid plus uniform noise in [0,1]; halos added as ``add_halo`` does."""
    P, G = S.parameters, S.grid
    Tprog = canonical_type(Tprog or P.Tprog)
    η = smoothed_noise(seed, G.nx, G.ny, σ, passes)
    u, v = geostrophic_uv(η, S)
    xx, yy = meshgrid(G.x_T, G.y_T)
    ph = uniform01(seed ^ 0x5DEECE66D, 3)
    kx = 1 + int(ph[2] * 4)
    base = 0.5 + 0.25 * np.sin(2 * math.pi * (kx * xx / G.Lx + ph[0])) * np.cos(2 * math.pi * (yy / G.Ly + ph[1]))
    noise = uniform01(seed ^ 0x2545F4914F6CDD1D, G.nx * G.ny).reshape((G.nx, G.ny), order="F")
    sst = np.clip(0.75 * base + 0.25 * noise, 0.0, 1.0) if sst_noise else base
    u, v, η, sst = (to_storage(a, Tprog) for a in (u, v, η, sst))
    u, v, η, sst = add_halo(u, v, η, sst, S, T=Tprog)
    return PrognosticVars(u=u, v=v, η=η, sst=sst, T=Tprog)
