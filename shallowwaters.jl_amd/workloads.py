"""Named workloads of the time step: the parameter sets the benchmark and the full-size tests
run, and the seeded batch of ensemble members.

Each returns ``Parameter`` values the reference accepts as they are (``src/default_parameters.jl``);
the states come from ``seeded_initial_conditions`` (smoothed seeded η noise with geostrophic
u, v and a seeded tracer), since the reference's own initial states are at rest.
"""
from __future__ import annotations

import numpy as np

from .initial_conditions import seeded_initial_conditions
from .model_setup import make_setup
from .parameter import Parameter


def double_gyre(nx: int = 4096, T="Float32", Tprog=None, **kw) -> Parameter:
    """The nonperiodic double gyre on an nx² grid of Lx = Ly = 3840 km over a seamount,
    double-gyre wind Fx0 = 0.12 Pa, Smagorinsky diffusion, quadratic drag, tracer on."""
    Tprog = Tprog or T
    return Parameter(**{**dict(T=T, Tprog=Tprog, Tcomm=Tprog, nx=nx, Lx=3840e3, L_ratio=1, bc="nonperiodic", α=0.0,
                               topography="seamount", wind_forcing_x="double_gyre", Fx0=0.12,
                               diffusion="Smagorinsky", bottom_drag="quadratic", tracer_advection=True, Ndays=10),
                        **kw})


def readme_double_gyre(T="Float32") -> Parameter:
    """The README's 100x100 run (README.md:55-62): Lx = 2000 km, H = 500 m, g = 10,
    nonperiodic over a seamount, double-gyre wind 0.12 Pa, Smagorinsky, quadratic drag, no
    tracer, 100 days from rest at CFL 1 (30,639 steps of 282 s)."""
    return Parameter(T=T, nx=100, L_ratio=1, Lx=2000e3, H=500.0, g=10.0, bc="nonperiodic", α=0.0,
                     topography="seamount", wind_forcing_x="double_gyre", Fx0=0.12, diffusion="Smagorinsky",
                     bottom_drag="quadratic", tracer_advection=False, Ndays=100, cfl=1.0)


def channel(T="Float32", topography="flat", nx: int = 2048) -> Parameter:
    """The periodic channel nx x nx/2 (default 2048x1024: Δ = 976.5625 m, dt = 13 s), channel
    wind 0.12 Pa, tracer every 300 steps."""
    return Parameter(T=T, nx=nx, L_ratio=2, bc="periodic", topography=topography, wind_forcing_x="channel", Fx0=0.12)


def ensemble_members(nmem: int = 64, seed: int = 4, T="Float32"):
    """The batch of ensemble members: grid sizes drawn from {128, 256, 512, 1024}² by the
    seed (L_ratio = 1), Fx0 ~ U(0.08, 0.16) Pa, H ~ U(400, 600) m, seeded smoothed η noise of
    σ = 0.05 m, the other parameters at the reference's defaults (periodic channel over the
    ridge, tracer on).  Returns a list of (ModelSetup, PrognosticVars)."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([128, 256, 512, 1024], size=nmem)
    out = []
    for k, n in enumerate(sizes):
        P = Parameter(T=T, nx=int(n), L_ratio=1, Fx0=float(rng.uniform(0.08, 0.16)), H=float(rng.uniform(400, 600)))
        S = make_setup(P)
        out.append((S, seeded_initial_conditions(S, seed=1000 + k, σ=0.05)))
    return out
