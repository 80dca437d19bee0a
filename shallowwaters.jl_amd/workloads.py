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


# ------------------------------------------------------------------ the adjoint workload's loss

def observation_twin(S, prog, nsteps: int, seed: int = 17, fraction: float = 0.01, σ: float = 0.02):
    """Synthetic observations for the adjoint workload: a twin run of ``nsteps`` steps from
    ``prog`` with η perturbed by seeded noise of σ metres, observed in η at a seeded
    ``fraction`` of the interior cells.  Returns (η of the twin's final state, mask), both
    of the halo-padded η shape."""
    from .numtypes import to_float64, to_storage
    from .initial_conditions import PrognosticVars
    from .time_integration import Simulation
    G, T = S.grid, S.parameters.Tprog
    rng = np.random.default_rng(seed)
    mask = np.zeros((G.nx + 2, G.ny + 2), dtype=bool)
    cells = rng.choice(G.nx * G.ny, size=max(1, int(round(fraction * G.nx * G.ny))), replace=False)
    mask[1 + cells % G.nx, 1 + cells // G.nx] = True
    e = to_float64(prog.η, T).copy(order="F")
    e[1:-1, 1:-1] += σ * rng.standard_normal((G.nx, G.ny))
    twin = PrognosticVars(u=prog.u, v=prog.v, η=to_storage(e, T), sst=prog.sst, T=T)
    with Simulation(S) as sim:
        sim.upload(twin)
        done, nan = sim.step(nsteps, 1, 0)
        if (done, nan) != (nsteps, 0):
            raise RuntimeError(f"twin run stopped at step {done} (NaN at {nan})")
        obs = to_float64(sim.download().η, T)
    return obs, mask


def final_state_loss(S, y, obs, mask):
    """The adjoint workload's loss of a final state y (halo-padded PrognosticVars): the
    kinetic energy ½Σu² + ½Σv² over the interior u and v points plus the η misfit
    ½Σ(η - η_obs)² over the observed cells.  Returns (loss, ȳ) with ȳ its gradient with
    respect to y (halo cells zero), in y's storage format."""
    from .numtypes import to_float64, to_storage
    from .initial_conditions import PrognosticVars
    T = S.parameters.Tprog
    u, v, e = (to_float64(a, T) for a in (y.u, y.v, y.η))
    ui, vi = u[2:-2, 2:-2], v[2:-2, 2:-2]
    r = np.where(mask, e - obs, 0.0)
    L = 0.5 * float(np.sum(ui * ui)) + 0.5 * float(np.sum(vi * vi)) + 0.5 * float(np.sum(r * r))
    gu = np.zeros_like(u, order="F")
    gv = np.zeros_like(v, order="F")
    gu[2:-2, 2:-2] = ui
    gv[2:-2, 2:-2] = vi
    ybar = PrognosticVars(u=to_storage(gu, T), v=to_storage(gv, T), η=to_storage(np.asfortranarray(r), T),
                          sst=to_storage(np.zeros_like(to_float64(y.sst, T), order="F"), T), T=T)
    return L, ybar
