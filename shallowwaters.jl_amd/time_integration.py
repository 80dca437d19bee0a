"""``time_integration`` and ``run_model`` on the MI355X library.

``time_integration(Prog, Diag, S)`` keeps the signature and contract of
``src/time_integration.jl:2-250``: it integrates ``S.grid.nt`` steps from the halo-padded
``Prog`` (mutated in place, as the reference mutates its arrays), stops early when NaNs are
detected (``feedback.jl:136-148``) and returns a new ``PrognosticVars`` without halos
(``remove_halo``).  The loop itself runs on the GPU through ``include/swm.h``.

``run_model`` mirrors ``src/run_model.jl:12-39``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .initial_conditions import PrognosticVars, initial_conditions, remove_halo
from .model_setup import ModelSetup, make_config, make_setup
from .numtypes import canonical_type, storage_dtype
from .parameter import Parameter


def setup_of(S: ModelSetup) -> "_abi.SwmSetup":
    """The ``swm_setup`` of a model setup: the parameters of its winds and topography and the
    grid scalars the fields are evaluated from (Δ, Lx, Ly, f₀ and β from ``src/grid.jl:21-23,91-92``)."""
    P, G = S.parameters, S.grid
    return _abi.SwmSetup(topography=_abi.TOPOGRAPHY[P.topography], wind_x=_abi.WIND[P.wind_forcing_x],
                         wind_y=_abi.WIND[P.wind_forcing_y], Lx=G.Lx, Ly=G.Ly, Delta=G.Δ, H=P.H,
                         topo_height=P.topo_height, topo_width=P.topo_width, Fx0=P.Fx0, Fy0=P.Fy0, rho=P.ρ,
                         f0=G.f0, beta=G.β, eta_refh=P.η_refh, eta_refw=P.η_refw, A=P.A, phik=P.ϕk, wk=P.wk,
                         R=G.R, phi=G.ϕ)


@dataclass
class DiagnosticVars:
    """``DiagnosticVars{T,Tprog}`` (``src/preallocate.jl:439-451``).  On the GPU the
    intermediate fields are fused away inside the kernels; what remains host-visible is
    the Runge-Kutta state, downloaded on request (``Simulation.download_rk``)."""
    T: str
    Tprog: str
    RungeKutta: dict = field(default_factory=dict)


def preallocate(T, Tprog, G) -> DiagnosticVars:
    """``preallocate(T,Tprog,G)``, ``src/preallocate.jl:454-471``."""
    return DiagnosticVars(T=canonical_type(T), Tprog=canonical_type(Tprog))


class Simulation:
    """One library context: device buffers, static fields and the prognostic state.

    ``statics="host"`` uploads the static fields ``S.forcing`` and ``S.grid`` hold;
    ``statics="device"`` computes them on the GPU from the parameters (``Forcing{T}(P,G)``,
    ``src/forcing.jl:11-53``, and ``f_q``, ``f_u``, ``f_v``, ``src/grid.jl:91-108``)."""

    def __init__(self, S: ModelSetup, statics: str = "host"):
        self.S = S
        self.cfg = make_config(S)
        L = _abi.load()
        ctx = ctypes.c_void_p()
        _abi.check(L.swm_create(ctypes.byref(self.cfg), ctypes.byref(ctx)), None, "swm_create")
        self.ctx = ctx
        if statics == "device":
            self._static = []
            _abi.check(L.swm_set_static_device(ctx, ctypes.byref(setup_of(S))), ctx, "swm_set_static_device")
            return
        if statics != "host":
            raise ValueError(f"statics must be 'host' or 'device', got {statics!r}")
        F, G = S.forcing, S.grid
        self._static = [np.asfortranarray(a) for a in (F.H, F.Fx, F.Fy, G.f_q, F.η_ref, F.Fη)]
        _abi.check(L.swm_set_static(ctx, *[_abi.ptr(a) for a in self._static]), ctx, "swm_set_static")
        if S.parameters.dynamics == "linear":
            self._static += [np.asfortranarray(a) for a in (G.f_u, G.f_v)]
            _abi.check(L.swm_set_static_linear(ctx, *[_abi.ptr(a) for a in self._static[-2:]]), ctx,
                       "swm_set_static_linear")

    def download_static(self):
        """The static fields as the context holds them: dict of H, Fx, Fy, f_q (and η_ref or
        Fη where continuity reads them) in the shapes ``Forcing`` and ``Grid`` hold."""
        G, P = self.S.grid, self.S.parameters
        dt = storage_dtype(P.T)
        z = lambda m, n: np.zeros((m, n), dtype=dt, order="F")
        out = {"H": z(G.nx + 2, G.ny + 2), "Fx": z(G.nux, G.nuy), "Fy": z(G.nvx, G.nvy), "f_q": z(G.nx + 1, G.ny + 1)}
        if P.surface_relax:
            out["η_ref"] = z(G.nx, G.ny)
        if P.surface_forcing:
            out["Fη"] = z(G.nx, G.ny)
        args = [out[k] if k in out else None for k in ("H", "Fx", "Fy", "f_q", "η_ref", "Fη")]
        _abi.check(_abi.load().swm_download_static(self.ctx, *[None if a is None else _abi.ptr(a) for a in args]),
                   self.ctx, "swm_download_static")
        return out

    def close(self):
        if getattr(self, "ctx", None):
            _abi.load().swm_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, Prog: PrognosticVars):
        arrs = [np.asfortranarray(a) for a in (Prog.u, Prog.v, Prog.η, Prog.sst)]
        dt = storage_dtype(self.S.parameters.Tprog)
        for a in arrs:
            if a.dtype != dt:
                raise TypeError(f"Prog arrays must be stored as {dt}, got {a.dtype}")
        _abi.check(_abi.load().swm_upload_prog(self.ctx, *[_abi.ptr(a) for a in arrs]), self.ctx, "swm_upload_prog")

    def download(self, Prog: PrognosticVars | None = None) -> PrognosticVars:
        if Prog is None:
            G, dt = self.S.grid, storage_dtype(self.S.parameters.Tprog)
            z = lambda m, n: np.zeros((m, n), dtype=dt, order="F")
            Prog = PrognosticVars(u=z(G.nux + 4, G.nuy + 4), v=z(G.nvx + 4, G.nvy + 4), η=z(G.nx + 2, G.ny + 2),
                                  sst=z(G.nx + 2, G.ny), T=self.S.parameters.Tprog)
        _abi.check(_abi.load().swm_download_prog(self.ctx, *[_abi.ptr(a) for a in (Prog.u, Prog.v, Prog.η, Prog.sst)]),
                   self.ctx, "swm_download_prog")
        return Prog

    def step(self, nsteps: int, i0: int = 1, t0: int = 0):
        """Steps ``i0 .. i0+nsteps-1``; returns ``(steps_done, nan_step)``."""
        sd, ns = ctypes.c_int64(0), ctypes.c_int64(0)
        _abi.check(_abi.load().swm_step(self.ctx, nsteps, i0, t0, ctypes.byref(sd), ctypes.byref(ns)),
                   self.ctx, "swm_step")
        return sd.value, ns.value

    def time_steps(self, nsteps: int, i0: int = 1, t0: int = 0):
        """Steps ``i0 .. i0+nsteps-1`` with each kernel launch timed by HIP events inside the
        step: returns {kind: (average ms, launches)} for the kinds that ran (the state
        advances as by ``step``)."""
        ms = (ctypes.c_double * 8)()
        n = (ctypes.c_int32 * 8)()
        _abi.check(_abi.load().swm_time_steps(self.ctx, nsteps, i0, t0, ms, n), self.ctx, "swm_time_steps")
        kinds = ("fused12", "fused34", "dissipation", "tracer", "step", "stage1", "stage_mid", "stage4")
        return {k: (ms[j], n[j]) for j, k in enumerate(kinds) if n[j] > 0}

    def stages(self, n_stages: int, i0: int = 1, t0: int = 0):
        _abi.check(_abi.load().swm_stages(self.ctx, n_stages, i0, t0), self.ctx, "swm_stages")

    # ---- single operators of the step on the current state (parity tests)
    def ghost(self):
        """``ghost_points!`` of u, v, η and sst in place (``src/ghost_points.jl:136-199``)."""
        _abi.check(_abi.load().swm_ghost(self.ctx), self.ctx, "swm_ghost")

    def dissipation(self):
        """Drag, diffusion and their tendencies on the state as (u0, v0, η0); the result
        becomes the state (``src/time_integration.jl:216-221``)."""
        _abi.check(_abi.load().swm_dissipation(self.ctx), self.ctx, "swm_dissipation")

    def tracer(self, i: int):
        """``tracer!`` at step ``i`` with the state's velocities as the post-dissipation
        ones (``src/tracer_advection.jl:1-44``)."""
        _abi.check(_abi.load().swm_tracer(self.ctx, i), self.ctx, "swm_tracer")

    def download_rk(self):
        G, dt = self.S.grid, storage_dtype(self.S.parameters.Tprog)
        z = lambda m, n: np.zeros((m, n), dtype=dt, order="F")
        out = (z(G.nux + 4, G.nuy + 4), z(G.nvx + 4, G.nvy + 4), z(G.nx + 2, G.ny + 2),
               z(G.nux + 4, G.nuy + 4), z(G.nvx + 4, G.nvy + 4), z(G.nx + 2, G.ny + 2))
        _abi.check(_abi.load().swm_download_rk(self.ctx, *[_abi.ptr(a) for a in out]), self.ctx, "swm_download_rk")
        return out

    def time_kernel(self, which: int, reps: int) -> float:
        """Average duration (ms) of one kernel launched ``reps`` times on the context's
        stream, timed with HIP events (0 middle RK stage, 1 first stage, 2 last stage,
        3 dissipation, 4 tracer advection)."""
        ms = ctypes.c_double(0)
        _abi.check(_abi.load().swm_time_kernel(self.ctx, which, reps, ctypes.byref(ms)), self.ctx, "swm_time_kernel")
        return ms.value

    def reduce(self):
        m, e = ctypes.c_double(0), ctypes.c_double(0)
        _abi.check(_abi.load().swm_reduce(self.ctx, ctypes.byref(m), ctypes.byref(e)), self.ctx, "swm_reduce")
        return m.value, e.value

    # ---- reverse mode (swm_adj_*): every time scheme (SSPRK2 RKs <= 3, SSPRK3 RKn <= 3), T = Tprog = Tcomm in
    # {Float64, Float32}
    def adjoint_forward(self, nsteps: int, i0: int = 1, t0: int = 0, snapshots: int = -1):
        """Steps ``i0 .. i0+nsteps-1`` recorded for a pullback, with ``snapshots`` device
        checkpoints placed by the binomial (revolve) schedule; -1: as many as half of the
        device's free memory holds, at most one per step (nothing recomputed)."""
        _abi.check(_abi.load().swm_adj_forward(self.ctx, nsteps, i0, t0, snapshots), self.ctx, "swm_adj_forward")

    def pullback(self, ybar: PrognosticVars):
        """Gradient of <ybar, y> with respect to the state given to ``adjoint_forward`` (a
        PrognosticVars in Tprog storage, halo cells zero); also returns the number of
        forward steps the reverse sweep recomputed.  ``self.sweep_ms`` receives the device
        time of the reverse sweep."""
        G, dt = self.S.grid, storage_dtype(self.S.parameters.Tprog)
        ins = [np.asfortranarray(a, dtype=dt) for a in (ybar.u, ybar.v, ybar.η, ybar.sst)]
        z = lambda m, n: np.zeros((m, n), dtype=dt, order="F")
        out = PrognosticVars(u=z(G.nux + 4, G.nuy + 4), v=z(G.nvx + 4, G.nvy + 4), η=z(G.nx + 2, G.ny + 2),
                             sst=z(G.nx + 2, G.ny), T=self.S.parameters.Tprog)
        nf, ms = ctypes.c_int64(0), ctypes.c_double(0)
        _abi.check(_abi.load().swm_adj_pullback(self.ctx, *[_abi.ptr(a) for a in ins],
                                                *[_abi.ptr(a) for a in (out.u, out.v, out.η, out.sst)], ctypes.byref(nf),
                                                ctypes.byref(ms)),
                   self.ctx, "swm_adj_pullback")
        self.sweep_ms = ms.value
        return out, nf.value

    def adjoint_track_Fx(self, on: bool = True):
        """Let the following pullbacks also accumulate the gradient with respect to the
        zonal wind stress Fx (``swm_adj_track_Fx``)."""
        _abi.check(_abi.load().swm_adj_track_Fx(self.ctx, int(bool(on))), self.ctx, "swm_adj_track_Fx")

    def pullback_Fx(self):
        """The gradient with respect to Fx (shape (nux, nuy) of ``Forcing.Fx``, format T) from
        the last pullback."""
        G = self.S.grid
        out = np.zeros((G.nux, G.nuy), dtype=storage_dtype(self.S.parameters.T), order="F")
        _abi.check(_abi.load().swm_adj_download_Fx(self.ctx, _abi.ptr(out)), self.ctx, "swm_adj_download_Fx")
        return out

    def adjoint_time(self, which: int, reps: int) -> float:
        """Average duration (ms): 0 forward step, 1 taped step, 2 reverse step, 3 stage
        adjoint kernel, 4 dissipation adjoint kernel."""
        ms = ctypes.c_double(0)
        _abi.check(_abi.load().swm_adj_time(self.ctx, which, reps, ctypes.byref(ms)), self.ctx, "swm_adj_time")
        return ms.value


class Ensemble:
    """Several ``Simulation`` s advanced together (``swm_ens_*``): one batched launch per
    kernel of a step for all members and graph replays of step pairs, or (``persistent``)
    all steps in one persistent launch; an autotuner for the segment lengths.  The members stay usable on their own (``download`` ...)."""

    def __init__(self, members):
        self.members = list(members)
        L = _abi.load()
        arr = (ctypes.c_void_p * len(self.members))(*[m.ctx for m in self.members])
        ens = ctypes.c_void_p()
        rc = L.swm_ens_create(len(self.members), arr, ctypes.byref(ens))
        if rc != 0:
            raise _abi.SwmError(f"swm_ens_create failed with status {rc}")
        self.ens = ens

    def _check(self, rc, what):
        if rc != 0:
            m = _abi.load().swm_ens_last_error(self.ens)
            raise _abi.SwmError(f"{what} failed with status {rc}: {m.decode() if m else ''}")

    def tune(self, reps: int = 5):
        """Autotune the rows per wave segment; returns (stage, fused, dissipation)."""
        a, b, c = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        self._check(_abi.load().swm_ens_tune(self.ens, reps, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                    "swm_ens_tune")
        return a.value, b.value, c.value

    def step(self, nsteps: int, i0: int = 1):
        """Steps ``i0 .. i0+nsteps-1`` of every member; returns (steps_done, nan_steps, graph_steps)."""
        done, graph = ctypes.c_int64(0), ctypes.c_int64(0)
        nan = (ctypes.c_int64 * len(self.members))()
        self._check(_abi.load().swm_ens_step(self.ens, nsteps, i0, ctypes.byref(done), nan, ctypes.byref(graph)),
                    "swm_ens_step")
        return done.value, list(nan), graph.value

    @property
    def persistent(self):
        return getattr(self, "_persistent", False)

    @persistent.setter
    def persistent(self, on):
        """Run eligible steps in one persistent launch or as batched launches and graph
        replays (default, faster); the results are the same.  "timed": persistent, with its
        phases timed (``phase_ms``)."""
        mode = 2 if on == "timed" else (1 if on else 0)
        self._check(_abi.load().swm_ens_set_persistent(self.ens, mode), "swm_ens_set_persistent")
        self._persistent = on

    def phase_ms(self):
        """Average ms of the persistent launch's phases in the last timed call: first fused
        stage pair, second, dissipation, whole step (each phase up to the barrier after it)."""
        out = (ctypes.c_double * 4)()
        self._check(_abi.load().swm_ens_phase_ms(self.ens, out), "swm_ens_phase_ms")
        return {"fused12": out[0], "fused34": out[1], "dissipation": out[2], "step": out[3]}

    def persistent_steps(self) -> int:
        """Steps of the last ``step`` call that ran in the persistent launch."""
        return self.persistent_status()[0]

    def persistent_status(self):
        """(steps of the last ``step`` call that ran in the persistent launch, why it was not
        used when it was not)."""
        n, why = ctypes.c_int64(0), ctypes.c_char_p()
        self._check(_abi.load().swm_ens_persistent_steps(self.ens, ctypes.byref(n), ctypes.byref(why)),
                    "swm_ens_persistent_steps")
        return n.value, (why.value or b"").decode()

    def close(self):
        if getattr(self, "ens", None):
            _abi.load().swm_ens_destroy(self.ens)
            self.ens = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _stops(S, fb, ncs) -> list:
    """The steps at which the host takes part: NetCDF output (every ``nout`` steps, and the
    step before it when tendencies are written), the duration estimate (after the tracer ran
    once and twice, ``feedback.jl:125-131``) and the progress percentages
    (``feedback.jl:168-181``, after step 100); the last step."""
    G, P = S.grid, S.parameters
    nt = G.nt
    stops = {nt}
    if P.output:
        stops.update(k * G.nout for k in range(1, nt // G.nout + 1))
        if ncs.wants_tendencies:
            stops.update(k * G.nout - 1 for k in range(1, nt // G.nout + 1) if k * G.nout > 1)
        stops.update((G.nadvstep, 2 * G.nadvstep))
    stops.update(i for i in range(101, nt + 1) if ((i + 1) / nt * 100 % 1) < (i / nt * 100 % 1))
    return sorted(i for i in stops if 1 <= i <= nt)


def time_integration(Prog: PrognosticVars, Diag: DiagnosticVars, S: ModelSetup) -> PrognosticVars:
    """``src/time_integration.jl:2-250`` on the GPU: the steps run on the device in chunks
    between the steps where the host writes output (``output_nc!``) or feedback."""
    from .output import NcFiles, feedback, feedback_end, feedback_init
    G, P = S.grid, S.parameters
    fb = feedback_init(S)                                   # :38-41
    ncs = NcFiles(fb, S)
    try:
        with Simulation(S) as sim:
            sim.upload(Prog)
            cur = remove_halo(Prog.u, Prog.v, Prog.η, Prog.sst, S)
            ncs.output(0, cur, cur[:3], S)
            before, before_i = cur, 0                       # the last state downloaded, and its step
            i = 0
            for stop in _stops(S, fb, ncs):
                done, nan_step = sim.step(stop - i, i + 1, i * G.dtint)
                i += done
                fb.i = i
                if nan_step:
                    feedback(fb, S, nan_step)
                    break
                feedback(fb, S)
                if P.output and (i % G.nout == 0 or (ncs.wants_tendencies and (i + 1) % G.nout == 0)):
                    g = sim.download()
                    now = remove_halo(g.u, g.v, g.η, g.sst, S)
                    if i % G.nout == 0:                     # tendencies: the state after step i-1
                        ncs.output(i, now, before[:3] if before_i == i - 1 else now[:3], S)
                    before, before_i = now, i
            sim.download(Prog)
    finally:
        ncs.close()
    feedback_end(fb)
    u, v, η, sst = remove_halo(Prog.u, Prog.v, Prog.η, Prog.sst, S)
    return PrognosticVars(u=u, v=v, η=η, sst=sst, T=S.parameters.Tprog)


class NoTangent:
    """ChainRulesCore.NoTangent(): the argument has no derivative."""

    def __repr__(self):
        return "NoTangent()"


class Tangent:
    """ChainRulesCore.Tangent{P}(; fields...): the derivative with respect to a struct, one
    entry per differentiated field (the others have none)."""

    def __init__(self, **fields):
        self.__dict__.update(fields)

    def __repr__(self):
        return "Tangent(" + ", ".join(f"{k}=..." for k in self.__dict__) + ")"


def rrule(f, Prog: PrognosticVars, Diag: DiagnosticVars, S: ModelSetup, snapshots: int = -1, wrt_Fx: bool = False):
    """``ChainRulesCore.rrule(::typeof(time_integration), Prog, Diag, S)``: runs
    ``time_integration`` on the GPU keeping ``snapshots`` device checkpoints (revolve; -1: as
    many as half of the free device memory holds, at most one per step) and
    returns ``(y, pullback)``; ``pullback(ȳ)`` takes the cotangent of the returned interior
    fields (a PrognosticVars without halo) and returns ``(NoTangent(), Proḡ, NoTangent(),
    S̄)`` with Proḡ the gradient with respect to ``Prog`` (halo cells zero: the
    upload's ghost copies make the halo a function of the interior).  S̄ is ``NoTangent()``,
    or with ``wrt_Fx`` the ``Tangent(forcing=Tangent(Fx=...))`` of the zonal wind stress."""
    if f is not time_integration:
        raise TypeError("rrule is defined for time_integration only")
    G, T = S.grid, S.parameters.Tprog
    sim = Simulation(S)
    sim.upload(Prog)
    if wrt_Fx:
        sim.adjoint_track_Fx(True)
    sim.adjoint_forward(G.nt, 1, 0, snapshots)
    Y = sim.download()
    u, v, η, sst = remove_halo(Y.u, Y.v, Y.η, Y.sst, S)
    y = PrognosticVars(u=u, v=v, η=η, sst=sst, T=T)
    dt = storage_dtype(T)

    def pullback(ybar: PrognosticVars):
        z = lambda m, n: np.zeros((m, n), dtype=dt, order="F")
        yb = PrognosticVars(u=z(G.nux + 4, G.nuy + 4), v=z(G.nvx + 4, G.nvy + 4), η=z(G.nx + 2, G.ny + 2),
                            sst=z(G.nx + 2, G.ny), T=T)
        yb.u[2:-2, 2:-2] = ybar.u          # remove_halo keeps the interior (initial_conditions.jl:170-181)
        yb.v[2:-2, 2:-2] = ybar.v
        yb.η[1:-1, 1:-1] = ybar.η
        yb.sst[1:-1, :] = ybar.sst
        xbar, _ = sim.pullback(yb)
        sbar = Tangent(forcing=Tangent(Fx=sim.pullback_Fx())) if wrt_Fx else NoTangent()
        return NoTangent(), xbar, NoTangent(), sbar

    pullback.simulation = sim
    return y, pullback


def run_model(T=None, P: Parameter | None = None, **kwargs) -> PrognosticVars:
    """``run_model(T; kwargs...)``, ``run_model(P)`` and ``run_model(T,P)``
    (``src/run_model.jl:12-39``)."""
    if isinstance(T, Parameter):
        P, T = T, T.T
    if P is None:
        P = Parameter(T=T if T is not None else "Float32", **kwargs)
    elif kwargs:
        raise TypeError("pass either a Parameter or keyword arguments, not both")
    S = make_setup(P)
    Prog = initial_conditions(P.Tprog, S)
    Diag = preallocate(P.T, P.Tprog, S.grid)
    return time_integration(Prog, Diag, S)
