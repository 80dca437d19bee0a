"""The adjoint of the time step, checked on the CPU.

Chain of pins:
  * the oracle's tangent-linear model (the oracle instantiated on dual numbers, exact
    forward-mode derivatives of the same operations) against central finite differences of
    the oracle itself (16 seeded directions, both boundary conditions, tracer on);
  * the adjoint kernels (csrc/adj_core.h, run on an emulated wavefront by tests/emu) against
    the transposed Jacobian of the tangent-linear model, column by column, on small grids:
    each operator alone (stage right-hand side, dissipation, tracer) and whole steps;
  * on grids wider than one wave strip, with NaN shifted in at every wave edge and short
    row segments (so that any dependence outside a wave's halo or a segment's warm-up
    shows up), the dot-product identity <J dx, y> = <dx, J^T y>.
Bar: relative error <= 1e-11 (FP64 rounding of sums of up to ~100 terms).
"""
import copy

import numpy as np
import pytest

import adjemu

TOL = 1e-11


def setup(swm, bc, nx=12, **kw):
    d = dict(T="Float64", nx=nx, L_ratio=0.75, bc=bc, α=0.5, diffusion="Smagorinsky", bottom_drag="quadratic",
             topography="seamount", wind_forcing_x="double_gyre", Uadv=20.0, tracer_consumption=True)
    d.update(kw)
    P = swm.Parameter(**d)
    S = swm.make_setup(P)
    return P, S, swm.make_config(S), swm.seeded_initial_conditions(S, seed=5, σ=0.8)


def live_masks(S):
    """Cells of u, v, η that are state (interior or written by ghost_points!); the others
    are constants of the model (the normal-velocity halo, the periodic u y-halo)."""
    G = S.grid
    nx, ny = G.nx, G.ny
    per = S.parameters.bc == "periodic"
    ep = 1 if per else 0
    ii, jj = np.meshgrid(np.arange(G.nux + 4), np.arange(G.nuy + 4), indexing="ij")
    x, y = ii - 1 - ep, jj - 1
    mu = ((x >= -2) & (x <= nx + 1) & (y >= 1) & (y <= ny)) if per else \
         ((x >= 1) & (x <= nx - 1) & (y >= -1) & (y <= ny + 2))
    ii, jj = np.meshgrid(np.arange(G.nvx + 4), np.arange(G.nvy + 4), indexing="ij")
    x, y = ii - 1, jj - 1
    mv = (x >= -1) & (x <= nx + 2) & (y >= 1) & (y <= ny - 1)
    return mu, mv, np.ones((nx + 2, ny + 2), bool)


def interior_masks(S):
    G = S.grid
    m = [np.zeros((G.nux + 4, G.nuy + 4), bool), np.zeros((G.nvx + 4, G.nvy + 4), bool),
         np.zeros((G.nx + 2, G.ny + 2), bool), np.zeros((G.nx + 2, G.ny), bool)]
    m[0][2:-2, 2:-2] = m[1][2:-2, 2:-2] = m[2][1:-1, 1:-1] = m[3][1:-1, :] = True
    return m


def jacobian_T(orc, S, cfg, op, x, lam, masks, **kw):
    """(J^T lam) of the tangent-linear op, one column per input cell in masks."""
    out = []
    for a, (xa, m) in enumerate(zip(x, masks)):
        g = np.zeros_like(xa)
        for idx in zip(*np.nonzero(m)):
            dx = [np.zeros_like(b) for b in x]
            dx[a][idx] = 1.0
            _, dy = orc.tlm(S, cfg, op, x, dx, **kw)
            g[idx] = sum(float(np.sum(d * l)) for d, l in zip(dy, lam))
        out.append(g)
    return out


def assert_close(name, ref, got, mask=None):
    ref = np.asarray(ref)
    got = np.asarray(got)
    if mask is not None:
        ref = np.where(mask, ref, 0.0)
        got = np.where(mask, got, 0.0)
    assert not np.isnan(got).any(), f"{name}: NaN"
    scale = np.abs(ref).max()
    err = np.abs(ref - got).max() / (scale if scale > 0 else 1.0)
    assert err <= TOL, f"{name}: max relative error {err:.2e}"


def ghosted_state(swm, S, prog):
    from shallowwaters_amd.initial_conditions import ghost_points, ghost_points_sst
    u, v, e, s = (np.array(a, dtype=np.float64, order="F") for a in (prog.u, prog.v, prog.η, prog.sst))
    ghost_points(u, v, e, S)
    ghost_points_sst(s, S)
    return u, v, e, s


# ------------------------------------------------------------------ tangent-linear model

@pytest.mark.parametrize("bc", ["nonperiodic", "periodic"])
def test_tangent_linear_model_matches_central_differences(swm, orc, bc):
    P, S, cfg, prog = setup(swm, bc)
    G = S.grid
    n = 2 * G.nadvstep                     # two tracer periods
    x = list(ghosted_state(swm, S, prog))
    masks = interior_masks(S)
    rng = np.random.default_rng(11)
    eps = 1e-6
    for d in range(16):
        dx = [np.where(m, rng.standard_normal(a.shape), 0.0) for a, m in zip(x, masks)]
        dxg = list(ghosted_state(swm, S, type(prog)(*dx)))        # a perturbation the upload keeps
        y, dy = orc.tlm(S, cfg, 0, x, dxg, nsteps=n)
        fd = []
        for sgn in (1, -1):
            p = copy.copy(prog)
            p.u, p.v, p.η, p.sst = (a + sgn * eps * b for a, b in zip(x, dxg))
            fd.append(orc.run(S, cfg, p, n)[:4])
        for k, name in enumerate(("u", "v", "η", "sst")):
            c = (fd[0][k] - fd[1][k]) / (2 * eps)
            err = np.linalg.norm(c - dy[k]) / np.linalg.norm(dy[k])
            assert err < 1e-8, f"direction {d} {name}: {err:.2e}"


@pytest.mark.parametrize("bc,seasonal", [("nonperiodic", False), ("periodic", True)])
def test_wind_stress_tangent_matches_central_differences(swm, orc, bc, seasonal):
    """The tangent of the state with respect to the wind stress field Fx (rhs.jl:242, a
    seasonal factor sin(ωt) when seasonal_wind_x) against central differences of the oracle
    with Fx ± ε·δFx (8 seeded directions, two tracer periods)."""
    P, S, cfg, prog = setup(swm, bc, seasonal_wind_x=seasonal, ωFx=300.0)
    G = S.grid
    n = 2 * G.nadvstep
    x = list(ghosted_state(swm, S, prog))
    rng = np.random.default_rng(12)
    eps = 1e-6
    Fx0 = S.forcing.Fx.copy()
    for d in range(8):
        dF = rng.standard_normal(Fx0.shape) * np.abs(Fx0).max()
        _, dy = orc.tlm(S, cfg, 0, x, None, nsteps=n, dFx=dF)
        fd = []
        for sgn in (1, -1):
            S2 = copy.copy(S)
            S2.forcing = copy.copy(S.forcing)
            S2.forcing.Fx = np.asfortranarray(Fx0 + sgn * eps * dF)
            fd.append(orc.run(S2, cfg, prog, n)[:4])
        for k, name in enumerate(("u", "v", "η", "sst")):
            c = (fd[0][k] - fd[1][k]) / (2 * eps)
            # the differences' rounding, ~1e-16·|y|/ε per cell, bounds what they can resolve:
            # the tracer's tangent (Fx acts on it only through the departure points) is small
            floor = 1e-16 * np.linalg.norm(fd[0][k]) / eps
            diff, nrm = np.linalg.norm(c - dy[k]), np.linalg.norm(dy[k])
            assert diff <= 1e-7 * nrm + 10 * floor, f"direction {d} {name}: |fd - tlm| {diff:.2e}, |tlm| {nrm:.2e}"


def test_tangent_linear_values_are_the_oracle(swm, orc):
    P, S, cfg, prog = setup(swm, "nonperiodic")
    x = list(ghosted_state(swm, S, prog))
    y, _ = orc.tlm(S, cfg, 0, x, nsteps=4)
    r = orc.run(S, cfg, prog, 4)
    for a, b in zip(y, r[:4]):
        assert np.array_equal(a, b)


# ------------------------------------------------------------------ operators

def linear_statics(S):
    """f_u, f_v appended to the emulated ops' inputs under the linear dynamics."""
    return [S.grid.f_u, S.grid.f_v] if S.parameters.dynamics == "linear" else []


@pytest.mark.parametrize("bc,cont", [("nonperiodic", {}), ("periodic", {}),
                                      ("nonperiodic", {"surface_relax": True, "t_relax": 2.0}),
                                      ("nonperiodic", {"adv_scheme": "Sadourny"}),
                                      ("periodic", {"adv_scheme": "Sadourny"}),
                                      ("nonperiodic", {"dynamics": "linear"}),
                                      ("periodic", {"dynamics": "linear"}),
                                      ("nonperiodic", {"dynamics": "linear", "surface_relax": True, "t_relax": 2.0})])
def test_stage_adjoint_is_the_transposed_jacobian(swm, orc, bc, cont):
    P, S, cfg, prog = setup(swm, bc, **cont)
    rng = np.random.default_rng(3)
    mu, mv, me = live_masks(S)
    x = [np.asfortranarray(a, dtype=np.float64) for a in (prog.u, prog.v, prog.η)]
    x = [np.where(m, a + 0.01 * (np.abs(a).max() + 1) * rng.standard_normal(a.shape), 0.0)
         for a, m in zip(x, (mu, mv, me))]
    im = interior_masks(S)
    lam = [np.where(m, rng.standard_normal(a.shape), 0.0) for a, m in zip(x, im[:3])]
    ref = jacobian_T(orc, S, cfg, 1, x, lam, (mu, mv, me))
    k = adjemu.config(S, cfg, L=5, nan_edges=True)
    got = adjemu.run(S, k, 1, [x[0], x[1], x[2], S.forcing.H, S.grid.f_q] + lam + linear_statics(S))
    for name, r, g, m in zip("uvη", ref, got, (mu, mv, me)):
        assert_close(f"{bc} μ{name}", r, g, m)


@pytest.mark.parametrize("bc", ["nonperiodic", "periodic"])
@pytest.mark.parametrize("drag,diff", [("quadratic", "Smagorinsky"), ("linear", "constant"), ("none", "Smagorinsky")])
def test_dissipation_adjoint_is_the_transposed_jacobian(swm, orc, bc, drag, diff):
    P, S, cfg, prog = setup(swm, bc, bottom_drag=drag, diffusion=diff)
    rng = np.random.default_rng(4)
    mu, mv, me = live_masks(S)
    x = [np.asfortranarray(a, dtype=np.float64) for a in (prog.u, prog.v, prog.η)]
    x = [np.where(m, a + 0.01 * (np.abs(a).max() + 1) * rng.standard_normal(a.shape), 0.0)
         for a, m in zip(x, (mu, mv, me))]
    lam = [np.where(m, rng.standard_normal(a.shape), 0.0) for a, m in zip(x[:2], (mu, mv))]
    ref = jacobian_T(orc, S, cfg, 2, x, lam, (mu, mv, me))
    k = adjemu.config(S, cfg, L=5, nan_edges=True)
    got = adjemu.run(S, k, 2, [x[0], x[1], x[2], S.forcing.H, lam[0], lam[1], np.zeros_like(x[2])])
    for name, r, g, m in zip("uvη", ref, got, (mu, mv, me)):
        assert_close(f"{bc} {drag} {diff} λ{name}", r, g, m)


@pytest.mark.parametrize("bc", ["nonperiodic", "periodic"])
def test_tracer_adjoint_is_the_transposed_jacobian(swm, orc, bc):
    P, S, cfg, prog = setup(swm, bc)
    G = S.grid
    rng = np.random.default_rng(6)
    u, v = (np.asfortranarray(a, dtype=np.float64) for a in (prog.u, prog.v))
    um, vm = 0.9 * u + 0.01, 1.1 * v - 0.01
    sst = np.asfortranarray(prog.sst, dtype=np.float64)
    x = [u, v, um, vm, sst]
    for i in (G.nadvstep, G.nadvstep - G.nadvstep_half):       # advection, snapshot
        lam = [rng.standard_normal(a.shape) for a in (u, v, sst)]
        ref = jacobian_T(orc, S, cfg, 3, x, lam, [np.ones(a.shape, bool) for a in x], i0=i)
        k = adjemu.config(S, cfg, L=5, nan_edges=True)
        got = adjemu.run(S, k, 3, x + lam, i=i)
        for name, r, g in zip(("u", "v", "um", "vm", "sst"), ref, got):
            assert_close(f"{bc} step {i} λ{name}", r, g)


# ------------------------------------------------------------------ whole steps

def reverse_steps(swm, orc, S, cfg, prog, n, lam, k):
    """Tape n oracle steps, run the emulated reverse sweep, fold: the gradient of <lam, y>."""
    G = S.grid
    u, v, e, s = ghosted_state(swm, S, prog)
    st = (u, v, e, s, np.zeros_like(u), np.zeros_like(v))
    tapes = []
    for i in range(1, n + 1):
        tp = orc.tape(S, cfg, i, (i - 1) * G.dtint, *st)
        tapes.append((st, tp))
        st = tp["new"] + tp["mid"]
    L = list(lam) + [np.zeros_like(u), np.zeros_like(v)]
    for i in range(n, 0, -1):
        x0, tp = tapes[i - 1]
        ins = [a for sk in tp["s"] for a in sk] + [a for pk in tp["p"] for a in pk] + list(tp["pre"]) + \
              list(tp["post"]) + list(tp["mid"]) + \
              [x0[3], S.forcing.H, S.grid.f_q] + L + linear_statics(S)
        L = adjemu.run(S, k, 4, ins, i=i)
    return adjemu.run(S, k, 5, L[:4]), [u, v, e, s], st


@pytest.mark.parametrize("bc,kw", [("nonperiodic", {}), ("periodic", {}), ("nonperiodic", {"RKo": 3}),
                                   ("periodic", {"RKo": 2, "cfl": 0.5, "Uadv": 40.0}),
                                   ("periodic", {"nstep_diff": 2, "diffusion": "constant", "bottom_drag": "linear"}),
                                   ("periodic", {"adv_scheme": "Sadourny"}),
                                   ("nonperiodic", {"adv_scheme": "Sadourny", "RKo": 3}),
                                   ("nonperiodic", {"dynamics": "linear"}),
                                   ("periodic", {"dynamics": "linear", "RKo": 3, "bottom_drag": "linear"}),
                                   ("nonperiodic", {"time_scheme": "SSPRK2"}),
                                   ("periodic", {"time_scheme": "SSPRK2", "RKs": 2, "cfl": 0.5, "Uadv": 40.0}),
                                   ("periodic", {"time_scheme": "SSPRK2", "adv_scheme": "Sadourny",
                                                 "surface_relax": True}),
                                   ("nonperiodic", {"time_scheme": "SSPRK2", "dynamics": "linear", "RKs": 2,
                                                    "cfl": 0.5, "Uadv": 40.0}),
                                   ("nonperiodic", {"time_scheme": "4SSPRK3"}),
                                   ("periodic", {"time_scheme": "4SSPRK3", "adv_scheme": "Sadourny"}),
                                   ("periodic", {"time_scheme": "4SSPRK3", "dynamics": "linear",
                                                 "surface_relax": True}),
                                   ("nonperiodic", {"time_scheme": "SSPRK3"}),
                                   ("periodic", {"time_scheme": "SSPRK3", "RKn": 3, "adv_scheme": "Sadourny"}),
                                   ("periodic", {"time_scheme": "SSPRK3", "dynamics": "linear",
                                                 "surface_relax": True})])
def test_reverse_steps_are_the_transposed_jacobian(swm, orc, bc, kw):
    """Three steps: the tracer's snapshot (step 2) and advection (step 3) included."""
    P, S, cfg, prog = setup(swm, bc, **kw)
    assert (S.grid.nadvstep, S.grid.nadvstep_half) == (3, 1)
    rng = np.random.default_rng(8)
    im = interior_masks(S)
    shapes = [m.shape for m in im]
    lam = [rng.standard_normal(sh) for sh in shapes]
    k = adjemu.config(S, cfg, L=5, nan_edges=True)
    g, x, _ = reverse_steps(swm, orc, S, cfg, prog, 3, lam, k)
    ref = jacobian_T(orc, S, cfg, 0, x, lam, im, nsteps=3)
    for name, r, gg, m in zip(("u", "v", "η", "sst"), ref, g, im):
        assert_close(f"{bc} {kw} ∇{name}", r, gg, m)
        assert np.all(gg[~m] == 0), f"{name}: halo of the gradient not cleared"


@pytest.mark.parametrize("bc,kw", [("nonperiodic", {}), ("periodic", {}), ("nonperiodic", {"dynamics": "linear"}),
                                   ("periodic", {"time_scheme": "SSPRK2"}),
                                   ("nonperiodic", {"time_scheme": "4SSPRK3"}),
                                   ("periodic", {"time_scheme": "SSPRK3", "RKn": 3})])
def test_dot_product_identity_across_wave_strips(swm, orc, bc, kw):
    """A grid of several wave strips and short segments: <J dx, y> = <dx, J^T y>."""
    P, S, cfg, prog = setup(swm, bc, nx=130, L_ratio=0.5, **kw)
    n = 4
    rng = np.random.default_rng(9)
    im = interior_masks(S)
    lam = [rng.standard_normal(m.shape) for m in im]
    k = adjemu.config(S, cfg, L=7, nan_edges=True)
    g, x, _ = reverse_steps(swm, orc, S, cfg, prog, n, lam, k)
    for d in range(3):
        dx = [np.where(m, rng.standard_normal(m.shape), 0.0) for m in im]
        dxg = list(ghosted_state(swm, S, type(prog)(*dx)))
        _, dy = orc.tlm(S, cfg, 0, x, dxg, nsteps=n)
        lhs = sum(float(np.sum(a * b)) for a, b in zip(dy, lam))
        rhs = sum(float(np.sum(a * b)) for a, b in zip(dx, g))
        assert abs(lhs - rhs) <= TOL * max(abs(lhs), 1.0), f"<J dx, y> {lhs!r} != <dx, J^T y> {rhs!r}"
