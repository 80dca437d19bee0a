"""Host setup (Parameter, Grid, Constants, Forcing, initial conditions) against values the
reference fixes: the README's own runs, closed forms and the reference's formulas at
special points."""
import math

import numpy as np
import pytest


def test_readme_step_counts(swm):
    # README.md:44-46: run_model(Float32,Ndays=10,...) on the default 100x50 grid, dt = 282 s
    S = swm.make_setup(swm.Parameter(T="Float32", Ndays=10, g=10, H=500, Fx0=0.12))
    G = S.grid
    assert (G.nx, G.ny, G.dtint, G.nt) == (100, 50, 282, 3064)
    # README.md:60-62: 100x100 nonperiodic double gyre over 100 days
    G = swm.make_setup(swm.Parameter(Ndays=100, nx=100, L_ratio=1, bc="nonperiodic",
                                     wind_forcing_x="double_gyre", topography="seamount")).grid
    assert (G.nx, G.ny, G.dtint, G.nt) == (100, 100, 282, 30639)
    # the 4096^2 workload: Δ = 937.5 m, dt = 13 s, 10 days = 66462 steps, tracer every 288 steps
    G = swm.make_setup(swm.Parameter(nx=4096, Lx=3840e3, L_ratio=1, bc="nonperiodic", Ndays=10)).grid
    assert (G.Δ, G.ny, G.dtint, G.nt, G.nadvstep) == (937.5, 4096, 13, 66462, 288)
    # the 2048x1024 channel: dt = 13 s, tracer every 300 steps
    G = swm.make_setup(swm.Parameter(nx=2048, L_ratio=2, bc="periodic")).grid
    assert (G.ny, G.dtint, G.nadvstep, G.nadvstep_half) == (1024, 13, 300, 150)


def test_grid_shapes_periodic_and_not(swm):
    for bc, ep in (("periodic", 1), ("nonperiodic", 0)):
        G = swm.make_setup(swm.Parameter(nx=20, L_ratio=2, bc=bc)).grid
        assert G.ep == ep
        assert G.nux == (20 if ep else 19) and G.nvy == G.ny - 1 == 9
        assert G.f_q.shape == (21, 11)
        assert G.x_u[0] == (0.0 if ep else G.Δ)


def test_constants_rounding(swm):
    P = swm.Parameter(T="Float32", nx=100)
    S = swm.make_setup(P)
    C, G = S.constants, S.grid
    dt = G.dtint / G.Δ
    assert C.RKaΔt == [float(np.float32(x * G.dtint / G.Δ)) for x in (1 / 6, 1 / 3, 1 / 3, 1 / 6)]
    assert C.RKbΔt == [float(np.float32(x * G.dtint / G.Δ)) for x in (.5, .5, 1.)]
    assert C.Δt_Δ == float(np.float32(dt))
    assert C.one_minus_α == -1.0                      # α = 2 default (no slip)
    assert C.cD == float(np.float32(-G.Δ * 1e-5))
    assert C.νB == float(np.float32(-500.0 / 30000))
    assert C.cSmag == float(np.float32(-0.15))
    ssp = C.SSPRK3c                                    # RKn = 2: s = 4, kn = 4, mn = 1
    assert (ssp.s, ssp.kn, ssp.mn) == (4, 4, 1)


def test_coriolis_closed_form(swm):
    # sind(45) = cosd(45) = √2/2 correctly rounded; f_q varies linearly in y, constant in x
    G = swm.make_setup(swm.Parameter(T="Float64", nx=40, L_ratio=1)).grid
    ω = 2 * math.pi / 86400
    assert G.f0 == 2 * ω * 0.7071067811865476
    assert G.β == 2 * ω / 6.371e6 * 0.7071067811865476
    fq = G.f_q
    assert np.all(fq == fq[:1, :])                     # x-invariant
    d = np.diff(fq[0, :])
    assert np.allclose(d, G.Δ * G.β * G.Δ, rtol=1e-9)
    # equator: f0 = 0 exactly, β = 2ω/R
    G0 = swm.make_setup(swm.Parameter(T="Float64", nx=40, ϕ=0.0)).grid
    assert G0.f0 == 0.0 and G0.β == 2 * ω / 6.371e6
    # pole: β = 0 exactly (cosd(90) = 0 as in Julia)
    assert swm.make_setup(swm.Parameter(T="Float64", nx=40, ϕ=90.0)).grid.β == 0.0


def test_wind_and_topography_shapes_and_symmetry(swm):
    P = swm.Parameter(T="Float64", nx=64, L_ratio=1, bc="nonperiodic", wind_forcing_x="channel",
                      topography="seamount")
    S = swm.make_setup(P)
    F, G = S.forcing, S.grid
    assert F.Fx.shape == (G.nux, G.nuy) and F.Fy.shape == (G.nvx, G.nvy) and F.H.shape == (G.nx + 2, G.ny + 2)
    # channel wind cos²(π(y/Ly-1/2)) is symmetric about mid-domain, largest there, zero-free
    fx = F.Fx[0, :]
    assert np.allclose(fx, fx[::-1], rtol=1e-14)
    assert np.isclose(fx.max(), G.Δ * 0.12 / 1e3 / 500, rtol=2e-3)
    # seamount: H = 500 - 50 exp(...) exp(...): deepest far away, 450 at the top
    assert F.H.max() < 500 and F.H.max() > 499.9 and F.H.min() > 450 and F.H.min() < 451
    assert np.allclose(F.H, F.H[::-1, :], rtol=1e-14) and np.allclose(F.H, F.H[:, ::-1], rtol=1e-14)
    # double gyre: cos(2πs)+2 sin(πs), s = y/Ly - 1/2: -3 at the south edge, +1 at the north edge
    P2 = P.replace(wind_forcing_x="double_gyre")
    fx2 = swm.make_setup(P2).forcing.Fx[0, :] / (G.Δ * 0.12 / 1e3 / 500)
    s = (G.y_u / G.Ly - 0.5)
    assert np.allclose(fx2, np.cos(2 * math.pi * s) + 2 * np.sin(math.pi * s), rtol=1e-12)
    assert fx2[0] < -2.9 and fx2[-1] > 0.9


def test_parameter_validation(swm):
    with pytest.raises(AssertionError, match="boundary condition"):
        swm.Parameter(bc="closed")
    with pytest.raises(AssertionError, match="RKo"):
        swm.Parameter(RKo=5)
    with pytest.raises(AssertionError):
        swm.Parameter(wind_forcing_y="channel")        # u-grid shaped Fy (forcing.jl:65)
    P = swm.Parameter(T="Float16", Tprog="Float32")
    assert (P.T, P.Tprog, P.Tcomm) == ("Float16", "Float32", "Float32")


def test_add_halo_and_ghosts(swm):
    for bc in ("periodic", "nonperiodic"):
        S = swm.make_setup(swm.Parameter(T="Float64", nx=8, L_ratio=1, bc=bc, α=0.5))
        G = S.grid
        rng = np.random.default_rng(0)
        u = rng.standard_normal((G.nux, G.nuy))
        v = rng.standard_normal((G.nvx, G.nvy))
        e = rng.standard_normal((G.nx, G.ny))
        s = rng.standard_normal((G.nx, G.ny))
        U, V, E, Ss = swm.add_halo(u, v, e, s, S)
        if bc == "periodic":
            assert np.array_equal(U[0, 2:-2], U[-4, 2:-2]) and np.array_equal(U[-1, 2:-2], U[3, 2:-2])
            assert np.all(U[:, :2] == 0) and np.all(U[:, -2:] == 0)          # u y-halo never written
            assert np.array_equal(E[0, :], E[-2, :]) or np.array_equal(E[0, 1:-1], E[-2, 1:-1])
        else:
            assert np.array_equal(U[2:-2, 0], 0.5 * U[2:-2, 3]) and np.array_equal(U[2:-2, -1], 0.5 * U[2:-2, -4])
            assert np.all(U[:2, 2:-2] == 0) and np.all(U[-2:, 2:-2] == 0)    # normal-velocity halo
            assert np.array_equal(V[0, 2:-2], 0.5 * V[3, 2:-2])
            assert np.array_equal(E[0, 1:-1], E[1, 1:-1]) and E[0, 0] == E[1, 1]
        assert np.all(V[2:-2, :2] == 0) and np.all(V[2:-2, -2:] == 0)        # v y-halo never written
        r = swm.remove_halo(U, V, E, Ss, S)
        for a, b in zip(r, (u, v, e, s)):
            assert np.array_equal(a, b)


def test_seeded_initial_conditions_are_deterministic(swm):
    S = swm.make_setup(swm.Parameter(T="Float32", nx=64, L_ratio=2, bc="periodic"))
    a = swm.seeded_initial_conditions(S, seed=3, σ=0.1)
    b = swm.seeded_initial_conditions(S, seed=3, σ=0.1)
    c = swm.seeded_initial_conditions(S, seed=4, σ=0.1)
    for x, y in zip(a.as_float64(), b.as_float64()):
        assert np.array_equal(x, y)
    assert not np.array_equal(a.η, c.η)
    e = a.as_float64()[2][1:-1, 1:-1]
    assert abs(e.std() - 0.1) < 1e-6 and 0 <= a.sst.min() and a.sst.max() <= 1
