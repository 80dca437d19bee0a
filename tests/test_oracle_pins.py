"""Closed-form pins of the oracle's dissipation, time-scheme combinations and continuity
variants.

Each ``pin_*`` function runs the oracle on a state whose result mathematics fixes and
returns the list of disagreements (empty when the restatement is right):

- quadratic drag on flows varying along one axis: Bu = cD·|u|·u/H, Bv = cD·|v|·v/H
  (bottom_drag.jl:20-76);
- linear drag on a ramp, through add_drag_diff_tendencies!: u ← u + Δt_diff·rD·u
  (bottom_drag.jl:80-107, diffusion.jl:221-251);
- constant biharmonic diffusion of a separable Fourier mode: the 5-point Laplacian has the
  eigenvalue λ = (2cos θ − 2) + (2cos φ − 2), so LLu1 + LLu2 = νB·λ²·u (diffusion.jl:20-35,
  175-218);
- the Smagorinsky coefficient on shear and strain: DS = a², DT = 0 for u = a·y; DT = 0 for
  u = a·x, v = −a·y and DT = 4a² for u = a·x, v = a·y under the reference's ``+``
  (diffusion.jl:68-111); and its viscous tensor on a weak mode riding a uniform shear, where
  ν = cS·|a| to first order (diffusion.jl:126-171);
- every time scheme on the f-plane inertial oscillation of the linear dynamics: one step
  multiplies u + i·v by the scheme's stability polynomial R(W), W = −i·f·Δt, in closed form
  (time_integration.jl:53-195, constants.jl:4-27,63-75);
- surface relaxation, surface forcing and seasonal wind from rest: dη = γ·(η_ref − η),
  dη = sin(ωt)·Fη, du = sin(ωt)·Fx (continuity.jl:3-49, forcing.jl:215-217).

``test_planted_mistakes_fail_their_pin`` plants each named mistake in the oracle
(``oracle.mutation``) and requires the pin that targets it to report a disagreement.
"""
import math

import numpy as np
import pytest


def _state(swm, S, u, v, e, sst=None):
    G = S.grid
    if sst is None:
        sst = np.zeros((G.nx, G.ny))
    return swm.PrognosticVars(*swm.add_halo(u, v, e, sst, S, T="Float64"), T="Float64")


def _setup(swm, **kw):
    base = dict(T="Float64", topography="flat", wind_forcing_x="none", wind_forcing_y="none", Fx0=0.0,
                tracer_advection=False)
    P = swm.Parameter(**{**base, **kw})
    return P, swm.make_setup(P)


def _grids(G):
    """Interior index grids (0-based x, y) of the u, v and T points."""
    iu, ju = np.meshgrid(np.arange(G.nux, dtype=float), np.arange(G.nuy, dtype=float), indexing="ij")
    iv, jv = np.meshgrid(np.arange(G.nvx, dtype=float), np.arange(G.nvy, dtype=float), indexing="ij")
    ie, je = np.meshgrid(np.arange(G.nx, dtype=float), np.arange(G.ny, dtype=float), indexing="ij")
    return (iu, ju), (iv, jv), (ie, je)


# ------------------------------------------------------------------ bottom drag

def pin_quadratic_drag(swm, orc, bc):
    """Quadratic drag of u = u(y), v = 0 and of v = v(x), u = 0 on a flat bottom, η = 0, with
    eighths so every square and root is exact: Bu = cD·|u|·u/H at each u point, Bv likewise
    (cD = −Δ·cD, constants.jl:90).  Targets: Bu without cD (MUT_BU_NO_CD), sqrtKE taking
    KEu from the row below (MUT_SQRTKE_ROW)."""
    bad = []
    _, S = _setup(swm, nx=20, L_ratio=1, bc=bc, bottom_drag="quadratic", diffusion="constant")
    G, cfg = S.grid, swm.make_config(S)
    ep, H, cD = (1 if bc == "periodic" else 0), S.parameters.H, cfg.cD
    (iu, ju), (iv, jv), (ie, _) = _grids(G)
    zu, zv, ze = np.zeros_like(iu), np.zeros_like(iv), np.zeros_like(ie)
    u = (ju - G.ny / 2 + 0.5) / 8                        # u(y): rows of distinct speeds, both signs
    st = _state(swm, S, u, zv, ze)
    aux = orc.op(S, cfg, "dissipation", st.u, st.v, st.η, st.sst)[4]
    U = st.u                                             # Bu[i,j] belongs to u[i+1+ep, j+1] (bottom_drag.jl:62)
    xs = range(0, G.nx + 1) if ep else range(2, G.nx - 2)   # nonperiodic: the never-written x-halo is 0
    for i in xs:
        for j in range(3, G.ny - 1):
            uu = U[i + 1 + ep, j + 1]
            want = ((cD * abs(uu)) * uu) / H
            if aux["Bu"][i, j] != want:
                bad.append(f"Bu[{i},{j}] = {aux['Bu'][i, j]!r}, closed form {want!r}")
    v = (iv - G.nx / 2 + 0.5) / 8                        # v(x)
    st = _state(swm, S, zu, v, ze)
    aux = orc.op(S, cfg, "dissipation", st.u, st.v, st.η, st.sst)[4]
    V = st.v                                             # Bv[i,j] belongs to v[i+1, j+1] (bottom_drag.jl:73)
    for i in range(2, G.nx):
        for j in range(2, G.ny - 2):
            vv = V[i + 1, j + 1]
            want = ((cD * abs(vv)) * vv) / H
            if aux["Bv"][i, j] != want:
                bad.append(f"Bv[{i},{j}] = {aux['Bv'][i, j]!r}, closed form {want!r}")
    return bad


def pin_linear_drag_update(swm, orc, bc):
    """Linear drag on ramps with no viscosity (νB = 0, so LLu = LLv = 0): each interior u
    becomes u + Δt_diff·(rD·u) of its own point (diffusion.jl:237,248).  Targets: the update
    reading Bu of the neighbouring column (MUT_ADD_BU_SHIFT)."""
    bad = []
    _, S = _setup(swm, nx=18, L_ratio=1.5, bc=bc, bottom_drag="linear", diffusion="constant")
    G, cfg = S.grid, swm.make_config(S)
    cfg.nuB = 0.0
    (iu, ju), (iv, jv), (ie, _) = _grids(G)
    u = (iu + 2 * ju + 1) / 16
    v = (3 * iv - jv + 2) / 16
    st = _state(swm, S, u, v, np.zeros_like(ie))
    nu, nv = orc.op(S, cfg, "dissipation", st.u, st.v, st.η, st.sst)[:2]
    dtd, rD = cfg.Dt_diff, cfg.rD
    want_u = u + dtd * (rD * u)
    want_v = v + dtd * (rD * v)
    got_u, got_v = nu[2:-2, 2:-2], nv[2:-2, 2:-2]
    if not np.array_equal(got_u, want_u):
        bad.append(f"u: {np.count_nonzero(got_u != want_u)} points differ from u + Δt_diff·rD·u")
    if not np.array_equal(got_v, want_v):
        bad.append(f"v: {np.count_nonzero(got_v != want_v)} points differ from v + Δt_diff·rD·v")
    return bad


# ------------------------------------------------------------------ biharmonic diffusion

def _mode(G, θ, φ):
    """A separable mode on the u and v points and the 5-point Laplacian's eigenvalue."""
    (iu, ju), (iv, jv), _ = _grids(G)
    u = np.cos(θ * iu + 0.4) * np.cos(φ * ju + 0.1)
    v = 0.7 * np.sin(θ * iv - 0.2) * np.cos(φ * jv + 0.3)
    lam = (2 * math.cos(θ) - 2) + (2 * math.cos(φ) - 2)
    return u, v, lam


def _clear(G, bc, r=3):
    """Slices of the interior u and v points at least r points from walls."""
    ys = slice(r, -r)
    xs = slice(None) if bc == "periodic" else slice(r, -r)
    return (xs, ys)


def pin_constant_biharmonic(swm, orc, bc):
    """Constant biharmonic diffusion of u = cos(θx+a)·cos(φy+b) (and a v mode): away from the
    walls the increment is Δt_diff·νB·λ²·u with λ = (2cos θ − 2) + (2cos φ − 2), the
    eigenvalue of the 5-point Laplacian, since LLu1 + LLu2 = νB(δxx + δyy)∇²u.  Targets:
    S12 without the ep shift (MUT_S12_NO_EP, periodic), S11 read from the next row
    (MUT_S11_ROW)."""
    bad = []
    _, S = _setup(swm, nx=32, L_ratio=1, bc=bc, bottom_drag="none", diffusion="constant")
    G, cfg = S.grid, swm.make_config(S)
    cfg.nuB, cfg.Dt_diff = -0.05, 1.0
    u, v, lam = _mode(G, 2 * math.pi * 3 / G.nx, 0.61)
    st = _state(swm, S, u, v, np.zeros((G.nx, G.ny)))
    nu, nv = orc.op(S, cfg, "dissipation", st.u, st.v, st.η, st.sst)[:2]
    c = _clear(G, bc)
    for name, new, old in (("u", nu[2:-2, 2:-2], u), ("v", nv[2:-2, 2:-2], v)):
        want = cfg.Dt_diff * cfg.nuB * lam ** 2 * old[c]
        err = np.max(np.abs((new - old)[c] - want)) / np.max(np.abs(want))
        if not err < 1e-12:
            bad.append(f"{name}: biharmonic increment off its closed form by {err:.2e} (relative)")
    return bad


def pin_smagorinsky_coefficient(swm, orc, bc):
    """Smagorinsky's coefficient on linear flows (a = 1/8, exact squares and roots):
    shear u = a·y gives DS = a², DT = 0, ν = cS·a; strain u = a·x, v = −a·y gives DT = 0 with
    the reference's tension (dudx + dvdy)² (diffusion.jl:95, which the docstring at :66
    writes with −); u = a·x, v = a·y gives DT = 4a², ν = 2·cS·a.  Targets: DT with − (MUT_DT_MINUS)."""
    bad = []
    _, S = _setup(swm, nx=24, L_ratio=1, bc=bc, bottom_drag="none", diffusion="Smagorinsky")
    G, cfg = S.grid, swm.make_config(S)
    a, cS = 0.125, cfg.cSmag
    (iu, ju), (iv, jv), (ie, _) = _grids(G)
    zu, zv, ze = np.zeros_like(iu), np.zeros_like(iv), np.zeros_like(ie)
    # T points clear of walls, and of the seam where an x-ramp jumps (periodic)
    xs = slice(5, -5) if bc == "periodic" else slice(3, -3)
    cases = (("shear", a * ju, zv, a * a, 0.0, cS * a, slice(3, -3)),
             ("strain", a * iu, -a * jv, 0.0, 0.0, None, xs),
             ("dilation", a * iu, a * jv, 0.0, 4 * a * a, cS * (2 * a), xs))
    for name, u, v, ds, dt, nu, x in cases:
        st = _state(swm, S, u, v, ze)
        aux = orc.op(S, cfg, "dissipation", st.u, st.v, st.η, st.sst)[4]
        box = (x, slice(3, -3))
        if not np.all(aux["DT"][box] == dt):
            bad.append(f"{name}: DT = {np.unique(aux['DT'][box])[:4]}, closed form {dt}")
        if not np.all(aux["DS"][box] == ds):
            bad.append(f"{name}: DS = {np.unique(aux['DS'][box])[:4]}, closed form {ds}")
        want_nu = nu if nu is not None else 0.0
        if not np.all(aux["nuSmag"][box] == want_nu):
            bad.append(f"{name}: νSmag = {np.unique(aux['nuSmag'][box])[:4]}, closed form {want_nu}")
    return bad


def pin_smagorinsky_tensor(swm, orc, bc):
    """Smagorinsky's viscous tensor on a weak mode ε·m riding the shear u = a·y: the shear
    fixes ν = cS·a up to O(ε/a), so the increment is Δt_diff·cS·a·λ²·ε·m to first order in ε
    (the shear itself has no biharmonic).  Targets: S12 without the ep shift
    (MUT_S12_NO_EP, periodic)."""
    bad = []
    _, S = _setup(swm, nx=32, L_ratio=1, bc=bc, bottom_drag="none", diffusion="Smagorinsky")
    G, cfg = S.grid, swm.make_config(S)
    cfg.cSmag, cfg.Dt_diff = -0.5, 1.0
    a, eps = 0.125, 1e-6
    mu, mv, lam = _mode(G, 2 * math.pi * 3 / G.nx, 0.61)
    (iu, ju), _, _ = _grids(G)
    u = a * ju + eps * mu
    v = eps * mv
    st = _state(swm, S, u, v, np.zeros((G.nx, G.ny)))
    nu, nv = orc.op(S, cfg, "dissipation", st.u, st.v, st.η, st.sst)[:2]
    c = _clear(G, bc, r=4)
    nu0 = abs(cfg.cSmag) * a * np.sign(cfg.cSmag)
    for name, new, old, m in (("u", nu[2:-2, 2:-2], u, mu), ("v", nv[2:-2, 2:-2], v, mv)):
        want = cfg.Dt_diff * nu0 * lam ** 2 * eps * m[c]
        err = np.max(np.abs((new - old)[c] - want)) / np.max(np.abs(want))
        if not err < 1e-3:
            bad.append(f"{name}: Smagorinsky increment off its first-order form by {err:.2e} (relative)")
    return bad


# ------------------------------------------------------------------ time schemes

def _poly(coef):
    return np.polynomial.Polynomial(coef)


def stability_polynomial(scheme, k):
    """Published stability polynomials R(W) of the reference's schemes (W = λΔt):
    classic RK of order p = k (2, 3, 4 stages): the truncated exponential;
    SSPRK(s,2) (Ketcheson 2008, s = k stages of Δt/(s−1)): 1/s + (s−1)/s·(1 + W/(s−1))^s;
    SSPRK(n²,3) (Ketcheson 2008; Ketcheson, Lóczi & Parsani 2014, n = k):
    (n−1)/(2n−1)·a^{n²} + n/(2n−1)·a^{(n−1)²}, a = 1 + W/(n² − n);
    SSPRK(4,3) (Kraaijevanger 1991): 1 + W + W²/2 + W³/6 + W⁴/48."""
    W = _poly([0.0, 1.0])
    if scheme == "RK":
        return sum((W ** j) * (1.0 / math.factorial(j)) for j in range(k + 1)), k
    if scheme == "SSPRK2":
        s = k
        return (1.0 / s) + ((s - 1.0) / s) * (1 + W * (1.0 / (s - 1))) ** s, 2
    if scheme == "SSPRK3":
        n = k
        a = 1 + W * (1.0 / (n * n - n))
        return ((n - 1.0) / (2 * n - 1)) * a ** (n * n) + (n / (2.0 * n - 1)) * a ** ((n - 1) ** 2), 3
    return _poly([1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 48]), 3


SCHEMES = [("RK", 2), ("RK", 3), ("RK", 4), ("SSPRK2", 2), ("SSPRK2", 3), ("SSPRK2", 4), ("SSPRK3", 2),
           ("SSPRK3", 3), ("4SSPRK3", 0)]


def pin_time_scheme(swm, orc, scheme, k):
    """Inertial oscillation on the f-plane (ϕ = 90°, linear dynamics, no wind, uniform u, v,
    η = 0): away from the walls du = f·v, dv = −f·u, so one step multiplies u + i·v by R(W)
    with W = −i·f·Δt, R the scheme's stability polynomial.  Targets: u1 built from u1 instead
    of u (MUT_RK_U1_FROM_U1), SSPRK2's weights swapped (MUT_SSPRK2_WEIGHTS), SSPRK3's kn
    weights swapped (MUT_SSPRK3_KN), 4SSPRK3's thirds swapped (MUT_4SSPRK3_THIRDS)."""
    bad = []
    R, order = stability_polynomial(scheme, k)
    c = R.coef
    for j in range(order + 1):                       # the closed form has the scheme's order
        if abs(c[j] - 1.0 / math.factorial(j)) > 1e-15:
            bad.append(f"closed form of {scheme}{k} fails the order condition at W^{j}")
    kw = {"RK": {"RKo": k}, "SSPRK2": {"RKs": k}, "SSPRK3": {"RKn": k}, "4SSPRK3": {}}[scheme]
    _, S = _setup(swm, nx=16, L_ratio=0.25, bc="periodic", ϕ=90.0, dynamics="linear", bottom_drag="none",
                  nstep_diff=2, time_scheme=scheme, **kw)
    G, cfg = S.grid, swm.make_config(S)
    assert G.ny == 64 and np.ptp(G.f_u) == 0 and np.ptp(G.f_v) == 0
    U0, V0 = 0.3, -0.2
    (iu, _), (iv, _), (ie, _) = _grids(G)
    st = _state(swm, S, np.full_like(iu, U0), np.full_like(iv, V0), np.zeros_like(ie))
    u, v = orc.run(S, cfg, st, 1)[:2]
    W = -1j * float(G.f_u.flat[0]) * S.constants.Δt_Δ
    z = R(W) * complex(U0, V0)
    rows = slice(2 + 26, 2 + 38)                     # 26 rows clear of either wall
    for name, got, want in (("u", u[2:-2, rows], z.real), ("v", v[2:-2, rows], z.imag)):
        err = np.max(np.abs(got - want)) / abs(z)
        if not err < 1e-13:
            bad.append(f"{scheme}{k}: {name} after one step off R(W)·(u0 + i·v0) by {err:.2e}")
    return bad


# ------------------------------------------------------------------ continuity variants, seasonal wind

def pin_relaxation(swm, orc, bc):
    """Surface relaxation from rest with η = η_ref + 1/4 (η_ref random sixteenths): the stage
    tendency is dη = γ·(η_ref − η) = −γ/4 at every interior point (continuity.jl:21).
    Targets: γ·(η − η_ref) (MUT_RELAX_SIGN)."""
    _, S = _setup(swm, nx=20, L_ratio=1.25, bc=bc, surface_relax=True)
    G, cfg = S.grid, swm.make_config(S)
    rng = np.random.default_rng(4)
    S.forcing.η_ref = np.asfortranarray(rng.integers(-16, 17, (G.nx, G.ny)) / 16.0)
    (iu, _), (iv, _), _ = _grids(G)
    st = _state(swm, S, np.zeros_like(iu), np.zeros_like(iv), S.forcing.η_ref + 0.25)
    de = orc.run(S, cfg, st, 1, stop_stage=1, want_rk=True)[4][8]
    want = cfg.gamma * -0.25
    if not np.all(de[1:-1, 1:-1] == want):
        return [f"dη = {np.unique(de[1:-1, 1:-1])[:4]}, closed form γ·(η_ref − η) = {want}"]
    return []


def pin_surface_forcing(swm, orc, bc):
    """Surface forcing from rest at time t: dη = sin(ωt)·Fη (continuity.jl:27-49 with
    Ftime, forcing.jl:215-217).  Targets: cos in place of sin (MUT_FTIME_COS)."""
    _, S = _setup(swm, nx=20, L_ratio=1.25, bc=bc, surface_forcing=True, ωFη=450.0)   # ωt ≈ −0.9 at t = 7Δt
    G, cfg = S.grid, swm.make_config(S)
    rng = np.random.default_rng(5)
    S.forcing.Fη = np.asfortranarray(rng.standard_normal((G.nx, G.ny)))
    (iu, _), (iv, _), (ie, _) = _grids(G)
    st = _state(swm, S, np.zeros_like(iu), np.zeros_like(iv), np.zeros_like(ie))
    t = 7 * G.dtint
    de = orc.run(S, cfg, st, 1, t0=t, stop_stage=1, want_rk=True)[4][8]
    want = math.sin(cfg.omegaFeta * t) * S.forcing.Fη
    err = np.max(np.abs(de[1:-1, 1:-1] - want)) / np.max(np.abs(want))
    return [] if err < 1e-15 else [f"dη off sin(ωt)·Fη by {err:.2e}"]


def pin_seasonal_wind(swm, orc, bc):
    """Seasonal zonal wind from rest at time t: du = sin(ωt)·Fx at every interior u point
    (rhs.jl:233-242, forcing.jl:215-217).  Targets: cos in place of sin (MUT_FTIME_COS)."""
    _, S = _setup(swm, nx=20, L_ratio=1.25, bc=bc, wind_forcing_x="double_gyre", wind_forcing_y="constant",
                  Fx0=0.12, seasonal_wind_x=True, ωFx=300.0)   # ωt ≈ 0.9 at t = 11Δt
    G, cfg = S.grid, swm.make_config(S)
    (iu, _), (iv, _), (ie, _) = _grids(G)
    st = _state(swm, S, np.zeros_like(iu), np.zeros_like(iv), np.zeros_like(ie))
    t = 11 * G.dtint
    du = orc.run(S, cfg, st, 1, t0=t, stop_stage=1, want_rk=True)[4][6]
    want = math.sin(cfg.omegaFx * t) * S.forcing.Fx
    err = np.max(np.abs(du[2:-2, 2:-2] - want)) / np.max(np.abs(want))
    return [] if err < 1e-15 else [f"du off sin(ωt)·Fx by {err:.2e}"]


# ------------------------------------------------------------------ the tests

BCS = ["periodic", "nonperiodic"]


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("pin", [pin_quadratic_drag, pin_linear_drag_update, pin_constant_biharmonic,
                                 pin_smagorinsky_coefficient, pin_smagorinsky_tensor, pin_relaxation,
                                 pin_surface_forcing, pin_seasonal_wind], ids=lambda f: f.__name__[4:])
def test_dissipation_and_forcing_pins(swm, orc, pin, bc):
    bad = pin(swm, orc, bc)
    assert not bad, "\n".join(bad[:10])


@pytest.mark.parametrize("scheme,k", SCHEMES, ids=[f"{s}{k or ''}" for s, k in SCHEMES])
def test_time_scheme_pins(swm, orc, scheme, k):
    bad = pin_time_scheme(swm, orc, scheme, k)
    assert not bad, "\n".join(bad)


# (mutation, pin, arguments): every planted mistake must make its pin fail
PLANTED = [
    (1, pin_smagorinsky_coefficient, ("nonperiodic",)),
    (1, pin_smagorinsky_coefficient, ("periodic",)),
    (2, pin_constant_biharmonic, ("periodic",)),
    (2, pin_smagorinsky_tensor, ("periodic",)),
    (3, pin_quadratic_drag, ("periodic",)),
    (3, pin_quadratic_drag, ("nonperiodic",)),
    (4, pin_quadratic_drag, ("nonperiodic",)),
    (5, pin_linear_drag_update, ("periodic",)),
    (5, pin_linear_drag_update, ("nonperiodic",)),
    (6, pin_time_scheme, ("RK", 4)),
    (6, pin_time_scheme, ("RK", 3)),
    (7, pin_time_scheme, ("SSPRK2", 3)),
    (8, pin_time_scheme, ("SSPRK3", 2)),
    (8, pin_time_scheme, ("SSPRK3", 3)),
    (9, pin_time_scheme, ("4SSPRK3", 0)),
    (10, pin_relaxation, ("nonperiodic",)),
    (11, pin_surface_forcing, ("periodic",)),
    (11, pin_seasonal_wind, ("nonperiodic",)),
    (12, pin_constant_biharmonic, ("nonperiodic",)),
]


@pytest.mark.parametrize("mut,pin,args", PLANTED, ids=[f"m{m}-{p.__name__[4:]}-{'-'.join(map(str, a))}"
                                                        for m, p, a in PLANTED])
def test_planted_mistakes_fail_their_pin(swm, orc, mut, pin, args):
    with orc.mutation(mut):
        bad = pin(swm, orc, *args)
    assert bad, f"mistake {mut} planted in the oracle went unnoticed by {pin.__name__}{args}"
    assert not pin(swm, orc, *args)        # and the pin passes again once it is removed
