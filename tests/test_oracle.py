"""Pins of the CPU oracle (the restatement every GPU result is checked against).

Julia is not available, so the reference cannot be run.  Its only value-level known answer
is test/runtests.jl:4-9 (no forcing from rest stays exactly zero); everything else here is
fixed by mathematics the reference's algorithm implies: closed forms of each stencil,
telescoping mass conservation, the Coriolis balance of a uniform flow, exact tracer
shifts, and the IEEE formats the mixed-precision paths emulate.  Each check is chosen so a
dropped term, a wrong sign, a shifted index or a transposed operand fails it.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2


# ------------------------------------------------------------------ stencils (gradients.jl, interpolations.jl)

def test_stencils_closed_forms(orc):
    m, n = 9, 7
    i = np.arange(1, m + 1, dtype=np.float64)[:, None]
    j = np.arange(1, n + 1, dtype=np.float64)[None, :]
    ramp = np.asfortranarray(3.0 * i - 5.0 * j + 0.0 * i * j)
    assert np.all(orc.stencil("dx", ramp) == 3.0)
    assert np.all(orc.stencil("dy", ramp) == -5.0)
    # Ix/Iy of a ramp: the midpoint value
    assert np.array_equal(orc.stencil("Ix", ramp), 3.0 * (i[:-1] + 0.5) - 5.0 * j)
    assert np.array_equal(orc.stencil("Iy", ramp), 3.0 * i - 5.0 * (j[:, :-1] + 0.5))
    # Ixy of a bilinear function is its value at the cell centre
    f = np.asfortranarray(i * j)
    assert np.array_equal(orc.stencil("Ixy", f), (i[:-1] + 0.5) * (j[:, :-1] + 0.5))
    # ∇² (resilient form) of x² + y² is 4, of x·y is 0; the operator is asymmetric-safe
    q = np.asfortranarray(i ** 2 + j ** 2)
    assert np.all(orc.stencil("lap", q) == 4.0)
    assert np.all(orc.stencil("lap", np.asfortranarray(i * j + 0 * q)) == 0.0)
    # orientation: ∂x differentiates along the first (column-major fastest) index
    g = np.asfortranarray(i ** 2 + 0 * j)
    assert np.array_equal(orc.stencil("dx", g), (2 * i[:-1] + 1) + 0 * j)
    assert np.all(orc.stencil("dy", g) == 0)


# ------------------------------------------------------------------ emulated formats

def test_binary16_matches_ieee(orc):
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-9, 5, 4000),
                        [65504.0, 65519.99, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8, -0.0, 0.0]])
    assert np.array_equal(orc.round_fmt(2, x), x.astype(np.float16).astype(np.float64))
    a = rng.standard_normal(4000).astype(np.float16).astype(np.float64) * 3
    b = rng.standard_normal(4000).astype(np.float16).astype(np.float64) + 4
    a16, b16 = a.astype(np.float16), b.astype(np.float16)
    for op, ref in ((0, a16 + b16), (1, a16 - b16), (2, a16 * b16), (3, a16 / b16)):
        assert np.array_equal(orc.arith(2, op, a, b), ref.astype(np.float64)), op
    assert np.array_equal(orc.arith(2, 4, b, b), np.sqrt(b16).astype(np.float64))


def _bf16_exact(x: float) -> float:
    """Round-to-nearest-even of a float64 to bfloat16 with exact rational arithmetic."""
    from fractions import Fraction
    if x == 0:
        return x
    fx = Fraction(x)
    e = math.floor(math.log2(abs(x)))
    while Fraction(2) ** e > abs(fx):
        e -= 1
    while Fraction(2) ** (e + 1) <= abs(fx):
        e += 1
    q = Fraction(2) ** (max(e, -126) - 7)
    k = fx / q
    fl = math.floor(k)
    r = k - fl
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return float(fl * q)


def test_bfloat16_rounding_is_correct(orc):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(3000) * 10.0 ** rng.integers(-30, 30, 3000)
    # values on and next to the bf16 ties, where double rounding via float32 would fail
    t = np.array([1.0 + 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -30, 1.0 + 3 * 2.0 ** -8, 3.0 + 2.0 ** -7 + 2.0 ** -40])
    x = np.concatenate([x, t, -t])
    got = orc.round_fmt(3, x)
    want = np.array([_bf16_exact(float(v)) for v in x])
    assert np.array_equal(got, want)


def test_host_bfloat16_storage_matches_oracle(swm, orc):
    from shallowwaters_amd.numtypes import to_storage, to_float64
    rng = np.random.default_rng(3)
    x = rng.standard_normal(5000) * 10.0 ** rng.integers(-20, 20, 5000)
    assert np.array_equal(to_float64(to_storage(x, "BFloat16"), "BFloat16"), orc.round_fmt(3, x))
    assert np.array_equal(to_float64(to_storage(x, "Float16"), "Float16"), orc.round_fmt(2, x))


# ------------------------------------------------------------------ whole step

def _run(swm, orc, P, prog=None, nsteps=1, **kw):
    S = swm.make_setup(P)
    if prog is None:
        prog = swm.initial_conditions(P.Tprog, S)
    return S, orc.run(S, swm.make_config(S), prog, nsteps, **kw)


def test_known_answer_no_forcing_stays_zero(swm, orc):
    """test/runtests.jl:4-9 (run_model(Ndays=1,Fx0=0), Float32, periodic 100x50)."""
    P = swm.Parameter(Ndays=1, Fx0=0)
    S, r = _run(swm, orc, P, nsteps=swm.make_setup(P).grid.nt)
    assert r[5] == S.grid.nt
    for a in r[:3]:
        assert np.all(a == 0)


@pytest.mark.parametrize("bc,T", [("nonperiodic", "Float64"), ("periodic", "Float64"), ("nonperiodic", "Float32")])
def test_rest_state_tendencies_are_the_wind(swm, orc, bc, T):
    """From rest every tendency but the wind vanishes: du = Fx at every interior u point
    (momentum_u!, rhs.jl:242), dv = Fy, dη = 0 -- pins the forcing's index mapping."""
    P = swm.Parameter(T=T, nx=23, L_ratio=0.7, bc=bc, wind_forcing_x="double_gyre", topography="seamount",
                      Fy0=0.3, wind_forcing_y="constant")
    S, r = _run(swm, orc, P, stop_stage=1, want_rk=True)
    du, dv, de = r[4][6:9]
    F = S.forcing
    assert np.array_equal(du[2:-2, 2:-2], F.Fx) and np.all(du[:2] == 0) and np.all(du[:, -2:] == 0)
    assert np.array_equal(dv[2:-2, 2:-2], F.Fy) and np.any(F.Fy != 0)
    assert np.all(de == 0)


def test_uniform_zonal_flow_coriolis_balance(swm, orc):
    """u = U0, v = η = 0, flat bottom, f constant (ϕ = 90°), no wind: the Arakawa-Hsu
    weights sum to q, so dv = -f·U0·Δ (Coriolis), du = 0 and dη = 0 exactly."""
    P = swm.Parameter(T="Float64", nx=24, L_ratio=1, bc="periodic", ϕ=90.0, topography="flat",
                      wind_forcing_x="none", wind_forcing_y="none", Fx0=0.0)
    S = swm.make_setup(P)
    G = S.grid
    assert G.β == 0.0
    U0 = 0.37
    u = np.full((G.nux, G.nuy), U0)
    prog = swm.PrognosticVars(*swm.add_halo(u, np.zeros((G.nvx, G.nvy)), np.zeros((G.nx, G.ny)),
                                            np.zeros((G.nx, G.ny)), S, T="Float64"), T="Float64")
    r = orc.run(S, swm.make_config(S), prog, 1, stop_stage=1, want_rk=True)
    du, dv, de = r[4][6:9]
    assert np.all(du == 0) and np.all(de == 0)
    # rows next to the walls of the periodic channel see the never-written u y-halo (0)
    inner = dv[2:-2, 3:-3]
    want = -G.f0 * G.Δ * U0
    assert np.allclose(inner, want, rtol=1e-14, atol=0)
    assert rel_l2(np.full_like(inner, want), inner) < 1e-15


@pytest.mark.parametrize("bc,topo", [("nonperiodic", "seamount"), ("periodic", "flat")])
def test_mass_is_conserved(swm, orc, bc, topo):
    """continuity! telescopes: Σ dη over the interior is -(boundary fluxes), zero through
    the never-written normal-velocity halo (nonperiodic) or cancelling through the periodic
    copies (flat bottom).  Dissipation does not touch η.  Drift stays at rounding level."""
    P = swm.Parameter(T="Float64", nx=40, L_ratio=1.25, bc=bc, topography=topo, wind_forcing_x="double_gyre",
                      diffusion="Smagorinsky", α=0.5, Uadv=10.0)
    S = swm.make_setup(P)
    prog = swm.seeded_initial_conditions(S, seed=5, σ=0.5)
    m0 = prog.η[1:-1, 1:-1].sum()
    r = orc.run(S, swm.make_config(S), prog, 60)
    m1 = r[2][1:-1, 1:-1].sum()
    scale = np.abs(r[2][1:-1, 1:-1]).sum()
    assert abs(m1 - m0) < 1e-12 * scale, (m0, m1, scale)
    # and the flow evolved
    assert rel_l2(prog.η, r[2]) > 1e-2


def shifted_channel(swm, k, nx=96, seed=3):
    """Periodic channel on a flat bottom: every static field (H, Fx(y), Fy, f_q(y)) is
    x-invariant.  Returns the seeded state and the same state circularly shifted by k
    columns."""
    P = swm.Parameter(T="Float64", nx=nx, L_ratio=3, bc="periodic", topography="flat", wind_forcing_x="channel",
                      diffusion="Smagorinsky", bottom_drag="quadratic")
    S = swm.make_setup(P)
    prog = swm.seeded_initial_conditions(S, seed=seed, σ=0.6)
    ui, vi, ei, si = swm.remove_halo(prog.u, prog.v, prog.η, prog.sst, S)
    roll = lambda a: np.roll(a, k, axis=0)
    U, V, E, SS = swm.add_halo(roll(ui), roll(vi), roll(ei), roll(si), S, T="Float64")
    shifted = swm.PrognosticVars(u=U, v=V, η=E, sst=SS, T="Float64")
    return P, S, prog, shifted


def check_shift_equivariance(swm, S, a, b, k, far=20):
    """One step from the shifted state equals the shifted step, bitwise, in every column
    at least `far` columns from both periodic seams; near a seam it does not.  The
    reference's AHβ!/AHγ! take the left neighbour of their first column from q[end,·]
    (PV_advection.jl:110,126), the corner x = nx rather than x = nx-1, so the seam is not
    translation invariant.  Any other position-dependent error (a wrong ghost copy, a
    mishandled strip or segment boundary of the GPU tiling) breaks the far columns."""
    nx = S.grid.nx
    ai = swm.remove_halo(*a, S)
    bi = swm.remove_halo(*b, S)
    # interior column c of u is x = c (ep = 1), of v, η, sst x = c + 1; seams at x = 0 and x = -k
    seams = (0, (-k) % nx)
    near = 0
    for name, x, y, off in zip(("u", "v", "η"), ai, bi, (0, 1, 1)):
        want = np.roll(x, k, axis=0)
        for c in range(nx):
            xo = (c + off - k) % nx                     # column in the unshifted frame
            dist = min(min((xo - s0) % nx, (s0 - xo) % nx) for s0 in seams)
            same = np.array_equal(want[c], y[c])
            if dist >= far:
                assert same, (name, c, xo, dist)
            near += not same
    assert near > 0        # the seam quirk is reproduced, not smoothed away
    return near


def test_x_shift_equivariance(swm, orc):
    k = 48
    P, S, prog, shifted = shifted_channel(swm, k)
    cfg = swm.make_config(S)
    a = orc.run(S, cfg, prog, 1)[:4]
    b = orc.run(S, cfg, shifted, 1)[:4]
    check_shift_equivariance(swm, S, a, b, k)


# ------------------------------------------------------------------ tracer (tracer_advection.jl)

def _tracer_setup(swm, bc):
    # c = sqrt(g H) = 64 m/s, Δ = 4096 m -> dt = 64 s; Uadv = 16 -> nadvstep = 4,
    # dtadvu = 256/4096 = 1/16: a uniform u = 16 m/s moves the tracer exactly one cell
    P = swm.Parameter(T="Float64", nx=16, Lx=16 * 4096.0, L_ratio=2, g=8.192, H=500.0, bc=bc, Uadv=16.0)
    S = swm.make_setup(P)
    G = S.grid
    assert (G.dtint, G.nadvstep, G.dtadvu, G.half_dtadvu) == (64, 4, 1 / 16, 1 / 32)
    return P, S, G


@pytest.mark.parametrize("bc", ["periodic", "nonperiodic"])
def test_tracer_uniform_flow_is_an_exact_shift(swm, orc, bc):
    P, S, G = _tracer_setup(swm, bc)
    rng = np.random.default_rng(7)
    sst0 = rng.random((G.nx, G.ny))
    u = np.full((G.nux, G.nuy), 16.0)
    U, V, E, SS = swm.add_halo(u, np.zeros((G.nvx, G.nvy)), np.zeros((G.nx, G.ny)), sst0, S, T="Float64")
    cfg = swm.make_config(S)
    um, vm, sst = orc.tracer(cfg, 2, U, V, np.zeros_like(U), np.zeros_like(V), SS)   # snapshot step
    assert np.array_equal(um, U) and np.array_equal(sst, SS)
    um, vm, sst = orc.tracer(cfg, 4, U, V, um, vm, sst)                                 # advection step
    if bc == "periodic":
        assert np.array_equal(sst[1:-1, :], np.roll(sst0, 1, axis=0))
    else:   # walls: the face velocity is 0, so only columns 3..nx-1 see a uniform flow
        assert np.array_equal(sst[3:-2, :], sst0[1:-2, :])
    # zero velocity leaves the tracer bitwise unchanged (floor(-0.0) = 0, x0 = -0.0)
    Z = np.zeros_like(U)
    _, _, s2 = orc.tracer(cfg, 4, Z, np.zeros_like(V), Z, np.zeros_like(V), SS)
    assert np.array_equal(s2, SS)


def test_tracer_linear_field_reproduced_by_bilinear(swm, orc):
    """A tracer linear in x and y is reproduced exactly along a uniform shift of half a cell
    in both directions (bilin of a linear function is exact)."""
    P, S, G = _tracer_setup(swm, "periodic")
    x = np.arange(G.nx, dtype=np.float64)[:, None]
    y = np.arange(G.ny, dtype=np.float64)[None, :]
    sst0 = np.asfortranarray(0.25 * x + 0.5 * y)
    u = np.full((G.nux, G.nuy), 8.0)          # half a cell per advection step
    v = np.full((G.nvx, G.nvy), 8.0 * G.dtadvu / G.dtadvv)
    U, V, E, SS = swm.add_halo(u, v, np.zeros((G.nx, G.ny)), sst0, S, T="Float64")
    cfg = swm.make_config(S)
    um, vm, _ = orc.tracer(cfg, 2, U, V, np.zeros_like(U), np.zeros_like(V), SS)
    _, _, sst = orc.tracer(cfg, 4, U, V, um, vm, SS)
    # interior columns away from the periodic seam; rows away from the clipped walls
    got = sst[3:-3, 3:-3]
    want = sst0[2:-2, 3:-3] - 0.25 * 0.5 - 0.5 * 0.5
    assert np.allclose(got, want, atol=1e-13)


# ------------------------------------------------------------------ committed fixtures (regression pin)

def test_golden_fixtures(swm, orc):
    import os
    from golden.make_fixtures import CASES, run_case
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for name in CASES:
        ref = np.load(os.path.join(here, f"{name}.npz"))
        got = run_case(swm, orc, name)
        for k in ("u", "v", "eta", "sst"):
            assert np.array_equal(ref[k], got[k]), (name, k)


# ------------------------------------------------------------------ low-precision options

@pytest.mark.parametrize("bc", ["nonperiodic", "periodic"])
@pytest.mark.parametrize("T", ["Float64", "Float32"])
def test_scaled_velocities_change_nothing_at_float32_and_float64(swm, orc, bc, T):
    """u, v carried as 2^6·u, 2^6·v: every factor 1/s sits on an exact operation (powers of
    two), so the run is bitwise the unscaled one wherever nothing under- or overflows.  A
    misplaced or missing factor breaks this (PV weights, Bernoulli ½ and g, divergence,
    drag speed, Smagorinsky ν, tracer departure, forcing, f)."""
    kw = dict(T=T, nx=40, L_ratio=0.8, bc=bc, α=0.5, diffusion="Smagorinsky", bottom_drag="quadratic",
              topography="seamount", wind_forcing_x="double_gyre", Uadv=20.0, tracer_consumption=True)
    runs = []
    for sc in (1.0, 64.0):
        S = swm.make_setup(swm.Parameter(scale=sc, **kw))
        prog = swm.seeded_initial_conditions(S, seed=3, σ=0.5)
        assert S.grid.nadvstep <= 5
        runs.append(orc.run(S, swm.make_config(S), prog, 10)[:4])
    for name, a, b in zip(("u", "v", "η", "sst"), *runs):
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("bc", ["nonperiodic", "periodic"])
def test_compensated_increments_track_float64_closer(swm, orc, bc):
    """Kahan-compensated RK increments: a Float32 run stays closer to the Float64 run than
    the plain Float32 run (200 steps; the remaining error is the Float32 right-hand side)."""
    kw = dict(nx=48, L_ratio=0.8, bc=bc, α=0.5, diffusion="Smagorinsky", bottom_drag="quadratic",
              topography="seamount", wind_forcing_x="double_gyre", Uadv=20.0)
    n = 200
    S64 = swm.make_setup(swm.Parameter(T="Float64", **kw))
    r64 = orc.run(S64, swm.make_config(S64), swm.seeded_initial_conditions(S64, seed=3, σ=0.5), n)[:3]
    err = {}
    for comp in (False, True):
        S = swm.make_setup(swm.Parameter(T="Float32", compensated=comp, **kw))
        r = orc.run(S, swm.make_config(S), swm.seeded_initial_conditions(S, seed=3, σ=0.5), n)[:3]
        err[comp] = [np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b) for a, b in zip(r, r64)]
    for k, name in enumerate(("u", "v", "η")):
        assert err[True][k] < err[False][k], f"{name}: compensated {err[True][k]:.2e} vs plain {err[False][k]:.2e}"
    assert err[True][0] < 0.8 * err[False][0]


# ------------------------------------------------------------------ PV advection schemes (PV_advection.jl)

def _nondivergent_fluxes(nx, ny, seed, margin=3):
    """q at the corners (nx+1, ny+1) and mass fluxes from a streamfunction ψ at the corners,
    U = -δy ψ on the u-points (nx+1, ny+2), V = δx ψ on the v-points (nx+2, ny+1), so the
    discrete divergence δx U + δy V vanishes exactly in every cell.  ψ is integer-valued and
    zero within ``margin`` cells of the walls, so U, V and both PV flux terms vanish there."""
    rng = np.random.default_rng(seed)
    P = np.zeros((nx + 5, ny + 6))                     # ψ at corner x = -2..nx+2, y = -3..ny+2
    P[2 + margin:nx + 3 - margin, 3 + margin:ny + 4 - margin] = rng.integers(-8, 9, (nx + 1 - 2 * margin, ny + 1 - 2 * margin))
    psi = lambda x, y: P[x + 2, y + 3]
    iU, jU = np.meshgrid(np.arange(1, nx + 2), np.arange(1, ny + 3), indexing="ij")   # U(i,j) at (i-1, j-1.5)
    U = -(psi(iU - 1, jU - 1) - psi(iU - 1, jU - 2))
    iV, jV = np.meshgrid(np.arange(1, nx + 3), np.arange(1, ny + 2), indexing="ij")   # V(i,j) at (i-1.5, j-1)
    V = psi(iV - 1, jV - 1) - psi(iV - 2, jV - 1)
    q = rng.standard_normal((nx + 1, ny + 1))                                          # q(i,j) at (i-1, j-1)
    return q, U, V


def _pv_invariants(q, qhv, qhu, U, V, ep):
    """(Σ q ζ_t, scale) and (Σ U·qhv - Σ V·qhu, scale) for du = qhv, dv = -qhu: the PV flux
    term's enstrophy and energy tendencies.  qhv(i,j) sits at u-point (i-ep, j-0.5), qhu(i,j)
    at v-point (i-0.5, j)."""
    nx, ny = q.shape[0] - 1, q.shape[1] - 1
    a, b = np.meshgrid(np.arange(1, nx), np.arange(1, ny), indexing="ij")             # interior corners
    zt = -(qhu[a, b - 1] - qhu[a - 1, b - 1]) - (qhv[a + ep - 1, b] - qhv[a + ep - 1, b - 1])
    qc = q[a, b]
    ens = (np.sum(qc * zt), np.sum(np.abs(qc * zt)))
    Uu = U[1 - ep:1 - ep + qhv.shape[0], 1:1 + qhv.shape[1]]                           # U at qhv's points
    Vv = V[1:1 + qhu.shape[0], 1:1 + qhu.shape[1]]                                     # V at qhu's points
    en = (np.sum(Uu * qhv) - np.sum(Vv * qhu), np.sum(np.abs(Uu * qhv)) + np.sum(np.abs(Vv * qhu)))
    return ens, en


@pytest.mark.parametrize("bc", ["nonperiodic", "periodic"])
def test_pv_advection_conservation(swm, orc, bc):
    """Sadourny (1975)'s scheme conserves enstrophy, Arakawa & Hsu (1990)'s conserves enstrophy
    and energy, for a non-divergent mass flux: Σ q ζ_t = 0 (and Σ u·(qhv, -qhu) = 0) up to
    rounding, exact identities of each scheme's stencil.  Sadourny's does not conserve energy,
    and a q combination shifted by one point breaks the enstrophy identity; both are checked so
    the sums are shown to see the stencils' indexing."""
    nx, ny = 24, 18
    q, U, V = _nondivergent_fluxes(nx, ny, seed=11)
    ep = 1 if bc == "periodic" else 0
    assert np.all(np.diff(U, axis=0)[:, 1:-1] + np.diff(V, axis=1)[1:-1, :] == 0)   # δx U + δy V = 0
    out = {}
    for scheme in ("Sadourny", "ArakawaHsu"):
        S = swm.make_setup(swm.Parameter(T="Float64", nx=nx, L_ratio=nx / ny, bc=bc, adv_scheme=scheme))
        assert (S.grid.nx, S.grid.ny) == (nx, ny)
        qhv, qhu = orc.pv_advection(swm.make_config(S), q, U, V)
        assert qhv.shape == (S.grid.nux, S.grid.nuy) and qhu.shape == (S.grid.nvx, S.grid.nvy)
        out[scheme] = (qhv, qhu)
        (e, es), (k, ks) = _pv_invariants(q, qhv, qhu, U, V, ep)
        assert es > 1 and abs(e) < 1e-13 * es, (scheme, e, es)
        if scheme == "ArakawaHsu":
            assert ks > 1 and abs(k) < 1e-13 * ks, (k, ks)
        else:
            assert abs(k) > 1e-3 * ks                                                  # energy not conserved
    qhv, qhu = out["Sadourny"]
    (e, es), _ = _pv_invariants(q, np.roll(qhv, 1, axis=1), qhu, U, V, ep)
    assert abs(e) > 1e-3 * es
    # Sadourny's closed form on a uniform q and uniform fluxes: qhv = q·V, qhu = q·U
    S = swm.make_setup(swm.Parameter(T="Float64", nx=nx, L_ratio=nx / ny, bc=bc, adv_scheme="Sadourny"))
    qhv, qhu = orc.pv_advection(swm.make_config(S), np.full_like(q, 3.0), np.full_like(U, 2.0), np.full_like(V, -5.0))
    assert np.all(qhv == -15.0) and np.all(qhu == 6.0)


# ------------------------------------------------------------------ linear dynamics (rhs.jl:55-86)

@pytest.mark.parametrize("bc", ["periodic", "nonperiodic"])
def test_linear_dynamics_closed_forms(swm, orc, bc):
    """rhs_linear!: du = f_u·Ixy(v) - ∂x(gη) + Fx, dv = -f_v·Ixy(u) - ∂y(gη) + Fy on the
    β-plane.  A uniform v (u = η = 0) gives du = f_u·V0 exactly away from the walls, a uniform
    u gives dv = -f_v·U0 exactly, a tilted η gives the pressure gradient -g·a; continuity!
    under the classic RK scheme reads the zero mass fluxes it was allocated with (rhs_linear!
    forms no U, V), so dη = 0, while SSPRK2's continuity! forms them and moves η."""
    P = swm.Parameter(T="Float64", nx=24, L_ratio=1.5, bc=bc, dynamics="linear", topography="flat",
                      wind_forcing_x="none", wind_forcing_y="none", Fx0=0.0)
    S = swm.make_setup(P)
    G = S.grid
    assert G.β != 0 and np.ptp(G.f_u) > 0

    def rk1(u, v, e, P=P, S=S):
        prog = swm.PrognosticVars(*swm.add_halo(u, v, e, np.zeros((G.nx, G.ny)), S, T="Float64"), T="Float64")
        r = orc.run(S, swm.make_config(S), prog, 1, stop_stage=1, want_rk=True)
        return r[4][6:9]

    z = lambda m, n: np.zeros((m, n))
    V0, U0 = 0.41, -0.23
    du, dv, de = rk1(z(G.nux, G.nuy), np.full((G.nvx, G.nvy), V0), z(G.nx, G.ny))
    assert np.array_equal(du[2:-2, 3:-3], G.f_u[:, 1:-1] * V0)          # rows 2..ny-1: interior v either side
    assert np.all(dv == 0) and np.all(de == 0)
    # v = j/2 on v row j (y = j): Ixy!(v) on u row j (y = j - 1/2) is j/2 - 1/4 exactly
    jv = np.arange(1, G.nvy + 1)[None, :] * 0.5 + np.zeros((G.nvx, 1))
    du, _, _ = rk1(z(G.nux, G.nuy), jv, z(G.nx, G.ny))
    ju = np.arange(1, G.nuy + 1)[None, :]
    assert np.array_equal(du[2:-2, 3:-3], (G.f_u * (0.5 * ju - 0.25))[:, 1:-1])
    if bc == "nonperiodic":                                             # u = x/2: Ixy!(u) at x - 1/2
        iu = np.arange(1, G.nux + 1)[:, None] * 0.5 + np.zeros((1, G.nuy))
        _, dv, _ = rk1(iu, z(G.nvx, G.nvy), z(G.nx, G.ny))
        iv = np.arange(1, G.nvx + 1)[:, None]
        assert np.array_equal(dv[2:-3, 2:-2], -(G.f_v * (0.5 * iv - 0.25))[:-1])
    du, dv, de = rk1(np.full((G.nux, G.nuy), U0), z(G.nvx, G.nvy), z(G.nx, G.ny))
    c = slice(0, None) if bc == "periodic" else slice(1, -1)            # next to a wall u_v sees u = 0
    assert np.array_equal(dv[2:-2, 2:-2][c], -(G.f_v * U0)[c]) and np.all(de == 0)
    assert np.all(du == 0)
    a, b = 1.7e-3, -0.6e-3
    i, j = np.meshgrid(np.arange(G.nx, dtype=float), np.arange(G.ny, dtype=float), indexing="ij")
    du, dv, de = rk1(z(G.nux, G.nuy), z(G.nvx, G.nvy), a * i + b * j)
    g = S.constants.g
    seam = 1 if bc == "periodic" else 0                                  # the periodic u column x = 0 spans the wrap
    assert np.allclose(du[2 + seam:-2, 2:-2], -g * a, rtol=1e-9, atol=0)
    assert np.allclose(dv[2:-2, 2:-2], -g * b, rtol=1e-9, atol=0)
    assert np.all(de == 0)
    # SSPRK2 forms U, V in continuity!: a uniform u over a tilted η converges mass
    P2 = swm.Parameter(T="Float64", nx=24, L_ratio=1.5, bc=bc, dynamics="linear", topography="flat",
                       wind_forcing_x="none", wind_forcing_y="none", Fx0=0.0, time_scheme="SSPRK2", RKs=2)
    S2 = swm.make_setup(P2)
    r = orc.run(S2, swm.make_config(S2), swm.PrognosticVars(*swm.add_halo(
        np.full((G.nux, G.nuy), U0), z(G.nvx, G.nvy), a * i, np.zeros((G.nx, G.ny)), S2, T="Float64"), T="Float64"), 1)
    assert np.any(r[2][1:-1, 1:-1] != a * i)


# ------------------------------------------------------------------ nstep_advcor > 0 (rhs.jl:36-38, time_integration.jl:204-211)

def test_nstep_advcor_schedule(swm, orc):
    """With nstep_advcor = N the stages read q and KEu, KEv as last formed: from the initial
    state (time_integration.jl:30), after the stages of every N-th step (:208-211), and KEu,
    KEv again by bottom_drag_quadratic! every diffusion step (bottom_drag.jl:35-38, the same
    arrays).  Exact consequences: stage 1 of step 1 equals the every-stage form; N = 2 and
    N = 7 agree through step 2 and N = 1 does not; a quadratic drag with cD = 0 (no force)
    differs from no drag only through the KE refresh, so only when N > 1."""
    base = dict(T="Float64", nx=40, L_ratio=1.25, bc="nonperiodic", topography="seamount",
                wind_forcing_x="double_gyre", diffusion="Smagorinsky", α=0.5, tracer_advection=False)
    def run(nsteps, stop_stage=0, **kw):
        P = swm.Parameter(**{**base, **kw})
        S = swm.make_setup(P)
        prog = swm.seeded_initial_conditions(S, seed=9, σ=0.6)
        return orc.run(S, swm.make_config(S), prog, nsteps, stop_stage=stop_stage, want_rk=stop_stage > 0)

    d0 = run(1, stop_stage=1, nstep_advcor=0)[4][6:9]
    d5 = run(1, stop_stage=1, nstep_advcor=5)[4][6:9]
    assert all(np.array_equal(a, b) for a, b in zip(d0, d5))
    d0 = run(1, stop_stage=2, nstep_advcor=0)[4][6:9]
    d5 = run(1, stop_stage=2, nstep_advcor=5)[4][6:9]
    assert not np.array_equal(d0[0], d5[0])                             # stage 2 reads the kept q
    same = lambda a, b: all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3]))
    r1, r2, r7 = (run(2, nstep_advcor=n) for n in (1, 2, 7))
    assert same(r2, r7) and not same(r1, r7)
    assert not same(run(1, nstep_advcor=0), run(1, nstep_advcor=7))
    assert same(run(1, nstep_advcor=1), run(1, nstep_advcor=7))         # the update follows the stages
    for n, differ in ((0, False), (1, False), (3, True)):
        a = run(2, nstep_advcor=n, bottom_drag="quadratic", cD=0.0)
        b = run(2, nstep_advcor=n, bottom_drag="none")
        assert same(a, b) != differ, n


# ------------------------------------------------------------------ the two oracle builds

@pytest.mark.parametrize("T,Tprog", [("Float32", "Float32"), ("Float64", "Float64"), ("BFloat16", "Float32")])
def test_vectorised_build_gives_the_same_bits(swm, orc, T, Tprog):
    """The -O3 vectorised build (the CPU baseline beside the GPU) evaluates the same
    operations in the same order, without FMA contraction: its results equal the -O2
    build's bit for bit."""
    P = swm.Parameter(T=T, Tprog=Tprog, nx=36, L_ratio=1.2, bc="nonperiodic", topography="seamount",
                      wind_forcing_x="double_gyre", diffusion="Smagorinsky", Uadv=20.0)
    S = swm.make_setup(P)
    prog = swm.seeded_initial_conditions(S, seed=4, σ=0.5)
    runs = []
    try:
        for variant in ("o2", "o3"):
            orc.use(variant)
            runs.append(orc.run(S, swm.make_config(S), prog, 8)[:4])
    finally:
        orc.use("o2")
    for name, a, b in zip(("u", "v", "η", "sst"), *runs):
        assert np.array_equal(a, b), name
