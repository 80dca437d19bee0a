"""The adjoint of the time step on the GPU (swm_adj_forward / swm_adj_pullback).

Pins:
  * bitwise: the GPU gradient equals the host build of the same adjoint kernels (tests/emu)
    driven by the oracle's taped steps, which is itself exact against the transposed
    Jacobian of the oracle's tangent-linear model (tests/test_adjoint_cpu.py);
  * the FP64 dot-product identity <J dx, ȳ> = <dx, J^T ȳ> with J from the CPU
    tangent-linear model, relative error <= 1e-11;
  * central finite differences of the GPU forward run (FP64), relative error <= 1e-6;
  * revolve: the gradient is bitwise the same for any number of snapshots, and the forward
    steps recomputed follow the schedule (0 snapshots: n(n+1)/2).
"""
import numpy as np
import pytest

import adjemu
from test_adjoint_cpu import ghosted_state, interior_masks, reverse_steps

pytestmark = pytest.mark.gpu


def setup(swm, bc, T="Float64", nx=70, **kw):
    d = dict(T=T, nx=nx, L_ratio=0.6, bc=bc, α=0.5, diffusion="Smagorinsky", bottom_drag="quadratic",
             topography="seamount", wind_forcing_x="double_gyre", Uadv=20.0, tracer_consumption=True)
    d.update(kw)
    P = swm.Parameter(**d)
    S = swm.make_setup(P)
    return P, S, swm.make_config(S), swm.seeded_initial_conditions(S, seed=5, σ=0.8)


def form(adv):
    """The right-hand side of a test case: a PV scheme of the nonlinear dynamics, or "linear"."""
    return dict(dynamics="linear") if adv == "linear" else dict(adv_scheme=adv)


def cotangent(S, seed, dtype):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(rng.standard_normal(m.shape).astype(dtype)) for m in interior_masks(S)]


def gpu_gradient(swm, S, prog, ybar, n, snapshots):
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        sim.adjoint_forward(n, 1, 0, snapshots)
        y = sim.download()
        g, nf = sim.pullback(swm.PrognosticVars(u=ybar[0], v=ybar[1], η=ybar[2], sst=ybar[3], T=S.parameters.Tprog))
        y2 = sim.download()
    for a, b in zip((y.u, y.v, y.η, y.sst), (y2.u, y2.v, y2.η, y2.sst)):
        assert np.array_equal(a, b), "the context must be left at the final state"
    return [g.u, g.v, g.η, g.sst], nf


@pytest.mark.parametrize("bc", ["nonperiodic", "periodic"])
@pytest.mark.parametrize("T", ["Float64", "Float32"])
@pytest.mark.parametrize("adv", ["ArakawaHsu", "Sadourny", "linear"])
def test_pullback_bitwise_equals_host_kernels(swm, orc, gpu, monkeypatch, bc, T, adv):
    monkeypatch.setenv("SWM_ROWS_PER_WAVE_ADJ", "8")        # many segments on a small grid
    P, S, cfg, prog = setup(swm, bc, T=T, **form(adv))
    assert (S.grid.nadvstep, S.grid.nadvstep_half) == (3, 1)
    n = 3                                                     # snapshot at step 2, advection at step 3
    dt = np.float64 if T == "Float64" else np.float32
    ybar = cotangent(S, 1, dt)
    print(f"pullback {bc} {T} {adv} nx={S.grid.nx} ny={S.grid.ny}", flush=True)
    g, _ = gpu_gradient(swm, S, prog, ybar, n, snapshots=1)
    k = adjemu.config(S, cfg, L=8, fmt=8 if T == "Float64" else 4)
    ref, _, _ = reverse_steps(swm, orc, S, cfg, prog, n, [a.astype(np.float64) for a in ybar], k)
    for name, r, a in zip(("u", "v", "η", "sst"), ref, g):
        a = np.asarray(a, dtype=np.float64)
        bad = np.argwhere(r != a)
        assert len(bad) == 0, f"{name}: {len(bad)} cells differ, first {bad[0].tolist()}: {r[tuple(bad[0])]!r} vs {a[tuple(bad[0])]!r}"


@pytest.mark.parametrize("bc", ["nonperiodic", "periodic"])
@pytest.mark.parametrize("T", ["Float64", "Float32"])
def test_tracer_gather_global_scan_bitwise(swm, orc, gpu, monkeypatch, bc, T):
    """The tracer adjoint's gather with every corner scanning in global memory (the path taken
    for departure distances beyond the LDS tiles' reach) equals the host build bitwise, as the
    tiled gather does in the test above."""
    monkeypatch.setenv("SWM_ROWS_PER_WAVE_ADJ", "8")
    monkeypatch.setenv("SWM_TRACER_GATHER_GLOBAL", "1")
    P, S, cfg, prog = setup(swm, bc, T=T)
    assert (S.grid.nadvstep, S.grid.nadvstep_half) == (3, 1)
    n = 3
    dt = np.float64 if T == "Float64" else np.float32
    ybar = cotangent(S, 2, dt)
    print(f"global-scan gather {bc} {T}", flush=True)
    g, _ = gpu_gradient(swm, S, prog, ybar, n, snapshots=1)
    k = adjemu.config(S, cfg, L=8, fmt=8 if T == "Float64" else 4)
    ref, _, _ = reverse_steps(swm, orc, S, cfg, prog, n, [a.astype(np.float64) for a in ybar], k)
    for name, r, a in zip(("u", "v", "η", "sst"), ref, g):
        a = np.asarray(a, dtype=np.float64)
        assert np.array_equal(r, a), f"{name}: {np.count_nonzero(r != a)} cells differ"


@pytest.mark.parametrize("case", ["nonperiodic-F32-RK4", "nonperiodic-F64-RK3", "periodic-F32-RK4", "periodic-F64-RK2",
                                  "nonperiodic-F32-RK4-relax"])
def test_taped_fused_pairs_equal_taped_stages(swm, gpu, monkeypatch, case):
    """The taped step on fused stage pairs (each keeping its first stage's output) gives the
    gradient of the taped step on one kernel per stage, bitwise; periodic over a flat bottom,
    whose seam the pairs' lanes wrap around."""
    bc, T, rk = case.split("-")[:3]
    kw = dict(RKo=int(rk[2]), topography="flat" if bc == "periodic" else "seamount")
    if case.endswith("relax"):
        kw.update(surface_relax=True, t_relax=2.0)
    P, S, cfg, prog = setup(swm, bc, T=T, **kw)
    dt = np.float64 if T == "Float64" else np.float32
    ybar = cotangent(S, 3, dt)
    out = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("SWM_FUSE", fuse)
        print(f"taped {case} fuse={fuse}", flush=True)
        out[fuse], _ = gpu_gradient(swm, S, prog, ybar, 4, snapshots=2)
    for name, a, b in zip(("u", "v", "η", "sst"), out["1"], out["0"]):
        assert np.array_equal(np.asarray(a), np.asarray(b)), f"{case} {name}: {np.count_nonzero(a != b)} cells differ"


@pytest.mark.parametrize("bc,adv", [("nonperiodic", "ArakawaHsu"), ("periodic", "ArakawaHsu"),
                                    ("nonperiodic", "Sadourny"), ("periodic", "Sadourny"),
                                    ("nonperiodic", "linear"), ("periodic", "linear")])
def test_dot_product_identity_fp64(swm, orc, gpu, bc, adv):
    P, S, cfg, prog = setup(swm, bc, nx=90, **form(adv))
    n = 7
    ybar = cotangent(S, 2, np.float64)
    print(f"dot product {bc} nx={S.grid.nx} ny={S.grid.ny} n={n}", flush=True)
    g, _ = gpu_gradient(swm, S, prog, ybar, n, snapshots=2)
    x = list(ghosted_state(swm, S, prog))
    rng = np.random.default_rng(3)
    for m in range(3):
        dx = [np.where(mk, rng.standard_normal(mk.shape), 0.0) for mk in interior_masks(S)]
        dxg = list(ghosted_state(swm, S, type(prog)(*dx)))
        _, dy = orc.tlm(S, cfg, 0, x, dxg, nsteps=n)
        lhs = sum(float(np.sum(a * b)) for a, b in zip(dy, ybar))
        rhs = sum(float(np.sum(a * b)) for a, b in zip(dx, g))
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs), f"<J dx, ȳ> = {lhs!r}, <dx, J^T ȳ> = {rhs!r}"


def test_central_differences_of_the_gpu_forward(swm, gpu):
    P, S, cfg, prog = setup(swm, "nonperiodic", nx=64)
    n = 6
    ybar = cotangent(S, 4, np.float64)
    g, _ = gpu_gradient(swm, S, prog, ybar, n, snapshots=3)
    rng = np.random.default_rng(5)
    dx = [np.where(mk, rng.standard_normal(mk.shape), 0.0) for mk in interior_masks(S)]
    eps = 1e-6

    def J(sign):
        p = type(prog)(*(a + sign * eps * b for a, b in zip((prog.u, prog.v, prog.η, prog.sst), dx)))
        with swm.Simulation(S) as sim:
            sim.upload(p)
            assert sim.step(n, 1, 0) == (n, 0)
            y = sim.download()
        return sum(float(np.sum(a * b)) for a, b in zip((y.u, y.v, y.η, y.sst), ybar))

    fd = (J(1) - J(-1)) / (2 * eps)
    ad = sum(float(np.sum(a * b)) for a, b in zip(dx, g))
    assert abs(fd - ad) <= 1e-6 * abs(ad), f"finite differences {fd!r} vs adjoint {ad!r}"


def test_revolve_snapshots_do_not_change_the_gradient(swm, gpu):
    P, S, cfg, prog = setup(swm, "periodic", T="Float32", nx=66)
    n = 7
    ybar = cotangent(S, 6, np.float32)
    runs = {s: gpu_gradient(swm, S, prog, ybar, n, snapshots=s) for s in (0, 1, 2, 3, 7)}
    g0, nf0 = runs[0]
    assert nf0 == n * (n + 1) // 2
    counts = [runs[s][1] for s in (0, 1, 2, 3, 7)]
    assert counts == sorted(counts, reverse=True) and counts[-1] <= 2 * n, counts
    for s, (g, _) in runs.items():
        for name, a, b in zip(("u", "v", "η", "sst"), g0, g):
            assert np.array_equal(a, b), f"{s} snapshots: ∇{name} differs"


def test_rrule_of_time_integration(swm, gpu):
    P, S, cfg, prog = setup(swm, "nonperiodic", T="Float32", nx=64, Ndays=0.05)
    G = S.grid
    ref = swm.time_integration(type(prog)(*(a.copy(order="F") for a in (prog.u, prog.v, prog.η, prog.sst))),
                               None, S)
    y, pullback = swm.rrule(swm.time_integration, prog, None, S, snapshots=4)
    for a, b in zip((ref.u, ref.v, ref.η, ref.sst), (y.u, y.v, y.η, y.sst)):
        assert np.array_equal(a, b)
    ybar = type(y)(*(np.ones_like(a) for a in (y.u, y.v, y.η, y.sst)))
    nt, pbar, nd, ns = pullback(ybar)
    assert isinstance(nt, swm.NoTangent) and isinstance(nd, swm.NoTangent) and isinstance(ns, swm.NoTangent)
    for a, m in zip((pbar.u, pbar.v, pbar.η, pbar.sst), interior_masks(S)):
        assert a.shape == m.shape and np.all(a[~m] == 0) and np.all(np.isfinite(a))
    assert G.nt >= 2


def test_adjoint_rejects_unsupported_configurations(swm, gpu):
    for kw in (dict(T="BFloat16", Tprog="Float32"), dict(T="Float32", time_scheme="SSPRK2", RKs=4),
               dict(T="Float32", time_scheme="SSPRK3", RKn=4)):
        P, S, cfg, prog = setup(swm, "nonperiodic", **kw)
        with swm.Simulation(S) as sim:
            sim.upload(prog)
            with pytest.raises(swm.SwmError if hasattr(swm, "SwmError") else RuntimeError, match="adjoint"):
                sim.adjoint_forward(2)
    P, S, cfg, prog = setup(swm, "nonperiodic", T="Float32")
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        with pytest.raises(RuntimeError, match="swm_adj_forward first"):
            sim.pullback(prog)


@pytest.mark.parametrize("bc,seasonal", [("nonperiodic", False), ("periodic", True)])
def test_wind_stress_gradient_dot_product_fp64(swm, orc, gpu, bc, seasonal):
    """The gradient with respect to the wind stress field Fx (swm_adj_track_Fx): FP64
    dot-product identity <J_Fx δFx, ȳ> = <δFx, ∇_Fx> with J_Fx from the oracle's
    tangent-linear model, with a seasonal factor sin(ωt) on the wind in the periodic case."""
    P, S, cfg, prog = setup(swm, bc, nx=80, seasonal_wind_x=seasonal, ωFx=300.0)
    n = 6
    ybar = cotangent(S, 3, np.float64)
    print(f"wind stress dot product {bc} seasonal={seasonal} nx={S.grid.nx} n={n}", flush=True)
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        sim.adjoint_track_Fx(True)
        sim.adjoint_forward(n, 1, 0, 2)
        sim.pullback(swm.PrognosticVars(u=ybar[0], v=ybar[1], η=ybar[2], sst=ybar[3], T="Float64"))
        gFx = sim.pullback_Fx()
    x = list(ghosted_state(swm, S, prog))
    rng = np.random.default_rng(9)
    for m in range(3):
        dF = rng.standard_normal(S.forcing.Fx.shape)
        _, dy = orc.tlm(S, cfg, 0, x, None, nsteps=n, dFx=dF)
        lhs = sum(float(np.sum(a * b)) for a, b in zip(dy, ybar))
        rhs = float(np.sum(dF * gFx))
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs), f"<J dFx, ȳ> = {lhs!r}, <dFx, ∇Fx> = {rhs!r}"


def test_adjoint_workload_gradients_by_central_differences(swm, gpu):
    """The adjoint workload at its size: 1024x1024 double gyre, 200 steps with the tracer
    advecting (Uadv = 20 m/s), revolve with 20 snapshots; the loss is the final kinetic energy
    plus the η misfit at a seeded 1% of the cells against a perturbed twin.  Gradients with
    respect to the initial η and the wind stress field Fx, in FP64, against central
    differences of the GPU forward run on 16 seeded directions in (η0, Fx)."""
    S = swm.make_setup(swm.workloads.double_gyre(1024, T="Float64", Uadv=20.0))
    G = S.grid
    n = 200
    prog = swm.seeded_initial_conditions(S, seed=5, σ=0.1)
    obs, mask = swm.workloads.observation_twin(S, prog, n)
    print(f"adjoint workload {G.nx}x{G.ny} n={n}, {int(mask.sum())} observed cells", flush=True)
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        sim.adjoint_track_Fx(True)
        sim.adjoint_forward(n, 1, 0, 20)
        L0, ybar = swm.workloads.final_state_loss(S, sim.download(), obs, mask)
        g, _ = sim.pullback(ybar)
        gFx = sim.pullback_Fx()
    assert np.any(g.η != 0) and np.any(gFx != 0)

    def loss(dη, dF, eps):
        import copy
        S2 = copy.copy(S)
        S2.forcing = copy.copy(S.forcing)
        S2.forcing.Fx = np.asfortranarray(S.forcing.Fx + eps * dF)
        p = swm.PrognosticVars(u=prog.u, v=prog.v, η=np.asfortranarray(prog.η + eps * dη), sst=prog.sst, T="Float64")
        with swm.Simulation(S2) as sim:
            sim.upload(p)
            assert sim.step(n, 1, 0) == (n, 0)
            return swm.workloads.final_state_loss(S2, sim.download(), obs, mask)[0]

    # The differences' own error is about the rounding of the loss (a sum over 10^6 cells)
    # over ε, the same in every direction; it is measured against the typical size of the
    # directional derivatives (their median), so that a direction with a small derivative
    # is held to the same absolute accuracy.
    rng = np.random.default_rng(16)
    eps = 3e-4
    pairs = []
    for d in range(16):
        dη = np.zeros_like(prog.η)
        dη[1:-1, 1:-1] = 0.1 * rng.standard_normal((G.nx, G.ny))
        dF = np.abs(S.forcing.Fx).max() * rng.standard_normal(S.forcing.Fx.shape)
        ad = float(np.sum(g.η * dη)) + float(np.sum(gFx * dF))
        fd = (loss(dη, dF, eps) - loss(dη, dF, -eps)) / (2 * eps)
        pairs.append((ad, fd))
        print(f"direction {d}: adjoint {ad:.10e} finite differences {fd:.10e}", flush=True)
    scale = float(np.median([abs(a) for a, _ in pairs]))
    worst = max(abs(f - a) / max(abs(a), scale) for a, f in pairs)
    print(f"worst relative difference {worst:.2e} (median |derivative| {scale:.3e}, loss {L0:.6e})", flush=True)
    assert worst <= 1e-6


@pytest.mark.parametrize("bc,kw", [("nonperiodic", dict(nstep_advcor=2)),
                                   ("periodic", dict(nstep_advcor=3, bottom_drag="linear")),
                                   ("nonperiodic", dict(nstep_advcor=2, adv_scheme="Sadourny", bottom_drag="none")),
                                   ("periodic", dict(nstep_advcor=1, adv_scheme="Sadourny"))])
def test_kept_fields_dot_product_fp64(swm, orc, gpu, bc, kw):
    """q and KEu+KEv kept over nstep_advcor steps (time_integration.jl:20-31, 208-211; KEu+KEv
    formed again by every quadratic-drag step): the FP64 dot-product identity with the
    oracle's tangent-linear model, whose kept fields are formed from the initial state as the
    uploaded context's are; revolve with 2 snapshots restores the kept fields it recomputes
    from."""
    P, S, cfg, prog = setup(swm, bc, nx=90, **kw)
    n = 7
    ybar = cotangent(S, 4, np.float64)
    print(f"kept fields {bc} {kw} n={n}", flush=True)
    g, _ = gpu_gradient(swm, S, prog, ybar, n, snapshots=2)
    g0, _ = gpu_gradient(swm, S, prog, ybar, n, snapshots=0)
    for a, b in zip(g, g0):
        assert np.array_equal(np.asarray(a), np.asarray(b)), "the gradient depends on the snapshots"
    x = list(ghosted_state(swm, S, prog))
    rng = np.random.default_rng(5)
    for m in range(3):
        dx = [np.where(mk, rng.standard_normal(mk.shape), 0.0) for mk in interior_masks(S)]
        dxg = list(ghosted_state(swm, S, type(prog)(*dx)))
        _, dy = orc.tlm(S, cfg, 0, x, dxg, nsteps=n)
        lhs = sum(float(np.sum(a * b)) for a, b in zip(dy, ybar))
        rhs = sum(float(np.sum(a * b)) for a, b in zip(dx, g))
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs), f"<J dx, ȳ> = {lhs!r}, <dx, J^T ȳ> = {rhs!r}"


SSP_CASES = [("nonperiodic", dict()),
             ("periodic", dict(surface_relax=True)),
             ("periodic", dict(adv_scheme="Sadourny", bottom_drag="linear")),
             ("nonperiodic", dict(dynamics="linear")),
             ("periodic", dict(nstep_advcor=2))]


@pytest.mark.parametrize("scheme", ["SSPRK2", "4SSPRK3", "SSPRK3"])
@pytest.mark.parametrize("bc,kw", SSP_CASES)
def test_ssp_schemes_dot_product_fp64(swm, orc, gpu, scheme, bc, kw):
    """The SSPRK2 scheme (time_integration.jl:80-110: momentum, then the continuity from the
    updated velocities, RKs stages, u0 = u/RKs + (RKs-1)/RKs·u1; two stages need a CFL number
    below one, so the two-stage cases run at cfl = 0.5), the 4SSPRK3 scheme (:158-195: four
    such stages with ½(u + (u + Δt_Δ·du)) updates, the third stage combined with the step's
    input by thirds) and the SSPRK3 scheme (:112-156: n² stages, the kn-th combined with the
    state kept after the mn-th; n = 2, whose kept state is the step's input, and n = 3): the
    FP64 dot-product identity with the oracle's tangent-linear model of the same steps; the
    gradient does not depend on the number of snapshots."""
    if scheme == "SSPRK2" and kw in (SSP_CASES[1][1], SSP_CASES[3][1]):
        kw = dict(kw, RKs=2, cfl=0.5)
    if scheme == "SSPRK3" and kw in (SSP_CASES[2][1], SSP_CASES[4][1]):
        kw = dict(kw, RKn=3)
    P, S, cfg, prog = setup(swm, bc, nx=90, time_scheme=scheme, **kw)
    n = 7
    ybar = cotangent(S, 6, np.float64)
    print(f"{scheme} {bc} {kw} n={n}", flush=True)
    g, _ = gpu_gradient(swm, S, prog, ybar, n, snapshots=2)
    g0, _ = gpu_gradient(swm, S, prog, ybar, n, snapshots=0)
    for a, b in zip(g, g0):
        assert np.array_equal(np.asarray(a), np.asarray(b)), "the gradient depends on the snapshots"
    x = list(ghosted_state(swm, S, prog))
    rng = np.random.default_rng(6)
    for m in range(3):
        dx = [np.where(mk, rng.standard_normal(mk.shape), 0.0) for mk in interior_masks(S)]
        dxg = list(ghosted_state(swm, S, type(prog)(*dx)))
        _, dy = orc.tlm(S, cfg, 0, x, dxg, nsteps=n)
        lhs = sum(float(np.sum(a * b)) for a, b in zip(dy, ybar))
        rhs = sum(float(np.sum(a * b)) for a, b in zip(dx, g))
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs), f"<J dx, ȳ> = {lhs!r}, <dx, J^T ȳ> = {rhs!r}"


def test_snapshots_from_free_memory_keep_every_state(swm, gpu):
    """snapshots < 0 (the default of adjoint_forward and rrule): as many snapshots as half of
    the free device memory holds, at most one per step.  The gradient equals that of a
    two-snapshot schedule bit for bit, the reverse sweep recomputes only its taped steps,
    and a second recording in the same context (the snapshot block reused) gives the same
    gradient."""
    P, S, cfg, prog = setup(swm, "nonperiodic", T="Float32", nx=70)
    n = 9
    ybar = cotangent(S, 7, np.float32)
    g2, nf2 = gpu_gradient(swm, S, prog, ybar, n, snapshots=2)
    yb = swm.PrognosticVars(u=ybar[0], v=ybar[1], η=ybar[2], sst=ybar[3], T=S.parameters.Tprog)
    with swm.Simulation(S) as sim:
        runs = []
        for _ in range(2):
            sim.upload(prog)
            sim.adjoint_forward(n, 1, 0)
            g, nf = sim.pullback(yb)
            runs.append(([g.u, g.v, g.η, g.sst], nf))
    assert nf2 > n
    for g, nf in runs:
        assert nf == n, nf
        for a, b in zip(g, g2):
            assert np.array_equal(np.asarray(a), np.asarray(b))
