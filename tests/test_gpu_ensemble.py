"""Ensembles (swm_ens_*): members of different grid sizes advanced by batched launches and
graph replays equal their own single-context runs and the oracle bit for bit."""
import numpy as np
import pytest

from conftest import first_mismatch

pytestmark = pytest.mark.gpu

SIZES = [(40, 1.0), (64, 2.0), (70, 0.8), (48, 1.5), (96, 1.0)]


def members(swm, bc, T="Float32", Tprog=None, tracer=True, Uadv=0.25, output_dt=1000.0, seed=5, topo=None):
    """bc "periodic-flat": periodic members on a flat bottom, whose stage pairs fuse around the
    seam; topo: one topography per member (default seamount)."""
    rng = np.random.default_rng(seed)
    out = []
    if bc == "periodic-flat":
        bc, topo = "periodic", ["flat"] * len(SIZES)
    for k, (nx, ratio) in enumerate(SIZES):
        P = swm.Parameter(T=T, Tprog=Tprog or T, nx=nx, L_ratio=ratio, bc=bc, Fx0=float(rng.uniform(0.08, 0.16)),
                          H=float(rng.uniform(400, 600)), topography=topo[k] if topo else "seamount",
                          wind_forcing_x="double_gyre",
                          diffusion="Smagorinsky", bottom_drag="quadratic", tracer_advection=tracer, Uadv=Uadv,
                          output_dt=output_dt)
        S = swm.make_setup(P)
        out.append((S, swm.seeded_initial_conditions(S, seed=100 + k, σ=0.6)))
    return out


def run_alone(swm, S, prog, n):
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        assert sim.step(n, 1, 0) == (n, 0)
        return sim.download()


def assert_oracle(swm, orc, S, prog, n, got, tag):
    """The member's state after n steps from prog equals the oracle's, bitwise."""
    r = orc.run(S, swm.make_config(S), prog, n)
    for name, x, y in zip(("u", "v", "η", "sst"), r[:4], (got.u, got.v, got.η, got.sst)):
        mm = first_mismatch(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64))
        assert mm is None, f"{tag} {name} against the oracle: {mm}"


def assert_equal(S, a, b, tag):
    T = S.parameters.Tprog
    import shallowwaters_amd as _  # noqa: F401  (package import for to_float64)
    for name, x, y in zip(("u", "v", "η", "sst"), (a.u, a.v, a.η, a.sst), (b.u, b.v, b.η, b.sst)):
        mm = first_mismatch(np.asarray(x, dtype=np.float64) if x.dtype != np.uint16 else x.astype(np.float64),
                            np.asarray(y, dtype=np.float64) if y.dtype != np.uint16 else y.astype(np.float64))
        assert mm is None, f"{tag} {name}: {mm}"


@pytest.mark.parametrize("bc", ["nonperiodic", "periodic", "periodic-flat", "periodic-mixed-seams"])
@pytest.mark.parametrize("events", [False, True])
@pytest.mark.parametrize("persistent", [True, False])
def test_ensemble_matches_single_runs(swm, orc, gpu, bc, events, persistent):
    """In the persistent launch every step runs inside one cooperative kernel, tracer and
    NaN-check steps included.  Without it: without tracer / NaN-check steps the ensemble runs
    step pairs as graph replays; with them (tracer every few steps, NaN check every step) it
    runs them eagerly in between.  Periodic members on a flat bottom run fused stage pairs
    wrapping around the seam; with one member whose seam does not allow that, every member
    takes the form with the seam strips on their own kernel (as do the seamount members of
    "periodic")."""
    n = 14
    topo = ["flat", "flat", "ridge", "flat", "flat"] if bc == "periodic-mixed-seams" else None
    ms = members(swm, "periodic" if bc == "periodic-mixed-seams" else bc, Uadv=30.0 if events else 0.25,
                 output_dt=0.02 if events else 1000.0, topo=topo)
    if events:
        assert any(S.grid.nadvstep <= 5 for S, _ in ms)
    ref = [run_alone(swm, S, p, n) for S, p in ms]
    sims = [swm.Simulation(S) for S, _ in ms]
    try:
        for sim, (S, p) in zip(sims, ms):
            sim.upload(p)
        with swm.Ensemble(sims) as ens:
            ens.persistent = persistent
            done, nan, graph = ens.step(n, 1)
            assert done == n and not any(nan)
            used, why = ens.persistent_status()
            # a fresh state takes one batched step to settle its slot roles
            assert used == (n - 1 if persistent else 0), why
        for sim, (S, p), r in zip(sims, ms, ref):
            got = sim.download()
            assert_equal(S, r, got, f"{bc} nx={S.grid.nx}")
            assert_oracle(swm, orc, S, p, n, got, f"{bc} nx={S.grid.nx}")
        if not events and not persistent:   # the first steps settle the slot roles; then step pairs replay the graph
            assert graph >= n - 4, graph
    finally:
        for sim in sims:
            sim.close()


@pytest.mark.parametrize("bc,T,Tprog", [("nonperiodic", "Float32", "Float32"), ("periodic", "Float64", "Float64"),
                                        ("nonperiodic", "BFloat16", "Float32")])
def test_tuned_ensemble_in_two_calls(swm, gpu, bc, T, Tprog):
    """After the autotuner picks segment lengths, steps issued in two calls still equal the
    single runs."""
    ms = members(swm, bc, T=T, Tprog=Tprog, tracer=True, Uadv=20.0)
    ref = [run_alone(swm, S, p, 9) for S, p in ms]
    sims = [swm.Simulation(S) for S, _ in ms]
    try:
        for sim, (S, p) in zip(sims, ms):
            sim.upload(p)
        with swm.Ensemble(sims) as ens:
            Ls, Lf, Ld = ens.tune(reps=2)
            assert Ls >= 4 and Lf >= 8 and Ld >= 8
            assert ens.step(4, 1)[0] == 4
            assert ens.step(5, 5)[0] == 5
        for sim, (S, _), r in zip(sims, ms, ref):
            assert_equal(S, r, sim.download(), f"{bc} {T} nx={S.grid.nx}")
    finally:
        for sim in sims:
            sim.close()


@pytest.mark.parametrize("kb", [dict(bc="nonperiodic"), dict(bc="periodic", adv_scheme="Sadourny"),
                                dict(bc="periodic", dynamics="linear"), dict(bc="periodic", nstep_advcor=2)])
def test_ensemble_rejects_mixed_options(swm, gpu, kb):
    """Members must share formats, boundary condition and right-hand-side form (scheme,
    dynamics, nstep_advcor > 0): one batched kernel sequence serves them all."""
    a = swm.Simulation(swm.make_setup(swm.Parameter(nx=40, bc="periodic")))
    b = swm.Simulation(swm.make_setup(swm.Parameter(nx=40, **kb)))
    try:
        for sim in (a, b):
            sim.upload(swm.seeded_initial_conditions(sim.S, seed=1, σ=0.1))
        with pytest.raises(Exception):
            swm.Ensemble([a, b])
    finally:
        a.close()
        b.close()


def test_ensemble_rejects_different_nstep_advcor(swm, gpu):
    """Members keeping q and KEu+KEv update them on the same steps: one kernel sequence."""
    sims = [swm.Simulation(swm.make_setup(swm.Parameter(nx=40, nstep_advcor=n))) for n in (2, 3)]
    try:
        for sim in sims:
            sim.upload(swm.seeded_initial_conditions(sim.S, seed=1, σ=0.1))
        with pytest.raises(Exception):
            swm.Ensemble(sims)
    finally:
        for sim in sims:
            sim.close()


@pytest.mark.parametrize("kw", [dict(adv_scheme="Sadourny"), dict(dynamics="linear"), dict(nstep_advcor=2),
                                dict(adv_scheme="Sadourny", nstep_advcor=3)])
@pytest.mark.parametrize("bc", ["nonperiodic", "periodic", "periodic-flat"])
def test_ensemble_of_other_right_hand_sides(swm, orc, gpu, kw, bc):
    """Members with Sadourny's scheme, the linear dynamics or q and KEu+KEv kept between
    nstep_advcor updates equal their own single-context runs bit for bit: fused batched
    pairs (nonperiodic, and periodic on a flat bottom, whose seams the lanes wrap around),
    batched single stages (periodic over a seamount); the kept fields' updates (from the
    initial state, every nstep_advcor-th step, KEu+KEv after each quadratic-drag step) run
    per member inside the step's plan."""
    n = 8
    rng = np.random.default_rng(3)
    ms = []
    topo = "flat" if bc == "periodic-flat" else "seamount"
    for k, (nx, ratio) in enumerate(SIZES[:3]):
        P = swm.Parameter(T="Float32", nx=nx, L_ratio=ratio, bc=bc.split("-")[0], Fx0=float(rng.uniform(0.08, 0.16)),
                          topography=topo, wind_forcing_x="double_gyre", Uadv=30.0, **kw)
        S = swm.make_setup(P)
        ms.append((S, swm.seeded_initial_conditions(S, seed=200 + k, σ=0.6)))
    ref = [run_alone(swm, S, p, n) for S, p in ms]
    sims = [swm.Simulation(S) for S, _ in ms]
    try:
        for sim, (S, p) in zip(sims, ms):
            sim.upload(p)
        with swm.Ensemble(sims) as ens:
            ens.tune(reps=2)                 # before the first step (kept fields not yet formed)
            done, nan, _ = ens.step(3, 1)
            assert done == 3 and not any(nan)
            ens.tune(reps=2)                 # between steps: the timed steps leave every field as it was
            done, nan, _ = ens.step(n - 3, 4)
            assert done == n - 3 and not any(nan)
        for sim, (S, p), r in zip(sims, ms, ref):
            got = sim.download()
            assert_equal(S, r, got, f"{kw} {bc} nx={S.grid.nx}")
            assert_oracle(swm, orc, S, p, n, got, f"{kw} {bc} nx={S.grid.nx}")
    finally:
        for sim in sims:
            sim.close()


def kept_field_members(swm, bc, nstep_advcor, k=4):
    """Members keeping q and KEu+KEv between updates, with no host work between steps (no
    tracer, no NaN check within the run), so that step pairs replay the captured graph."""
    rng = np.random.default_rng(8)
    ms = []
    for j, (nx, ratio) in enumerate(SIZES[:k]):
        P = swm.Parameter(T="Float32", nx=nx, L_ratio=ratio, bc=bc.split("-")[0], Fx0=float(rng.uniform(0.08, 0.16)),
                          topography="flat" if bc == "periodic-flat" else "seamount", wind_forcing_x="double_gyre",
                          tracer_advection=False, output_dt=1000.0, nstep_advcor=nstep_advcor)
        S = swm.make_setup(P)
        ms.append((S, swm.seeded_initial_conditions(S, seed=300 + j, σ=0.6)))
    return ms


@pytest.mark.parametrize("nstep_advcor", [1, 2])
@pytest.mark.parametrize("bc", ["nonperiodic", "periodic-flat"])
def test_graph_replays_of_kept_field_updates(swm, orc, gpu, bc, nstep_advcor):
    """The kept fields' updates (per member, inside the step's plan) replayed in captured
    graphs: bitwise equal to single runs and the oracle, most steps as graph replays."""
    n = 12
    ms = kept_field_members(swm, bc, nstep_advcor)
    ref = [run_alone(swm, S, p, n) for S, p in ms]
    sims = [swm.Simulation(S) for S, _ in ms]
    try:
        for sim, (S, p) in zip(sims, ms):
            sim.upload(p)
        with swm.Ensemble(sims) as ens:
            done, nan, graph = ens.step(n, 1)
        assert done == n and not any(nan)
        assert graph >= n - 4, graph
        for sim, (S, p), r in zip(sims, ms, ref):
            got = sim.download()
            assert_equal(S, r, got, f"{bc} N={nstep_advcor} nx={S.grid.nx}")
            assert_oracle(swm, orc, S, p, n, got, f"{bc} N={nstep_advcor} nx={S.grid.nx}")
    finally:
        for sim in sims:
            sim.close()


@pytest.mark.parametrize("nstep_advcor", [1, 2])
def test_member_stepped_alone_between_ensemble_calls(swm, gpu, nstep_advcor):
    """A member stepped on its own between two ensemble calls changes its slot roles, so the
    next call plans other kernels' arguments than the graph captured: the graph must not be
    replayed stale (regression: it was)."""
    ms = kept_field_members(swm, "nonperiodic", nstep_advcor)
    dt = [S.grid.dtint for S, _ in ms]
    ref = []
    for k, (S, p) in enumerate(ms):                       # the same sequence of steps, one member at a time
        with swm.Simulation(S) as sim:
            sim.upload(p)
            assert sim.step(6, 1, 0) == (6, 0)
            if k == 0:
                assert sim.step(1, 7, 6 * dt[k]) == (1, 0)
            assert sim.step(8, 8, 7 * dt[k]) == (8, 0)
            ref.append(sim.download())
    sims = [swm.Simulation(S) for S, _ in ms]
    try:
        for sim, (S, p) in zip(sims, ms):
            sim.upload(p)
        with swm.Ensemble(sims) as ens:
            assert ens.step(6, 1)[0] == 6
            assert sims[0].step(1, 7, 6 * dt[0]) == (1, 0)
            done, nan, graph = ens.step(8, 8)
            assert done == 8 and not any(nan)
        for sim, (S, _), r in zip(sims, ms, ref):
            assert_equal(S, r, sim.download(), f"N={nstep_advcor} nx={S.grid.nx}")
    finally:
        for sim in sims:
            sim.close()


def test_refused_ensemble_step_keeps_the_pending_update(swm, gpu):
    """A member given a new state between ensemble calls still has to form its kept q and
    KEu+KEv from it; the others do not, so the ensemble refuses the step.  The refusal must
    leave that update pending: the member run on its own then equals a fresh run from the
    new state (regression: planning had cleared it)."""
    ms = kept_field_members(swm, "nonperiodic", 2)
    S0 = ms[0][0]
    fresh = swm.seeded_initial_conditions(S0, seed=77, σ=0.6)
    ref = run_alone(swm, S0, fresh, 5)
    sims = [swm.Simulation(S) for S, _ in ms]
    try:
        for sim, (S, p) in zip(sims, ms):
            sim.upload(p)
        with swm.Ensemble(sims) as ens:
            assert ens.step(2, 1)[0] == 2
            sims[0].upload(fresh)
            with pytest.raises(Exception, match="kernel sequence"):
                ens.step(2, 3)
        assert sims[0].step(5, 1, 0) == (5, 0)
        assert_equal(S0, ref, sims[0].download(), "re-uploaded member")
    finally:
        for sim in sims:
            sim.close()


@pytest.mark.parametrize("bc", ["nonperiodic", "periodic"])
def test_persistent_launch_in_calls_with_a_nan_stop(swm, gpu, bc):
    """The persistent launch against the batched launches and graph replays, call by call:
    odd and even step counts (the slot roles and tracer slot carried between calls), events
    on each member's own schedule, then a member whose tracer holds a NaN: every member stops
    after the first step whose check sees it, the member keeping its state from before that
    step (feedback.jl:136-148, time_integration.jl:235-237)."""
    ms = members(swm, bc, Uadv=10.0, output_dt=0.5)
    nouts = [S.grid.nout for S, _ in ms]
    assert min(nouts) >= 2 and len(set(nouts)) > 1, nouts
    assert len(set(S.grid.nadvstep for S, _ in ms)) > 1
    results = {}
    for persistent in (True, False):
        sims = [swm.Simulation(S) for S, _ in ms]
        try:
            for sim, (S, p) in zip(sims, ms):
                sim.upload(p)
            with swm.Ensemble(sims) as ens:
                ens.persistent = persistent
                calls = []
                for nsteps, i0 in ((3, 1), (4, 4), (1, 8)):
                    done, nan, _ = ens.step(nsteps, i0)
                    calls.append((done, tuple(nan), ens.persistent_steps()))
                mid = [sim.download() for sim in sims]
                bad = mid[2]
                bad.sst[5, 7] = np.nan
                sims[2].upload(bad)
                done, nan, _ = ens.step(20, 9)
                calls.append((done, tuple(nan), ens.persistent_steps()))
            results[persistent] = (calls, [sim.download() for sim in sims])
        finally:
            for sim in sims:
                sim.close()
    (cp, sp), (cg, sg) = results[True], results[False]
    assert [c[:2] for c in cp] == [c[:2] for c in cg], (cp, cg)
    assert all(c[0] - 1 <= c[2] <= c[0] for c in cp) and all(c[2] == 0 for c in cg), (cp, cg)
    done, nan = cp[3][:2]
    stop = 9 + done - 1
    assert stop % nouts[2] == 0 and nan[2] == stop and sum(1 for x in nan if x) == 1, (done, nan, nouts)
    for (S, _), a, b in zip(ms, sp, sg):
        assert_equal(S, a, b, f"{bc} nx={S.grid.nx}")
