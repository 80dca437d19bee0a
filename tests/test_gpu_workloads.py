"""The benchmark's workloads on the GPU against the oracle, at their own sizes.

  * the README's 100x100 double gyre (Float32, from rest): bitwise after 10 steps and after
    the whole 100-day run (30,639 steps);
  * the 2048x1024 periodic channel through its first tracer advection (snapshot at step 150,
    advection at step 300): bitwise;
  * the 4096² double gyre with the reduced-precision right-hand sides (BFloat16 and Float16
    tendencies over a Float32 state, velocities carried scaled by 2^6, compensated RK
    increments): one step bitwise; and the Float64 path, the long-run reference of those
    forms, two steps bitwise;
  * the 64-member ensemble (sizes 128²-1024² drawn from the seed): two batched steps, every
    member bitwise equal to the oracle;
  * the 1024² adjoint with the tracer advecting within the steps (Uadv = 20 m/s: snapshot at
    step 2, advection at step 3): the FP64 dot-product identity with the oracle's
    tangent-linear model.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import first_mismatch

pytestmark = pytest.mark.gpu


def gpu_run(swm, S, prog, n, i0=1, t0=0):
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        assert sim.step(n, i0, t0) == (n, 0)
        return sim.download()


def assert_bitwise(tag, ref, got):
    for name, a, b in zip(("u", "v", "η", "sst"), ref, (got.u, got.v, got.η, got.sst)):
        mm = first_mismatch(np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64))
        assert mm is None, f"{tag} {name}: {mm}"


def test_readme_double_gyre_10_steps_and_100_days(swm, orc, gpu):
    S = swm.make_setup(swm.workloads.readme_double_gyre())
    G = S.grid
    assert (G.nx, G.ny, G.dtint, G.nt) == (100, 100, 282, 30639)
    prog = swm.initial_conditions("Float32", S)
    cfg = swm.make_config(S)
    print(f"README double gyre {G.nx}x{G.ny}: 10 steps, then {G.nt}", flush=True)
    with ThreadPoolExecutor(1) as ex:
        full = ex.submit(orc.run, S, cfg, prog, G.nt)            # the oracle's 100 days beside the GPU work
        assert_bitwise("10 steps", orc.run(S, cfg, prog, 10)[:4], gpu_run(swm, S, prog, 10))
        got = gpu_run(swm, S, prog, G.nt)
        ref = full.result()
    assert ref[5] == G.nt
    assert_bitwise(f"{G.nt} steps", ref[:4], got)
    assert np.abs(got.u).max() > 0.05                            # the gyres spun up


def test_channel_through_the_first_tracer_advection(swm, orc, gpu):
    S = swm.make_setup(swm.workloads.channel("Float32"))
    G = S.grid
    assert (G.nx, G.ny, G.nadvstep, G.nadvstep_half) == (2048, 1024, 300, 150)
    prog = swm.seeded_initial_conditions(S, seed=2024, σ=0.1)
    i0, n = 148, 154                       # steps 148 .. 301: snapshot at 150, advection at 300
    t0 = (i0 - 1) * G.dtint
    print(f"channel {G.nx}x{G.ny}: steps {i0}..{i0 + n - 1}", flush=True)
    with ThreadPoolExecutor(1) as ex:
        fut = ex.submit(orc.run, S, swm.make_config(S), prog, n, i0, t0)
        got = gpu_run(swm, S, prog, n, i0, t0)
        ref = fut.result()
    assert not np.array_equal(ref[3], prog.sst)                  # the tracer moved
    assert_bitwise("channel", ref[:4], got)


FORMS = {"bf16-s64-kahan": dict(T="BFloat16", Tprog="Float32", scale=64.0, compensated=True),
         "f16-s64-kahan": dict(T="Float16", Tprog="Float32", scale=64.0, compensated=True)}


def test_double_gyre_4096_reduced_precision_and_float64(swm, orc, gpu):
    """One step of each reduced-precision form, two of Float64; the oracle runs them in
    parallel threads beside the GPU."""
    cases = [(name, swm.make_setup(swm.workloads.double_gyre(4096, **kw)), 1) for name, kw in FORMS.items()]
    cases.append(("f64", swm.make_setup(swm.workloads.double_gyre(4096, T="Float64")), 2))
    progs = [swm.seeded_initial_conditions(S, seed=2024, σ=0.1) for _, S, _ in cases]
    with ThreadPoolExecutor(len(cases)) as ex:
        futs = [ex.submit(orc.run, S, swm.make_config(S), p, n) for (_, S, n), p in zip(cases, progs)]
        for (name, S, n), p, f in zip(cases, progs, futs):
            print(f"double gyre 4096² {name}: {n} step(s)", flush=True)
            got = gpu_run(swm, S, p, n)
            assert_bitwise(name, f.result()[:4], got)


@pytest.mark.parametrize("persistent", [False, True])
def test_ensemble_of_64_members_against_the_oracle(swm, orc, gpu, persistent):
    """Batched launches and graph replays, or the persistent launch (after the one batched
    step that settles a fresh state's slot roles)."""
    ms = swm.workloads.ensemble_members(64, seed=4)
    assert sorted({S.grid.nx for S, _ in ms}) == [128, 256, 512, 1024]
    n = 3
    print(f"ensemble of {len(ms)}: {n} steps, persistent={persistent}", flush=True)
    with ThreadPoolExecutor(8) as ex:
        refs = [ex.submit(orc.run, S, swm.make_config(S), p, n) for S, p in ms]
        sims = [swm.Simulation(S) for S, _ in ms]
        try:
            for sim, (_, p) in zip(sims, ms):
                sim.upload(p)
            with swm.Ensemble(sims) as ens:
                ens.persistent = persistent
                done, nan, _ = ens.step(n, 1)
                used, why = ens.persistent_status()
            assert done == n and not any(nan)
            assert used == (n - 1 if persistent else 0), why
            for k, (sim, (S, _), r) in enumerate(zip(sims, ms, refs)):
                assert_bitwise(f"member {k} ({S.grid.nx}²)", r.result()[:4], sim.download())
        finally:
            for sim in sims:
                sim.close()


def test_adjoint_1024_dot_product_with_the_tracer(swm, orc, gpu):
    from test_adjoint_cpu import ghosted_state, interior_masks
    S = swm.make_setup(swm.workloads.double_gyre(1024, T="Float64", Uadv=20.0))
    G = S.grid
    assert (G.nadvstep, G.nadvstep_half) == (3, 1)               # snapshot at step 2, advection at step 3
    cfg = swm.make_config(S)
    prog = swm.seeded_initial_conditions(S, seed=5, σ=0.1)
    n = 3
    rng = np.random.default_rng(2)
    ybar = [np.asfortranarray(rng.standard_normal(m.shape)) for m in interior_masks(S)]
    print(f"adjoint {G.nx}x{G.ny} Float64, {n} steps", flush=True)
    dx = [np.where(mk, rng.standard_normal(mk.shape), 0.0) for mk in interior_masks(S)]
    x = list(ghosted_state(swm, S, prog))
    dxg = list(ghosted_state(swm, S, type(prog)(*dx)))
    with ThreadPoolExecutor(1) as ex:
        tl = ex.submit(orc.tlm, S, cfg, 0, x, dxg, n)
        with swm.Simulation(S) as sim:
            sim.upload(prog)
            sim.adjoint_forward(n, 1, 0, 1)
            g, _ = sim.pullback(swm.PrognosticVars(u=ybar[0], v=ybar[1], η=ybar[2], sst=ybar[3], T="Float64"))
        _, dy = tl.result()
    lhs = sum(float(np.sum(a * b)) for a, b in zip(dy, ybar))
    rhs = sum(float(np.sum(a * b)) for a, b in zip(dx, (g.u, g.v, g.η, g.sst)))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), f"<J dx, ȳ> = {lhs!r}, <dx, J^T ȳ> = {rhs!r}"
    assert float(np.sum(dy[3] * ybar[3])) != 0.0                 # the tracer's tangent took part


def test_float32_arithmetic_forms_against_float64(swm, gpu):
    """BFloat16 / Float16 values with Float32 arithmetic against the Float64 run of the same
    seeded double gyre (512², one model day, tracer on).  Bounds: the mean surface height of
    the closed basin (conserved exactly in exact arithmetic) moves by at most 1e-8 m, and by at
    most 1e-3 of what the reference's BFloat16 arithmetic moves it; the relative L2 error of η
    is at most half the BFloat16 arithmetic's and below 0.1."""
    def run(T, **kw):
        S = swm.make_setup(swm.workloads.double_gyre(nx=512, T=T, Tprog="Float64" if T == "Float64" else "Float32",
                                                     **kw))
        prog = swm.seeded_initial_conditions(S, seed=2024, σ=0.1)
        n = int(np.ceil(86400 / S.grid.dtint))
        with swm.Simulation(S) as sim:
            sim.upload(prog)
            m0, _ = sim.reduce()
            assert sim.step(n, 1, 0) == (n, 0)
            m1, _ = sim.reduce()
            y = sim.download()
        e = swm.to_float64(y.η, S.parameters.Tprog)[1:-1, 1:-1]
        return e, (m1 - m0) / (S.grid.nx * S.grid.ny)
    e64, _ = run("Float64")
    rel = lambda e: float(np.linalg.norm(e - e64) / np.linalg.norm(e64))
    e_b, dm_b = run("BFloat16")
    for T in ("BFloat16", "Float16"):
        e, dm = run(T, arithmetic="Float32")
        print(f"{T} values, Float32 arithmetic: mass drift {dm:.3e} m (BFloat16 arithmetic {dm_b:.3e}), "
              f"η rel L2 {rel(e):.3e} (BFloat16 arithmetic {rel(e_b):.3e})", flush=True)
        assert abs(dm) <= 1e-8 and abs(dm) <= 1e-3 * abs(dm_b), (T, dm, dm_b)
        assert rel(e) <= 0.5 * rel(e_b) and rel(e) < 0.1, (T, rel(e), rel(e_b))
