"""Parity of the HIP time step with the CPU oracle, through the C ABI.

Bar: Float64 and Float32 results are bitwise equal to the oracle (same operations in the
same order, IEEE division and sqrt, no FMA contraction); the tolerances of the
specification (relative L2 <= 1e-12 at Float64, <= 1e-5 at Float32, per field after 10
steps) are asserted as well.
"""
import numpy as np
import pytest

from conftest import first_mismatch, rel_l2

pytestmark = pytest.mark.gpu

TOL = {"Float64": 1e-12, "Float32": 1e-5}

# small grids that cross strip (60 columns) and segment (32 rows) boundaries
CASES = {
    "np-f64-smag-quad": dict(T="Float64", nx=70, L_ratio=0.9, bc="nonperiodic", α=0.0, diffusion="Smagorinsky",
                             bottom_drag="quadratic", topography="seamount", wind_forcing_x="double_gyre"),
    "np-f32-const-lin": dict(T="Float32", nx=64, L_ratio=1.3, bc="nonperiodic", α=2.0, diffusion="constant",
                             bottom_drag="linear", topography="ridge", wind_forcing_x="channel"),
    "np-f32-smag-quad-partial": dict(T="Float32", nx=61, L_ratio=0.8, bc="nonperiodic", α=0.5, diffusion="Smagorinsky",
                                     bottom_drag="quadratic", topography="seamount", wind_forcing_x="shear"),
    "per-f64-smag-quad": dict(T="Float64", nx=72, L_ratio=1.1, bc="periodic", α=1.0, diffusion="Smagorinsky",
                              bottom_drag="quadratic", topography="seamount", wind_forcing_x="double_gyre"),
    "per-f32-const-none": dict(T="Float32", nx=123, L_ratio=2.0, bc="periodic", diffusion="constant",
                               bottom_drag="none", topography="flat", wind_forcing_x="channel"),
    "per-f32-smag-quad-rk3": dict(T="Float32", nx=66, L_ratio=1.0, bc="periodic", diffusion="Smagorinsky",
                                  bottom_drag="quadratic", topography="ridges", wind_forcing_x="constant", RKo=3,
                                  cfl=0.6),
    "np-f64-const-quad-rk2": dict(T="Float64", nx=50, L_ratio=0.5, bc="nonperiodic", α=1.5, diffusion="constant",
                                  bottom_drag="quadratic", topography="seamount", wind_forcing_x="double_gyre", RKo=2,
                                  cfl=0.5),
    # the strong-stability-preserving schemes with their semi-implicit continuity
    # (test/runtests.jl:48-82 pairs of RKs/RKn and cfl)
    "np-f64-ssprk2-s3": dict(T="Float64", nx=66, L_ratio=1.0, bc="nonperiodic", α=0.0, diffusion="Smagorinsky",
                             topography="seamount", wind_forcing_x="double_gyre", time_scheme="SSPRK2", RKs=3, cfl=1.2),
    "per-f32-ssprk2-s2": dict(T="Float32", nx=64, L_ratio=2.0, bc="periodic", time_scheme="SSPRK2", RKs=2, cfl=0.7),
    "np-f32-ssprk2-s4": dict(T="Float32", nx=62, L_ratio=0.9, bc="nonperiodic", α=1.0, diffusion="Smagorinsky",
                             topography="ridge", time_scheme="SSPRK2", RKs=4, cfl=1.4),
    "per-f64-ssprk3-n2": dict(T="Float64", nx=70, L_ratio=1.2, bc="periodic", diffusion="Smagorinsky",
                              topography="seamount", time_scheme="SSPRK3", RKn=2, cfl=1.2),
    "np-f32-ssprk3-n3": dict(T="Float32", nx=64, L_ratio=1.0, bc="nonperiodic", α=2.0, topography="seamount",
                             wind_forcing_x="double_gyre", time_scheme="SSPRK3", RKn=3, cfl=2.5),
    "np-f64-4ssprk3": dict(T="Float64", nx=61, L_ratio=1.0, bc="nonperiodic", α=0.5, diffusion="Smagorinsky",
                           topography="seamount", wind_forcing_x="double_gyre", time_scheme="4SSPRK3", cfl=1.0),
    "per-f32-4ssprk3": dict(T="Float32", nx=64, L_ratio=2.0, bc="periodic", time_scheme="4SSPRK3", cfl=1.0),
    # continuity variants and seasonal winds (continuity.jl:3-49, rhs.jl:233-238)
    "np-f64-relax": dict(T="Float64", nx=64, L_ratio=1.0, bc="nonperiodic", surface_relax=True, t_relax=2.0,
                         topography="seamount", wind_forcing_x="double_gyre"),
    "per-f32-forcing-seasonal": dict(T="Float32", nx=64, L_ratio=2.0, bc="periodic", surface_forcing=True,
                                     seasonal_wind_x=True, ωFx=400.0, A=3e-3),
    "np-f32-ssprk2-relax": dict(T="Float32", nx=64, L_ratio=1.0, bc="nonperiodic", surface_relax=True, t_relax=2.0,
                                time_scheme="SSPRK2", RKs=3, cfl=1.2),
    # mixed precision: T = Float16 / BFloat16 right-hand side with Tprog = Float32 state
    # (time_integration.jl:57-64, 200-228), and ghost copies through a narrower Tcomm
    "np-f16-f32-smag-quad": dict(T="Float16", Tprog="Float32", nx=64, L_ratio=1.0, bc="nonperiodic", α=0.5,
                                 diffusion="Smagorinsky", bottom_drag="quadratic", topography="seamount",
                                 wind_forcing_x="double_gyre"),
    "per-bf16-f32-comm-bf16": dict(T="BFloat16", Tprog="Float32", Tcomm="BFloat16", nx=66, L_ratio=1.5, bc="periodic",
                                   diffusion="Smagorinsky", bottom_drag="quadratic", topography="seamount",
                                   wind_forcing_x="channel"),
    "per-f16-f32-const-lin": dict(T="Float16", Tprog="Float32", Tcomm="Float16", nx=62, L_ratio=1.0, bc="periodic",
                                  diffusion="constant", bottom_drag="linear", topography="ridge", RKo=3, cfl=0.6),
    "np-bf16-f32-relax": dict(T="BFloat16", Tprog="Float32", nx=64, L_ratio=1.0, bc="nonperiodic", surface_relax=True,
                              t_relax=2.0, topography="seamount", wind_forcing_x="double_gyre",
                              tracer_consumption=True),
    "per-f32-comm-f16": dict(T="Float32", Tcomm="Float16", nx=64, L_ratio=2.0, bc="periodic", diffusion="Smagorinsky",
                             topography="seamount", wind_forcing_x="channel"),
    "per-f64-comm-bf16": dict(T="Float64", Tcomm="BFloat16", nx=60, L_ratio=1.0, bc="periodic", time_scheme="SSPRK2",
                              RKs=3, cfl=1.2),
    # low-precision options: u, v carried scaled (Float16: s·u·h must stay below 65504, so
    # s = 8 for these velocities), Kahan-compensated RK increments
    "np-f16-f32-scaled-comp": dict(T="Float16", Tprog="Float32", nx=64, L_ratio=1.0, bc="nonperiodic", α=0.5,
                                   diffusion="Smagorinsky", bottom_drag="quadratic", topography="seamount",
                                   wind_forcing_x="double_gyre", scale=8.0, compensated=True),
    "per-bf16-f32-scaled": dict(T="BFloat16", Tprog="Float32", nx=66, L_ratio=1.5, bc="periodic",
                                diffusion="Smagorinsky", bottom_drag="quadratic", topography="seamount",
                                wind_forcing_x="channel", scale=64.0, tracer_consumption=True),
    # Float16 / BFloat16 values with Float32 arithmetic (arithmetic="Float32", a form of this
    # library): the stage inputs, statics and constants in T, the right-hand side in Float32
    "np-bf16a-smag-quad": dict(T="BFloat16", Tprog="Float32", arithmetic="Float32", nx=64, L_ratio=1.0,
                               bc="nonperiodic", α=0.5, diffusion="Smagorinsky", bottom_drag="quadratic",
                               topography="seamount", wind_forcing_x="double_gyre"),
    "per-bf16a-comm-bf16": dict(T="BFloat16", Tprog="Float32", Tcomm="BFloat16", arithmetic="Float32", nx=66,
                                L_ratio=1.5, bc="periodic", diffusion="Smagorinsky", bottom_drag="quadratic",
                                topography="seamount", wind_forcing_x="channel"),
    "per-f16a-flat-lin-rk3": dict(T="Float16", Tprog="Float32", Tcomm="Float16", arithmetic="Float32", nx=62,
                                  L_ratio=1.0, bc="periodic", diffusion="constant", bottom_drag="linear",
                                  topography="flat", RKo=3, cfl=0.6),
    "np-bf16a-scaled-comp-relax": dict(T="BFloat16", Tprog="Float32", arithmetic="Float32", nx=64, L_ratio=1.0,
                                       bc="nonperiodic", surface_relax=True, t_relax=2.0, topography="seamount",
                                       wind_forcing_x="double_gyre", scale=64.0, compensated=True,
                                       tracer_consumption=True),
    "np-f32-comp-rk3": dict(T="Float32", nx=61, L_ratio=0.8, bc="nonperiodic", α=0.5, diffusion="Smagorinsky",
                            topography="seamount", wind_forcing_x="double_gyre", compensated=True, RKo=3, cfl=0.6),
    "per-f64-scaled-comp-ssp": dict(T="Float64", nx=60, L_ratio=1.0, bc="periodic", scale=64.0,
                                    time_scheme="SSPRK2", RKs=3, cfl=1.2),
    # tracer consumption (tracer_advection.jl:262-275)
    "np-f32-consumption": dict(T="Float32", nx=64, L_ratio=1.0, bc="nonperiodic", tracer_consumption=True,
                               topography="seamount", wind_forcing_x="double_gyre"),
    "per-f64-consumption": dict(T="Float64", nx=60, L_ratio=1.0, bc="periodic", tracer_consumption=True),
    # Sadourny's enstrophy-conserving PV advection (PV_advection.jl:31-59, rhs.jl:117-120)
    "np-f64-sadourny": dict(T="Float64", nx=64, L_ratio=1.0, bc="nonperiodic", α=0.5, diffusion="Smagorinsky",
                            bottom_drag="quadratic", topography="seamount", wind_forcing_x="double_gyre",
                            adv_scheme="Sadourny"),
    "per-f32-sadourny": dict(T="Float32", nx=66, L_ratio=1.5, bc="periodic", diffusion="Smagorinsky",
                             bottom_drag="quadratic", topography="seamount", wind_forcing_x="channel",
                             adv_scheme="Sadourny"),
    "np-f32-sadourny-rk3": dict(T="Float32", nx=61, L_ratio=0.8, bc="nonperiodic", α=2.0, diffusion="constant",
                                bottom_drag="linear", topography="ridge", adv_scheme="Sadourny", RKo=3, cfl=0.6),
    "per-f64-sadourny-ssprk3": dict(T="Float64", nx=60, L_ratio=1.0, bc="periodic", time_scheme="SSPRK3", RKn=2,
                                    cfl=1.2, adv_scheme="Sadourny"),
    "np-f16-f32-sadourny-scaled": dict(T="Float16", Tprog="Float32", nx=64, L_ratio=1.0, bc="nonperiodic", α=0.5,
                                       diffusion="Smagorinsky", bottom_drag="quadratic", topography="seamount",
                                       wind_forcing_x="double_gyre", scale=8.0, adv_scheme="Sadourny"),
    # linear dynamics (rhs.jl:55-86): under the classic RK scheme η moves by relaxation or
    # forcing only (no mass fluxes are formed), under SSPRK continuity! forms them
    "np-f64-linear": dict(T="Float64", nx=64, L_ratio=1.0, bc="nonperiodic", α=0.5, diffusion="Smagorinsky",
                          bottom_drag="quadratic", topography="seamount", wind_forcing_x="double_gyre",
                          dynamics="linear", nstep_advcor=2),
    "per-f32-linear-relax": dict(T="Float32", nx=66, L_ratio=1.5, bc="periodic", diffusion="constant",
                                 bottom_drag="linear", topography="ridge", wind_forcing_x="channel", dynamics="linear",
                                 surface_relax=True, t_relax=2.0),
    "np-f32-linear-ssprk2": dict(T="Float32", nx=62, L_ratio=0.9, bc="nonperiodic", α=1.0, diffusion="Smagorinsky",
                                 topography="seamount", wind_forcing_x="double_gyre", time_scheme="SSPRK2", RKs=3,
                                 cfl=1.2, dynamics="linear"),
    "per-f64-linear-4ssprk3": dict(T="Float64", nx=60, L_ratio=1.0, bc="periodic", time_scheme="4SSPRK3", cfl=1.0,
                                   dynamics="linear", surface_forcing=True),
    "per-bf16-f32-linear-scaled": dict(T="BFloat16", Tprog="Float32", nx=66, L_ratio=1.5, bc="periodic",
                                       diffusion="Smagorinsky", bottom_drag="quadratic", topography="seamount",
                                       wind_forcing_x="channel", scale=64.0, dynamics="linear"),
    # periodic fused stage pairs: the lanes wrap around the seam (flat bottom, exact copies)
    "per-f64-flat-fused": dict(T="Float64", nx=70, L_ratio=1.1, bc="periodic", diffusion="Smagorinsky",
                               bottom_drag="quadratic", topography="flat", wind_forcing_x="channel"),
    "per-f32-flat-comp-relax": dict(T="Float32", nx=64, L_ratio=1.0, bc="periodic", topography="flat",
                                    wind_forcing_x="shear", compensated=True, surface_relax=True, t_relax=2.0),
    "per-f64-flat-rk2-forcing": dict(T="Float64", nx=58, L_ratio=0.9, bc="periodic", topography="flat",
                                     wind_forcing_x="double_gyre", RKo=2, cfl=0.5, surface_forcing=True),
    "per-f32-flat-narrow": dict(T="Float32", nx=24, L_ratio=0.5, bc="periodic", topography="flat",
                                wind_forcing_x="channel", diffusion="Smagorinsky"),
    # periodic fused stage pairs whose seam the lanes cannot wrap around (H of the halo grid
    # differs across it, or Tcomm rounds the ghost copies): seam strips on their own kernel
    "per-f32-ridge-comp": dict(T="Float32", nx=130, L_ratio=1.2, bc="periodic", topography="ridge",
                               wind_forcing_x="channel", diffusion="Smagorinsky", compensated=True),
    "per-f64-ridges-rk2-relax": dict(T="Float64", nx=118, L_ratio=1.0, bc="periodic", topography="ridges",
                                     wind_forcing_x="double_gyre", RKo=2, cfl=0.5, surface_relax=True, t_relax=2.0),
    "per-f32-ridge-narrow": dict(T="Float32", nx=30, L_ratio=0.6, bc="periodic", topography="ridge",
                                 wind_forcing_x="channel", diffusion="Smagorinsky", bottom_drag="quadratic"),
    "per-f64-seamount-comm-f32": dict(T="Float64", Tcomm="Float32", nx=124, L_ratio=1.0, bc="periodic",
                                      topography="seamount", wind_forcing_x="double_gyre", diffusion="Smagorinsky"),
    "per-f16-f32-ridges": dict(T="Float16", Tprog="Float32", nx=126, L_ratio=1.0, bc="periodic", topography="ridges",
                               wind_forcing_x="channel", diffusion="Smagorinsky", bottom_drag="quadratic"),
    # fused stage pairs of the other right-hand-side forms (RK4, T = Tprog, a wrapping seam)
    "per-f32-flat-sadourny-advcor2": dict(T="Float32", nx=66, L_ratio=1.0, bc="periodic", topography="flat",
                                          wind_forcing_x="channel", diffusion="Smagorinsky", bottom_drag="quadratic",
                                          adv_scheme="Sadourny", nstep_advcor=2),
    "per-f64-flat-linear": dict(T="Float64", nx=70, L_ratio=1.2, bc="periodic", topography="flat",
                                wind_forcing_x="channel", dynamics="linear"),
    "np-f32-sadourny-advcor3": dict(T="Float32", nx=118, L_ratio=0.8, bc="nonperiodic", α=1.0, topography="seamount",
                                    wind_forcing_x="double_gyre", adv_scheme="Sadourny", nstep_advcor=3),
    # advection and Coriolis terms kept between updates every nstep_advcor steps
    # (time_integration.jl:204-211; the quadratic drag refreshes the kinetic part each
    # diffusion step)
    "np-f64-advcor3": dict(T="Float64", nx=64, L_ratio=1.0, bc="nonperiodic", α=0.5, diffusion="Smagorinsky",
                           bottom_drag="quadratic", topography="seamount", wind_forcing_x="double_gyre",
                           nstep_advcor=3),
    "per-f64-advcor1": dict(T="Float64", nx=72, L_ratio=1.1, bc="periodic", α=1.0, diffusion="Smagorinsky",
                            bottom_drag="quadratic", topography="seamount", wind_forcing_x="double_gyre",
                            nstep_advcor=1),
    "per-f32-advcor2-sadourny": dict(T="Float32", nx=66, L_ratio=1.5, bc="periodic", diffusion="constant",
                                     bottom_drag="linear", topography="ridge", wind_forcing_x="channel",
                                     adv_scheme="Sadourny", nstep_advcor=2),
    "np-f32-advcor4-ssprk2": dict(T="Float32", nx=62, L_ratio=0.9, bc="nonperiodic", α=1.0, diffusion="Smagorinsky",
                                  topography="seamount", wind_forcing_x="double_gyre", time_scheme="SSPRK2", RKs=3,
                                  cfl=1.2, nstep_advcor=4, nstep_diff=2),
    "np-f16-f32-advcor2": dict(T="Float16", Tprog="Float32", nx=64, L_ratio=1.0, bc="nonperiodic", α=0.5,
                               diffusion="Smagorinsky", bottom_drag="quadratic", topography="seamount",
                               wind_forcing_x="double_gyre", nstep_advcor=2),
    "per-bf16-f32-advcor3-scaled": dict(T="BFloat16", Tprog="Float32", nx=66, L_ratio=1.5, bc="periodic",
                                        diffusion="Smagorinsky", bottom_drag="none", topography="seamount",
                                        wind_forcing_x="channel", scale=64.0, nstep_advcor=3),
}


def n_stages(P):
    return {"RK": P.RKo, "SSPRK2": P.RKs, "SSPRK3": P.RKn ** 2, "4SSPRK3": 4}[P.time_scheme]


def setup_case(swm, name, seed=11, σ=0.8, **over):
    kw = dict(CASES[name])
    kw.update(over)
    kw.setdefault("Uadv", 20.0)          # tracer every few steps
    P = swm.Parameter(**kw)
    S = swm.make_setup(P)
    prog = swm.seeded_initial_conditions(S, seed=seed, σ=σ)
    return P, S, prog


def ghosted(swm, S, u, v, e):
    """Apply the host ghost points (the reference's ghost_points!) to copies."""
    T = S.parameters.Tprog
    u, v, e = (swm.to_float64(a, T).copy(order="F") for a in (u, v, e))
    from shallowwaters_amd.initial_conditions import ghost_points
    ghost_points(u, v, e, S)
    return u, v, e


def assert_same(name, ref, got, T):
    ref64 = np.asarray(ref, dtype=np.float64)
    got64 = np.asarray(got, dtype=np.float64)
    mm = first_mismatch(ref64, got64)
    err = rel_l2(ref64, got64)
    assert err <= TOL[T], f"{name}: rel L2 {err:.3e}; {mm}"
    assert mm is None, f"{name}: not bitwise equal ({err:.3e}): {mm}"


@pytest.mark.parametrize("case", list(CASES))
def test_rk_stages_bitwise(swm, orc, gpu, case):
    P, S, prog = setup_case(swm, case)
    cfg = swm.make_config(S)
    T = P.Tprog
    ns = n_stages(P)
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        for k in range(1, ns + 1):
            sim.stages(k, 1, 0)
            g_u0, g_v0, g_e0, g_u1, g_v1, g_e1 = sim.download_rk()
            *_, rk, _, _ = orc.run(S, cfg, prog, 1, stop_stage=k, want_rk=True)
            r_u0, r_v0, r_e0, r_u1, r_v1, r_e1 = rk[:6]
            # u1 enters the next stage after ghost_points! (time_integration.jl:57,84,118,162)
            r_u1, r_v1, r_e1 = ghosted(swm, S, r_u1, r_v1, r_e1)
            if P.T != P.Tprog:      # u1rhs = T(u1): the GPU keeps only that (time_integration.jl:62-64)
                # (velocities carried scaled are rounded to T as s·u, then unscaled)
                sc = P.scale
                r_u1, r_v1 = (swm.to_float64(swm.to_storage(a * sc, P.T), P.T) / sc for a in (r_u1, r_v1))
                r_e1 = swm.to_float64(swm.to_storage(r_e1, P.T), P.T)
            assert_same(f"{case} stage {k} u1", r_u1, g_u1, T)
            assert_same(f"{case} stage {k} v1", r_v1, g_v1, T)
            assert_same(f"{case} stage {k} η1", r_e1, g_e1, T)
            if k < ns:
                if P.time_scheme == "RK":   # u0 halo is not maintained between stages; compare the interior
                    assert_same(f"{case} stage {k} u0", r_u0[2:-2, 2:-2], g_u0[2:-2, 2:-2], T)
                    assert_same(f"{case} stage {k} v0", r_v0[2:-2, 2:-2], g_v0[2:-2, 2:-2], T)
                    assert_same(f"{case} stage {k} η0", r_e0[1:-1, 1:-1], g_e0[1:-1, 1:-1], T)
            else:           # ghost_points!(u0,v0,η0) after the stages (:198)
                r_u0, r_v0, r_e0 = ghosted(swm, S, r_u0, r_v0, r_e0)
                assert_same(f"{case} stage {k} u0", r_u0, g_u0, T)
                assert_same(f"{case} stage {k} v0", r_v0, g_v0, T)
                assert_same(f"{case} stage {k} η0", r_e0, g_e0, T)


@pytest.mark.parametrize("case", list(CASES))
def test_ten_steps_bitwise(swm, orc, gpu, case):
    P, S, prog = setup_case(swm, case)
    assert S.grid.nadvstep <= 8, S.grid.nadvstep     # the tracer fires within the run
    nsteps = 10
    r = orc.run(S, swm.make_config(S), prog, nsteps)
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        done, nan_step = sim.step(nsteps, 1, 0)
        g = sim.download()
    assert (done, nan_step) == (nsteps, 0)
    for name, a, b in zip(("u", "v", "η", "sst"), r[:4], (g.u, g.v, g.η, g.sst)):
        assert_same(f"{case} {name}", swm.to_float64(a, P.Tprog), swm.to_float64(b, P.Tprog), P.Tprog)


def test_nstep_diff_and_chunked_calls(swm, orc, gpu):
    """Dissipation every 3rd step, steps issued in two calls with i0/t0 carried over."""
    P, S, prog = setup_case(swm, "per-f64-smag-quad", nstep_diff=3)
    r = orc.run(S, swm.make_config(S), prog, 11)
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        assert sim.step(4, 1, 0) == (4, 0)
        assert sim.step(7, 5, 4 * S.grid.dtint) == (7, 0)
        g = sim.download()
    for name, a, b in zip(("u", "v", "η", "sst"), r[:4], (g.u, g.v, g.η, g.sst)):
        assert_same(f"chunked {name}", a, b, P.Tprog)


def test_exact_zero_known_answer(swm, gpu):
    """test/runtests.jl:4-9: no forcing from rest stays exactly zero (1 model day)."""
    P = swm.Parameter(T="Float32", Ndays=1, Fx0=0)
    prog = swm.run_model(P)
    assert np.all(prog.u == 0) and np.all(prog.v == 0) and np.all(prog.η == 0)


def test_nan_detection_stops_before_copy_back(swm, orc, gpu):
    """feedback.jl:136-148 + time_integration.jl:235-237: NaNs seen at i % nout == 0 stop
    the loop and Prog keeps the state of the previous step."""
    P, S, prog = setup_case(swm, "np-f64-smag-quad", tracer_advection=False, output_dt=0.25)
    nout = S.grid.nout
    assert 1 <= nout <= 4, nout
    u = prog.u.copy(order="F")
    u[10, 12] = np.nan
    prog.u = u
    r = orc.run(S, swm.make_config(S), prog, 10)
    assert r[6] == nout
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        done, nan_step = sim.step(10, 1, 0)
        g = sim.download()
    assert (done, nan_step) == (nout - 1, nout)
    for name, a, b in zip(("u", "v", "η", "sst"), r[:4], (g.u, g.v, g.η, g.sst)):
        assert np.array_equal(a, b, equal_nan=True), name


def test_upload_rejects_nonzero_unwritten_halo(swm, gpu):
    P, S, prog = setup_case(swm, "np-f32-const-lin")
    u = prog.u.copy(order="F")
    u[0, 10] = 1.0          # normal-velocity halo of a nonperiodic u
    prog.u = u
    with swm.Simulation(S) as sim:
        with pytest.raises(Exception, match="halo"):
            sim.upload(prog)


def test_unsupported_options_fail_loudly(swm, gpu):
    for kw in (dict(T="Float16", Tprog="Float16"), dict(T="BFloat16", Tprog="Float32", time_scheme="SSPRK2")):
        P = swm.Parameter(nx=40, **kw)
        S = swm.make_setup(P)
        with pytest.raises(Exception):
            swm.Simulation(S)


def test_x_shift_equivariance_gpu(swm, gpu):
    """The GPU step commutes with a circular x shift of the periodic flat channel away from
    the seams (tests/test_oracle.py::check_shift_equivariance)."""
    from test_oracle import check_shift_equivariance, shifted_channel
    k = 48
    P, S, prog, shifted = shifted_channel(swm, k)
    out = []
    for p in (prog, shifted):
        with swm.Simulation(S) as sim:
            sim.upload(p)
            assert sim.step(1, 1, 0) == (1, 0)
            g = sim.download()
        out.append((g.u, g.v, g.η, g.sst))
    check_shift_equivariance(swm, S, out[0], out[1], k)


@pytest.mark.parametrize("case", ["np-f64-smag-quad", "per-f32-const-none"])
def test_reduce_matches_fp64_sums(swm, gpu, case):
    """swm_reduce: FP64 mass and energy of the current Prog against numpy FP64 sums over
    the downloaded fields (mass = Σ η, energy = Σ ½gη² + ½h u² + ½h v² per T-cell with
    u at the cell's right face, v at its top face, h = η + H)."""
    P, S, prog = setup_case(swm, case)
    G, F = S.grid, S.forcing
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        sim.step(3, 1, 0)
        mass, energy = sim.reduce()
        g = sim.download()
    T = P.Tprog
    u, v, e = (swm.to_float64(a, T) for a in (g.u, g.v, g.η))
    H = swm.to_float64(F.H, P.T)
    ep = 1 if P.bc == "periodic" else 0
    eta = e[1:G.nx + 1, 1:G.ny + 1]
    h = eta + H[1:G.nx + 1, 1:G.ny + 1]
    # Julia u[x+2+ep, y+2] and v[x+2, y+2] are the faces of T-cell (x, y)
    uf = u[1 + ep + 1:1 + ep + 1 + G.nx, 2:2 + G.ny]                # x = 1..nx (x = nx: zero wall halo if ep = 0)
    vf = np.zeros_like(eta)
    vf[:, :G.ny - 1] = v[2:2 + G.nx, 2:2 + G.ny - 1]                # y = 1..ny-1
    g_ = float(swm.make_config(S).g)
    m_ref = eta.sum()
    e_ref = (0.5 * g_ * eta * eta + 0.5 * h * uf * uf + 0.5 * h * vf * vf).sum()
    assert abs(mass - m_ref) <= 1e-12 * np.abs(eta).sum(), (mass, m_ref)
    assert abs(energy - e_ref) <= 1e-12 * abs(e_ref), (energy, e_ref)


@pytest.mark.parametrize("T,Tprog", [("Float32", "Float32"), ("Float64", "Float64"), ("BFloat16", "Float32")])
def test_fused_stage_pairs_on_a_wider_grid(swm, orc, gpu, T, Tprog, monkeypatch):
    """Consecutive RK stages run as fused pairs (nonperiodic); on a grid of several strips
    and segments both the fused and the one-kernel-per-stage path equal the oracle bitwise."""
    P = swm.Parameter(T=T, Tprog=Tprog, nx=181, L_ratio=1.7, bc="nonperiodic", α=0.7, diffusion="Smagorinsky",
                      bottom_drag="quadratic", topography="seamount", wind_forcing_x="double_gyre", Uadv=20.0)
    S = swm.make_setup(P)
    prog = swm.seeded_initial_conditions(S, seed=23, σ=0.8)
    r = orc.run(S, swm.make_config(S), prog, 6)
    for fuse in ("1", "0"):
        monkeypatch.setenv("SWM_FUSE", fuse)
        with swm.Simulation(S) as sim:
            sim.upload(prog)
            assert sim.step(6, 1, 0) == (6, 0)
            g = sim.download()
        for name, a, b in zip(("u", "v", "η", "sst"), r[:4], (g.u, g.v, g.η, g.sst)):
            assert_same(f"fuse={fuse} {name}", swm.to_float64(a, Tprog), swm.to_float64(b, Tprog), Tprog)


def test_linear_dynamics_needs_f_on_the_u_and_v_grids(swm, gpu):
    """The C ABI: a step of the linear dynamics without swm_set_static_linear fails loudly
    (it would read f_u, f_v), and a nonlinear context refuses the call."""
    import ctypes
    from shallowwaters_amd import _abi
    L = _abi.load()
    for dyn in ("linear", "nonlinear"):
        P, S, prog = setup_case(swm, "np-f32-const-lin", dynamics=dyn)
        cfg = swm.make_config(S)
        ctx = ctypes.c_void_p()
        _abi.check(L.swm_create(ctypes.byref(cfg), ctypes.byref(ctx)), None, "swm_create")
        try:
            F, G = S.forcing, S.grid
            st = [np.asfortranarray(a) for a in (F.H, F.Fx, F.Fy, G.f_q, F.η_ref, F.Fη)]
            _abi.check(L.swm_set_static(ctx, *[_abi.ptr(a) for a in st]), ctx, "swm_set_static")
            arrs = [np.asfortranarray(a) for a in (prog.u, prog.v, prog.η, prog.sst)]
            _abi.check(L.swm_upload_prog(ctx, *[_abi.ptr(a) for a in arrs]), ctx, "swm_upload_prog")
            fuv = [np.asfortranarray(a) for a in (G.f_u, G.f_v)]
            if dyn == "linear":
                done, nan = ctypes.c_int64(0), ctypes.c_int64(0)
                with pytest.raises(_abi.SwmError, match="swm_set_static_linear first"):
                    _abi.check(L.swm_step(ctx, 1, 1, 0, ctypes.byref(done), ctypes.byref(nan)), ctx, "swm_step")
                _abi.check(L.swm_set_static_linear(ctx, *[_abi.ptr(a) for a in fuv]), ctx, "swm_set_static_linear")
                _abi.check(L.swm_step(ctx, 1, 1, 0, ctypes.byref(done), ctypes.byref(nan)), ctx, "swm_step")
                assert (done.value, nan.value) == (1, 0)
            else:
                with pytest.raises(_abi.SwmError, match="linear dynamics only"):
                    _abi.check(L.swm_set_static_linear(ctx, *[_abi.ptr(a) for a in fuv]), ctx, "swm_set_static_linear")
        finally:
            L.swm_destroy(ctx)


def test_time_steps_advances_like_step(swm, gpu):
    """swm_time_steps runs the same steps as swm_step (bitwise the same state) and times
    each kernel launch inside them: one fused pair of each kind, one dissipation and one
    step per step, a tracer advection every nadvstep steps."""
    P, S, prog = setup_case(swm, "np-f64-smag-quad")
    n = 2 * S.grid.nadvstep + 1
    with swm.Simulation(S) as a, swm.Simulation(S) as b:
        a.upload(prog)
        b.upload(prog)
        t = a.time_steps(n, 1, 0)
        assert b.step(n, 1, 0) == (n, 0)
        ya, yb = a.download(), b.download()
    for x, y in zip((ya.u, ya.v, ya.η, ya.sst), (yb.u, yb.v, yb.η, yb.sst)):
        assert np.array_equal(x, y)
    assert {k: c for k, (_, c) in t.items()} == {"fused12": n, "fused34": n, "dissipation": n, "step": n, "tracer": 2}
    assert all(ms > 0 for ms, _ in t.values())
    assert t["step"][0] > t["fused12"][0] + t["fused34"][0]
