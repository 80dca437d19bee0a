"""Single operators of the step on the GPU against the oracle's same operators, bitwise:
ghost points (swm_ghost, and through the upload), the dissipation (swm_dissipation: drag,
diffusion, their tendencies and the velocity ghost points) and the tracer (swm_tracer:
mid-point snapshot, then the semi-Lagrangian advection), each on its own so a mismatch is
located in one operator."""
import numpy as np
import pytest

from conftest import first_mismatch

pytestmark = pytest.mark.gpu

CASES = {
    "np-f32-smag-quad": dict(T="Float32", bc="nonperiodic", diffusion="Smagorinsky", bottom_drag="quadratic", α=0.5),
    "per-f64-smag-quad": dict(T="Float64", bc="periodic", diffusion="Smagorinsky", bottom_drag="quadratic"),
    "np-f64-const-lin": dict(T="Float64", bc="nonperiodic", diffusion="constant", bottom_drag="linear", α=1.0),
    "per-f32-const-none": dict(T="Float32", bc="periodic", diffusion="constant", bottom_drag="none"),
    "np-bf16-f32-smag-quad": dict(T="BFloat16", Tprog="Float32", bc="nonperiodic", diffusion="Smagorinsky",
                                  bottom_drag="quadratic", α=0.5),
    "per-f16-f32-scaled": dict(T="Float16", Tprog="Float32", bc="periodic", diffusion="Smagorinsky",
                               bottom_drag="quadratic", scale=4.0),
    "np-f32-consume": dict(T="Float32", bc="nonperiodic", diffusion="Smagorinsky", tracer_consumption=True),
}


def setup(swm, kw, nx=74):
    P = swm.Parameter(nx=nx, L_ratio=0.8, topography="seamount", wind_forcing_x="double_gyre", Uadv=30.0, **kw)
    S = swm.make_setup(P)
    return S, swm.make_config(S), swm.seeded_initial_conditions(S, seed=21, σ=0.6)


def fields(p):
    return (p.u, p.v, p.η, p.sst)


def assert_bitwise(tag, ref, got):
    for name, a, b in zip(("u", "v", "η", "sst"), ref, got):
        mm = first_mismatch(np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64))
        assert mm is None, f"{tag} {name}: {mm}"


@pytest.mark.parametrize("case", list(CASES))
def test_ghost_points(swm, orc, gpu, case):
    S, cfg, prog = setup(swm, CASES[case])
    # the halo cells ghost_points! writes, filled with junk: the upload (and swm_ghost) rewrite them
    junk = swm.PrognosticVars(*(np.array(a, copy=True, order="F") for a in fields(prog)), T=S.parameters.Tprog)
    junk.η[0, :] = 7.0
    junk.η[:, -1] = -3.0
    junk.sst[0, :] = 5.0
    ref = orc.op(S, cfg, "ghost", *fields(junk))[:4]
    with swm.Simulation(S) as sim:
        sim.upload(junk)
        assert_bitwise(f"{case} upload", ref, fields(sim.download()))
        sim.ghost()                                 # ghost_points! is idempotent
        assert_bitwise(f"{case} swm_ghost", ref, fields(sim.download()))


@pytest.mark.parametrize("case", list(CASES))
def test_dissipation(swm, orc, gpu, case):
    S, cfg, prog = setup(swm, CASES[case])
    g = orc.op(S, cfg, "ghost", *fields(prog))[:4]
    ref = orc.op(S, cfg, "dissipation", *g)[:4]
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        sim.dissipation()
        got = sim.download()
    assert_bitwise(case, ref, fields(got))
    assert not np.array_equal(ref[0], g[0])          # the operator changed the velocities


@pytest.mark.parametrize("case", list(CASES))
def test_tracer(swm, orc, gpu, case):
    """Snapshot step with one state, advection step with another (um, vm kept between)."""
    S, cfg, prog = setup(swm, CASES[case])
    G = S.grid
    assert G.nadvstep >= 2
    i_snap = G.nadvstep - G.nadvstep_half          # (i + nadvstep_half) % nadvstep == 0
    i_adv = G.nadvstep
    second = swm.seeded_initial_conditions(S, seed=22, σ=0.6)
    a = orc.op(S, cfg, "ghost", *fields(prog))[:4]
    b = orc.op(S, cfg, "ghost", *fields(second))[:4]
    T = S.parameters.T
    zu = np.zeros((G.nux + 4, G.nuy + 4))
    zv = np.zeros((G.nvx + 4, G.nvy + 4))
    _, _, _, _, mid = orc.op(S, cfg, "tracer", *a, i=i_snap, um=zu, vm=zv)
    _, _, _, ref_sst, _ = orc.op(S, cfg, "tracer", *b, i=i_adv, um=mid["um"], vm=mid["vm"])
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        sim.tracer(i_snap)
        sim.upload(second)
        sim.tracer(i_adv)
        got = sim.download()
    mm = first_mismatch(np.asarray(ref_sst, dtype=np.float64), np.asarray(got.sst, dtype=np.float64))
    assert mm is None, f"{case} {T} sst: {mm}"
    assert not np.array_equal(ref_sst, b[3])          # the tracer moved
