"""The GPU step at the benchmark's full sizes: the 4096x4096 double gyre (Float32, fused
stage pairs) and the 2048x1024 periodic channel with the tracer (Float32 and Float64).

The oracle finishes one 4096^2 step in about two seconds on one core, so two steps are
compared bitwise at full size; longer runs are checked through properties that do not
depend on the size: the fused and the one-kernel-per-stage paths agree bitwise, continuity
conserves mass (FP64 sum of η through swm_reduce) up to the rounding of the Float32 state,
and the periodic channel commutes with a circular shift of its columns away from the seam.
"""
import numpy as np
import pytest

from conftest import first_mismatch

pytestmark = pytest.mark.gpu


def double_gyre_4096(swm):
    """The benchmark workload: 4096^2, Lx = Ly = 3840 km, nonperiodic, seamount, double gyre
    wind (Fx0 = 0.12 Pa), Smagorinsky, quadratic drag, tracer on."""
    return swm.Parameter(T="Float32", nx=4096, Lx=3840e3, L_ratio=1, bc="nonperiodic", α=0.0,
                         topography="seamount", wind_forcing_x="double_gyre", Fx0=0.12, diffusion="Smagorinsky",
                         bottom_drag="quadratic", tracer_advection=True, Ndays=10)


def channel_2048(swm, T):
    """The periodic channel: 2048x1024, flat bottom, channel wind, tracer on (every 300 steps)."""
    return swm.Parameter(T=T, nx=2048, L_ratio=2, bc="periodic", topography="flat", wind_forcing_x="channel",
                         Fx0=0.12, tracer_advection=True)


def run_gpu(swm, S, prog, n, fuse, monkeypatch):
    monkeypatch.setenv("SWM_FUSE", fuse)
    with swm.Simulation(S) as sim:
        sim.upload(prog)
        assert sim.step(n, 1, 0) == (n, 0)
        g = sim.download()
        mass, _ = sim.reduce()
    return g, mass


def assert_bitwise(tag, ref, got):
    for name, a, b in zip(("u", "v", "η", "sst"), ref, got):
        mm = first_mismatch(np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64))
        assert mm is None, f"{tag} {name}: {mm}"


@pytest.mark.parametrize("which", ["double_gyre_4096", "channel_2048_f32", "channel_2048_f64"])
def test_full_size_two_steps_bitwise(swm, orc, gpu, which, monkeypatch):
    P = double_gyre_4096(swm) if which == "double_gyre_4096" else channel_2048(swm, which[-3:].replace("f", "Float"))
    S = swm.make_setup(P)
    G = S.grid
    print(f"{which}: {G.nx}x{G.ny} {P.T} {P.bc}", flush=True)
    prog = swm.seeded_initial_conditions(S, seed=2024, σ=0.1)
    r = orc.run(S, swm.make_config(S), prog, 2)
    for fuse in ("1", "0"):
        g, _ = run_gpu(swm, S, prog, 2, fuse, monkeypatch)
        assert_bitwise(f"{which} fuse={fuse}", r[:4], (g.u, g.v, g.η, g.sst))


@pytest.mark.parametrize("which", ["double_gyre_4096", "channel_2048_f32"])
def test_full_size_fused_equals_per_stage_and_conserves_mass(swm, gpu, which, monkeypatch):
    P = double_gyre_4096(swm) if which == "double_gyre_4096" else channel_2048(swm, "Float32")
    S = swm.make_setup(P)
    G = S.grid
    n = 50
    print(f"{which}: {G.nx}x{G.ny}, {n} steps", flush=True)
    prog = swm.seeded_initial_conditions(S, seed=7, σ=0.1)
    m0 = prog.η[1:-1, 1:-1].astype(np.float64).sum()
    scale = np.abs(prog.η[1:-1, 1:-1].astype(np.float64)).sum()
    a, ma = run_gpu(swm, S, prog, n, "1", monkeypatch)
    b, mb = run_gpu(swm, S, prog, n, "0", monkeypatch)
    assert_bitwise(f"{which} fused vs per stage", (a.u, a.v, a.η, a.sst), (b.u, b.v, b.η, b.sst))
    assert ma == mb
    # 4 stages x n steps of Float32 updates of every cell; unbiased rounding keeps the drift
    # of the sum far below this bound, a missing or doubled flux term does not
    assert abs(ma - m0) <= 1e-6 * scale, (m0, ma, scale)
    assert not np.array_equal(a.η, prog.η)


def test_full_size_channel_shift_equivariance(swm, gpu, monkeypatch):
    """A circular shift of the periodic channel's columns by k commutes with 2 steps, away
    from the seam (the AHβ/AHγ wrap sits there; its effect spreads by the stencils' reach,
    about 14 columns a step)."""
    from test_oracle import check_shift_equivariance
    P = channel_2048(swm, "Float32")
    S = swm.make_setup(P)
    G = S.grid
    k = 512
    prog = swm.seeded_initial_conditions(S, seed=3, σ=0.1)
    sh = swm.PrognosticVars(*swm.add_halo(*(np.roll(a, k, axis=0) for a in
                                            swm.remove_halo(prog.u, prog.v, prog.η, prog.sst, S)), S, T="Float32"),
                            T="Float32")
    print(f"channel 2048x1024 shift {k}", flush=True)
    out = []
    for p in (prog, sh):
        g, _ = run_gpu(swm, S, p, 2, "1", monkeypatch)
        out.append((g.u, g.v, g.η, g.sst))
    check_shift_equivariance(swm, S, out[0], out[1], k, far=40)
