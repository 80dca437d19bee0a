"""The static fields computed on the device (swm_set_static_device) against the host's
Forcing{T} and Grid fields (src/forcing.jl:11-217, src/grid.jl:91-108), which the host
evaluates in Float64 with numpy's libm and rounds to T.

f_q, f_u, f_v, the flat bottom and the constant winds use + - * / only, in the same order on
both sides, so they agree bitwise.  The fields that go through exp, cos, sin or tanh use the
device library's double functions on one side and numpy's on the other; these may differ in
the last bit of a double, which moves a value of T only when it lies within that bit of a
rounding boundary of T: for T narrower than Float64 at most one unit in the last place of
T in a handful of elements; for Float64 a few units in the last place of the field's largest
value (the libm difference carried through cos², sums and products).  Where the fields agree bitwise, a run stepped with
them agrees bitwise with a run stepped with the uploaded ones."""
import numpy as np
import pytest

from conftest import first_mismatch

pytestmark = pytest.mark.gpu

CASES = {
    "np-f32-seamount-gyre": dict(T="Float32", bc="nonperiodic", topography="seamount", wind_forcing_x="double_gyre",
                                 Fx0=0.12, diffusion="Smagorinsky", bottom_drag="quadratic"),
    "per-f64-ridge-channel-relax": dict(T="Float64", bc="periodic", topography="ridge", wind_forcing_x="channel",
                                        surface_relax=True, t_relax=5.0),
    "np-f32-ridges-shear-forcing": dict(T="Float32", bc="nonperiodic", topography="ridges", wind_forcing_x="shear",
                                        surface_forcing=True),
    "per-f64-flat-const": dict(T="Float64", bc="periodic", topography="flat", wind_forcing_x="constant", Fy0=0.03),
    "np-bf16-f32-seamount-gyre": dict(T="BFloat16", Tprog="Float32", bc="nonperiodic", topography="seamount",
                                      wind_forcing_x="double_gyre", Fx0=0.12),
    "per-f16-f32-scaled-ridge": dict(T="Float16", Tprog="Float32", bc="periodic", topography="ridge",
                                     wind_forcing_x="channel", scale=4.0),
    "np-f32-none": dict(T="Float32", bc="nonperiodic", topography="flat", wind_forcing_x="none",
                        wind_forcing_y="none"),
}


def setup(swm, kw, nx=96):
    P = swm.Parameter(nx=nx, L_ratio=1, **kw)
    return swm.make_setup(P)


def ulps(a, b, T):
    """|a - b| in units of the last place of T at max(|a|, |b|) (element-wise)."""
    e = {"Float64": 52, "Float32": 23, "Float16": 10, "BFloat16": 7}[T]
    m = np.maximum(np.abs(a), np.abs(b))
    unit = np.where(m > 0, np.exp2(np.floor(np.log2(np.where(m > 0, m, 1.0))) - e), 1.0)
    return np.abs(a - b) / unit


@pytest.mark.parametrize("case", list(CASES))
def test_device_statics_match_host(swm, gpu, case):
    S = setup(swm, CASES[case])
    T = S.parameters.T
    to64 = lambda a: swm.to_float64(np.asarray(a), T)
    with swm.Simulation(S) as host, swm.Simulation(S, statics="device") as dev:
        h, d = host.download_static(), dev.download_static()
        assert set(h) == set(d)
        # the host context hands back what it was given (scaled where velocities are scaled)
        s = 2.0 ** int(np.log2(S.parameters.scale)) if S.parameters.scale != 1 else 1.0
        F, G = S.forcing, S.grid
        given = {"H": F.H, "Fx": F.Fx, "Fy": F.Fy, "f_q": G.f_q, "η_ref": F.η_ref, "Fη": F.Fη}
        for k in h:
            sk = s if k in ("Fx", "Fy", "f_q") else 1.0
            assert first_mismatch(to64(given[k]) * sk, to64(h[k])) is None, f"{case} {k}: upload round trip"
        inexact = 0
        for k in h:
            a, b = to64(h[k]), to64(d[k])
            assert a.shape == b.shape
            if k in ("f_q",) or (k == "H" and S.parameters.topography == "flat") or \
                    (k in ("Fx", "Fy") and S.parameters.wind_forcing_x in ("constant", "none")):
                assert first_mismatch(a, b) is None, f"{case} {k}: arithmetic-only field differs"
            inexact += int(np.count_nonzero(a != b))
            if T == "Float64":
                # the libm difference itself, carried through cos², sums and products: a few
                # units in the last place of the field's largest value
                scale = float(np.max(np.abs(a))) or 1.0
                err = float(np.max(np.abs(a - b))) / (scale * 2.0 ** -52)
                assert err <= 8.0, f"{case} {k}: {err} ulp of max|{k}|"
            else:
                u = ulps(a, b, T)
                assert np.all(u <= 1.0), f"{case} {k}: max {u.max()} ulp of {T}"
                assert np.count_nonzero(a != b) <= max(4, a.size // 2000), f"{case} {k}: {np.count_nonzero(a != b)} differ"
        if inexact == 0:       # identical statics: identical runs
            prog = swm.seeded_initial_conditions(S, seed=5, σ=0.3)
            out = []
            for sim in (host, dev):
                sim.upload(prog)
                assert sim.step(12, 1, 0) == (12, 0)
                y = sim.download()
                out.append((y.u, y.v, y.η, y.sst))
            for name, a, b in zip(("u", "v", "η", "sst"), *out):
                mm = first_mismatch(np.asarray(a, np.float64), np.asarray(b, np.float64))
                assert mm is None, f"{case} {name}: {mm}"


@pytest.mark.parametrize("T", ["Float32", "Float64"])
def test_linear_dynamics_f_u_f_v(swm, gpu, T):
    """The linear dynamics read f_u and f_v; with a flat bottom and constant winds every
    device field is arithmetic-only, so 20 steps agree bitwise with the uploaded fields."""
    S = setup(swm, dict(T=T, bc="nonperiodic", dynamics="linear", topography="flat", wind_forcing_x="constant",
                        Fy0=0.02, bottom_drag="linear", diffusion="constant"))
    prog = swm.seeded_initial_conditions(S, seed=9, σ=0.3)
    out = []
    for statics in ("host", "device"):
        with swm.Simulation(S, statics=statics) as sim:
            sim.upload(prog)
            assert sim.step(20, 1, 0) == (20, 0)
            y = sim.download()
            out.append((y.u, y.v, y.η))
    for name, a, b in zip(("u", "v", "η"), *out):
        mm = first_mismatch(np.asarray(a, np.float64), np.asarray(b, np.float64))
        assert mm is None, f"{T} linear {name}: {mm}"


def test_device_setup_errors(swm, gpu):
    from importlib import import_module
    import ctypes
    abi = import_module("shallowwaters_amd._abi")
    ti = import_module("shallowwaters_amd.time_integration")
    S = setup(swm, dict(T="Float32"))
    with swm.Simulation(S) as sim:
        with pytest.raises(abi.SwmError, match="already set"):
            abi.check(abi.load().swm_set_static_device(sim.ctx, ctypes.byref(ti.setup_of(S))), sim.ctx, "dev")
    st = ti.setup_of(S)
    st.wind_y = abi.WIND["channel"]
    cfg = swm.make_config(S)
    ctx = ctypes.c_void_p()
    L = abi.load()
    abi.check(L.swm_create(ctypes.byref(cfg), ctypes.byref(ctx)), None, "swm_create")
    try:
        with pytest.raises(abi.SwmError, match="wind_y must be constant or none"):
            abi.check(L.swm_set_static_device(ctx, ctypes.byref(st)), ctx, "dev")
        with pytest.raises(abi.SwmError, match="not set"):
            abi.check(L.swm_download_static(ctx, None, None, None, None, None, None), ctx, "download")
    finally:
        L.swm_destroy(ctx)
    with pytest.raises(ValueError, match="statics"):
        swm.Simulation(S, statics="gpu")
