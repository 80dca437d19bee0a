"""swm_reduce_noalloc: the reduction of swm_reduce with its scratch in the library's static
device array instead of per-context memory; equal to swm_reduce bitwise, on two calls in
a row, the second on a smaller grid (the array keeps the first call's larger contents)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def reductions(swm, sim):
    from shallowwaters_amd import _abi
    L = _abi.load()
    out = []
    for fn in (L.swm_reduce, L.swm_reduce_noalloc):
        m, e = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _abi.check(fn(sim.ctx, ctypes.byref(m), ctypes.byref(e)), sim.ctx, fn.__name__)
        out.append((m.value, e.value))
    return out


@pytest.mark.parametrize("T", ["Float32", "Float64"])
def test_noalloc_reduce_equals_reduce(swm, gpu, T):
    cases = [dict(nx=300, L_ratio=0.75, bc="nonperiodic"), dict(nx=96, L_ratio=2.0, bc="periodic")]
    for kw in cases:                                   # the second call: a smaller grid
        S = swm.make_setup(swm.Parameter(T=T, topography="seamount", wind_forcing_x="double_gyre", **kw))
        print(f"noalloc reduce {T} {S.grid.nx}x{S.grid.ny}", flush=True)
        with swm.Simulation(S) as sim:
            sim.upload(swm.seeded_initial_conditions(S, seed=11, σ=0.4))
            assert sim.step(3, 1, 0) == (3, 0)
            (m0, e0), (m1, e1) = reductions(swm, sim)
            assert (m0, e0) == (m1, e1), (kw, m0, m1, e0, e1)
            assert np.isfinite(m0) and e0 > 0


def test_noalloc_reduce_rejects_too_many_rows(swm, gpu):
    from shallowwaters_amd import _abi
    S = swm.make_setup(swm.Parameter(T="Float32", nx=3, L_ratio=3 / 32769, bc="nonperiodic"))
    assert S.grid.ny == 32769
    with swm.Simulation(S) as sim:
        sim.upload(swm.seeded_initial_conditions(S, seed=1, σ=0.1))
        m, e = ctypes.c_double(0.0), ctypes.c_double(0.0)
        rc = _abi.load().swm_reduce_noalloc(sim.ctx, ctypes.byref(m), ctypes.byref(e))
        assert rc == -1
        assert b"32768" in _abi.load().swm_last_error(sim.ctx)
