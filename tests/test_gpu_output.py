"""run_model with NetCDF output on the GPU: the time loop runs in chunks between the steps
where the host writes output or feedback; the records and the final state equal single
uninterrupted runs of the same number of steps bitwise, and a restart from the last record
starts from that state."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

pytestmark = pytest.mark.gpu


def test_run_model_output_records_and_restart(swm, gpu, tmp_path):
    kw = dict(T="Float32", nx=64, L_ratio=2, bc="periodic", Ndays=0.25, output_dt=1.0, Uadv=5.0)
    P = swm.Parameter(output=True, outpath=str(tmp_path), output_vars=["u", "v", "η", "sst", "du", "dv", "dη"], **kw)
    S = swm.make_setup(P)
    G = S.grid
    assert G.nt // G.nout >= 3
    print(f"run_model nx={G.nx} nt={G.nt} nout={G.nout} nadvstep={G.nadvstep}", flush=True)
    y = swm.run_model(P)
    run = tmp_path / "run0000"
    assert (run / "progress.txt").read_text().strip().endswith("s.")      # "Integration done in ...s."
    k = G.nt // G.nout                                                    # the last record

    def state_after(n):
        S0 = swm.make_setup(swm.Parameter(**kw))
        prog = swm.initial_conditions("Float32", S0)
        with swm.Simulation(S0) as sim:
            sim.upload(prog)
            if n:
                assert sim.step(n, 1, 0) == (n, 0)
            g = sim.download()
        return swm.remove_halo(g.u, g.v, g.η, g.sst, S0)

    end = state_after(G.nt)
    for a, b in zip((y.u, y.v, y.η, y.sst), end):
        assert np.array_equal(a, b)
    at, before = state_after(k * G.nout), state_after(k * G.nout - 1)
    for name, var, i in (("u", "u", 0), ("v", "v", 1), ("eta", "eta", 2), ("sst", "sst", 3)):
        with netcdf_file(str(run / f"{name}.nc"), "r", mmap=False) as nc:
            assert list(nc.variables["t"][:]) == [j * G.nout * G.dtint for j in range(k + 1)]
            assert np.array_equal(nc.variables[var][k].T, at[i].astype(np.float32))
    for name, i in (("du", 0), ("dv", 1), ("deta", 2)):
        with netcdf_file(str(run / f"{name}.nc"), "r", mmap=False) as nc:
            d = at[i].astype(np.float32) - before[i].astype(np.float32)
            assert np.array_equal(nc.variables[name][k].T, d) and np.all(nc.variables[name][0] == 0)
    R = swm.make_setup(swm.Parameter(initial_cond="ncfile", initpath=str(tmp_path), init_run_id=0,
                                     sst_initial="restart", **kw))
    p = swm.initial_conditions("Float32", R)
    for a, b in zip(swm.remove_halo(p.u, p.v, p.η, p.sst, R), at):
        assert np.array_equal(a, b)
