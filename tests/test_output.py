"""Run folders, NetCDF output and restarts (output.jl, feedback.jl, the ncfile branch of
initial_conditions.jl), on the host: the files' layout and attributes, round trips through
the restart reader, the run-id modes and the restart interpolation on fields it reproduces
exactly (bilinear on linear fields, flat beyond the edges)."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file


def setup(swm, tmp_path, **kw):
    d = dict(T="Float32", nx=40, L_ratio=2, bc="periodic", output=True, outpath=str(tmp_path), Ndays=0.2)
    d.update(kw)
    return swm.make_setup(swm.Parameter(**d))


def out(swm):
    import importlib
    return importlib.import_module(swm.__name__ + ".output")


def test_readable_secs_known_answers(swm):
    o = out(swm)
    assert o.readable_secs(3.14159) == "3.14s"
    assert o.readable_secs(42.25) == "42.2s"
    assert o.readable_secs(125) == "2min, 5s"
    assert o.readable_secs(3 * 3600 + 7 * 60) == "3h, 7min"
    assert o.readable_secs(2 * 86400 + 5 * 3600) == "2d, 5h"


def test_run_id_modes(swm, tmp_path):
    o = out(swm)
    S = setup(swm, tmp_path)
    assert o.get_run_id_path(S) == (0, os.path.join(str(tmp_path), "run0000"))
    for k in (1, 3):
        os.mkdir(tmp_path / f"run{k:04d}")
    assert o.get_run_id_path(S)[0] == 4                                   # continue: largest + 1
    assert o.get_run_id_path(setup(swm, tmp_path, get_id_mode="fill"))[0] == 2
    assert o.gap([0, 1, 2]) == 3 and o.gap([2, 5]) == 3
    (tmp_path / "run0007").mkdir()
    (tmp_path / "run0007" / "old.txt").write_text("x")
    rid, path = o.get_run_id_path(setup(swm, tmp_path, get_id_mode="specific", run_id=7))
    assert rid == 7 and os.listdir(path) == []                             # specific: recreated empty


def fields(S, k):
    G = S.grid
    rng = np.random.default_rng(k)
    return (rng.standard_normal((G.nux, G.nuy)), rng.standard_normal((G.nvx, G.nvy)),
            rng.standard_normal((G.nx, G.ny)), rng.standard_normal((G.nx, G.ny)))


def test_files_layout_and_restart_round_trip(swm, tmp_path):
    o = out(swm)
    S = setup(swm, tmp_path, output_vars=["u", "v", "η", "sst", "du", "dv", "dη"])
    G = S.grid
    fb = o.feedback_init(S)
    assert fb.run_id == 0 and os.path.isfile(os.path.join(fb.runpath, "progress.txt"))
    assert "nx=40" in open(os.path.join(fb.runpath, "parameter.txt")).read()
    ncs = o.NcFiles(fb, S)
    states = [fields(S, k) for k in range(3)]
    for k, st in enumerate(states):
        ncs.output(k * G.nout, st, states[k - 1][:3] if k else st[:3], S)
    ncs.close()
    o.feedback_end(fb)
    for name, var, x, y in (("u", "u", G.x_u, G.y_u), ("v", "v", G.x_v, G.y_v), ("eta", "eta", G.x_T, G.y_T),
                            ("sst", "sst", G.x_T, G.y_T), ("du", "du", G.x_u, G.y_u), ("deta", "deta", G.x_T, G.y_T)):
        with netcdf_file(os.path.join(fb.runpath, name + ".nc"), "r", mmap=False) as nc:
            v = nc.variables[var]
            assert v.dimensions == ("t", "y", "x") and v.typecode() == "f"
            assert v.shape == (3, len(y), len(x))
            assert np.array_equal(nc.variables["x"][:], x) and np.array_equal(nc.variables["y"][:], y)
            assert list(nc.variables["t"][:]) == [0, G.nout * G.dtint, 2 * G.nout * G.dtint]
            assert nc.variables["t"].typecode() == "i"
            assert v.missing_value == np.float32(-999999.0) and nc.variables["x"].units == b"m"
            if name == "u":
                assert np.array_equal(v[2].T, states[2][0].astype(np.float32))
            if name == "du":
                assert np.all(v[0] == 0)
                assert np.array_equal(v[1].T, states[1][0].astype(np.float32) - states[0][0].astype(np.float32))
    # restart from the second record and from the last one (init_starti = -1)
    for starti, k in ((2, 1), (-1, 2)):
        R = setup(swm, tmp_path, output=False, initial_cond="ncfile", init_run_id=0, init_starti=starti,
                  sst_initial="restart")
        P = swm.initial_conditions("Float32", R)
        u, v, e, s = swm.remove_halo(P.u, P.v, P.η, P.sst, R)
        for a, b in zip((u, v, e, s), states[k]):
            assert np.array_equal(a, b.astype(np.float32))


def test_restart_interpolation_reproduces_linear_fields(swm, tmp_path):
    o = out(swm)
    S = setup(swm, tmp_path, nx=40)                     # periodic: the reader's u points
    G = S.grid
    fb = o.feedback_init(S)
    ncs = o.NcFiles(fb, S)
    lin = lambda X, Y: 0.25 * X - 0.5 * Y + 1.0
    xu, yu = np.meshgrid(np.arange(G.nux, dtype=float), np.arange(G.nuy) + 0.5, indexing="ij")
    xv, yv = np.meshgrid(np.arange(G.nvx) + 0.5, np.arange(1, G.nvy + 1, dtype=float), indexing="ij")
    xT, yT = np.meshgrid(np.arange(G.nx) + 0.5, np.arange(G.ny) + 0.5, indexing="ij")
    ncs.output(0, (lin(xu, yu), lin(xv, yv), lin(xT, yT), lin(xT, yT)), None, S)
    ncs.close()
    o.feedback_end(fb)
    R = setup(swm, tmp_path, nx=20, output=False, initial_cond="ncfile", init_run_id=0)
    Gn = R.grid
    u, v, e, _ = o.read_initial_conditions(R)
    d = G.nx / Gn.nx
    X, Y = np.meshgrid(d * np.arange(Gn.nux), d / 2 + d * np.arange(Gn.ny), indexing="ij")
    assert np.allclose(u, lin(np.minimum(X, G.nux - 1), Y), rtol=0, atol=1e-5)
    X, Y = np.meshgrid(d / 2 + d * np.arange(Gn.nx), d * np.arange(1, Gn.ny), indexing="ij")
    assert np.allclose(v, lin(np.clip(X, 0.5, G.nvx - 0.5), Y), rtol=0, atol=1e-5)
    X, Y = np.meshgrid(d / 2 + d * np.arange(Gn.nx), d / 2 + d * np.arange(Gn.ny), indexing="ij")
    assert np.allclose(e, lin(X, Y), rtol=0, atol=1e-5)
    # without interpolation a size mismatch is an error, as in the reference
    R2 = setup(swm, tmp_path, nx=20, output=False, initial_cond="ncfile", init_run_id=0, init_interpolation=False)
    with pytest.raises(ValueError, match="doesn't match"):
        o.read_initial_conditions(R2)


def test_unavailable_outputs_and_formats(swm, tmp_path):
    o = out(swm)
    S = setup(swm, tmp_path, output_vars=["u", "q"])
    fb = o.feedback_init(S)
    with pytest.raises(NotImplementedError, match="potential vorticity"):
        o.NcFiles(fb, S)
    o.feedback_end(fb)
    bad = tmp_path / "run0009"
    bad.mkdir()
    (bad / "u.nc").write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    R = setup(swm, tmp_path, output=False, initial_cond="ncfile", init_run_id=9)
    with pytest.raises(ValueError, match="NETCDF4"):
        o.read_initial_conditions(R)
