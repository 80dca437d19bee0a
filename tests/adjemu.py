"""ctypes front end of the host build of the adjoint kernels (``tests/emu/adj_emu.cpp``).

TEST INFRASTRUCTURE ONLY: the adjoint marches of ``csrc/adj_core.h`` run on an emulated
wavefront on the CPU, so that tests can check them against the oracle's tangent-linear
model without a GPU, and the GPU build against them.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "emu")
LIB_PATH = os.path.join(EMU_DIR, "libadj_emu.so")
_lib = None


class EmuCfg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("nx", "ny", "per", "RKo", "cont", "drag", "smag", "consume", "tracer", "nadvstep",
                 "nadvstep_half", "nstep_diff", "L", "nan_edges", "fmt", "sad", "lin", "ssp", "s4", "s3", "kn", "mn")] + \
               [("a", ctypes.c_double * 4), ("b", ctypes.c_double * 4)] + \
               [(n, ctypes.c_double) for n in ("cs", "sa", "sb", "cs2")] + \
               [(n, ctypes.c_double) for n in
                ("g", "gamma", "k12", "oma", "cD", "rD", "cS", "nuB", "dtd", "hdu", "hdv", "du", "dv", "jSST")]


def build(force: bool = False) -> str:
    """Build the emulation library (make rebuilds it when a source is newer)."""
    subprocess.run(["make", "-C", EMU_DIR, "-B" if force else "all"], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.emu_run.argtypes = [vp, ctypes.c_int32, ctypes.c_int64, vp, vp]
        L.emu_run.restype = ctypes.c_int
        L.emu_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def config(S, cfg, L=8, nan_edges=False, fmt=8):
    """The adjoint's scalars from a setup and its swm_config, as the GPU driver derives them."""
    P = S.parameters
    from shallowwaters_amd.numtypes import round_scalar
    k = EmuCfg()
    k.nx, k.ny = cfg.nx, cfg.ny
    k.per = 1 if cfg.bc == 0 else 0
    k.RKo = cfg.RKo
    k.cont, k.drag, k.smag = cfg.continuity, cfg.bottom_drag, cfg.diffusion
    k.consume, k.tracer = cfg.tracer_consumption, cfg.tracer_advection
    k.nadvstep, k.nadvstep_half, k.nstep_diff = cfg.nadvstep, cfg.nadvstep_half, cfg.nstep_diff
    k.L, k.nan_edges, k.fmt = L, int(nan_edges), fmt
    for j in range(4):
        k.a[j] = cfg.RKaDt[j]
        k.b[j] = cfg.RKbDt[j] if j < 3 else 0.0
    k.g, k.gamma, k.oma = cfg.g, cfg.gamma, cfg.one_minus_alpha
    k.sad = int(cfg.adv_scheme == 1)                       # Sadourny: PV weight ½
    k.lin = int(cfg.dynamics == 1)                         # the linear dynamics
    if cfg.time_scheme == 1:                               # SSPRK2: stages, Δt_Δs, weights of u0
        k.ssp = cfg.RKs
        k.cs = round_scalar(cfg.Dt_Ds, P.Tprog)
        k.sa = round_scalar(1.0 / cfg.RKs, P.Tprog)
        k.sb = round_scalar((cfg.RKs - 1) / cfg.RKs, P.Tprog)
    elif cfg.time_scheme == 2:                             # SSPRK3: stages from 1, kn-th combines with the mn-th's state
        k.s3, k.kn, k.mn = cfg.ssp_s, cfg.ssp_kn - 1, cfg.ssp_mn - 1
        k.cs, k.cs2 = round_scalar(cfg.ssp_Dt_Dn, P.Tprog), round_scalar(cfg.ssp_Dt_Dnc, P.Tprog)
        k.sa, k.sb = round_scalar(cfg.ssp_kna, P.Tprog), round_scalar(cfg.ssp_knb, P.Tprog)
    elif cfg.time_scheme == 3:                             # 4SSPRK3: ½·Δt_Δ on the tendency, thirds of stage 3
        k.s4 = 1
        k.cs = 0.5 * round_scalar(cfg.Dt_D, P.Tprog)
        k.sa = round_scalar(2.0 / 3.0, P.Tprog)
        k.sb = round_scalar(1.0 / 3.0, P.Tprog)
    k.k12 = round_scalar(0.5 if k.sad else 1.0 / 12.0, P.T)
    k.cD, k.rD, k.cS, k.nuB, k.dtd = cfg.cD, cfg.rD, cfg.cSmag, cfg.nuB, cfg.Dt_diff
    k.hdu, k.hdv, k.du, k.dv, k.jSST = cfg.half_dtadvu, cfg.half_dtadvv, cfg.dtadvu, cfg.dtadvv, cfg.jSST
    return k


def _shapes(S):
    G = S.grid
    return {"u": (G.nux + 4, G.nuy + 4), "v": (G.nvx + 4, G.nvy + 4), "e": (G.nx + 2, G.ny + 2),
            "sst": (G.nx + 2, G.ny)}


_OUT = {1: "uve", 2: "uve", 3: ["u", "v", "u", "v", "sst"], 4: ["u", "v", "e", "sst", "u", "v"],
        5: ["u", "v", "e", "sst"]}


def run(S, k, op, ins, i=1):
    """Run emulated adjoint op ``op`` (see adj_emu.cpp) on float64 Julia-layout arrays."""
    sh = _shapes(S)
    names = _OUT[op]
    outs = [np.zeros(sh[n], order="F") for n in names]
    ins = [np.asfortranarray(a, dtype=np.float64) for a in ins]
    vp = ctypes.c_void_p
    pin = (vp * len(ins))(*[a.ctypes.data for a in ins])
    pout = (vp * len(outs))(*[a.ctypes.data for a in outs])
    rc = lib().emu_run(ctypes.byref(k), op, i, pin, pout)
    if rc != 0:
        raise RuntimeError(f"emulated adjoint op {op} failed ({rc}): {lib().emu_last_error().decode()}")
    return outs
