"""ctypes front end of the CPU oracle (``oracle/swm_oracle.cpp``).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the cpu_baseline
leg of bench.py, as the checker.  The product path never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the two builds of the same source (oracle/Makefile): "o2" checks the tests, "o3" (vectorised)
# is the CPU baseline; both evaluate every operation in the same order with no FMA contraction
LIB_PATHS = {"o2": os.path.join(HERE, "libswm_oracle.so"), "o3": os.path.join(HERE, "libswm_oracle_o3.so")}
LIB_PATH = LIB_PATHS["o2"]
_libs = {}
_variant = "o2"


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (strict IEEE flags), both builds."""
    if force or not all(os.path.exists(p) for p in LIB_PATHS.values()):
        subprocess.run(["make", "-C", HERE, "-j2", "-B" if force else "all"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return LIB_PATH


def use(variant: str) -> None:
    """Select the build later calls go to ("o2", the default, or "o3")."""
    global _variant
    if variant not in LIB_PATHS:
        raise ValueError(variant)
    _variant = variant


def lib():
    if _variant not in _libs:
        path = LIB_PATHS[_variant]
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library {path} is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        L.swmo_run.argtypes = [vp] * 13 + [i64, i64, i64, i32, ctypes.POINTER(vp),
                                           ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.swmo_run.restype = ctypes.c_int
        L.swmo_stencil.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, vp]
        L.swmo_round.argtypes = [ctypes.c_int, vp, vp, i64]
        L.swmo_arith.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, i64]
        _libs[_variant] = L
    return _libs[_variant]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _fortran(a):
    assert a.flags.f_contiguous, "arrays must be in Julia (Fortran) layout"
    return a


class OracleError(RuntimeError):
    pass


def run(S, cfg, prog, nsteps, i0=1, t0=0, stop_stage=0, want_rk=False):
    """Run ``nsteps`` steps of the reference loop on copies of ``prog`` (u, v, η, sst in
    Tprog host storage).  Returns ``(u, v, η, sst, rk, steps_done, nan_step)`` where ``rk``
    is ``None`` or the tuple u0 v0 η0 u1 v1 η1 du dv dη of the last step."""
    F, G = S.forcing, S.grid
    u, v, e, s = (np.array(x, order="F", copy=True) for x in (prog.u, prog.v, prog.η, prog.sst))
    rk = None
    rk_ptrs = None
    if want_rk:
        rk = tuple(np.zeros_like(x, order="F") for x in (u, v, e, u, v, e, u, v, e))
        rk_ptrs = (ctypes.c_void_p * 9)(*[x.ctypes.data for x in rk])
    sd, ns = ctypes.c_int64(0), ctypes.c_int64(0)
    statics = [_fortran(F.H), _fortran(F.Fx), _fortran(F.Fy), _fortran(G.f_q),
               _fortran(F.η_ref), _fortran(F.Fη), _fortran(G.f_u), _fortran(G.f_v)]
    rc = lib().swmo_run(ctypes.byref(cfg), *[_ptr(a) for a in statics], _ptr(u), _ptr(v), _ptr(e), _ptr(s),
                        nsteps, i0, t0, stop_stage, rk_ptrs, ctypes.byref(sd), ctypes.byref(ns))
    if rc != 0:
        raise OracleError("oracle run failed (shape/option error or an index the reference throws on)")
    return u, v, e, s, rk, sd.value, ns.value


def tracer(cfg, i, u, v, um, vm, sst):
    """``tracer!`` alone at step ``i`` (Float64): returns updated copies of um, vm, sst."""
    um, vm, sst = (np.array(x, dtype=np.float64, order="F", copy=True) for x in (um, vm, sst))
    u = np.asfortranarray(u, dtype=np.float64)
    v = np.asfortranarray(v, dtype=np.float64)
    L = lib()
    L.swmo_tracer.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 5
    if L.swmo_tracer(ctypes.byref(cfg), i, _ptr(u), _ptr(v), _ptr(um), _ptr(vm), _ptr(sst)) != 0:
        raise OracleError("tracer failed (NaN or departure outside the arrays)")
    return um, vm, sst


# outputs of the dissipation operator (swmo_op op 1), in order
DISSIPATION_AUX = ("Bu", "Bv", "LLu1", "LLu2", "LLv1", "LLv2", "DT", "DS", "nuSmag")


def _aux_shapes(G):
    nx, ny = G.nx, G.ny
    return {"Bu": (nx + 1, ny + 2), "Bv": (nx + 2, ny + 1), "LLu1": (G.nux, G.nuy + 2), "LLu2": (nx + 1, ny),
            "LLv1": (nx, ny + 1), "LLv2": (G.nvx + 2, G.nvy), "DT": (nx + 2, ny + 2), "DS": (nx + 2, ny + 2),
            "nuSmag": (nx + 2, ny + 2)}


def op(S, cfg, which, u, v, eta, sst, i=0, um=None, vm=None):
    """One operator of the step on copies of a state (Tprog storage, Julia layout):
    ``"ghost"`` ghost_points! of u, v, η and sst; ``"dissipation"`` drag, diffusion,
    add_drag_diff_tendencies! and ghost_points_uv! of (u, v, η) as (u0, v0, η0);
    ``"tracer"`` tracer! at step ``i`` with (u, v) as the post-dissipation velocities and the
    mid-point snapshot (um, vm).  Returns ``(u, v, η, sst, aux)``: aux is a dict of the
    dissipation's intermediates (float64) or, for the tracer, the updated um, vm."""
    F, G = S.forcing, S.grid
    code = {"ghost": 0, "dissipation": 1, "tracer": 2}[which]
    u, v, e, s = (np.array(x, order="F", copy=True) for x in (u, v, eta, sst))
    vp = ctypes.c_void_p
    aux, names = None, ()
    if code == 1:
        names = DISSIPATION_AUX
        shp = _aux_shapes(G)
        aux = {k: np.zeros(shp[k], order="F") for k in names}
    elif code == 2:
        names = ("um", "vm")
        aux = {"um": np.array(um, dtype=np.float64, order="F", copy=True),
               "vm": np.array(vm, dtype=np.float64, order="F", copy=True)}
    ptrs = (vp * 9)(*[aux[k].ctypes.data for k in names]) if aux else None
    L = lib()
    L.swmo_op.argtypes = [vp] * 5 + [ctypes.c_int32] + [vp] * 4 + [ctypes.c_int64, vp]
    statics = [_fortran(F.H), _fortran(F.Fx), _fortran(F.Fy), _fortran(G.f_q)]
    rc = L.swmo_op(ctypes.byref(cfg), *[_ptr(a) for a in statics], code, _ptr(u), _ptr(v), _ptr(e), _ptr(s), i,
                   ctypes.cast(ptrs, vp) if ptrs is not None else None)
    if rc != 0:
        raise OracleError(f"oracle operator {which} failed")
    return u, v, e, s, aux


class mutation:
    """Context manager planting mistake ``k`` in the oracle (the MUT_* list of
    swm_oracle.cpp) for the pins' own tests: inside it the restatement is deliberately
    wrong, and the pin that targets the mistake must fail."""

    def __init__(self, k):
        self.k = k

    def __enter__(self):
        L = lib()
        L.swmo_set_mutation.argtypes = [ctypes.c_int32]
        L.swmo_set_mutation(self.k)
        return self

    def __exit__(self, *exc):
        lib().swmo_set_mutation(0)
        return False


def pv_advection(cfg, q, U, V):
    """The PV advection terms alone (Float64) of the configured scheme: from q at the corners
    (nx+1, ny+1) and the mass fluxes U (nx+1, ny+2), V (nx+2, ny+1) -> (qhv, qhu) of shapes
    (nux, nuy), (nvx, nvy).  q's scheme-specific combinations are formed inside."""
    q, U, V = (np.asfortranarray(a, dtype=np.float64) for a in (q, U, V))
    nx, ny = q.shape[0] - 1, q.shape[1] - 1
    nux = nx if cfg.bc == 0 else nx - 1                     # SWM_PERIODIC = 0
    qhv = np.zeros((nux, ny), order="F")
    qhu = np.zeros((nx, ny - 1), order="F")
    L = lib()
    L.swmo_pv_advection.argtypes = [ctypes.c_void_p] * 6
    if L.swmo_pv_advection(ctypes.byref(cfg), *[_ptr(a) for a in (q, U, V, qhv, qhu)]) != 0:
        raise OracleError("shape error")
    return qhv, qhu


_STENCIL = {"dx": (0, -1, 0), "dy": (1, 0, -1), "lap": (2, -2, -2), "Ix": (3, -1, 0), "Iy": (4, 0, -1),
            "Ixy": (5, -1, -1)}


def stencil(name, a):
    """Apply one reference stencil (∂x, ∂y, ∇², Ix, Iy, Ixy) to a float64 array."""
    op, dm, dn = _STENCIL[name]
    a = np.asfortranarray(a, dtype=np.float64)
    out = np.zeros((a.shape[0] + dm, a.shape[1] + dn), order="F")
    if lib().swmo_stencil(op, _ptr(a), a.shape[0], a.shape[1], _ptr(out)) != 0:
        raise OracleError("shape error")
    return out


def round_fmt(fmt_code, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().swmo_round(fmt_code, _ptr(x), _ptr(out), x.size)
    return out


def arith(fmt_code, op, a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty_like(a)
    lib().swmo_arith(fmt_code, op, _ptr(a), _ptr(b), _ptr(out), a.size)
    return out


# ------------------------------------------------------------------ tangent-linear model

def _statics64(S):
    F, G = S.forcing, S.grid
    return [np.asfortranarray(a, dtype=np.float64) if a is not None else None
            for a in (F.H, F.Fx, F.Fy, G.f_q, F.η_ref, F.Fη, G.f_u, G.f_v)]


def _out_shapes(S, op):
    G = S.grid
    su, sv, se, ss = (G.nux + 4, G.nuy + 4), (G.nvx + 4, G.nvy + 4), (G.nx + 2, G.ny + 2), (G.nx + 2, G.ny)
    return {0: [su, sv, se, ss], 1: [su, sv, se], 2: [su, sv], 3: [su, sv, ss]}[op]


def tlm(S, cfg, op, x, dx=None, nsteps=1, i0=1, t0=0, dFx=None):
    """Tangent-linear model of the oracle (Float64 values with one directional derivative).

    ``op`` 0: ``nsteps`` steps of the loop from x = (u, v, η, sst) after the upload's ghost
    copies; 1: rhs + continuity of a classic RK stage at x = (u, v, η) -> (du, dv, dη);
    2: drag + diffusion + ghost_points_uv! of x = (u0, v0, η0) -> (u0, v0);
    3: tracer! at step i0 of x = (u, v, um, vm, sst) -> (um, vm, sst).
    ``dFx``: a tangent of the wind stress Fx (shape (nux, nuy)), for any op.
    Returns (y, dy) as lists of float64 Julia-layout arrays."""
    x = [np.asfortranarray(a, dtype=np.float64) for a in x]
    dx = None if dx is None else [np.asfortranarray(a, dtype=np.float64) for a in dx]
    shapes = _out_shapes(S, op)
    y = [np.zeros(s, order="F") for s in shapes]
    dy = [np.zeros(s, order="F") for s in shapes]
    vp = ctypes.c_void_p
    L = lib()
    L.swmo_tlm.argtypes = [vp] * 9 + [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64] + [vp] * 5
    arr = lambda lst: (vp * len(lst))(*[a.ctypes.data if a is not None else None for a in lst])
    xs, ys, dys = arr(x), arr(y), arr(dy)
    dxs = arr(dx) if dx is not None else None
    st = _statics64(S)
    dfx = np.asfortranarray(dFx, dtype=np.float64) if dFx is not None else None
    rc = L.swmo_tlm(ctypes.byref(cfg), *[_ptr(a) for a in st], op, nsteps, i0, t0, xs, dxs, ys, dys, _ptr(dfx))
    if rc != 0:
        raise OracleError("tangent-linear run failed")
    return y, dy


def tape(S, cfg, i, t, u, v, e, sst, um, vm):
    """Intermediate states of one Float64 step (any scheme) from the given state: returns a
    dict with 's' (list of the RK stage inputs (u1, v1, η1); under SSPRK2 / SSPRK3 the stage
    inputs and (u1, v1, η1) after the last stage), 'p' (4SSPRK3: list of every
    stage's (u1, v1) after its momentum update and ghost copies; else empty), 'pre' (u0, v0,
    η0 after the stages' ghost copies), 'post' (u0, v0 after the dissipation), 'mid' (um, vm
    after the step) and 'new' (u, v, η, sst)."""
    G = S.grid
    su, sv, se, ss = (G.nux + 4, G.nuy + 4), (G.nvx + 4, G.nvy + 4), (G.nx + 2, G.ny + 2), (G.nx + 2, G.ny)
    P = S.parameters
    RKo = {"SSPRK2": P.RKs + 1, "SSPRK3": P.RKn ** 2 + 1, "4SSPRK3": 4}.get(P.time_scheme, P.RKo)
    npair = 4 if P.time_scheme == "4SSPRK3" else 0
    shapes = [su, sv, se] * RKo + [su, sv] * npair + [su, sv, se] + [su, sv] + [su, sv] + [su, sv, se, ss]
    out = [np.zeros(s, order="F") for s in shapes]
    vp = ctypes.c_void_p
    L = lib()
    L.swmo_tape.argtypes = [vp] * 9 + [ctypes.c_int64, ctypes.c_int64] + [vp] * 7
    ins = [np.asfortranarray(a, dtype=np.float64) for a in (u, v, e, sst, um, vm)]
    ptrs = (vp * len(out))(*[a.ctypes.data for a in out])
    st = _statics64(S)
    if L.swmo_tape(ctypes.byref(cfg), *[_ptr(a) for a in st], i, t, *[_ptr(a) for a in ins], ptrs) != 0:
        raise OracleError("tape failed")
    p = [tuple(out[3 * RKo + 2 * j:3 * RKo + 2 * j + 2]) for j in range(npair)]
    k = 3 * RKo + 2 * npair
    return {"s": [tuple(out[3 * j:3 * j + 3]) for j in range(RKo)], "p": p, "pre": tuple(out[k:k + 3]),
            "post": tuple(out[k + 3:k + 5]), "mid": tuple(out[k + 5:k + 7]), "new": tuple(out[k + 7:k + 11])}
