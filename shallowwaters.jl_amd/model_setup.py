"""``ModelSetup`` and its resolution into the C-ABI ``swm_config``.

``ModelSetup{T,Tprog}(parameters, grid, constants, forcing)`` mirrors
``src/model_setup.jl:1-6``.  ``make_config`` turns the string options of ``Parameter`` into
the enums of ``include/swm.h`` and copies the pre-rounded constants; the GPU kernels are
specialised on those enums instead of branching on strings per cell.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

from .constants import Constants, make_constants
from .forcing import Forcing, make_forcing
from .grid import Grid, make_grid
from .numtypes import TYPE_CODES
from .parameter import Parameter

SWM_ABI_VERSION = 2
SWM_PERIODIC, SWM_NONPERIODIC = 0, 1
SCHEMES = {"RK": 0, "SSPRK2": 1, "SSPRK3": 2, "4SSPRK3": 3}
DRAGS = {"quadratic": 0, "linear": 1, "none": 2}
DIFFUSIONS = {"constant": 0, "Smagorinsky": 1}
ADV = {"ArakawaHsu": 0, "Sadourny": 1}
DYN = {"nonlinear": 0, "linear": 1}


class swm_config(ctypes.Structure):
    """Field-for-field copy of ``swm_config`` in include/swm.h."""
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("bc", ctypes.c_int32),
        ("T", ctypes.c_int32), ("Tprog", ctypes.c_int32), ("Tcomm", ctypes.c_int32),
        ("time_scheme", ctypes.c_int32), ("RKo", ctypes.c_int32), ("RKs", ctypes.c_int32), ("RKn", ctypes.c_int32),
        ("nstep_diff", ctypes.c_int32), ("nstep_advcor", ctypes.c_int32),
        ("adv_scheme", ctypes.c_int32), ("dynamics", ctypes.c_int32), ("bottom_drag", ctypes.c_int32),
        ("diffusion", ctypes.c_int32), ("continuity", ctypes.c_int32),
        ("tracer_advection", ctypes.c_int32), ("tracer_consumption", ctypes.c_int32),
        ("seasonal_wind_x", ctypes.c_int32), ("seasonal_wind_y", ctypes.c_int32),
        ("scale_log2", ctypes.c_int32),
        ("dtint", ctypes.c_int64), ("nadvstep", ctypes.c_int64), ("nadvstep_half", ctypes.c_int64),
        ("nout", ctypes.c_int64),
        ("RKaDt", ctypes.c_double * 4), ("RKbDt", ctypes.c_double * 3),
        ("Dt_Ds", ctypes.c_double), ("Dt_D", ctypes.c_double), ("Dt_D_half", ctypes.c_double),
        ("ssp_n", ctypes.c_int32), ("ssp_s", ctypes.c_int32), ("ssp_kn", ctypes.c_int32), ("ssp_mn", ctypes.c_int32),
        ("ssp_Dt_Dn", ctypes.c_double), ("ssp_kna", ctypes.c_double), ("ssp_knb", ctypes.c_double),
        ("ssp_Dt_Dnc", ctypes.c_double),
        ("one_minus_alpha", ctypes.c_double), ("Dt_diff", ctypes.c_double),
        ("g", ctypes.c_double), ("cD", ctypes.c_double), ("rD", ctypes.c_double), ("gamma", ctypes.c_double),
        ("cSmag", ctypes.c_double), ("nuB", ctypes.c_double),
        ("dtadvu", ctypes.c_double), ("dtadvv", ctypes.c_double),
        ("half_dtadvu", ctypes.c_double), ("half_dtadvv", ctypes.c_double),
        ("jSST", ctypes.c_double), ("SSTmin", ctypes.c_double),
        ("omegaFx", ctypes.c_double), ("omegaFy", ctypes.c_double), ("omegaFeta", ctypes.c_double),
        ("compensated", ctypes.c_int32), ("t_arith32", ctypes.c_int32),
    ]


@dataclass
class ModelSetup:
    parameters: Parameter
    grid: Grid
    constants: Constants
    forcing: Forcing


def make_setup(P: Parameter) -> ModelSetup:
    """The setup half of ``run_model(T,P)``, ``src/run_model.jl:25-33``."""
    G = make_grid(P)
    C = make_constants(P, G)
    F = make_forcing(P, G)
    return ModelSetup(parameters=P, grid=G, constants=C, forcing=F)


def make_config(S: ModelSetup) -> swm_config:
    P, G, C = S.parameters, S.grid, S.constants
    c = swm_config()
    c.abi_version = SWM_ABI_VERSION
    c.nx, c.ny = G.nx, G.ny
    c.bc = SWM_PERIODIC if P.bc == "periodic" else SWM_NONPERIODIC
    c.T, c.Tprog, c.Tcomm = TYPE_CODES[P.T], TYPE_CODES[P.Tprog], TYPE_CODES[P.Tcomm]
    c.time_scheme = SCHEMES[P.time_scheme]
    c.RKo, c.RKs, c.RKn = P.RKo, P.RKs, P.RKn
    c.nstep_diff, c.nstep_advcor = P.nstep_diff, P.nstep_advcor
    c.adv_scheme, c.dynamics = ADV[P.adv_scheme], DYN[P.dynamics]
    c.bottom_drag, c.diffusion = DRAGS[P.bottom_drag], DIFFUSIONS[P.diffusion]
    c.continuity = 1 if P.surface_relax else (2 if P.surface_forcing else 0)
    c.tracer_advection, c.tracer_consumption = int(P.tracer_advection), int(P.tracer_consumption)
    c.seasonal_wind_x, c.seasonal_wind_y = int(P.seasonal_wind_x), int(P.seasonal_wind_y)
    c.dtint, c.nadvstep, c.nadvstep_half, c.nout = G.dtint, G.nadvstep, G.nadvstep_half, G.nout
    for k, x in enumerate(C.RKaΔt):
        c.RKaDt[k] = x
    for k, x in enumerate(C.RKbΔt):
        c.RKbDt[k] = x
    c.Dt_Ds, c.Dt_D, c.Dt_D_half = C.Δt_Δs, C.Δt_Δ, C.Δt_Δ_half
    s = C.SSPRK3c
    c.ssp_n, c.ssp_s, c.ssp_kn, c.ssp_mn = s.n, s.s, s.kn, s.mn
    c.ssp_Dt_Dn, c.ssp_kna, c.ssp_knb, c.ssp_Dt_Dnc = s.Δt_Δn, s.kna, s.knb, s.Δt_Δnc
    c.one_minus_alpha, c.Dt_diff = C.one_minus_α, G.Δt_diff
    c.g, c.cD, c.rD, c.gamma, c.cSmag, c.nuB = C.g, C.cD, C.rD, C.γ, C.cSmag, C.νB
    c.dtadvu, c.dtadvv, c.half_dtadvu, c.half_dtadvv = G.dtadvu, G.dtadvv, G.half_dtadvu, G.half_dtadvv
    c.jSST, c.SSTmin = C.jSST, C.SSTmin
    c.omegaFx, c.omegaFy, c.omegaFeta = C.ωFx, C.ωFy, C.ωFη
    c.scale_log2 = int(round(math.log2(P.scale)))
    c.compensated = int(P.compensated)
    c.t_arith32 = int(P.arithmetic == "Float32")
    return c
