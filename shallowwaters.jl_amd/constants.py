"""``Constants{T,Tprog}`` and ``SSPRK3coeff``: the scalars the time step consumes.

Mirrors ``src/constants.jl:1-115``.  Each value is computed in float64 in the reference's
order and rounded once to the type the reference stores it in; the rounded value is kept
as a python float (exactly representable in that type).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from .numtypes import round_scalar


@dataclass
class SSPRK3coeff:
    """``src/constants.jl:4-27``"""
    n: int
    s: int
    kn: int
    mn: int
    Δt_Δn: float
    kna: float
    knb: float
    Δt_Δnc: float


def make_ssprk3coeff(P, Δt_Δ: float, Tprog: str) -> SSPRK3coeff:
    n = P.RKn
    s = n ** 2
    kn = n * (n + 1) // 2 + 1
    mn = (n - 1) * (n - 2) // 2 + 1
    return SSPRK3coeff(
        n=n, s=s, kn=kn, mn=mn,
        Δt_Δn=round_scalar(Δt_Δ / (n ** 2 - n), Tprog),
        kna=round_scalar((n - 1) / (2 * n - 1), Tprog),
        knb=round_scalar(n / (2 * n - 1), Tprog),
        Δt_Δnc=round_scalar(Δt_Δ / (n * (2 * n - 1)), Tprog),
    )


@dataclass
class Constants:
    RKaΔt: list
    RKbΔt: list
    Δt_Δs: float
    Δt_Δ: float
    Δt_Δ_half: float
    SSPRK3c: SSPRK3coeff
    one_minus_α: float
    g: float
    cD: float
    rD: float
    γ: float
    cSmag: float
    νB: float
    rSST: float
    jSST: float
    SSTmin: float
    ωFη: float
    ωFx: float
    ωFy: float


def make_constants(P, G) -> Constants:
    """``Constants{T,Tprog}(P,G)``, ``src/constants.jl:58-115``."""
    T, Tprog = P.T, P.Tprog
    dtint, Δ = G.dtint, G.Δ
    if P.RKo == 2:
        a, b = [1 / 2, 1 / 2], [1.0]
    elif P.RKo == 3:
        a, b = [1 / 4, 0., 3 / 4], [1 / 3, 2 / 3]
    else:
        a, b = [1 / 6, 1 / 3, 1 / 3, 1 / 6], [.5, .5, 1.]
    # Tprog.([...]*G.dtint/G.Δ): the vector times an Int, then divided by Δ
    RKaΔt = [round_scalar(x * dtint / Δ, Tprog) for x in a]
    RKbΔt = [round_scalar(x * dtint / Δ, Tprog) for x in b]

    Δt_Δs = round_scalar(dtint / Δ / (P.RKs - 1), Tprog)
    Δt_Δ = round_scalar(dtint / Δ, Tprog)
    Δt_Δ_half = round_scalar(dtint / Δ / 2, Tprog)
    # SSPRK3coeff receives the already rounded Δt_Δ (constants.jl:81)
    SSPRK3c = make_ssprk3coeff(P, Δt_Δ, Tprog)

    return Constants(
        RKaΔt=RKaΔt, RKbΔt=RKbΔt, Δt_Δs=Δt_Δs, Δt_Δ=Δt_Δ, Δt_Δ_half=Δt_Δ_half, SSPRK3c=SSPRK3c,
        one_minus_α=round_scalar(1 - P.α, Tprog),
        g=round_scalar(P.g, T),
        cD=round_scalar(-Δ * P.cD, T),
        rD=round_scalar(-Δ / (P.τD * 24 * 3600), T),
        γ=round_scalar(Δ / (P.t_relax * 3600 * 24), T),
        cSmag=round_scalar(-P.cSmag, T),
        νB=round_scalar(-P.νB / 30000, T),
        rSST=round_scalar(G.dtadvint / (P.τSST * 3600 * 24), T),
        jSST=round_scalar(G.dtadvint / (P.jSST * 3600 * 24), T),
        SSTmin=round_scalar(P.SSTmin, T),
        ωFη=-2 * math.pi * P.ωFη / 24 / 365.25 / 3600,
        ωFx=2 * math.pi * P.ωFx / 24 / 365.25 / 3600,
        ωFy=2 * math.pi * P.ωFy / 24 / 365.25 / 3600,
    )
