"""The ``Parameter`` keyword struct of ShallowWaters.jl, mirrored for the MI355X path.

Field names, defaults and the validation messages follow
``src/default_parameters.jl:1-148`` (struct ``:1-114``, asserts ``:116-147``).  Two
deliberate differences, both rejections of inputs the reference accepts but then
mishandles:

* ``wind_forcing_y`` in {"channel", "shear", "double_gyre"} builds ``Fy`` on the u-grid
  (``src/forcing.jl:65,80,96``) but ``momentum_v!`` reads it with v-grid indices
  (``src/rhs.jl:271``) -- an out-of-bounds read under ``@inbounds``.
* ``wind_forcing_y="none"`` with ``wind_forcing_x!="none"`` leaves ``Fy`` undefined
  (typo at ``src/forcing.jl:35``).

Both raise ``AssertionError`` here, at construction.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

from .numtypes import canonical_type


@dataclass(kw_only=True)
class Parameter:
    T: object = "Float32"                 # number format of the right-hand side
    Tprog: object = None                  # number format of the prognostic variables (default T)
    Tcomm: object = None                  # number format of ghost-point copies (default Tprog)

    # DOMAIN RESOLUTION AND RATIO
    nx: int = 100
    Lx: float = 2000e3
    L_ratio: float = 2

    # PHYSICAL CONSTANTS
    g: float = 10.
    H: float = 500.
    ρ: float = 1e3
    ϕ: float = 45.
    ω: float = 2 * math.pi / (24 * 3600)
    R: float = 6.371e6

    # WIND FORCING OPTIONS
    wind_forcing_x: str = "channel"
    wind_forcing_y: str = "constant"
    Fx0: float = 0.12
    Fy0: float = 0.0
    seasonal_wind_x: bool = False
    seasonal_wind_y: bool = False
    ωFx: float = 1.0
    ωFy: float = 1.0

    # BOTTOM TOPOGRAPHY OPTIONS
    topography: str = "ridge"
    topo_height: float = 50.
    topo_width: float = 300e3

    # SURFACE RELAXATION
    surface_relax: bool = False
    t_relax: float = 100.
    η_refh: float = 5.
    η_refw: float = 50e3

    # SURFACE FORCING
    surface_forcing: bool = False
    ωFη: float = 1.0
    A: float = 3e-5
    ϕk: float = None                      # default ϕ
    wk: float = 10e3

    # TIME STEPPING OPTIONS
    time_scheme: str = "RK"
    RKo: int = 4
    RKs: int = 3
    RKn: int = 2
    cfl: float = 1.0
    Ndays: float = 10.0
    nstep_diff: int = 1
    nstep_advcor: int = 0

    # BOUNDARY CONDITION OPTIONS
    bc: str = "periodic"
    α: float = 2.

    # MOMENTUM ADVECTION OPTIONS
    adv_scheme: str = "ArakawaHsu"
    dynamics: str = "nonlinear"

    # BOTTOM FRICTION OPTIONS
    bottom_drag: str = "quadratic"
    cD: float = 1e-5
    τD: float = 300.

    # DIFFUSION OPTIONS
    diffusion: str = "constant"
    νB: float = 500.0
    cSmag: float = 0.15

    # LOW-PRECISION OPTIONS (not in this version of the reference; the name `scale` follows
    # later ShallowWaters.jl versions): u, v are carried as scale·u, scale·v inside the time
    # step (a power of two, so Float32 / Float64 results are unchanged); compensated sums
    # the RK increments apart and adds them with Kahan summation (classic RK only)
    scale: float = 1.0
    compensated: bool = False
    # arithmetic="Float32" with T Float16 / BFloat16 and Tprog Float32: T's values (stage
    # inputs, static fields, T constants) stay in T, the right-hand side is computed in Float32
    # and its results are not rounded back to T (None: the arithmetic of T itself)
    arithmetic: object = None

    # TRACER ADVECTION
    tracer_advection: bool = True
    tracer_relaxation: bool = False
    tracer_consumption: bool = False
    tracer_pumping: bool = False
    injection_region: str = "west"
    sst_initial: str = "south"
    sst_rect_coords: list = field(default_factory=lambda: [0., 0.15, 0., 1.0])
    Uadv: float = 0.25
    SSTmax: float = 1.
    SSTmin: float = 0.
    τSST: float = 500.
    jSST: float = 365.
    SST_λ0: float = 222e3
    SST_λs: float = 111e3
    SST_γ0: float = 8.35
    SSTw: float = 5e3
    SSTϕ: float = 0.5

    # OUTPUT OPTIONS
    output: bool = False
    output_vars: list = field(default_factory=lambda: ["u", "v", "η", "sst"])
    output_dt: float = 6
    outpath: str = "."

    # INITIAL CONDITIONS
    initial_cond: str = "rest"
    initpath: str = None                  # default outpath
    init_run_id: int = 0
    init_starti: int = -1
    get_id_mode: str = "continue"
    run_id: int = -1
    init_interpolation: bool = True

    def __post_init__(self):
        self.T = canonical_type(self.T)
        self.Tprog = self.T if self.Tprog is None else canonical_type(self.Tprog)
        self.Tcomm = self.Tprog if self.Tcomm is None else canonical_type(self.Tcomm)
        if self.ϕk is None:
            self.ϕk = self.ϕ
        if self.initpath is None:
            self.initpath = self.outpath
        self._validate()

    def _validate(self):
        # the assertions of src/default_parameters.jl:116-147, same order and messages
        nx, Lx, L_ratio = self.nx, self.Lx, self.L_ratio
        assert all(x > 0. for x in (nx, Lx, L_ratio)), "nx, Lx, L_ratio have to be >0"
        assert all(x > 0. for x in (self.g, self.H, self.ρ, self.ω, self.R)), "g,H,ρ,ω,R have to be >0"
        assert -90.0 <= self.ϕ <= 90.0, f"ϕ has to be in (-90,90), {self.ϕ} given."
        winds = ["channel", "double_gyre", "shear", "constant", "none"]
        assert self.wind_forcing_x in winds, f"Wind forcing '{self.wind_forcing_x}' unsupported"
        assert self.wind_forcing_y in winds, f"Wind forcing '{self.wind_forcing_y}' unsupported"
        assert self.topography in ["ridge", "seamount", "flat", "ridges"], \
            f"Topography '{self.topography}' unsupported"
        assert self.topo_width > 0.0, f"topo_width has to be >0, {self.topo_width} given."
        assert self.t_relax > 0.0, f"t_relax has to be >0, {self.t_relax} given."
        assert self.η_refw > 0.0, f"η_refw has to be >0, {self.η_refw} given."
        assert self.time_scheme in ["RK", "SSPRK2", "SSPRK3", "4SSPRK3"], \
            f"Time scheme {self.time_scheme} unsupported."
        assert self.RKo in [2, 3, 4], f"RKo has to be 2,3 or 4; {self.RKo} given."
        assert self.RKs > 1, f"RKs has to be >= 2; {self.RKs} given."
        assert self.Ndays > 0.0, f"Ndays has to be >0, {self.Ndays} given."
        assert self.nstep_diff > 0, f"nstep_diff has to be >0, {self.nstep_diff} given."
        assert self.nstep_advcor >= 0, f"nstep_advcor has to be >=0, {self.nstep_advcor} given."
        assert self.bc in ["periodic", "nonperiodic"], f"boundary condition '{self.bc}' unsupported."
        assert 0.0 <= self.α <= 2.0, f"Tangential boundary condition α has to be in (0,2), {self.α} given."
        assert self.adv_scheme in ["Sadourny", "ArakawaHsu"], f"Advection scheme '{self.adv_scheme}' unsupported"
        assert self.dynamics in ["linear", "nonlinear"], f"Dynamics '{self.dynamics}' unsupported."
        assert self.bottom_drag in ["quadratic", "linear", "none"], f"Bottom drag '{self.bottom_drag}' unsupported."
        assert self.cD >= 0.0, f"Bottom drag coefficient cD has to be >=0, {self.cD} given."
        assert self.τD >= 0.0, f"Bottom drag coefficient τD has to be >=0, {self.τD} given."
        assert self.diffusion in ["Smagorinsky", "constant"], f"Diffusion '{self.diffusion}' unsupported."
        assert self.νB > 0.0, f"Diffusion scaling constant νB has to be >0, {self.νB} given."
        assert self.cSmag > 0.0, f"Smagorinsky coefficient cSmag has to be >0, {self.cSmag} given."
        assert self.injection_region in ["west", "south"], f"Injection region '{self.injection_region}' unsupported."
        assert self.Uadv > 0.0, f"Advection velocity scale Uadv has to be >0, {self.Uadv} given."
        assert self.output_dt > 0, f"Output time step has to be >0, {self.output_dt} given."
        assert self.initial_cond in ["rest", "ncfile"], f"Initial conditions '{self.initial_cond}' unsupported."
        assert self.init_run_id >= 0, \
            f"Initial condition run id, init_run_id, has to be >= 0, {self.init_run_id} given."
        assert self.init_starti > 0 or self.init_starti == -1, \
            f"Start index, init_starti, has to be >0 || -1, {self.init_starti} given."
        assert self.get_id_mode in ["continue", "fill", "specific"], f"get_id_mode {self.get_id_mode} unsupported."
        k = math.log2(self.scale) if self.scale > 0 else 0.5
        assert self.scale > 0 and k == int(k) and -20 <= k <= 20, \
            f"scale has to be a power of two in 2^-20 .. 2^20, {self.scale} given."
        assert not self.compensated or self.time_scheme == "RK", "compensated requires the classic RK scheme"
        if self.arithmetic is not None:
            self.arithmetic = canonical_type(self.arithmetic)
            assert self.arithmetic == "Float32" and self.T in ("Float16", "BFloat16") and self.Tprog == "Float32", \
                "arithmetic='Float32' applies to T in {Float16, BFloat16} with Tprog = Float32"
            assert self.time_scheme == "RK" and self.nstep_advcor == 0, \
                "arithmetic='Float32' is implemented for the classic RK scheme with nstep_advcor = 0"
        # rejected here although the reference accepts them (see module docstring)
        assert self.wind_forcing_y in ["constant", "none"], \
            f"wind_forcing_y '{self.wind_forcing_y}' builds Fy on the u-grid; only 'constant' and 'none' are supported"
        assert not (self.wind_forcing_y == "none" and self.wind_forcing_x != "none"), \
            "wind_forcing_y='none' requires wind_forcing_x='none' (Fy is otherwise left undefined)"

    def replace(self, **kw) -> "Parameter":
        d = {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}
        d.update(kw)
        return Parameter(**d)
