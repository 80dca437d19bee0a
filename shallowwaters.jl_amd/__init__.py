"""ShallowWaters.jl's time-stepping path on AMD MI355X (gfx950).

Host-side mirror of the reference's public surface (``src/ShallowWaters.jl:3`` exports
``run_model`` and ``Parameter``) with the hot loop ``time_integration`` replaced by
hand-written HIP kernels behind the C ABI of ``include/swm.h`` (``csrc/``).

The directory name is not a python identifier; load it with ``load_package()`` of the
repository's ``__graft_entry__`` or ``importlib`` under the name ``shallowwaters_amd``.
"""
from .parameter import Parameter
from .numtypes import canonical_type, to_storage, to_float64
from .grid import Grid, make_grid
from .constants import Constants, make_constants
from .forcing import Forcing, make_forcing
from .model_setup import ModelSetup, make_setup, make_config, swm_config
from .initial_conditions import (PrognosticVars, initial_conditions, add_halo, remove_halo,
                                 seeded_initial_conditions, zero_prognostic)
from .time_integration import (time_integration, run_model, preallocate, DiagnosticVars, Simulation, Ensemble,
                               rrule, NoTangent, Tangent)
from . import workloads

__all__ = ["Parameter", "run_model", "time_integration", "preallocate", "DiagnosticVars",
           "PrognosticVars", "ModelSetup", "make_setup", "make_config", "initial_conditions",
           "add_halo", "remove_halo", "seeded_initial_conditions", "Simulation", "Ensemble", "Grid", "Constants",
           "Forcing", "canonical_type", "to_storage", "to_float64", "swm_config", "zero_prognostic",
           "make_grid", "make_constants", "make_forcing", "workloads", "rrule", "NoTangent", "Tangent"]
