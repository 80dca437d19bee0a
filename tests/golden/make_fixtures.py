"""Writes the golden fixtures of the oracle: small grids, 10 steps, both boundary
conditions, Float64 and Float32 (u, v, η, sst after the run, in Julia layout).

They pin the oracle against unintended change; the reference-derived pins live in
tests/test_oracle.py.  Regenerate with `python tests/golden/make_fixtures.py`.
"""
import os
import sys

import numpy as np

CASES = {
    "np12x8_f64": dict(T="Float64", nx=12, L_ratio=1.5, bc="nonperiodic", α=0.5),
    "per16x8_f64": dict(T="Float64", nx=16, L_ratio=2.0, bc="periodic"),
    "np12x8_f32": dict(T="Float32", nx=12, L_ratio=1.5, bc="nonperiodic", α=0.0),
    "per16x8_f32": dict(T="Float32", nx=16, L_ratio=2.0, bc="periodic"),
}
COMMON = dict(topography="seamount", wind_forcing_x="double_gyre", diffusion="Smagorinsky", bottom_drag="quadratic",
              Uadv=200.0, topo_width=400e3)


def run_case(swm, orc, name, nsteps=10):
    kw = dict(COMMON)
    kw.update(CASES[name])
    P = swm.Parameter(**kw)
    S = swm.make_setup(P)
    prog = swm.seeded_initial_conditions(S, seed=99, σ=0.3)
    u, v, e, s, *_ = orc.run(S, swm.make_config(S), prog, nsteps)
    return {"u": u, "v": v, "eta": e, "sst": s}


if __name__ == "__main__":
    root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    sys.path.insert(0, root)
    from __graft_entry__ import load_oracle, load_package
    swm, orc = load_package(), load_oracle()
    here = os.path.dirname(os.path.abspath(__file__))
    for name in CASES:
        np.savez(os.path.join(here, f"{name}.npz"), **run_case(swm, orc, name))
        print("wrote", name)
