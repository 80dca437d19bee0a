"""Pins of the oracle's Float16 / BFloat16 values with Float32 arithmetic (``Parameter(T=...,
Tprog="Float32", arithmetic="Float32")``, a form of this library, not of the reference):

- same conversions as the reference's low-precision T: with every operation rounded to the
  format (the planted mistake MUT_WIDE_ROUNDS) the form is bitwise the reference's
  ``T = BFloat16`` / ``Float16`` run, tracer advection, diffusion and drag included; so the
  values that enter T (stage inputs, static fields, constants) are rounded where, and only
  where, the reference rounds them;
- Float32 arithmetic: on a state, static fields and constants that the format holds exactly
  (Sadourny's scheme, whose PV weight is ½), the first stage's tendencies are bitwise those of
  the Float32 model.  Rounding any intermediate to the format (MUT_WIDE_ROUNDS) breaks this.
"""
import numpy as np
import pytest


def _model(swm, T, arithmetic=None, **kw):
    d = dict(T=T, Tprog="Float32", nx=40, L_ratio=1, bc="nonperiodic", topography="seamount",
             wind_forcing_x="double_gyre", Fx0=0.12, diffusion="Smagorinsky", bottom_drag="quadratic",
             Uadv=25.0, α=0.5, arithmetic=arithmetic)
    d.update(kw)
    P = swm.Parameter(**d)
    S = swm.make_setup(P)
    return S, swm.make_config(S)


def _same(a, b):
    return all(np.array_equal(np.asarray(x, np.float64), np.asarray(y, np.float64)) for x, y in zip(a, b))


def pin_conversions_as_reference(swm, orc, T, tcomm):
    """Returns the disagreements between the form with rounding operations and the
    reference's T run (empty when the conversions agree)."""
    S, cfg = _model(swm, T, "Float32", Tcomm=tcomm)
    S0, cfg0 = _model(swm, T, None, Tcomm=tcomm)
    prog = swm.seeded_initial_conditions(S, seed=3, σ=0.5)
    n = 2 * S.grid.nadvstep + 1                       # through two tracer advections
    a = orc.run(S, cfg, prog, n)[:4]
    b = orc.run(S0, cfg0, prog, n)[:4]
    return [] if _same(a, b) else [f"{T} Tcomm={tcomm}: the form with rounding operations is not the T run"]


def pin_float32_arithmetic(swm, orc, T):
    """Returns the disagreements between the form's first-stage tendencies and the Float32
    model's on inputs the format holds exactly."""
    kw = dict(adv_scheme="Sadourny", topography="flat", H=512.0, wind_forcing_x="none", wind_forcing_y="none",
              Fx0=0.0, tracer_advection=False)
    S, cfg = _model(swm, T, "Float32", **kw)
    S32, cfg32 = _model(swm, "Float32", None, **kw)
    # the (T) constants as the form holds them; f_q as the form holds it
    for name in ("g", "cD", "rD", "gamma", "cSmag", "nuB", "dtadvu", "dtadvv", "half_dtadvu", "half_dtadvv", "jSST",
                 "SSTmin"):
        setattr(cfg32, name, getattr(cfg, name))
    S32.grid.f_q = swm.to_float64(S.grid.f_q, T).astype(np.float32, order="F")
    assert np.array_equal(np.asarray(S32.forcing.H, np.float64), swm.to_float64(S.forcing.H, T))
    # a state the format holds exactly
    prog = swm.seeded_initial_conditions(S, seed=7, σ=0.4)
    r = [swm.to_float64(swm.to_storage(np.asarray(x, np.float64), T), T).astype(np.float32, order="F")
         for x in (prog.u, prog.v, prog.η, prog.sst)]
    st = swm.PrognosticVars(*r, T="Float32")
    ra = orc.run(S, cfg, st, 1, stop_stage=1, want_rk=True)[4]
    rb = orc.run(S32, cfg32, st, 1, stop_stage=1, want_rk=True)[4]
    bad = []
    for name, x, y in zip(("du", "dv", "dη"), ra[6:], rb[6:]):
        if not np.array_equal(x, y):
            bad.append(f"{T} {name}: {int(np.count_nonzero(x != y))} cells differ from Float32")
    if not np.any(ra[6]):
        bad.append("the tendencies are all zero: the pin compares nothing")
    return bad


@pytest.mark.parametrize("T,tcomm", [("BFloat16", "Float32"), ("BFloat16", "BFloat16"), ("Float16", "Float32")])
def test_form_rounds_where_the_reference_rounds(swm, orc, T, tcomm):
    with orc.mutation(13):                            # MUT_WIDE_ROUNDS
        assert not pin_conversions_as_reference(swm, orc, T, tcomm)
    # and without it the form differs from the reference's T run
    assert pin_conversions_as_reference(swm, orc, T, tcomm)


@pytest.mark.parametrize("T", ["BFloat16", "Float16"])
def test_form_computes_in_float32(swm, orc, T):
    assert not pin_float32_arithmetic(swm, orc, T)
    with orc.mutation(13):
        assert pin_float32_arithmetic(swm, orc, T), "rounded operations went unnoticed"


def test_form_is_validated(swm):
    with pytest.raises(AssertionError, match="arithmetic"):
        swm.Parameter(T="Float32", arithmetic="Float32")
    with pytest.raises(AssertionError, match="nstep_advcor"):
        swm.Parameter(T="BFloat16", Tprog="Float32", arithmetic="Float32", nstep_advcor=2)
    S, cfg = _model(swm, "BFloat16", "Float32")
    assert cfg.t_arith32 == 1 and cfg.T == 3 and cfg.Tprog == 1
