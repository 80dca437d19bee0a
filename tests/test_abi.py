"""The C ABI library without a GPU: it loads, exports every entry point include/swm.h
declares, and the product path never reaches the oracle."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "swm.h")
PKG = os.path.join(ROOT, "shallowwaters.jl_amd")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*\*?(swm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    for s in ("swm_create", "swm_destroy", "swm_set_static", "swm_upload_prog", "swm_download_prog", "swm_step",
              "swm_stages", "swm_download_rk", "swm_reduce", "swm_last_error", "swm_time_kernel"):
        assert s in syms


def test_library_loads_and_exports_every_declared_symbol(swm):
    from shallowwaters_amd import _abi
    if not os.path.exists(_abi.LIB_PATH):
        pytest.fail("libswm_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_abi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_abi.EXPORTS) == declared_symbols()
    v = ctypes.c_int32(0)
    assert L.swm_version(ctypes.byref(v)) == 0 and v.value == 2


def struct_fields(name):
    text = open(HEADER).read()
    body = text[text.index(f"typedef struct {name} {{"):text.index(f"}} {name};")]
    names = []
    for line in body.splitlines()[1:]:
        line = line.split("/*")[0].strip()
        if not line or not (line.startswith("int") or line.startswith("double")):
            continue
        decl = line.split(None, 1)[1].rstrip(";")
        for part in decl.split(","):
            names.append(part.strip().split("[")[0])
    return names


def test_config_struct_matches_header(swm):
    """The ctypes mirror of swm_config lists the header's fields in order."""
    assert [f[0] for f in swm.swm_config._fields_] == struct_fields("swm_config")


def test_setup_struct_and_enums_match_header(swm):
    """swm_setup's ctypes mirror, and the topography and wind codes, as the header has them."""
    from shallowwaters_amd import _abi
    assert [f[0] for f in _abi.SwmSetup._fields_] == struct_fields("swm_setup")
    assert ctypes.sizeof(_abi.SwmSetup) == 4 * 4 + 18 * 8
    text = open(HEADER).read()
    for prefix, table in (("SWM_TOPO_", _abi.TOPOGRAPHY), ("SWM_WIND_", _abi.WIND)):
        codes = {k.lower(): int(v) for k, v in re.findall(prefix + r"(\w+)\s*=\s*(\d+)", text)}
        assert codes == table


def test_setup_of_carries_the_grid_scalars(swm):
    from shallowwaters_amd.time_integration import setup_of
    S = swm.make_setup(swm.Parameter(nx=64, topography="seamount", wind_forcing_x="double_gyre", ϕ=30.0))
    st = setup_of(S)
    G, P = S.grid, S.parameters
    assert (st.Delta, st.Lx, st.Ly, st.f0, st.beta) == (G.Δ, G.Lx, G.Ly, G.f0, G.β)
    assert (st.topography, st.wind_x, st.wind_y) == (3, 2, 3)
    assert (st.H, st.rho, st.phi, st.phik) == (P.H, P.ρ, 30.0, 30.0)


def test_product_path_never_uses_the_oracle():
    for dirpath, _, files in os.walk(PKG):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f), encoding="utf-8").read()
                assert "oracle" not in text.lower(), os.path.join(dirpath, f)
