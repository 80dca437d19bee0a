"""Number formats of the time-stepping path: names, host storage and rounding.

The reference is generic in three formats (``src/default_parameters.jl:3-6``):
``T`` for the right-hand side, ``Tprog`` for the prognostic variables and ``Tcomm`` for
ghost-point copies.  Julia's ``Float64``, ``Float32`` and ``Float16`` map to numpy's IEEE
types; ``BFloat16`` (not a dependency of the reference) is held on the host as ``uint16``
bit patterns and rounded to nearest even from float32.
"""
from __future__ import annotations

import numpy as np

# codes shared with include/swm.h (SWM_F64 ...)
TYPE_CODES = {"Float64": 0, "Float32": 1, "Float16": 2, "BFloat16": 3}
_ALIASES = {
    "float64": "Float64", "f64": "Float64", "double": "Float64",
    "float32": "Float32", "f32": "Float32", "float": "Float32", "single": "Float32",
    "float16": "Float16", "f16": "Float16", "half": "Float16",
    "bfloat16": "BFloat16", "bf16": "BFloat16",
}


def canonical_type(t) -> str:
    """Map ``np.float32``, ``"Float32"``, ``"f32"``, ``torch.float32`` ... to a canonical name."""
    if isinstance(t, str):
        if t in TYPE_CODES:
            return t
        key = t.lower()
        if key in _ALIASES:
            return _ALIASES[key]
        raise ValueError(f"unsupported number format {t!r}")
    if t is None:
        raise ValueError("number format is None")
    for np_t, name in ((np.float64, "Float64"), (np.float32, "Float32"), (np.float16, "Float16")):
        if t is np_t or (isinstance(t, np.dtype) and t == np.dtype(np_t)):
            return name
    s = str(t).replace("torch.", "")
    if s.lower() in _ALIASES:
        return _ALIASES[s.lower()]
    raise ValueError(f"unsupported number format {t!r}")


def storage_dtype(name: str) -> np.dtype:
    name = canonical_type(name)
    return {"Float64": np.dtype(np.float64), "Float32": np.dtype(np.float32),
            "Float16": np.dtype(np.float16), "BFloat16": np.dtype(np.uint16)}[name]


def itemsize(name: str) -> int:
    return storage_dtype(name).itemsize


def bf16_bits_from_f32(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even); returns uint16 bits."""
    a = np.asarray(a, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    lsb = (b >> 16) & 1
    r = ((b + 0x7FFF + lsb) >> 16).astype(np.uint16)
    r = np.where(nan, ((b >> 16) | 0x40).astype(np.uint16), r)
    return r.astype(np.uint16)


def f32_from_bf16_bits(bits) -> np.ndarray:
    bits = np.asarray(bits, dtype=np.uint16)
    return (bits.astype(np.uint32) << 16).view(np.float32)


def f32_round_to_odd(x) -> np.ndarray:
    """float64 -> float32 rounding to odd: float32 then a second rounding to any format of
    at most 22 significand bits (bfloat16, binary16) is correctly rounded from float64."""
    x = np.asarray(x, dtype=np.float64)
    f = x.astype(np.float32)
    bits = f.view(np.uint32).copy()
    inexact = (f.astype(np.float64) != x) & np.isfinite(f) & np.isfinite(x)
    away = inexact & (np.abs(f.astype(np.float64)) > np.abs(x))
    bits[away] -= np.uint32(1)
    bits[inexact] |= np.uint32(1)
    return bits.view(np.float32)


def to_storage(x, name: str, order="F") -> np.ndarray:
    """Round float64/float32 values to format ``name`` (Julia's ``T.(x)``) in host storage."""
    name = canonical_type(name)
    x = np.asarray(x)
    if name == "BFloat16":
        out = bf16_bits_from_f32(x if x.dtype == np.float32 else f32_round_to_odd(x))
    else:
        out = x.astype(storage_dtype(name))
    return np.asarray(out, order=order)


def to_float64(x, name: str) -> np.ndarray:
    """Exact widening of host storage of format ``name`` to float64."""
    name = canonical_type(name)
    if name == "BFloat16":
        return f32_from_bf16_bits(x).astype(np.float64)
    return np.asarray(x).astype(np.float64)


def round_scalar(x: float, name: str) -> float:
    """``T(x)`` for a Float64 scalar, returned as the exactly representable float."""
    return float(to_float64(to_storage(np.float64(x), name), name))
