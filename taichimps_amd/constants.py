"""Precision selectors, in the place of the reference's Taichi float types.

``f64`` is the reference's FP64 path. ``f32split`` stores positions as an FP64
bin origin plus an FP32 offset and computes contacts and integration in FP32,
keeping the virial reduction, the pressure and the servo in FP64.
"""

from . import _native

f64 = "f64"
f32split = "f32split"

_CODES = {f64: _native.FP64, "fp64": _native.FP64, "float64": _native.FP64,
          f32split: _native.FP32_SPLIT, "fp32split": _native.FP32_SPLIT,
          "fp32-split": _native.FP32_SPLIT}


def precision_code(float_type) -> int:
    """The C ABI precision code of a float_type argument."""
    if float_type is None:
        return _native.FP64
    key = getattr(float_type, "__name__", float_type)
    if isinstance(key, str) and key.lower() in _CODES:
        return _CODES[key.lower()]
    raise ValueError(f"unknown float_type {float_type!r}; use taichimps_amd.f64 or .f32split")
