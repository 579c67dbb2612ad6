"""MI355X (gfx950) backend for the Fluide_Engine hybrid particle-grid substep.

`capi` binds the C ABI of include/fluide_hip.h; the HIP library it loads is the
only compute path (no CPU fallback).
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
