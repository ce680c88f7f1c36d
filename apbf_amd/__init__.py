"""apbf_amd — MI355X (gfx950) implementation of APBF's per-substep particle
pipeline behind a C ABI (include/apbf.h), with a thin Python front end.

The compute path is libapbf_hip.so; there is no CPU fallback.
"""
from ._lib import ApbfError, LIB_PATH, lib  # noqa: F401
from . import config  # noqa: F401

__all__ = ["ApbfError", "LIB_PATH", "lib", "config"]
