"""MI355X engine for simpimc's action-evaluation and sampling hot path.

The compute lives in the HIP library ``libspimc.so`` (C ABI in include/spimc.h);
this package holds its ctypes binding (``abi``) and the system/table builders
(``system``).
"""
from . import abi, system  # noqa: F401
from .abi import Engine, SpimcError  # noqa: F401

__all__ = ["abi", "system", "Engine", "SpimcError"]
