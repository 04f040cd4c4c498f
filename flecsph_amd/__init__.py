"""flecsph_amd: MI355X-native (gfx950) per-step SPH particle pipeline of
FleCSPH behind the C ABI of include/fsph.h (libfsph.so)."""
from . import _abi
from ._abi import default_params
from .system import BatchSystem, BodySystem, FsphError, autotune, set_tuning

__all__ = ["BatchSystem", "BodySystem", "FsphError", "autotune", "set_tuning",
           "default_params", "_abi"]
