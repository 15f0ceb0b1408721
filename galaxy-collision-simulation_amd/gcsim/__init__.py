"""gcsim — Python access to the gfx950 galaxy-collision force engine
(libgcsim.so, C ABI in include/gcsim.h)."""
from ._lib import GcError, load  # noqa: F401
from .engine import Engine, device_count, make_params, solver_available  # noqa: F401
from . import generators  # noqa: F401
from .writer import SnapshotWriter  # noqa: F401

__all__ = ["Engine", "GcError", "SnapshotWriter", "device_count", "generators", "load",
           "make_params", "solver_available"]
