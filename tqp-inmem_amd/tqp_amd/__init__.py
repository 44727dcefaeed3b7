"""MI355X-native relational hot path: filter + compaction, sort-based equi-join, group-by
aggregate and sort/top-k behind the reference engine's Python operator interface, with the
compute in hand-written HIP kernels for gfx950 (libtqhip.so, loaded through ctypes)."""
from . import _lib, constants, conversion, kernels, variable  # noqa: F401

__all__ = ["_lib", "constants", "conversion", "kernels", "variable"]
