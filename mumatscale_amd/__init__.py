"""MI355X-native cellular-automaton step of the muMatScale solidification model.

The step runs in libmmca.so (HIP kernels for gfx950 behind the C ABI of
include/mmca.h); this package holds its Python binding and a reader of the
reference's control-file format.
"""
from . import abi  # noqa: F401
from .ctrl import Config, read_ctrl  # noqa: F401
