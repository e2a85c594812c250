"""calaman_g_amd -- MI355X (gfx950) matrix exponential, a drop-in for calaman_G's expm path.

The arithmetic lives in libcalaman_expm.so (hand-written HIP kernels behind the
C ABI of include/calaman_expm.h); this package is its Python host mirror.
"""
from ._lib import CalamanError, LIB_PATH, lib  # noqa: F401
from .expm import *  # noqa: F401,F403
from . import cg  # noqa: F401

__version__ = "0.1.0"
