"""MI355X-native implementation of NcorpiON's mutual-interaction step.

The product is the C-ABI HIP library libncg.so (include/ncg.h) and the host
glue that puts it behind NcorpiON's own entry points (ncorpion_amd/glue/).
This package binds the library for Python (ctypes) and holds the synthetic
initial-condition generator.
"""
from ._abi import Context, NcgError, load, LIB_PATH, PRECISIONS  # noqa: F401
from . import ic  # noqa: F401

__all__ = ["Context", "NcgError", "load", "LIB_PATH", "PRECISIONS", "ic"]
