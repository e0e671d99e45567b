"""anabasis_amd -- MI355X-native implementation of the Anabasis SIMPLE hot path.

The numerical work lives in libanabasis_hip.so (HIP kernels for gfx950 behind
the C ABI of include/anabasis_hip.h).  This package is the Python view of that
ABI: ``libpoisson`` mirrors the libpoisson/libscalar call API and ``simple``
the generic_simple_v1 driver.
"""
from ._lib import AnError, LIB_PATH, lib  # noqa: F401

__all__ = ["AnError", "LIB_PATH", "lib"]
