"""sbh_amd: MI355X-native Barnes-Hut gravity and winding-number summation.

`sbh_amd.ffi` mirrors the reference's python/jax_sbh_ffi.py surface on
PyTorch device tensors; every entry point calls the HIP library
libsbh_amd.so through the C ABI declared in include/sbh_amd.h.
"""
from . import ffi  # noqa: F401
from ._lib import library_path, load_library  # noqa: F401

__all__ = ["ffi", "library_path", "load_library"]
from . import gravity  # noqa: E402,F401
from . import graph  # noqa: E402,F401
from . import synthetic  # noqa: E402,F401
from . import wn  # noqa: E402,F401
from . import fields  # noqa: E402,F401
from . import jax_sbh  # noqa: E402,F401
from . import stochastic  # noqa: E402,F401

__all__ += ["gravity", "graph", "synthetic", "wn", "fields", "jax_sbh", "stochastic"]
