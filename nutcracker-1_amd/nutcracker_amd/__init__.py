"""nutcracker_amd -- MI355X-native variational-MPS inner loop.

* ``core``: the reference's Nutcracker::Core entry points over the C-ABI of
  libnutcracker_hip.so (HIP kernels for gfx950).
* ``operators``: MPO sites in the reference's from/to block format and the
  bond-dimension helpers of utilities.cpp.
"""

from . import operators  # noqa: F401

__all__ = ["core", "operators"]


def __getattr__(name):
    if name == "core":
        from . import core
        return core
    if name == "chain":
        from . import chain
        return chain
    raise AttributeError(name)
