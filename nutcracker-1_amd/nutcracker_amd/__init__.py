"""nutcracker_amd -- MI355X-native variational-MPS inner loop.

* ``core``: the reference's Nutcracker::Core entry points over the C-ABI of
  libnutcracker_hip.so (HIP kernels for gfx950).
* ``chain``: the device-resident sweep driver.
* ``operators``: MPO sites in the reference's from/to block format and the
  bond-dimension helpers of utilities.cpp.
* ``fixtures``: seeded raw initial states.
"""

from . import fixtures, operators  # noqa: F401

__all__ = ["core", "chain", "fixtures", "operators"]
