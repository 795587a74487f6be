"""Seeded initial states.

The reference seeds its random MPS from gfortran's unseeded ``rand()``
(core.f95:2617-2671), which no other runtime reproduces.  Here every site gets an
independent stream ``numpy.random.default_rng([seed, site])`` of complex Gaussian
entries with Re, Im ~ N(0, 1/2), filled in the storage order of the site tensor
psi(r, l, s) (r fastest).  Both the device driver and the oracle canonicalise the
same raw tensors themselves (sites 2..N right-normalised by their polar factor,
site 1 divided by its norm), so the fixture is the raw data only.
"""

from __future__ import annotations

from typing import List, Sequence

import numpy as np

CPLX = np.complex128


def raw_site(seed: int, site: int, br: int, bl: int, d: int) -> np.ndarray:
    rng = np.random.default_rng([int(seed), int(site)])
    n = br * bl * d
    v = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * np.sqrt(0.5)
    return np.asfortranarray(v.reshape((br, bl, d), order="F"))


def raw_state(seed: int, physical_dimensions: Sequence[int], bonds: Sequence[int]) -> List[np.ndarray]:
    return [raw_site(seed, i, bonds[i + 1], bonds[i], d) for i, d in enumerate(physical_dimensions)]
