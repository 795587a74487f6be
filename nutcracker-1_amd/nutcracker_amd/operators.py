"""MPO sites in the reference's sparse from/to block format, and bond-dimension helpers.

An operator site is the reference's ``OperatorSite`` (includes/nutcracker/tensors.hpp:
1641-1854): ``idx`` is a (2, n) array of 1-based ``uint32`` (from, to) pairs and
``mats`` a (d, d, n) complex128 Fortran-order array.  ``construct_operator_site``
follows operators.cpp:83-112: each link's d x d matrix is given row-major (as a
ublas matrix or a YAML ``data`` list) and its storage is copied verbatim into the
column-major block, so the stored block is W_n(a, b) = M_n(b, a).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

CPLX = np.complex128


@dataclass
class OperatorSite:
    d: int
    cl: int
    cr: int
    idx: np.ndarray   # (2, n) uint32, 1-based (from, to)
    mats: np.ndarray  # (d, d, n) complex128, Fortran order, W_n(a, b)

    @property
    def nmat(self) -> int:
        return int(self.idx.shape[1])


def construct_operator_site(d: int, cl: int, cr: int,
                            links: Sequence[Tuple[int, int, np.ndarray]]) -> OperatorSite:
    """operators.cpp:83-112 -- links are (from, to, M) with M a row-major d x d matrix."""
    if not links:
        raise ValueError("an operator site needs at least one link")
    idx = np.zeros((2, len(links)), dtype=np.uint32)
    mats = np.zeros((d, d, len(links)), dtype=CPLX, order="F")
    for n, (f, t, m) in enumerate(links):
        m = np.asarray(m, dtype=CPLX)
        if m.shape != (d, d):
            raise ValueError("link matrix must be d x d")
        if not (1 <= f <= cl and 1 <= t <= cr):
            raise ValueError("link index out of range")
        idx[0, n] = f
        idx[1, n] = t
        # the row-major storage of M read back column-major is M transposed
        mats[:, :, n] = m.T
    return OperatorSite(d, cl, cr, idx, mats)


# Pauli matrices (utilities.cpp:103-116)
I2 = np.eye(2, dtype=CPLX)
X = np.array([[0, 1], [1, 0]], dtype=CPLX)
Y = np.array([[0, -1j], [1j, 0]], dtype=CPLX)
Z = np.array([[1, 0], [0, -1]], dtype=CPLX)


def external_field_operator(n: int, m: np.ndarray) -> List[OperatorSite]:
    """operators.cpp:25-82 -- H = sum_i M_i."""
    m = np.asarray(m, dtype=CPLX)
    d = m.shape[0]
    eye = np.eye(d, dtype=CPLX)
    if n == 1:
        return [construct_operator_site(d, 1, 1, [(1, 1, m)])]
    first = construct_operator_site(d, 1, 2, [(1, 1, eye), (1, 2, m)])
    middle = construct_operator_site(d, 2, 2, [(1, 1, eye), (1, 2, m), (2, 2, eye)])
    last = construct_operator_site(d, 2, 1, [(1, 1, m), (2, 1, eye)])
    return [first] + [middle] * (n - 2) + [last]


def transverse_ising_operator(n: int, coupling: float) -> List[OperatorSite]:
    """operators.cpp:114-165 -- the reference's TFIM: H = sum_i Z_i + coupling * sum_i X_i X_{i+1}."""
    if n < 2:
        raise ValueError("need at least two sites")
    x2 = coupling * X
    first = construct_operator_site(2, 1, 3, [(1, 1, I2), (1, 2, X), (1, 3, Z)])
    middle = construct_operator_site(2, 3, 3, [(1, 1, I2), (1, 2, X), (2, 3, x2), (1, 3, Z), (3, 3, I2)])
    last = construct_operator_site(2, 3, 1, [(1, 1, Z), (2, 1, x2), (3, 1, I2)])
    return [first] + [middle] * (n - 2) + [last]


def tfim_zz_x_operator(n: int, h: float) -> List[OperatorSite]:
    """H = -sum_i Z_i Z_{i+1} - h sum_i X_i as a w = 3 from/to state machine
    (state 1: nothing placed yet, 2: a Z placed, 3: done)."""
    if n < 2:
        raise ValueError("need at least two sites")
    mz = -Z
    hx = -h * X
    first = construct_operator_site(2, 1, 3, [(1, 1, I2), (1, 2, Z), (1, 3, hx)])
    middle = construct_operator_site(2, 3, 3, [(1, 1, I2), (1, 2, Z), (2, 3, mz), (1, 3, hx), (3, 3, I2)])
    last = construct_operator_site(2, 3, 1, [(1, 1, hx), (2, 1, mz), (3, 1, I2)])
    return [first] + [middle] * (n - 2) + [last]


def spin_matrices(d: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(S+, S-, Sz) for spin (d-1)/2 in the basis m = s, s-1, ..., -s."""
    s = (d - 1) / 2.0
    ms = s - np.arange(d)
    sz = np.diag(ms).astype(CPLX)
    sp = np.zeros((d, d), dtype=CPLX)
    for k in range(1, d):
        m = ms[k]
        sp[k - 1, k] = np.sqrt(s * (s + 1) - m * (m + 1))
    return sp, sp.conj().T.copy(), sz


def xxz_operator(n: int, d: int = 2, J: float = 1.0, delta: float = 1.0) -> List[OperatorSite]:
    """H = J sum_i (Sx Sx + Sy Sy + delta Sz Sz) with S = spin (d-1)/2, as a w = 5
    from/to state machine with the topology of the reference's
    examples/generate/yaml/Heisenberg-ising-model.py:16-87."""
    if n < 2:
        raise ValueError("need at least two sites")
    sp, sm, sz = spin_matrices(d)
    eye = np.eye(d, dtype=CPLX)
    a, b, c = 0.5 * J * sm, 0.5 * J * sp, J * delta * sz
    first = construct_operator_site(d, 1, 5, [(1, 1, eye), (1, 2, sp), (1, 3, sm), (1, 4, sz)])
    middle = construct_operator_site(d, 5, 5, [(1, 1, eye), (1, 2, sp), (1, 3, sm), (1, 4, sz),
                                               (2, 5, a), (3, 5, b), (4, 5, c), (5, 5, eye)])
    last = construct_operator_site(d, 5, 1, [(2, 1, a), (3, 1, b), (4, 1, c), (5, 1, eye)])
    return [first] + [middle] * (n - 2) + [last]


def _product_prefix(dims: Sequence[int]) -> List[int]:
    out, p = [], 1
    for d in dims:
        p *= d
        out.append(p)
    return out


class RequestedBandwidthDimensionTooLargeError(ValueError):
    pass


def compute_bandwidth_dimension_sequence(requested: int, physical_dimensions: Sequence[int]) -> List[int]:
    """utilities.cpp:32-91 -- bond dimensions 1, d1, d1 d2, ... capped at ``requested``
    from both ends; length N+1."""
    if requested == 0:
        raise ValueError("requested bandwidth for the sequence must be greater than zero")
    dims = list(physical_dimensions)
    middle = (len(dims) + 1) // 2
    fwd = [1]
    for p in _product_prefix(dims[:middle]):
        if p >= requested:
            break
        fwd.append(p)
    rev = [1]
    for p in _product_prefix(list(reversed(dims))[:middle]):
        if p >= requested:
            break
        rev.append(p)
    if len(fwd) == middle + 1 or len(rev) == middle + 1:
        raise RequestedBandwidthDimensionTooLargeError(requested)
    fwd += [requested] * (len(dims) + 1 - len(fwd) - len(rev))
    fwd += list(reversed(rev))
    return fwd


def maximum_bandwidth_dimension(physical_dimensions: Sequence[int]) -> int:
    """utilities.hpp:264-283."""
    dims = list(physical_dimensions)
    if len(dims) == 1:
        return 1
    middle = (len(dims) + 1) // 2
    a = int(np.prod(dims[:middle], dtype=object))
    b = int(np.prod(list(reversed(dims))[:middle], dtype=object))
    return min(a, b, 2 ** 32 - 1)


def outside_tolerance(a, b, tol) -> bool:
    """utilities.hpp:418-420."""
    return ((abs(a) + abs(b)) / 2 > tol) and (abs(a - b) / (abs(a) + abs(b) + tol) > tol)
