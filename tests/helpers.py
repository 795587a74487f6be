"""Shared fixtures for the parity tests: seeded random inputs in the reference's
layouts, following the generators of tests/python/test.py:175-196."""

import json
import os

import numpy as np

CPLX = np.complex128
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_goldens.json")


def goldens():
    with open(GOLDEN) as f:
        return json.load(f)


def z(pairs):
    return np.array([complex(a, b) for a, b in pairs], dtype=CPLX)


def crand(rng, *shape):
    """tests/python/test.py:175-177: uniform in [-1,1] + i [-1,1]."""
    return np.asfortranarray((rng.random(shape) * 2 - 1) + 1j * (rng.random(shape) * 2 - 1))


def random_sparse_matrices(rng, cl, cr, d, symmetric=False, nblocks=None):
    """tests/python/test.py:180-189 -- random (from, to) pairs and d x d blocks."""
    nb = int(rng.integers(2, cl + cr + 1)) if nblocks is None else nblocks
    idx = np.array([(int(rng.integers(1, cl + 1)), int(rng.integers(1, cr + 1))) for _ in range(nb)],
                   dtype=np.uint32).T.copy(order="F")
    mats = crand(rng, d, d, nb)
    if symmetric:
        mats = np.asfortranarray(mats + mats.transpose(1, 0, 2).conj())
    return idx, mats


def hermitian_env(rng, b, c):
    e = crand(rng, b, b, c)
    return np.asfortranarray(e + e.transpose(1, 0, 2).conj())


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = max(np.linalg.norm(b.ravel()), 1e-300)
    return float(np.linalg.norm((a - b).ravel()) / den)


def phase_overlap(a, b):
    """|<a|b>| / (|a||b|): 1 for vectors equal up to a phase."""
    a = np.asarray(a).ravel(order="F")
    b = np.asarray(b).ravel(order="F")
    return abs(np.vdot(a, b)) / (np.linalg.norm(a) * np.linalg.norm(b))


def dense_hamiltonian(ops):
    """The dense matrix of a chain's MPO (small N only): the blocks summed into
    W[from, to] and contracted along the chain, site 0 the most significant."""
    tensors = []
    for o in ops:
        W = np.zeros((o.cl, o.cr, o.d, o.d), complex)
        for k in range(o.nmat):
            f, t = int(o.idx[0][k]) - 1, int(o.idx[1][k]) - 1
            W[f, t] += np.asarray(o.mats)[:, :, k].T
        tensors.append(W)
    H = tensors[0]
    for W in tensors[1:]:
        H = np.einsum("aAij,ABkl->aBikjl", H, W).reshape(H.shape[0], W.shape[1], H.shape[2] * W.shape[2], -1)
    return H[0, 0]


def tfim_zz_x_ground_energy(n, h):
    """Ground energy of H = -sum Z_i Z_{i+1} - h sum X_i on an open chain of n
    sites: free fermions (Lieb, Schultz, Mattis), E0 = (tr A - sum_k eps_k)/2 - n h
    with eps_k^2 the eigenvalues of (A - B)(A + B); pinned against exact
    diagonalisation in test_oracle.py."""
    A = np.zeros((n, n))
    B = np.zeros((n, n))
    for i in range(n):
        A[i, i] = 2 * h
    for i in range(n - 1):
        A[i, i + 1] = A[i + 1, i] = -1.0
        B[i, i + 1] = -1.0
        B[i + 1, i] = 1.0
    lam = np.sqrt(np.clip(np.linalg.eigvalsh((A - B) @ (A + B)), 0, None))
    return 0.5 * (np.trace(A) - lam.sum()) - n * h


def dense_mps(sites):
    """A matrix product state (site tensors (br, bl, d)) as a dense vector."""
    v = np.ones((1, 1), complex)
    for A in sites:
        v = np.einsum("cl,rls->csr", v, A).reshape(-1, A.shape[0])
    return v[:, 0]

