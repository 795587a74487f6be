"""Pins of the oracle's projector algebra (CPU only).

* compute_overlap_with_projectors (core.f95:1266-1339) against its closed form:
  for overlap vectors V (n x p) turned into reflectors, the weights are
  P^T V^T v for the pivot permutation P, so the overlap is ||V^T v|| (bilinear,
  the reference's projector convention Pt x = 0, tests/python/test.py:953).
* tests/projectors.cpp:195-243 restated: the projector of one random state,
  formed at every site from overlap boundaries and overlap vectors, measures
  |<state 1|state 2>| of a second random state at every site (1e-12).
* Null-reflector swaps (core.f95:1417-1432) that are not the identity: a null
  row of the pivoted R followed by a non-null one.  The orthogonal space is then
  spanned by the Q columns of the null rows and those past min(n, p), which the
  filter and the strategy-1 solve (through project_matrix_into_orthog_space and
  its swap, core.f95:506-530, 1666-1762) must use.
"""

import numpy as np
import pytest
import scipy.linalg

from helpers import crand, dense_mps, hermitian_env, rel

import oracle.core as oc


def _reflectors(n, V):
    p = V.shape[1]
    rank, refl, tau, swaps = oc.convert_vectors_to_reflectors(n, p, V)
    return rank, refl, tau, swaps


@pytest.mark.parametrize("n,p,seed", [(7, 1, 0), (30, 2, 1), (64, 3, 2), (5, 4, 3), (200, 3, 4)])
def test_overlap_with_projectors_closed_form(n, p, seed):
    rng = np.random.default_rng(seed)
    V = crand(rng, n, p)
    rank, refl, tau, swaps = _reflectors(n, V)
    k = min(n, p)
    assert rank == k
    for _ in range(3):
        v = crand(rng, n)
        got = oc.compute_overlap_with_projectors(p, k, refl, tau, swaps, n, v)
        want = np.linalg.norm(V.T @ v)
        assert abs(got - want) <= 1e-12 * want, (got, want)
        # the sesquilinear form (a conjugated operand) is a different number
        assert abs(got - np.linalg.norm(V.conj().T @ v)) > 1e-6 * want
    # a vector in the bilinear orthogonal complement has no overlap
    null = scipy.linalg.null_space(V.T)
    if null.shape[1]:
        v0 = null @ crand(rng, null.shape[1])
        assert oc.compute_overlap_with_projectors(p, k, refl, tau, swaps, n, v0) < 1e-12 * np.linalg.norm(v0)


def _random_mps(rng, dims, chi):
    bonds = [1]
    for i in range(1, len(dims)):
        left = int(np.prod(dims[:i]))
        right = int(np.prod(dims[i:]))
        bonds.append(min(chi, left, right))
    bonds.append(1)
    return [crand(rng, bonds[i + 1], bonds[i], dims[i]) for i in range(len(dims))], bonds


def _gauges(sites, bonds, dims):
    """The three overlap-site sets of a projector state (projectors.hpp:324-359):
    left-canonical, centre and right-canonical forms of the same state, made with
    the oracle's polar moves."""
    N = len(sites)
    # right-canonical form: sweep from the right with norm_denorm_going_left
    right = [s.copy() for s in sites]
    for i in range(N - 1, 0, -1):
        info, den, nrm = oc.norm_denorm_going_left(bonds[i - 1], bonds[i], bonds[i + 1], dims[i - 1], dims[i],
                                                   right[i - 1], right[i])
        assert info == 0
        right[i - 1], right[i] = den, nrm
    # centre at site i with sites < i left-canonical: sweep right from the right form
    centre = [None] * N
    left = [None] * N
    cur = [s.copy() for s in right]
    for i in range(N):
        centre[i] = cur[i].copy()
        if i < N - 1:
            info, nrm, den = oc.norm_denorm_going_right(bonds[i], bonds[i + 1], bonds[i + 2], dims[i], dims[i + 1],
                                                        cur[i], cur[i + 1])
            assert info == 0
            cur[i], cur[i + 1] = nrm, den
            left[i] = nrm
    return left, centre, right


@pytest.mark.parametrize("dims,chi,seed", [((2, 3, 2, 2), 4, 0), ((3, 2, 4), 5, 1), ((2, 2, 2, 2, 2), 3, 2),
                                           ((4,), 1, 3), ((2, 5), 2, 4)])
def test_projector_overlap_at_every_site(dims, chi, seed):
    """tests/projectors.cpp:195-243 on the oracle."""
    rng = np.random.default_rng(seed)
    s1, bonds = _random_mps(rng, dims, chi)
    s2, bonds2 = _random_mps(rng, dims, chi)
    assert bonds == bonds2
    N = len(dims)
    want = abs(np.vdot(dense_mps(s1), dense_mps(s2)))
    left, centre, right = _gauges(s1, bonds, dims)
    # overlap sites (bl, d, br) = conj of the state site, form_overlap_site_tensor
    ov = lambda s, i: oc.form_overlap_site_tensor(bonds[i + 1], bonds[i], dims[i], s)
    # right overlap boundaries from the projector's right-canonical sites and state 2
    rb = [None] * (N + 1)
    rb[N] = np.ones((1, 1), complex, order="F")
    for i in range(N - 1, 0, -1):
        rb[i] = oc.contract_vs_right(bonds[i], bonds[i + 1], bonds[i], bonds[i + 1], dims[i], rb[i + 1],
                                     ov(right[i], i), s2[i])
    lb = np.ones((1, 1), complex, order="F")
    for i in range(N):
        n = bonds[i + 1] * bonds[i] * dims[i]
        vec = oc.form_overlap_vector(bonds[i], bonds[i + 1], bonds[i], bonds[i + 1], dims[i], lb, rb[i + 1],
                                     ov(centre[i], i))
        V = np.asfortranarray(vec.reshape(n, 1, order="F"))
        rank, refl, tau, swaps = _reflectors(n, V)
        got = oc.compute_overlap_with_projectors(1, min(n, 1), refl, tau, swaps, n, s2[i])
        assert abs(got - want) <= 1e-12 * max(want, 1e-300), (i, got, want)
        if i < N - 1:
            lb = oc.contract_vs_left(bonds[i], bonds[i + 1], bonds[i], bonds[i + 1], dims[i], lb, ov(left[i], i),
                                     s2[i])


def _swapped_vectors(rng, n):
    """V = Q R for a random n x 4 isometry Q and an upper-triangular R with
    diagonal (1, t, 0.9 t, 0.3 t), t = 0.98e-12, whose trailing column norms keep
    the columns in zgeqp3's pivot order.  Row 1 (0-based) has norm t, below the
    reflector null threshold 1e-12; row 2, with its off-diagonal 0.5 t, has norm
    1.03 t above it; row 3 is null again.  So a null reflector precedes a non-null
    one and the swaps are not the identity."""
    Q, _ = np.linalg.qr(crand(rng, n, 4))
    t = 0.98e-12
    R = np.array([[1.0, 0.3, -0.2j, 0.1],
                  [0, t, 0, 0],
                  [0, 0, 0.9 * t, 0.5 * t],
                  [0, 0, 0, 0.3 * t]], complex)
    return np.asfortranarray(Q @ R), Q


def _kept_basis(V, n):
    """Columns spanning the bilinear orthogonal space the reference keeps: conj of
    the Q columns of the null rows of the pivoted R and of those past min(n, p)."""
    p = V.shape[1]
    Qf, Rf, piv = scipy.linalg.qr(V, pivoting=True, mode="full")
    k = min(n, p)
    null_rows = [j for j in range(k) if np.linalg.norm(Rf[j, j:]) < np.float32(1e-12)]
    kept = null_rows + list(range(k, n))
    return np.conj(Qf[:, kept]), len(null_rows)


def test_non_identity_swaps_filter():
    rng = np.random.default_rng(5)
    n = 10
    V, _ = _swapped_vectors(rng, n)
    rank, refl, tau, swaps = _reflectors(n, V)
    assert rank == 2
    assert list(swaps) == [1, 3, 3, 4]  # the null reflector is moved past the non-null one
    B, nnull = _kept_basis(V, n)
    assert nnull == 2
    orth = n - rank
    assert B.shape[1] == orth
    for _ in range(3):
        x = crand(rng, n)
        f = oc.filter_components_outside_orthog(n, 4, 4, orth, refl, tau, swaps, x)
        # the filter is the orthogonal projector onto span(B)
        want = B @ (B.conj().T @ x)
        assert rel(f, want) < 1e-12
    # without the swap, the filter would keep the non-null direction instead
    ident = np.arange(1, 5, dtype=np.int32)
    f_wrong = oc.filter_components_outside_orthog(n, 4, 4, orth, refl, tau, ident, x)
    assert rel(f_wrong, want) > 1e-3


def _swap_case(seed):
    rng = np.random.default_rng(seed)
    bl, br, d, cl, cr = 1, 3, 2, 2, 2
    n = bl * br * d
    V, _ = _swapped_vectors(rng, n)
    rank, refl, tau, swaps = _reflectors(n, V)
    assert rank == 2 and list(swaps) == [1, 3, 3, 4]
    orth = n - rank
    assert orth == 4
    idx = np.array([[1, 1], [1, 2], [2, 2]], np.uint32).T.copy(order="F")
    mats = crand(rng, d, d, 3)
    mats = np.asfortranarray(mats + mats.transpose(1, 0, 2).conj())
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    guess = oc.filter_components_outside_orthog(n, 4, 4, orth, refl, tau, swaps, crand(rng, n))
    return dict(bl=bl, br=br, d=d, cl=cl, cr=cr, n=n, V=V, rank=rank, refl=refl, tau=tau, swaps=swaps, orth=orth,
                idx=idx, mats=mats, L=L, R=R, guess=guess)


def swap_case_expected(c):
    """Lowest eigenvalue of H_eff on the kept space span(B), and B."""
    H = oc.compute_optimization_matrix(c["bl"], c["br"], c["cl"], c["cr"], c["d"], c["L"], c["idx"], c["mats"],
                                       c["R"])
    B, _ = _kept_basis(c["V"], c["n"])
    return np.linalg.eigvalsh(B.conj().T @ H @ B)[0], B, H


def test_non_identity_swaps_strategy_1():
    """A strategy-1 site (orthogonal space of dimension 4) with those projectors.
    The vector projection that strategies 2 and 3 apply around every matvec
    (operate_on, core.f95:2529-2556) gives H_eff on span(B): its lowest
    eigenvalue is that of B^H H B.  The reference's strategy 1 builds the
    projected matrix with swap_matrix_inplace (core.f95:506-530), which swaps
    only the first n_swaps entries of the columns and strides the row swap by
    n_swaps instead of the leading dimension; with non-identity swaps its matrix
    is not H_eff on span(B), and the oracle, which restates it literally,
    returns another number.  The device path follows the vector projection."""
    c = _swap_case(6)
    w0, B, H = swap_case_expected(c)
    n, orth = c["n"], c["orth"]
    args = (c["refl"], c["tau"], c["swaps"])
    # the vector projection, column by column: project(H unproject(e_j))
    Hp = np.zeros((orth, orth), complex)
    for j in range(orth):
        e = np.zeros(orth, complex)
        e[j] = 1
        full = oc.unproject_from_orthogonal_space(n, 4, 4, orth, *args, e)
        Hp[:, j] = oc.project_into_orthogonal_space(n, 4, 4, orth, *args, H @ full)
    assert abs(np.linalg.eigvalsh(0.5 * (Hp + Hp.conj().T))[0] - w0) <= 1e-10 * max(1.0, abs(w0))
    # the literal strategy 1 of the reference
    info, it, res, ev, nrm = oc.optimize(c["bl"], c["br"], c["cl"], c["cr"], c["d"], c["L"], c["idx"], c["mats"],
                                         c["R"], 4, 4, orth, *args, "SR", 1e-12, 1000, c["guess"])
    assert info == 0
    assert abs(ev.real - w0) > 1e-3 * max(1.0, abs(w0))
    # with identity swaps (no null reflector before a non-null one) the two agree
    rng = np.random.default_rng(7)
    V = crand(rng, n, 2)
    rank, refl, tau, swaps = _reflectors(n, V)
    assert list(swaps) == [1, 2]
    B2 = np.conj(scipy.linalg.qr(V, mode="full")[0][:, 2:])
    w2 = np.linalg.eigvalsh(B2.conj().T @ H @ B2)[0]
    g2 = oc.filter_components_outside_orthog(n, 2, 2, n - 2, refl, tau, swaps, crand(rng, n))
    info, it, res, ev, nrm = oc.optimize(c["bl"], c["br"], c["cl"], c["cr"], c["d"], c["L"], c["idx"], c["mats"],
                                         c["R"], 2, 2, n - 2, refl, tau, swaps, "SR", 1e-12, 1000, g2)
    assert info == 0 and abs(ev.real - w2) <= 1e-10 * max(1.0, abs(w2))
