"""The remaining Fortran-core entry points of sources/common/core.cpp on the GPU,
the optimizer's error paths, physical dimensions above 4 and null-reflector
swaps that are not the identity.

* Deterministic routines (contract_matrix_left, extend_state_vector_fragment,
  construct_left/right_exp_boundary, norm_for_left/right) against their
  formulas (core.f95 lines in each test) and, for the normalisations, the
  oracle's zgesdd polar factor.
* Random routines (rand_norm/unnorm_state_site_tensor, random_projector_matrix,
  increase_bandwidth_between, increase_bandwidth_with_environment) draw from
  a device hash, not gfortran's rand(): they are checked by the properties
  their callers rely on -- normalisation, orthonormal columns, rank, the state
  (or the expectation value) unchanged by the bond growth.
* Error paths of optimize_ (core.f95:1962-1983, 2565-2607): info 10 and 11,
  -14 with maxiter = 1, the ncv x 3/2 retries up to the full space, the dense
  strategy-1 fallback when the space is smaller than the restart count, and
  the retry stopping with -14 at the solver's largest basis.
"""

import numpy as np
import pytest

from helpers import crand, hermitian_env, random_sparse_matrices, rel

import oracle.core as oc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from nutcracker_amd import core as c
    return c


def test_contract_matrix_left(core):
    """core.f95:1195-1234: X L X^H."""
    rng = np.random.default_rng(1)
    for bl, br in [(3, 5), (7, 2), (1, 1), (40, 33)]:
        L = crand(rng, bl, bl)
        X = crand(rng, br, bl)
        assert rel(core.contract_matrix_left(bl, br, L, X), X @ L @ X.conj().T) < 1e-13


def test_extend_state_vector_fragment(core):
    """core.f95:1163-1193: new(r, s, l) = sum_m site(r, m, s) old(m, l)."""
    rng = np.random.default_rng(2)
    for bm, br, dl, dr in [(2, 3, 4, 2), (1, 2, 1, 3), (5, 4, 8, 2)]:
        old = crand(rng, bm, dl)
        site = crand(rng, br, bm, dr)
        want = np.einsum("rms,ml->rsl", site, old)
        assert rel(core.extend_state_vector_fragment(bm, br, dl, dr, old, site), want) < 1e-13


def test_construct_exp_boundaries(core):
    """core.f95:3402-3436."""
    rng = np.random.default_rng(3)
    for b, c in [(1, 1), (3, 2), (6, 5)]:
        s = crand(rng, b)
        o = crand(rng, c)
        L = core.construct_left_exp_boundary(b, c, s, o)
        R = core.construct_right_exp_boundary(b, c, s, o)
        assert rel(L, np.einsum("i,j,k->ijk", s, s.conj(), o)) < 1e-15
        assert rel(R, np.einsum("i,j,k->ijk", s.conj(), s, o)) < 1e-15


@pytest.mark.parametrize("br,bl,d", [(4, 3, 2), (6, 9, 3), (16, 8, 4), (2, 1, 5)])
def test_norm_for_left_and_right(core, br, bl, d):
    """core.f95:2880-2960: the polar factor U Vh of the site as (br x bl d), or of
    the transposed site as (bl x br d), equal to the oracle's zgesdd factor."""
    rng = np.random.default_rng(br * 100 + bl * 10 + d)
    site = crand(rng, br, bl, d)
    if bl * d >= br:
        info, out = core.norm_for_left(br, bl, d, site)
        assert info == 0
        _, u, s, vt = oc.mysvd(br, bl * d, site)
        assert rel(out.reshape(br, bl * d, order="F"), u @ vt) < 1e-10
    if br * d >= bl:
        info, out = core.norm_for_right(br, bl, d, site)
        assert info == 0
        A = np.transpose(site, (1, 0, 2)).reshape(bl, br * d, order="F")
        _, u, s, vt = oc.mysvd(bl, br * d, A)
        want = np.transpose((u @ vt).reshape(bl, br, d, order="F"), (1, 0, 2))
        assert rel(out, want) < 1e-10


def test_norm_for_left_refuses_too_few_degrees_of_freedom(core):
    with pytest.raises(RuntimeError, match="degrees of freedom"):
        core.norm_for_left(8, 2, 2, np.ones((8, 2, 2), complex))


@pytest.mark.parametrize("br,bl,d", [(4, 8, 2), (3, 3, 3), (10, 4, 2), (1, 1, 1)])
def test_rand_norm_state_site_tensor(core, br, bl, d):
    a = core.rand_norm_state_site_tensor(br, bl, d)
    b = core.rand_norm_state_site_tensor(br, bl, d)
    # sum_{r,s} psi(r,l,s) conj(psi(r,l',s)) = delta (core.f95:2634-2659)
    G = np.einsum("rls,rms->lm", a, a.conj())
    assert np.abs(G - np.eye(bl)).max() < 1e-12
    if br * bl * d > 1:
        assert rel(a, b) > 1e-3  # a new draw per call, like the reference's rand() stream


def test_rand_unnorm_state_site_tensor(core):
    a = core.rand_unnorm_state_site_tensor(5, 3, 2)
    b = core.rand_unnorm_state_site_tensor(5, 3, 2)
    assert abs(np.linalg.norm(a) - 1) < 1e-14 and abs(np.linalg.norm(b) - 1) < 1e-14
    assert rel(a, b) > 1e-3


@pytest.mark.parametrize("n,p", [(20, 3), (7, 1), (3, 5)])
def test_random_projector_matrix(core, n, p):
    rank, refl, tau, swaps = core.random_projector_matrix(n, p)
    k = min(n, p)
    assert rank == k and list(swaps) == list(range(1, k + 1))
    if rank < n:
        x = crand(np.random.default_rng(n + p), n)
        f = core.filter_components_outside_orthog(n, p, k, n - rank, refl, tau, swaps, x)
        assert core.compute_overlap_with_projectors(p, k, refl, tau, swaps, n, f) < 1e-12 * np.linalg.norm(x)


@pytest.mark.parametrize("bl,bm,br,dl,dr,nb", [(2, 3, 4, 2, 2, 5), (4, 4, 4, 3, 3, 7), (1, 2, 2, 2, 2, 2)])
def test_increase_bandwidth_between(core, bl, bm, br, dl, dr, nb):
    """core.f95:3090-3144 (states.hpp:387-443): the two-site state is unchanged,
    the right site comes back right-normalised."""
    rng = np.random.default_rng(bl + 10 * bm + 100 * nb)
    A = crand(rng, bm, bl, dl)
    B = crand(rng, br, bm, dr)
    info, A2, B2 = core.increase_bandwidth_between(bl, bm, br, dl, dr, nb, A, B)
    assert info == 0
    assert A2.shape == (nb, bl, dl) and B2.shape == (br, nb, dr)
    before = np.einsum("mla,rmb->labr", A, B)
    after = np.einsum("mla,rmb->labr", A2, B2)
    assert rel(after, before) < 1e-12
    G = np.einsum("rls,rms->lm", B2, B2.conj())
    assert np.abs(G - np.eye(nb)).max() < 1e-12 if br * dr >= nb else True


def test_increase_bandwidth_with_environment(core):
    """core.f95:3246-3297: an isometry M with M^H M = 1 cancels between the
    widened site and the widened environments, so the site expectation value
    (core.f95:1045-1088) is unchanged; Hermitian environments stay Hermitian."""
    rng = np.random.default_rng(9)
    b, c, d, nb = 4, 3, 2, 7
    idx, mats = random_sparse_matrices(rng, c, c, d, symmetric=True)
    L = hermitian_env(rng, b, c)
    R = hermitian_env(rng, b, c)
    psi = crand(rng, b, b, d)
    e0 = oc.compute_expectation(b, b, c, c, d, L, psi, idx, mats, R)
    s2, L2, R2 = core.increase_bandwidth_with_environment(b, c, d, nb, psi, L, R)
    e1 = oc.compute_expectation(nb, nb, c, c, d, L2, s2, idx, mats, R2)
    assert abs(e1 - e0) < 1e-12 * abs(e0)
    assert np.abs(L2 - L2.transpose(1, 0, 2).conj()).max() < 1e-12 * np.abs(L2).max()
    assert abs(np.linalg.norm(s2) - np.linalg.norm(psi)) < 1e-12 * np.linalg.norm(psi)


# ---------------------------------------------------------------------------
# optimize_ error paths
# ---------------------------------------------------------------------------

def _hermitian_site(rng, bl, br, d, cl=2, cr=2):
    idx = np.array([[1, 1], [1, 2], [2, 2]], np.uint32).T.copy(order="F") if (cl, cr) == (2, 2) else \
        np.ones((2, 1), np.uint32)
    mats = crand(rng, d, d, idx.shape[1])
    mats = np.asfortranarray(mats + mats.transpose(1, 0, 2).conj())
    return idx, mats, hermitian_env(rng, bl, cl), hermitian_env(rng, br, cr)


def _opt(core, bl, br, cl, cr, d, L, idx, mats, R, guess, tol=1e-12, iterations=10000, nproj=0, nrefl=0, orth=None,
         refl=None, tau=None, swaps=None):
    n = bl * br * d
    return core.optimize(bl, br, cl, cr, d, L, idx, mats, R, nproj, nrefl, n if orth is None else orth, refl, tau,
                         swaps, "SR", tol, iterations, guess)


def test_info_10_empty_orthogonal_space(core):
    rng = np.random.default_rng(11)
    idx, mats, L, R = _hermitian_site(rng, 2, 3, 2)
    info, *_ = _opt(core, 2, 3, 2, 2, 2, L, idx, mats, R, crand(rng, 3, 2, 2), orth=0)
    assert info == 10
    # with projectors: as many reflectors as the space has dimensions
    n = 12
    rank, refl, tau, swaps = core.convert_vectors_to_reflectors(n, n, crand(rng, n, n))
    info, *_ = _opt(core, 2, 3, 2, 2, 2, L, idx, mats, R, crand(rng, 3, 2, 2), nproj=n, nrefl=n, orth=n - rank,
                    refl=refl, tau=tau, swaps=swaps)
    assert info == 10


def test_info_11_guess_inside_projector_space(core):
    rng = np.random.default_rng(12)
    idx, mats, L, R = _hermitian_site(rng, 2, 3, 2)
    n = 12
    V = crand(rng, n, 2)
    rank, refl, tau, swaps = core.convert_vectors_to_reflectors(n, 2, V)
    guess = np.asfortranarray(V[:, 0].conj().reshape(3, 2, 2, order="F"))
    info, *_ = _opt(core, 2, 3, 2, 2, 2, L, idx, mats, R, guess, nproj=2, nrefl=2, orth=n - rank, refl=refl,
                    tau=tau, swaps=swaps)
    assert info == 11
    info, *_ = _opt(core, 2, 3, 2, 2, 2, L, idx, mats, R, crand(rng, 3, 2, 2), orth=n + 1)
    assert info == 5


def test_minus_14_with_one_restart(core):
    """maxiter = 1 at a strategy-3 site: no convergence, and ncv = 3 exceeds the
    restart count, so no retry (core.f95:2588)."""
    rng = np.random.default_rng(13)
    bl, br, d = 12, 12, 2
    idx, mats, L, R = _hermitian_site(rng, bl, br, d)
    s0 = core.optimizer_stats()
    info, it, res, ev, nrm = _opt(core, bl, br, 2, 2, d, L, idx, mats, R, crand(rng, br, bl, d), iterations=1)
    s1 = core.optimizer_stats()
    assert info == -14
    assert s1["first_run_failures"] == s0["first_run_failures"] + 1
    assert s1["ncv_retries"] == s0["ncv_retries"] and s1["dense_fallbacks"] == s0["dense_fallbacks"]
    assert np.abs(res).max() == 0


def _dense_w0(bl, br, cl, cr, d, L, idx, mats, R):
    H = oc.compute_optimization_matrix(bl, br, cl, cr, d, L, idx, mats, R)
    return np.linalg.eigvalsh(0.5 * (H + H.conj().T))[0]


def test_ncv_retries_reach_the_full_space(core):
    """A tolerance no restart meets (1e-300): ncv 3 -> 4 -> 6 -> 8 = the full space
    (n = 8), where the basis spans everything and the Ritz pair is exact."""
    rng = np.random.default_rng(14)
    bl, br, d = 2, 2, 2
    idx, mats, L, R = _hermitian_site(rng, bl, br, d)
    s0 = core.optimizer_stats()
    info, it, res, ev, nrm = _opt(core, bl, br, 2, 2, d, L, idx, mats, R, crand(rng, br, bl, d), tol=1e-300,
                                  iterations=8)
    s1 = core.optimizer_stats()
    assert info == 0
    assert s1["ncv_retries"] - s0["ncv_retries"] == 3
    assert s1["dense_fallbacks"] == s0["dense_fallbacks"]
    w0 = _dense_w0(bl, br, 2, 2, d, L, idx, mats, R)
    assert abs(ev.real - w0) <= 1e-10 * max(1.0, abs(w0))


def test_dense_fallback_when_space_is_smaller_than_restarts(core):
    """n = 8 < maxiter = 20: after -14 the reference solves the dense H_eff
    (strategy 1, core.f95:2569-2587)."""
    rng = np.random.default_rng(15)
    bl, br, d = 2, 2, 2
    idx, mats, L, R = _hermitian_site(rng, bl, br, d)
    s0 = core.optimizer_stats()
    info, it, res, ev, nrm = _opt(core, bl, br, 2, 2, d, L, idx, mats, R, crand(rng, br, bl, d), tol=1e-300,
                                  iterations=20)
    s1 = core.optimizer_stats()
    assert info == 0
    assert s1["dense_fallbacks"] == s0["dense_fallbacks"] + 1
    w0 = _dense_w0(bl, br, 2, 2, d, L, idx, mats, R)
    assert abs(ev.real - w0) <= 1e-10 * max(1.0, abs(w0))
    assert abs(nrm - 1) < 1e-12


def test_ncv_retry_stops_at_the_largest_basis(core):
    """n = 2048 > 62: the retries grow ncv up to 62 and then return -14 (not a
    solver error).  H_eff = R (bl = 1, d = 1, one block) is the 1-D Laplacian
    plus the identity: its lowest eigenvalues are 2e-6 apart in a spread of 4,
    so 62 restarts of any basis up to 62 vectors (under 2000 matvecs) leave the
    Ritz estimate far above 1e-12.  (A problem that does converge to round-off
    ends through the invariant-subspace test, as ARPACK's would, whatever the
    tolerance; so the tolerance alone cannot force -14.)"""
    rng = np.random.default_rng(16)
    bl, br, d, n = 1, 2048, 1, 2048
    T = 3.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    R = np.asfortranarray(T.astype(complex).reshape(n, n, 1))
    L = np.ones((1, 1, 1), complex, order="F")
    idx = np.ones((2, 1), np.uint32, order="F")
    mats = np.ones((1, 1, 1), complex, order="F")
    s0 = core.optimizer_stats()
    info, *_ = _opt(core, bl, br, 1, 1, d, L, idx, mats, R, crand(rng, br, bl, d), tol=1e-12, iterations=62)
    s1 = core.optimizer_stats()
    assert info == -14
    assert s1["ncv_retries"] - s0["ncv_retries"] == 8  # 4, 6, 9, 13, 19, 28, 42, 62
    assert s1["unconverged"] == s0["unconverged"] + 1


def test_optimizer_ncv_option(core):
    """nc_set_optimizer_ncv: a larger basis gives the same eigenvalue."""
    rng = np.random.default_rng(17)
    bl, br, d = 10, 9, 2
    idx, mats, L, R = _hermitian_site(rng, bl, br, d)
    g = crand(rng, br, bl, d)
    w0 = _dense_w0(bl, br, 2, 2, d, L, idx, mats, R)
    prev = core.set_optimizer_ncv(16)
    try:
        info, it, res, ev, nrm = _opt(core, bl, br, 2, 2, d, L, idx, mats, R, g)
    finally:
        core.set_optimizer_ncv(prev)
    assert prev == 3
    assert info == 0 and abs(ev.real - w0) <= 1e-10 * max(1.0, abs(w0))
    with pytest.raises(ValueError):
        core.set_optimizer_ncv(100)


# ---------------------------------------------------------------------------
# physical dimension above 4 (tests/projectors.cpp:178 draws d up to 5)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("d", [1, 5, 6, 9])
def test_any_physical_dimension(core, d):
    rng = np.random.default_rng(20 + d)
    bl, br, cl, cr = 5, 4, 3, 2
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    psi = crand(rng, br, bl, d)
    nl = core.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi)
    assert rel(nl, oc.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi)) < 1e-12
    nr = core.contract_sos_right(bl, br, cl, cr, d, R, idx, mats, psi)
    assert rel(nr, oc.contract_sos_right(bl, br, cl, cr, d, R, idx, mats, psi)) < 1e-12
    H = core.compute_optimization_matrix(bl, br, cl, cr, d, L, idx, mats, R)
    assert rel(H, oc.compute_optimization_matrix(bl, br, cl, cr, d, L, idx, mats, R)) < 1e-12


# ---------------------------------------------------------------------------
# null-reflector swaps that are not the identity
# ---------------------------------------------------------------------------

def _q_from_reflectors(n, refl, tau):
    """Q = H_0 H_1 ... H_{k-1}, H_i = 1 - tau_i v_i v_i^H, from LAPACK storage."""
    Q = np.eye(n, dtype=complex)
    for i in range(len(tau)):
        v = np.zeros(n, complex)
        v[i] = 1
        v[i + 1:] = refl[i + 1:, i]
        Q = Q - tau[i] * np.outer(Q @ v, v.conj())
    return Q


def test_non_identity_swaps_on_the_device(core):
    """The case of tests/test_oracle_projectors.py on the device.  The Q columns
    of the null rows are fixed by entries of size 1e-12 carrying round-off of
    1e-16, so they are determined only to about 1e-4; the device's own
    reflectors define the kept space exactly: its filter is the projector onto
    conj(Q[:, kept]) with the swaps applied (null rows 1 and 3, and 4..n-1), and
    its strategy-1 solve gives H_eff's lowest eigenvalue there (the vector
    projection of operate_on, core.f95:2529-2556).  That space agrees with
    LAPACK's to the conditioning of those columns."""
    import test_oracle_projectors as top
    c = top._swap_case(6)
    w0, B, H = top.swap_case_expected(c)
    n, orth = c["n"], c["orth"]
    rank, refl, tau, swaps = core.convert_vectors_to_reflectors(n, 4, c["V"])
    assert rank == 2 and list(swaps) == [1, 3, 3, 4]
    Q = _q_from_reflectors(n, refl, tau)
    assert np.abs(Q.conj().T @ Q - np.eye(n)).max() < 1e-13
    Bd = np.conj(Q[:, [1, 3] + list(range(4, n))])
    x = crand(np.random.default_rng(1), n)
    f = core.filter_components_outside_orthog(n, 4, 4, orth, refl, tau, swaps, x)
    assert rel(f, Bd @ (Bd.conj().T @ x)) < 1e-12
    # the same space as LAPACK's, to the conditioning of the null-row columns
    assert rel(f, B @ (B.conj().T @ x)) < 1e-3
    wd = np.linalg.eigvalsh(Bd.conj().T @ H @ Bd)[0]
    assert abs(wd - w0) <= 1e-3 * abs(w0)
    g = core.filter_components_outside_orthog(n, 4, 4, orth, refl, tau, swaps, c["guess"])
    info, it, res, ev, nrm = core.optimize(c["bl"], c["br"], c["cl"], c["cr"], c["d"], c["L"], c["idx"], c["mats"],
                                           c["R"], 4, 4, orth, refl, tau, swaps, "SR", 1e-12, 1000,
                                           g.reshape(c["br"], c["bl"], c["d"], order="F"))
    assert info == 0
    assert abs(ev.real - wd) <= 1e-10 * max(1.0, abs(wd))
    r = res.reshape(n, order="F")
    assert rel(Bd @ (Bd.conj().T @ r), r) < 1e-10
