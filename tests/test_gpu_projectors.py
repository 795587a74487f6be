"""Projector formation on the GPU (core.f95:954-1005, 1341-1435, 3299-3373) and
excited levels of the device-resident chain (chain.cpp:60-73, 175-263, 301-310).

Formation routines: against the oracle (LAPACK zgeqp3 for the reflectors) and
the invariants of tests/python/test.py:1778-1808.  Levels: the reference's own
energies -- tests/stress.cpp:107-131 (four lowest TFIM levels, 1e-7),
tests/chain.cpp:448-541 (external-field levels, 1e-12; solveForEigenvalues, 1e-13)
-- with the overlap with the lower levels checked at every site solve
(tests/chain.cpp:451-456, 1e-12)."""

import numpy as np
import pytest

from helpers import crand, goldens, hermitian_env, random_sparse_matrices, rel

import oracle.core as oc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from nutcracker_amd import core as c
    return c


@pytest.fixture(scope="module")
def mods():
    from nutcracker_amd import chain, operators
    return chain, operators


@pytest.mark.parametrize("dims", [(2, 3, 2), (4, 2, 3), (3, 4, 4), (1, 5, 2)])
def test_form_overlap_site_tensor(core, dims):
    """test.py:1778-1784: overlap(bl, d, br) = conj(state) transposed, exactly."""
    br, bl, d = dims
    rng = np.random.default_rng(sum(dims))
    st = crand(rng, br, bl, d)
    ov = core.form_overlap_site_tensor(br, bl, d, st)
    assert np.array_equal(ov, st.transpose(1, 2, 0).conj())
    assert np.array_equal(ov, oc.form_overlap_site_tensor(br, bl, d, st))


@pytest.mark.parametrize("dims", [(2, 3, 4, 5, 2), (3, 3, 2, 2, 3), (5, 4, 6, 7, 2), (1, 1, 2, 3, 4)])
def test_form_overlap_vector(core, dims):
    blo, bro, bln, brn, d = dims
    rng = np.random.default_rng(7 * sum(dims))
    L = crand(rng, bln, blo)
    R = crand(rng, bro, brn)
    P = crand(rng, blo, d, bro)
    v = core.form_overlap_vector(blo, bro, bln, brn, d, L, R, P)
    assert rel(v, oc.form_overlap_vector(blo, bro, bln, brn, d, L, R, P)) < 1e-13
    # brute force: v(r, l, s) = sum L(l, l') P(l', s, r') R(r', r)
    bf = np.einsum("ab,bsc,cr->ras", L, P, R)
    assert rel(v, bf) < 1e-13


@pytest.mark.parametrize("dims", [(2, 3, 2, 2, 3), (3, 2, 4, 2, 2), (4, 4, 4, 3, 2), (1, 2, 3, 2, 2)])
def test_form_norm_overlap_tensors(core, dims):
    """test.py:1786-1807 invariants, and the oracle's tensors (polar uniqueness)."""
    br, bm, bl, dl, dr = dims
    rng = np.random.default_rng(11 * sum(dims))
    u1 = crand(rng, bm, bl, dl)
    r2 = crand(rng, br, bm, dr)
    lo1, uo1, u2, ro2 = core.form_norm_overlap_tensors(bl, bm, br, dl, dr, u1, r2)
    assert np.array_equal(ro2.transpose(2, 0, 1).conj(), r2)
    assert np.array_equal(uo1.transpose(2, 0, 1).conj(), u1)
    ln1 = lo1.transpose(2, 0, 1).conj()
    lhs = np.tensordot(u1, r2, (0, 1))
    rhs = np.tensordot(ln1, u2, (0, 1))
    assert rel(rhs, lhs) < 1e-12
    for got, want in zip((lo1, uo1, u2, ro2), oc.form_norm_overlap_tensors(bl, bm, br, dl, dr, u1, r2)):
        assert rel(got, want) < 1e-10


def _vectors(rng, n, m, kind):
    v = crand(rng, n, m)
    if kind == "dependent" and m >= 3:
        v[:, 2] = v[:, 0] - 0.5j * v[:, 1]
    if kind == "zero" and m >= 2:
        v[:, 1] = 0
    if kind == "units":
        v = np.zeros((n, m), complex)
        for j in range(m):
            v[n - 1 - j, j] = 1.0 if j % 2 == 0 else 1j
    return np.asfortranarray(v)


@pytest.mark.parametrize("n,m,kind", [(12, 1, "random"), (40, 3, "random"), (300, 4, "random"), (3, 4, "random"),
                                      (30, 3, "dependent"), (25, 3, "zero"), (5, 2, "units"), (9, 3, "units")])
def test_convert_vectors_to_reflectors(core, n, m, kind):
    """Against LAPACK zgeqp3 (the oracle): rank, swaps, coefficients and the
    factored matrix; and the projector built from the GPU reflectors removes the
    span of the vectors (bilinear, test.py:953)."""
    rng = np.random.default_rng(n * 31 + m)
    vec = _vectors(rng, n, m, kind)
    rank, refl, tau, swaps = core.convert_vectors_to_reflectors(n, m, vec)
    orank, orefl, otau, oswaps = oc.convert_vectors_to_reflectors(n, m, vec)
    assert rank == orank
    assert list(swaps) == list(oswaps)
    # reflectors past the rank are built from round-off; compare the determined part
    assert np.allclose(tau[:rank], otau[:rank], atol=1e-12, rtol=0)
    assert rel(refl[:, :rank], orefl[:, :rank]) < 1e-12
    assert rel(refl[:rank, rank:], orefl[:rank, rank:]) < 1e-12 if rank < m else True
    if rank > 0 and rank < n:
        k = min(n, m)
        orth = n - rank
        x = crand(rng, n)
        f = core.filter_components_outside_orthog(n, m, k, orth, refl, tau, swaps, x)
        indep = vec[:, [j for j in range(m) if np.linalg.norm(vec[:, j]) > 0]]
        assert np.abs(indep.T @ f).max() < 1e-12 * max(1.0, np.linalg.norm(x))


@pytest.mark.parametrize("complex_unit", [False, True])
def test_orthogonalization_kat_with_device_reflectors(core, complex_unit):
    """test.py:1003-1031 / 1033-1061: H = diag(4, 3, 2, 1) with e_4 (or i e_4)
    projected out gives eigenvalue 2 with eigenvector e_3."""
    L = np.ones((1, 1, 1), complex)
    R = np.ones((1, 1, 1), complex)
    idx = np.ones((2, 1), np.uint32)
    mats = np.diag([4, 3, 2, 1]).astype(complex).reshape(4, 4, 1)
    proj = np.array([0, 0, 0, 1j if complex_unit else 1], complex)
    rank, refl, tau, swaps = core.convert_vectors_to_reflectors(4, 1, proj.reshape(4, 1))
    assert rank == 1
    rng = np.random.default_rng(5)
    guess = core.unproject_from_orthogonal_space(4, 1, 1, 3, refl, tau, swaps, crand(rng, 3))
    guess /= np.linalg.norm(guess)
    info, it, res, ev, nrm = core.optimize(1, 1, 1, 1, 4, L, idx, mats, R, 1, 1, 3, refl, tau, swaps, "SR", 0.0,
                                           10000, guess)
    assert info == 0
    r = res.ravel()
    assert abs(np.vdot(proj, r)) < 1e-12
    r = r / r[2]
    assert np.allclose(r, [0, 0, 1, 0], atol=1e-12)
    assert abs(ev - 2) < 1e-12


# ---- excited levels of the chain ------------------------------------------

def _exact_levels(ops, levels):
    """Dense diagonalisation of the chain's MPO (small N only)."""
    n = len(ops)
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
    H = H[0, 0]
    assert np.allclose(H, H.conj().T)
    return np.sort(np.linalg.eigvalsh(H))[:levels]


@pytest.mark.parametrize("case", goldens()["tfim_four_levels"]["cases"])
def test_tfim_four_levels(mods, case):
    """tests/stress.cpp:107-131: the four lowest levels to 1e-7, at the exact
    bond dimension of the chain (the reference reaches it by growing chi)."""
    chain, ops_mod = mods
    ops = ops_mod.transverse_ising_operator(case["sites"], case["coupling"])
    print("case", case, flush=True)
    # ncv = 8 instead of the reference's 3: the same levels, far fewer restarts on
    # the near-degenerate pairs of the ordered phase
    ch = chain.Chain(ops, 2 ** (case["sites"] // 2), seed=3, rethrow=True, site_tol=1e-10, sweep_tol=1e-9, ncv=8)
    overlaps = []
    energies = []
    for level in range(4):
        if level:
            ch.construct_and_add_projector_from_state()
            ch.randomize(100 + level)
            ch.reset()
            overlaps.append(ch.compute_projector_overlap_at_current_site())
        ch.sweep_until_converged()
        overlaps.append(ch.compute_projector_overlap_at_current_site())
        energies.append(ch.energy)
    print("levels", energies, "max overlap", max(overlaps), flush=True)
    assert np.allclose(energies, case["levels"], atol=1e-7, rtol=0)
    assert max(overlaps) < 1e-12


@pytest.mark.parametrize("d,n,levels", [(2, 1, [0, 1]), (2, 2, [0, 1, 1, 2]), (2, 4, [0, 1, 1, 1]),
                                        (3, 1, [0, 1, 2]), (3, 2, [0, 1, 1, 2, 2]), (3, 3, [0, 1, 1, 1, 2])])
@pytest.mark.parametrize("grow", [True, False])
def test_external_field_levels(mods, d, n, levels, grow):
    """tests/chain.cpp:496-517: external field diag(0, 1, ..., d-1) on every site,
    levels to 1e-12 (the reference loosens the sanity threshold 10x for d = n = 3),
    with the reference's hard-wired ncv = 3, and either its bond-dimension schedule
    (chi0 = 1, grown by one per optimizeChain step) or the exact bond dimension.

    The levels are exactly degenerate multiplets, so a sweep often holds an exact
    eigenvector of some multiplet.  Two behaviours of the reference decide which
    level the chain reaches from there: znaitr keeps a residual that survives the
    second Gram-Schmidt pass as the next Krylov direction (however small; with
    the guess an exact eigenvector that round-off direction is how the solve
    leaves a higher multiplet), and zgesdd completes the null directions of a
    rank-deficient site (right after the bond grows) with low-index basis
    vectors, which for this operator are its low-energy states."""
    chain, ops_mod = mods
    ops = ops_mod.external_field_operator(n, np.diag(np.arange(d)).astype(complex))
    sanity = 1e-11 if (d == 3 and n == 3) else 1e-12
    if grow:
        ch = chain.Chain(ops, 1, seed=5, rethrow=True, sanity=sanity, ncv=3, grow=True)
    else:
        ch = chain.Chain(ops, d ** n, seed=5, rethrow=True, sanity=sanity, ncv=3)
    got = ch.solve_for_multiple_levels(len(levels), seed=9)
    print("levels", got, flush=True)
    assert np.allclose(got, levels, atol=1e-12, rtol=0)


@pytest.mark.parametrize("n", [3, 4, 5])
def test_solve_for_eigenvalues_local_field(mods, n):
    """tests/chain.cpp:528-543: LocalExternalField diag(0, 1) on every site,
    initial bond dimension 3 (grown as optimizeChain does): eigenvalues 0, 1, 1, 1
    to 1e-13."""
    chain, ops_mod = mods
    ops = ops_mod.external_field_operator(n, np.diag([0.0, 1.0]).astype(complex))
    ch = chain.Chain(ops, 3, seed=2, rethrow=True, grow=True)
    got = ch.solve_for_eigenvalues(4, seed=4)
    assert abs(got[0]) < 1e-13
    assert np.allclose(got[1:], 1.0, atol=1e-13, rtol=0)


def test_xxz_levels_match_exact_diagonalisation(mods):
    """Four lowest levels of an 8-site XXZ chain (Delta = 0.8) at the exact bond
    dimension against dense diagonalisation of the same MPO."""
    chain, ops_mod = mods
    ops = ops_mod.xxz_operator(8, 2, 1.0, 0.8)
    exact = _exact_levels(ops, 4)
    ch = chain.Chain(ops, 16, seed=8, rethrow=True)
    got = ch.solve_for_multiple_levels(4, seed=21)
    print("levels", got, "exact", exact, flush=True)
    assert np.allclose(got, exact, atol=1e-9, rtol=0)


def _dense_state(ch):
    """The chain's state as a dense vector (site tensors (br, bl, d))."""
    v = np.ones((1, 1), complex)  # (configurations, right bond)
    for i in range(ch.n):
        A = ch.site(i)  # (br, bl, d)
        v = np.einsum("cl,rls->csr", v, A).reshape(-1, A.shape[0])
    return v[:, 0]


def test_state_overlap_brute_force(mods):
    """computeStateOverlap (states.hpp:349-382) against the dense states; <a|a> = 1."""
    chain, ops_mod = mods
    ops = ops_mod.xxz_operator(6, 2, 1.0, 0.8)
    a = chain.Chain(ops, 4, seed=1, rethrow=True)
    b = chain.Chain(ops, 3, seed=2, rethrow=True)
    b.perform_optimization_sweep()
    va, vb = _dense_state(a), _dense_state(b)
    got = a.compute_state_overlap(b)
    assert abs(got - np.vdot(va, vb)) < 1e-12
    assert abs(a.compute_state_overlap(a) - 1) < 1e-12
    assert abs(b.compute_state_overlap(b) - 1) < 1e-12


CONFIG1_LEVELS = [-20.400217867027, -20.400216436515, -19.373376103501, -19.373374672990]


def test_config1_four_levels(mods):
    """Config 1: H = -sum ZZ - 0.5 sum X, N = 20, chi = 16, site tolerance 1e-12,
    four levels through projectors, against the free-fermion closed form
    (pinned in test_oracle.py).  chi = 16 truncates the chain; the levels agree to
    6e-11 (measured), the bound is 1e-9 absolute."""
    chain, ops_mod = mods
    ops = ops_mod.tfim_zz_x_operator(20, 0.5)
    ch = chain.Chain(ops, 16, seed=20, rethrow=True, ncv=8)
    got = ch.solve_for_multiple_levels(4, seed=40)
    err = np.abs(np.array(got) - CONFIG1_LEVELS)
    print("config 1 levels", got, "abs err", err, "stats", ch.stats(), flush=True)
    assert err.max() < 1e-9


def test_bandwidth_growth_keeps_state(mods):
    """increaseBandwidthDimension (chain.cpp:75-121): the state and its energy are
    unchanged, the bonds follow computeBandwidthDimensionSequence, and sweeping at
    the larger chi lowers the energy."""
    chain, ops_mod = mods
    ops = ops_mod.xxz_operator(8, 2, 1.0, 0.8)
    ch = chain.Chain(ops, 2, seed=4, rethrow=True)
    ch.sweep_until_converged()
    e0 = ch.energy
    v0 = _dense_state(ch)
    ch.increase_bandwidth_dimension(5)
    assert ch.bonds == [1, 2, 4, 5, 5, 5, 4, 2, 1]
    v1 = _dense_state(ch)
    assert abs(abs(np.vdot(v0, v1)) - 1) < 1e-12
    assert np.linalg.norm(v1 - v0) < 1e-10
    assert abs(ch.expectation().real - e0) < 1e-12 * abs(e0)
    ch.sweep_until_converged()
    assert ch.energy < e0 - 1e-6


def test_optimize_chain_grows_to_exact_ground_state(mods):
    """optimizeChain from chi = 1: grows until the energy converges to 1e-12 and
    reaches the dense ground state of an 8-site XXZ chain."""
    chain, ops_mod = mods
    ops = ops_mod.xxz_operator(8, 2, 1.0, 0.8)
    ch = chain.Chain(ops, 1, seed=6, rethrow=True, grow=True)
    ch.optimize_chain()
    exact = _exact_levels(ops, 1)[0]
    print("chi", max(ch.bonds), "E", ch.energy, "exact", exact, flush=True)
    assert abs(ch.energy - exact) < 1e-10
