"""Parity of the site optimizer and the site move (HIP path through the C-ABI)
with the oracle (ARPACK / LAPACK) and the reference's known answers."""

import numpy as np
import pytest

from helpers import crand, goldens, hermitian_env, phase_overlap, random_sparse_matrices, rel

import oracle.core as oc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from nutcracker_amd import core as c
    c.device_info()
    return c


def test_optimizer_external_field_kat(core):
    """tests/optimizer.cpp:22-76."""
    rng = np.random.default_rng(3)
    for d in goldens()["optimizer_external_field_one_site"]["physical_dimensions"]:
        diag = np.ones(d, complex)
        diag[-1] = -1
        mats = np.diag(diag).reshape(d, d, 1, order="F")
        guess = crand(rng, 1, 1, d)
        guess /= np.linalg.norm(guess)
        info, it, res, ev, nrm = core.optimize(1, 1, 1, 1, d, np.ones(1), np.array([[1], [1]]), mats, np.ones(1),
                                               0, 0, d, None, None, None, "SR", 0.0, 10000, guess)
        assert info == 0 and it == 0
        assert abs(ev - (-1)) < 1e-12
        assert np.all(np.abs(res.ravel()[:-1]) < 1e-15)
        assert abs(abs(res.ravel()[-1]) - 1) < 1e-12


@pytest.mark.parametrize("case", goldens()["optimizer_modes"]["cases"])
def test_optimizer_modes_kat(core, case):
    """tests/optimizer.cpp:77-136."""
    rng = np.random.default_rng(5)
    mats = np.diag(np.array(case["diag"], complex)).reshape(4, 4, 1, order="F")
    info, it, res, ev, nrm = core.optimize(1, 1, 1, 1, 4, np.ones(1), np.array([[1], [1]]), mats, np.ones(1),
                                           0, 0, 4, None, None, None, case["which"], 0.0, 10000, crand(rng, 4))
    assert info == 0 and it == 0
    assert abs(ev - case["eigenvalue"]) < 1e-12
    assert np.all(np.abs(res.ravel()[:3]) < 1e-15)


OPT_SHAPES = [(2, 2, 3, 2, 3), (2, 6, 7, 2, 3), (3, 5, 5, 3, 2), (2, 3, 4, 3, 3), (2, 16, 16, 3, 3),
              (2, 40, 40, 5, 5), (3, 20, 24, 5, 5), (4, 9, 8, 2, 2)]


@pytest.mark.parametrize("shape", OPT_SHAPES)
@pytest.mark.parametrize("which", ["SR", "LR"])
def test_optimize_random_hermitian(core, shape, which):
    """tests/python/test.py:878-922 -- eigenvalue agrees with ARPACK (oracle) and
    with eigh of the dense H_eff; eigenvector agrees up to a phase."""
    d, bl, br, cl, cr = shape
    rng = np.random.default_rng(1000 + sum(shape))
    print("shape", shape, which, flush=True)
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    guess = crand(rng, br, bl, d)
    n = d * bl * br
    info, it, res, ev, nrm = core.optimize(bl, br, cl, cr, d, L, idx, mats, R, 0, 0, n, None, None, None,
                                           which, 1e-12, 10000, guess)
    oinfo, oit, ores, oev, onrm = oc.optimize(bl, br, cl, cr, d, L, idx, mats, R, 0, 0, n, None, None, None,
                                              which, 1e-12, 10000, guess)
    assert info == 0 and oinfo == 0
    M = oc.compute_optimization_matrix(bl, br, cl, cr, d, L, idx, mats, R)
    w = np.linalg.eigvalsh(M)
    target = w[0] if which == "SR" else w[-1]
    assert abs(ev.real - target) <= 1e-10 * max(1.0, abs(target))
    assert abs(ev.real - oev.real) <= 1e-10 * max(1.0, abs(target))
    assert abs(nrm - 1) < 1e-12
    # eigenvector: same as the oracle's up to a phase (non-degenerate extremal eigenvalue)
    gap = np.min(np.abs(np.delete(w, 0 if which == "SR" else -1) - target))
    if gap > 1e-6 * max(1, abs(target)):
        assert phase_overlap(res, ores) > 1 - 1e-9


def test_optimize_info_codes(core):
    """core.f95:1962-1983: 5 when orth > n, 10 when orth <= 0."""
    rng = np.random.default_rng(9)
    d, bl, br, cl, cr = 2, 3, 3, 2, 2
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    n = d * bl * br
    info, *_ = core.optimize(bl, br, cl, cr, d, L, idx, mats, R, 0, 0, n + 1, None, None, None, "SR", 1e-12, 100,
                             crand(rng, br, bl, d))
    assert info == 5


@pytest.mark.parametrize("dims", [(1, 2, 2, 2, 2), (2, 3, 4, 2, 2), (4, 8, 16, 2, 2), (3, 9, 12, 3, 3),
                                  (16, 32, 40, 2, 2), (64, 128, 130, 2, 3)])
def test_site_move_right(core, dims):
    """norm_denorm_going_right: the polar factor is unique for full-rank inputs,
    so the tensors must equal the oracle's (zgesdd) ones."""
    bl, bm, br, dl, dr = dims
    rng = np.random.default_rng(sum(dims))
    tn = crand(rng, bm, bl, dl)
    td = crand(rng, br, bm, dr)
    info, on, od = core.norm_denorm_going_right(bl, bm, br, dl, dr, tn, td)
    oinfo, oon, ood = oc.norm_denorm_going_right(bl, bm, br, dl, dr, tn, td)
    assert info == 0 and oinfo == 0
    assert rel(on, oon) < 1e-10
    assert rel(od, ood) < 1e-10
    # rows of the normalized site are orthonormal
    A = on.reshape(bm, bl * dl, order="F")
    assert np.abs(A @ A.conj().T - np.eye(bm)).max() < 1e-12
    # the two-site state is unchanged
    before = np.einsum("mla,rmb->labr", tn, td)
    after = np.einsum("mla,rmb->labr", on, od)
    assert rel(after, before) < 1e-12


@pytest.mark.parametrize("dims", [(2, 2, 1, 2, 2), (3, 4, 2, 2, 2), (8, 16, 8, 2, 2), (40, 32, 16, 2, 2),
                                  (130, 128, 64, 3, 2)])
def test_site_move_left(core, dims):
    """norm_denorm_going_left (C-side names bll, bl, br, dl, d)."""
    bll, bl, br, dl, d = dims
    rng = np.random.default_rng(100 + sum(dims))
    td = crand(rng, bl, bll, dl)
    tn = crand(rng, br, bl, d)
    info, od, on = core.norm_denorm_going_left(bll, bl, br, dl, d, td, tn)
    oinfo, ood, oon = oc.norm_denorm_going_left(bll, bl, br, dl, d, td, tn)
    assert info == 0 and oinfo == 0
    assert rel(on, oon) < 1e-10
    assert rel(od, ood) < 1e-10
    before = np.einsum("mla,rmb->labr", td, tn)
    after = np.einsum("mla,rmb->labr", od, on)
    assert rel(after, before) < 1e-12


def _reflectors(rng, n, k):
    vecs = crand(rng, n, k)
    rank, refl, tau, swaps = oc.convert_vectors_to_reflectors(n, k, vecs)
    return vecs, rank, refl, tau, swaps


def test_projector_ops(core):
    rng = np.random.default_rng(21)
    for n, k in [(8, 1), (12, 2), (30, 3), (200, 3)]:
        vecs, rank, refl, tau, swaps = _reflectors(rng, n, k)
        orth = n - rank
        v = crand(rng, n)
        p = core.project_into_orthogonal_space(n, k, rank, orth, refl, tau, swaps, v)
        po = oc.project_into_orthogonal_space(n, k, rank, orth, refl, tau, swaps, v)
        assert rel(p, po) < 1e-12
        u = core.unproject_from_orthogonal_space(n, k, rank, orth, refl, tau, swaps, p)
        uo = oc.unproject_from_orthogonal_space(n, k, rank, orth, refl, tau, swaps, po)
        assert rel(u, uo) < 1e-12
        # the projected vector has no overlap with the projector span
        # the reference's projector convention is bilinear: P^T x = 0 (tests/python/test.py:953)
        assert np.abs(vecs.T @ u).max() < 1e-12 * np.linalg.norm(v)
        ov = core.compute_overlap_with_projectors(k, rank, refl, tau, swaps, n, v)
        assert abs(ov - oc.compute_overlap_with_projectors(k, rank, refl, tau, swaps, n, v)) < 1e-12 * max(1, ov)
        f = core.filter_components_outside_orthog(n, k, rank, orth, refl, tau, swaps, v)
        assert rel(f, oc.filter_components_outside_orthog(n, k, rank, orth, refl, tau, swaps, v)) < 1e-12


def test_optimize_with_projectors(core):
    """tests/python/test.py:924-976 -- the result lies in the orthogonal space and
    its eigenvalue is the smallest of the projected H_eff."""
    rng = np.random.default_rng(31)
    for (d, bl, br, cl, cr, k) in [(2, 3, 4, 2, 2, 2), (2, 5, 6, 3, 3, 3), (3, 4, 4, 2, 3, 1), (2, 1, 2, 1, 2, 2)]:
        L = hermitian_env(rng, bl, cl)
        R = hermitian_env(rng, br, cr)
        idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
        n = d * bl * br
        vecs, rank, refl, tau, swaps = _reflectors(rng, n, k)
        orth = n - rank
        guess = oc.unproject_from_orthogonal_space(n, k, rank, orth, refl, tau, swaps, crand(rng, orth))
        guess /= np.linalg.norm(guess)
        info, it, res, ev, nrm = core.optimize(bl, br, cl, cr, d, L, idx, mats, R, k, rank, orth, refl, tau, swaps,
                                               "SR", 1e-12, 10000, guess)
        oinfo, oit, ores, oev, onrm = oc.optimize(bl, br, cl, cr, d, L, idx, mats, R, k, rank, orth, refl, tau, swaps,
                                                  "SR", 1e-12, 10000, guess)
        assert info == 0 and oinfo == 0
        assert abs(ev.real - oev.real) <= 1e-10 * max(1, abs(oev))
        assert np.abs(vecs.T @ res.ravel(order="F")).max() < 1e-10
        # guess inside the projector space -> 11
        bad = vecs[:, 0] / np.linalg.norm(vecs[:, 0])
        if orth < n:
            info2, *_ = core.optimize(bl, br, cl, cr, d, L, idx, mats, R, k, rank, orth, refl, tau, swaps,
                                      "SR", 1e-12, 10000, bad)
            assert info2 == 11


def test_site_move_rejects_too_few_degrees_of_freedom(core):
    """core.f95:2740-2744 / 2823-2827 stop the process; the C-ABI reports an error instead."""
    rng = np.random.default_rng(0)
    with pytest.raises(RuntimeError, match="degrees of freedom"):
        core.norm_denorm_going_left(8, 16, 4, 2, 2, crand(rng, 16, 8, 2), crand(rng, 4, 16, 2))
    with pytest.raises(RuntimeError, match="degrees of freedom"):
        core.norm_denorm_going_right(2, 16, 4, 2, 2, crand(rng, 16, 2, 2), crand(rng, 4, 16, 2))


@pytest.mark.parametrize("rank", [1, 3])
def test_site_move_rank_deficient(core, rank):
    """A rank-deficient site (a product state, as the external-field chains reach):
    the polar factor is completed with orthonormal rows and the state is kept."""
    rng = np.random.default_rng(40 + rank)
    bl, bm, br, dl, dr = 4, 6, 5, 2, 2
    u = crand(rng, bm, rank)
    v = crand(rng, rank, bl * dl)
    tn = np.asfortranarray((u @ v).reshape(bm, bl, dl, order="F"))
    td = crand(rng, br, bm, dr)
    info, on, od = core.norm_denorm_going_right(bl, bm, br, dl, dr, tn, td)
    assert info == 0
    A = on.reshape(bm, bl * dl, order="F")
    assert np.abs(A @ A.conj().T - np.eye(bm)).max() < 1e-12
    before = np.einsum("mla,rmb->labr", tn, td)
    after = np.einsum("mla,rmb->labr", on, od)
    assert rel(after, before) < 1e-12


@pytest.mark.parametrize("dims", [(2, 4, 5, 2, 2), (4, 8, 3, 2, 2), (3, 6, 4, 2, 3)])
def test_site_move_unitary_input(core, dims):
    """A site that is already normalised (orthonormal rows, also square) keeps
    orthonormal rows and the state -- the case a bond-dimension increase and a
    second move in the same direction produce."""
    bl, bm, br, dl, dr = dims
    assert bm == bl * dl
    rng = np.random.default_rng(sum(dims))
    q, _ = np.linalg.qr(crand(rng, bm, bm))
    tn = np.asfortranarray(q.reshape(bm, bl, dl, order="F"))
    td = crand(rng, br, bm, dr)
    info, on, od = core.norm_denorm_going_right(bl, bm, br, dl, dr, tn, td)
    assert info == 0
    A = on.reshape(bm, bl * dl, order="F")
    assert np.abs(A @ A.conj().T - np.eye(bm)).max() < 1e-12
    before = np.einsum("mla,rmb->labr", tn, td)
    after = np.einsum("mla,rmb->labr", on, od)
    assert rel(after, before) < 1e-12


@pytest.fixture
def svd_only(core):
    """Every polar split completed by rocSOLVER's SVD (Newton-Schulz step limit 0)."""
    import ctypes as C
    h = core.lib()
    f = h.nc_set_polar_max_steps
    f.restype = C.c_int
    f.argtypes = [C.c_int]
    st = h.nc_polar_stats
    st.restype = C.c_int
    st.argtypes = [C.POINTER(C.c_int64)]

    def stats():
        out = (C.c_int64 * 3)()
        assert st(out) == 0
        return list(out)

    prev = f(0)
    assert prev >= 1 and f(-1) == -1 and f(1001) == -1
    try:
        yield stats
    finally:
        f(prev)


@pytest.mark.parametrize("dims", [(2, 3, 4, 2, 2), (16, 32, 40, 2, 2), (64, 128, 130, 2, 3)])
def test_site_move_svd_completion(core, svd_only, dims):
    """The rocSOLVER completion of the polar split (taken when the Newton-Schulz
    iteration does not converge, here forced for every split): on full-rank
    input the tensors equal the oracle's zgesdd ones (core.f95:1-55,
    2708-2878), on both sides, and every split is counted as completed by the
    SVD."""
    bl, bm, br, dl, dr = dims
    rng = np.random.default_rng(700 + sum(dims))
    before_stats = svd_only()
    tn = crand(rng, bm, bl, dl)
    td = crand(rng, br, bm, dr)
    info, on, od = core.norm_denorm_going_right(bl, bm, br, dl, dr, tn, td)
    oinfo, oon, ood = oc.norm_denorm_going_right(bl, bm, br, dl, dr, tn, td)
    assert info == 0 and oinfo == 0
    assert rel(on, oon) < 1e-10 and rel(od, ood) < 1e-10
    # the left move on mirrored shapes
    tdl = crand(rng, bl, br, dl)
    tnl = crand(rng, bm, bl, dr)
    info, od2, on2 = core.norm_denorm_going_left(br, bl, bm, dl, dr, tdl, tnl)
    oinfo, ood2, oon2 = oc.norm_denorm_going_left(br, bl, bm, dl, dr, tdl, tnl)
    assert info == 0 and oinfo == 0
    assert rel(on2, oon2) < 1e-10 and rel(od2, ood2) < 1e-10
    after_stats = svd_only()
    assert after_stats[2] - before_stats[2] == 2


@pytest.mark.parametrize("rank", [1, 3])
def test_site_move_svd_completion_rank_deficient(core, svd_only, rank):
    """The SVD completion on a rank-deficient site: the null directions get an
    orthonormal completion and the two-site state is kept."""
    rng = np.random.default_rng(80 + rank)
    bl, bm, br, dl, dr = 4, 6, 5, 2, 2
    u = crand(rng, bm, rank)
    v = crand(rng, rank, bl * dl)
    tn = np.asfortranarray((u @ v).reshape(bm, bl, dl, order="F"))
    td = crand(rng, br, bm, dr)
    info, on, od = core.norm_denorm_going_right(bl, bm, br, dl, dr, tn, td)
    assert info == 0
    A = on.reshape(bm, bl * dl, order="F")
    assert np.abs(A @ A.conj().T - np.eye(bm)).max() < 1e-12
    before = np.einsum("mla,rmb->labr", tn, td)
    after = np.einsum("mla,rmb->labr", on, od)
    assert rel(after, before) < 1e-12


@pytest.mark.parametrize("ncv", [5, 12, 20, 24, 30])
def test_optimize_basis_sizes(core, ncv):
    """The Krylov kernels by basis size: the fused Gram-Schmidt update-and-dots
    kernel has variants up to 4, 8, 16 and 24 columns (and two reads beyond),
    the one-pass thick restart up to 4, 8, 16 and 24 basis vectors (a GEMM
    beyond), and the deferred second update runs while the next step's fused
    kernel takes its columns.  Every size finds the eigenvalue of the dense
    H_eff (eigh) to 1e-10 and returns a unit eigenvector with a small residual."""
    d, bl, br, cl, cr = 2, 40, 36, 5, 5
    rng = np.random.default_rng(31 + ncv)
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    guess = crand(rng, br, bl, d)
    n = d * bl * br
    prev = core.set_optimizer_ncv(ncv)
    try:
        info, it, res, ev, nrm = core.optimize(bl, br, cl, cr, d, L, idx, mats, R, 0, 0, n, None, None, None,
                                               "SR", 1e-12, 10000, guess)
    finally:
        core.set_optimizer_ncv(prev)
    print("ncv", ncv, "restarts", it, "eigenvalue", ev, flush=True)
    assert info == 0
    M = oc.compute_optimization_matrix(bl, br, cl, cr, d, L, idx, mats, R)
    w = np.linalg.eigvalsh(M)
    assert abs(ev.real - w[0]) <= 1e-10 * max(1.0, abs(w[0]))
    assert abs(nrm - 1) < 1e-12
    x = res.reshape(-1, order="F")
    assert np.linalg.norm(M @ x - w[0] * x) <= 1e-8 * abs(w[0])
