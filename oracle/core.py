"""CPU restatement of Nutcracker's Fortran numeric core (test infrastructure only).

This module is the ORACLE.  It restates, routine by routine, the variational-MPS
inner loop of the reference's ``sources/common/core.f95`` in numpy, and calls the
same third-party routines the reference links: LAPACK ``zgesdd``, ``zheevr``,
``zgeqp3``, ``zunmqr`` (through ``scipy.linalg.lapack``) and ARPACK ``znaupd`` /
``zneupd`` (through scipy's bundled ARPACK, the reverse-communication interface
the reference drives in ``core.f95:2465-2527``).  ``zgemm`` calls become numpy
matrix products on the same column-major reshapes, which numpy hands to OpenBLAS.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker / the timed CPU baseline.
The product (``nutcracker-1_amd``) never imports or calls it.

Conventions (all from core.f95): arrays are complex128 in Fortran (column-major)
order; state site psi(r,l,s) has shape (br,bl,d); environments L(i,j,k) (bl,bl,cl)
and R(r',r,k) (br,br,cr); MPO blocks idx(2,n) hold 1-based (from,to) pairs and
mats(d,d,n) the block matrices W_n(a,b).

Pinning (what checks each routine; see tests/test_oracle_*.py):
* stages 1-3, contract_sos_left/right, compute_expectation,
  contract_expectation_boundaries, contract_vs_left/right: the exact integer
  goldens of tests/boundaries.cpp, plus brute-force einsum over the reference's
  edge-list specs (tests/python/test.py:245-766) and the dense H_eff.
* optimize (strategies 1/2/3): optimizer KATs (tests/optimizer.cpp:22-138),
  projector KATs (tests/python/test.py:978-1126) and scipy.linalg.eigh on the
  dense H_eff (library routine).
* compute_overlap_with_projectors, filter_components_outside_orthog: the closed
  form ||V^T v|| of the overlap, tests/projectors.cpp:195-243 restated (the
  overlap at every site equals |<state 1|state 2>|), and non-identity null
  swaps (tests/test_oracle_projectors.py).
* norm_denorm_going_left/right: polar-factor uniqueness (U Vh unitary rows,
  P U Vh == A with P Hermitian PSD) and state preservation.
* end-to-end: TFIM energies of tests/chain.cpp and tests/stress.cpp, checked
  against the free-fermion closed form, and external-field energies E = -N.
"""

from __future__ import annotations

import numpy as np
from scipy.linalg import lapack
from scipy.sparse.linalg._eigen.arpack import _arpack

CPLX = np.complex128


def _f(a, shape):
    """Fortran-order reshape (what Fortran does when it passes an array)."""
    return np.reshape(a, shape, order="F")


def _fz(shape):
    return np.zeros(shape, dtype=CPLX, order="F")


# ---------------------------------------------------------------------------
# Matvec stages (core.f95:532-621)
# ---------------------------------------------------------------------------

def iteration_stage_1(bl, cl, cr, d, L, idx, mats):
    """core.f95:532-563 -- I(i,a,k2,j,b) = sum_{n: to_n=k2} W_n(a,b) L(i,j,from_n)."""
    L = _f(L, (bl, bl, cl))
    mats = _f(mats, (d, d, -1))
    idx = np.asarray(idx).reshape(2, -1, order="F")
    out = _fz((bl, d, cr, bl, d))
    for n in range(idx.shape[1]):
        k1 = int(idx[0, n]) - 1
        k2 = int(idx[1, n]) - 1
        out[:, :, k2, :, :] += np.einsum("ij,ab->iajb", L[:, :, k1], mats[:, :, n])
    return out


def iteration_stage_2(bl, br, cr, d, I, x):
    """core.f95:565-592 -- zgemm NN br x (cr*bl*d) x (bl*d): T = x * I."""
    T = _f(x, (br, bl * d)) @ _f(I, (bl * d, cr * bl * d))
    return np.asfortranarray(_f(T, (br, cr, bl, d)))


def iteration_stage_3(bl, br, cr, d, T, R):
    """core.f95:594-621 -- zgemm NN br x (bl*d) x (br*cr): y = R * T."""
    y = _f(R, (br, br * cr)) @ _f(T, (br * cr, bl * d))
    return np.asfortranarray(_f(y, (br, bl, d)))


def matvec(bl, br, cl, cr, d, L, idx, mats, R, x, stage1=None):
    """One H_eff application as optimize_strategy_3's operate_on does it without
    projectors (core.f95:2529-2556): dense stage-1 tensor, then two zgemms."""
    if stage1 is None:
        stage1 = iteration_stage_1(bl, cl, cr, d, L, idx, mats)
    T = iteration_stage_2(bl, br, cr, d, stage1, _f(x, (br, bl, d)))
    return iteration_stage_3(bl, br, cr, d, T, R)


# ---------------------------------------------------------------------------
# Environment contractors (core.f95:623-848, 1045-1088, 1142-1161)
# ---------------------------------------------------------------------------

def contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi):
    """core.f95:623-677 -- stage 1, stage 2, zgemm NC (br*cr) x br x (bl*d),
    then reshape with order (1,3,2)."""
    psi = _f(psi, (br, bl, d))
    I = iteration_stage_1(bl, cl, cr, d, L, idx, mats)
    T = iteration_stage_2(bl, br, cr, d, I, psi)
    S3 = _f(T, (br * cr, bl * d)) @ _f(psi, (br, bl * d)).conj().T
    S3 = _f(S3, (br, cr, br))
    return np.asfortranarray(np.transpose(S3, (0, 2, 1)))


def contract_sos_right_stage_1(bl, br, cr, d, R, psi):
    """core.f95:679-708 -- zgemm CN (bl*d) x (br*cr) x br."""
    S1 = _f(psi, (br, bl * d)).conj().T @ _f(R, (br, br * cr))
    return np.asfortranarray(_f(S1, (bl, d, br, cr)))


def contract_sos_right_stage_2a(bl, br, d, W, psi):
    """core.f95:710-734 -- T2a(j,k,i) = sum_a psi(i,j,a) W(a,k)."""
    psi = _f(psi, (br, bl, d))
    return np.asfortranarray(np.einsum("ija,ak->jki", psi, _f(W, (d, d))))


def contract_sos_right_stage_2b(bl, br, d, S1_slice, T2a, N_slice):
    """core.f95:736-764 -- zgemm NT bl x bl x (d*br), accumulating."""
    N_slice += _f(S1_slice, (bl, d * br)) @ _f(T2a, (bl, d * br)).T
    return N_slice


def contract_sos_right_stage_2(bl, br, cl, cr, d, S1, idx, mats, psi):
    """core.f95:766-812 -- per block: stage 2a then stage 2b into N(:,:,from)."""
    S1 = _f(S1, (bl, d, br, cr))
    mats = _f(mats, (d, d, -1))
    idx = np.asarray(idx).reshape(2, -1, order="F")
    N = _fz((bl, bl, cl))
    for n in range(idx.shape[1]):
        T2a = contract_sos_right_stage_2a(bl, br, d, mats[:, :, n], psi)
        k_to = int(idx[1, n]) - 1
        k_from = int(idx[0, n]) - 1
        contract_sos_right_stage_2b(bl, br, d, S1[:, :, :, k_to], T2a, N[:, :, k_from])
    return N


def contract_sos_right(bl, br, cl, cr, d, R, idx, mats, psi):
    """core.f95:814-848."""
    S1 = contract_sos_right_stage_1(bl, br, cr, d, R, psi)
    return contract_sos_right_stage_2(bl, br, cl, cr, d, S1, idx, mats, psi)


def compute_expectation(bl, br, cl, cr, d, L, psi, idx, mats, R):
    """core.f95:1045-1088 -- contract_sos_right then sum N(i,j,k) L(j,i,k)."""
    N = contract_sos_right(bl, br, cl, cr, d, R, idx, mats, psi)
    L = _f(L, (bl, bl, cl))
    return complex(np.sum(N * np.transpose(L, (1, 0, 2))))


def contract_expectation_boundaries(b, c, L, R):
    """core.f95:1142-1161 -- sum(transpose(L,(2,1,3)) * R)."""
    L = _f(L, (b, b, c))
    R = _f(R, (b, b, c))
    return complex(np.sum(np.transpose(L, (1, 0, 2)) * R))


def compute_optimization_matrix(bl, br, cl, cr, d, L, idx, mats, R):
    """core.f95:1090-1140 -- dense H_eff(r1,l2,o2, r2,l1,o1) as an n x n matrix
    (n = br*bl*d) acting on column-major vectors x(r,l,s)."""
    L = _f(L, (bl, bl, cl))
    R = _f(R, (br, br, cr))
    mats = _f(mats, (d, d, -1))
    idx = np.asarray(idx).reshape(2, -1, order="F")
    M = _fz((br, bl, d, br, bl, d))
    for n in range(idx.shape[1]):
        l3 = int(idx[0, n]) - 1
        r3 = int(idx[1, n]) - 1
        # M(r1,l2,o2,r2,l1,o1) += L(l1,l2,l3) W(o1,o2) R(r1,r2,r3)
        M += np.einsum("ab,xy,AB->AbyBax", L[:, :, l3], mats[:, :, n], R[:, :, r3])
    n_full = br * bl * d
    return np.asfortranarray(_f(M, (n_full, n_full)))


# ---------------------------------------------------------------------------
# Overlap (projector) contractors (core.f95:850-1005, 3299-3373)
# ---------------------------------------------------------------------------

def contract_vs_left(blo, bro, bln, brn, d, L, proj_site, state_site):
    """core.f95:850-896."""
    inter = _f(L, (bln, blo)) @ _f(proj_site, (blo, d * bro))
    out = _f(state_site, (brn, bln * d)) @ _f(inter, (bln * d, bro))
    return np.asfortranarray(out)


def contract_vs_right(blo, bro, bln, brn, d, R, proj_site, state_site):
    """core.f95:898-952."""
    inter = _f(proj_site, (blo * d, bro)) @ _f(R, (bro, brn))
    st = np.transpose(_f(state_site, (brn, bln, d)), (2, 0, 1))  # (d, brn, bln)
    out = _f(inter, (blo, d * brn)) @ _f(st, (d * brn, bln))
    return np.asfortranarray(out)


def form_overlap_vector(blo, bro, bln, brn, d, L, R, proj_site):
    """core.f95:954-1005 -- returns the overlap vector as (brn, bln, d)."""
    inter = _f(proj_site, (blo * d, bro)) @ _f(R, (bro, brn))
    t = _f(L, (bln, blo)) @ _f(inter, (blo, d * brn))
    t = _f(t, (bln, d, brn))
    return np.asfortranarray(np.transpose(t, (2, 0, 1)))


def form_overlap_site_tensor(br, bl, d, state_site):
    """core.f95:3299-3313 -- overlap(bl,d,br) = conj(state(br,bl,d)) permuted."""
    s = _f(state_site, (br, bl, d))
    return np.asfortranarray(np.transpose(s.conj(), (1, 2, 0)))


# ---------------------------------------------------------------------------
# Projector algebra (core.f95:464-504, 1266-1664, 1866-1909)
# ---------------------------------------------------------------------------

def swap_inplace(swaps, v):
    """core.f95:464-483."""
    for i in range(len(swaps)):
        s = int(swaps[i]) - 1
        if s != i:
            v[i], v[s] = v[s], v[i]


def unswap_inplace(swaps, v):
    """core.f95:485-504."""
    for i in range(len(swaps) - 1, -1, -1):
        s = int(swaps[i]) - 1
        if s != i:
            v[i], v[s] = v[s], v[i]


def _zunmqr_row(trans, reflectors, nrefl, coeffs, vec):
    """zunmqr('R', trans) applied to the 1 x n row vector ``vec``."""
    n = vec.shape[0]
    if nrefl == 0:
        return vec.copy()
    a = np.asfortranarray(reflectors[:, :nrefl])
    c = np.asfortranarray(vec.reshape(1, n))
    cq, work, info = lapack.zunmqr("R", trans, a, np.asarray(coeffs[:nrefl], CPLX), c, max(1, 64 * n))
    if info != 0:
        raise RuntimeError("zunmqr info=%d" % info)
    return np.asarray(cq).reshape(n).copy()


def project_into_orthogonal_space(n, nproj, nrefl, orth, reflectors, coeffs, swaps, vec):
    """core.f95:1513-1589."""
    vec = np.asarray(vec, CPLX).reshape(n, order="F")
    if nproj == 0:
        return vec.copy()
    start = n - (orth - 1) - 1  # 0-based start of the orthogonal block
    inter = _zunmqr_row("N", _f(reflectors, (n, nproj)), nrefl, coeffs, vec)
    swap_inplace(swaps[:nrefl], inter)
    return inter[start:].copy()


def unproject_from_orthogonal_space(n, nproj, nrefl, orth, reflectors, coeffs, swaps, vec):
    """core.f95:1591-1664."""
    vec = np.asarray(vec, CPLX).reshape(orth, order="F")
    if nproj == 0:
        return vec.copy()
    start = n - (orth - 1) - 1
    full = np.zeros(n, CPLX)
    full[start:] = vec
    unswap_inplace(swaps[:nrefl], full)
    return _zunmqr_row("C", _f(reflectors, (n, nproj)), nrefl, coeffs, full)


def filter_components_outside_orthog(n, nproj, nrefl, orth, reflectors, coeffs, swaps, vec):
    """core.f95:1866-1909."""
    if nproj == 0:
        return np.asarray(vec, CPLX).reshape(n, order="F").copy()
    p = project_into_orthogonal_space(n, nproj, nrefl, orth, reflectors, coeffs, swaps, vec)
    return unproject_from_orthogonal_space(n, nproj, nrefl, orth, reflectors, coeffs, swaps, p)


def compute_overlap_with_projectors(nproj, nrefl, reflectors, coeffs, swaps, n, vec):
    """core.f95:1266-1339 -- norm of the projector-space weights of ``vec``."""
    if nproj == 0:
        return 0.0
    reflectors = _f(reflectors, (n, nproj))
    pv = project_into_orthogonal_space(n, nproj, nrefl, n, reflectors, coeffs, swaps, vec)
    unswap_inplace(swaps[:nrefl], pv[: len(swaps[:nrefl])])
    w = np.zeros(nproj, CPLX)
    w[:nrefl] = pv[:nrefl]
    # ztrmv('U','T','N'): w(1:nrefl) = transpose(upper(R)) * w(1:nrefl)
    U = np.triu(reflectors[:nrefl, :nrefl])
    w[:nrefl] = U.T @ w[:nrefl]
    if nproj > nrefl:
        w[nrefl:] = reflectors[:, nrefl:].T @ pv
    return float(np.linalg.norm(w))


def convert_vectors_to_reflectors(m, n, vectors):
    """core.f95:1341-1435 -- zgeqp3 then null-reflector swaps.

    Returns (rank, reflectors(m,n) overwritten, coefficients, swaps(1-based))."""
    if n == 0:
        return 0, _fz((m, 0)), np.zeros(0, CPLX), np.zeros(0, np.int32)
    a = np.asfortranarray(_f(np.asarray(vectors, CPLX), (m, n)))
    qr, jpvt, tau, work, info = lapack.zgeqp3(a, lwork=max(1, 64 * (n + 1)))
    if info != 0:
        raise RuntimeError("zgeqp3 info=%d" % info)
    k = min(m, n)
    norms = np.array([np.linalg.norm(qr[i, i:n]) for i in range(k)])
    swaps = np.arange(1, k + 1, dtype=np.int32)
    i = 0
    while i < k:
        if norms[i] < 1e-12:
            j = i
            while True:
                j += 1
                if j == k:
                    break
                if norms[j] > 1e-12:
                    break
            if j == k:
                break
            norms[i], norms[j] = norms[j], norms[i]
            swaps[i] = j + 1
        i += 1
    rank = i
    return rank, np.asfortranarray(qr), np.asarray(tau, CPLX), swaps


# ---------------------------------------------------------------------------
# Site move (core.f95:1-55, 2708-2878)
# ---------------------------------------------------------------------------

def mysvd(m, n, A):
    """core.f95:1-55 -- thin SVD with zgesdd('S')."""
    u, s, vt, info = lapack.zgesdd(np.asfortranarray(_f(A, (m, n))), compute_uv=1, full_matrices=0)
    return info, u, s, vt


def norm_denorm_going_left(bl, bm, br, dl, dr, to_denorm, to_norm):
    """core.f95:2708-2788.  ``to_denorm`` is (bm,bl,dl), ``to_norm`` is (br,bm,dr).
    Returns (info, denormalized(bm,bl,dl), normalized(br,bm,dr))."""
    if br * dr < bm:
        raise ValueError("Not enough degrees of freedom to normalize.")
    w = np.transpose(_f(to_norm, (br, bm, dr)), (1, 0, 2))  # (bm, br, dr)
    info, u, s, vt = mysvd(bm, br * dr, _f(w, (bm, br * dr)))
    normalized = np.asfortranarray(np.transpose(_f(u @ vt, (bm, br, dr)), (1, 0, 2)))
    uc = u.conj()
    dw = uc.conj().T @ _f(to_denorm, (bm, bl * dl))
    dw = dw * s[:, None]
    denormalized = np.asfortranarray(_f(uc @ dw, (bm, bl, dl)))
    return info, denormalized, normalized


def norm_denorm_going_right(bl, bm, br, dl, dr, to_norm, to_denorm):
    """core.f95:2790-2878.  ``to_norm`` is (bm,bl,dl), ``to_denorm`` is (br,bm,dr).
    Returns (info, normalized(bm,bl,dl), denormalized(br,bm,dr))."""
    if bl * dl < bm:
        raise ValueError("Not enough degrees of freedom to normalize.")
    info, u, s, vt = mysvd(bm, bl * dl, _f(to_norm, (bm, bl * dl)))
    normalized = np.asfortranarray(_f(u @ vt, (bm, bl, dl)))
    w1 = np.transpose(_f(to_denorm, (br, bm, dr)), (1, 0, 2))  # (bm, br, dr)
    uc = u.conj()
    w2 = uc.conj().T @ _f(w1, (bm, br * dr))
    w2 = w2 * s[:, None]
    w1 = _f(uc @ w2, (bm, br, dr))
    denormalized = np.asfortranarray(np.transpose(w1, (1, 0, 2)))
    return info, normalized, denormalized


# ---------------------------------------------------------------------------
# Eigen-solvers (core.f95:211-462, 1911-2609)
# ---------------------------------------------------------------------------

def lapack_eigenvalue_real_optimizer(n, M, which):
    """core.f95:211-325 -- zheevr range 'I', upper triangle."""
    if which == "SR":
        k = 1
    elif which == "LR":
        k = n
    else:
        raise ValueError("Invalid eigenvalue requested: " + which)
    w, z, m, isuppz, info = lapack.zheevr(np.asfortranarray(M), compute_v=1, range="I", lower=0, il=k, iu=k)
    if info != 0:
        raise RuntimeError("zheevr info=%d" % info)
    return complex(w[0]), np.asarray(z[:, 0], CPLX)


def lapack_eigenvalue_mag_maximizer(n, M):
    """core.f95:327-462 -- largest-magnitude of the two extreme eigenvalues."""
    w1, z1, _, _, info = lapack.zheevr(np.asfortranarray(M), compute_v=1, range="I", lower=0, il=n, iu=n)
    w0, z0, _, _, info2 = lapack.zheevr(np.asfortranarray(M), compute_v=1, range="I", lower=0, il=1, iu=1)
    if info or info2:
        raise RuntimeError("zheevr failed")
    if abs(w1[0]) > abs(w0[0]):
        return complex(w1[0]), np.asarray(z1[:, 0], CPLX)
    return complex(w0[0]), np.asarray(z0[:, 0], CPLX)


def project_matrix_into_orthog_space(n, nproj, nrefl, orth, reflectors, coeffs, swaps, M):
    """core.f95:1666-1762."""
    M = _f(M, (n, n))
    if nproj == 0:
        return np.asfortranarray(M.copy())
    start = n - (orth - 1) - 1
    refl = np.asfortranarray(_f(reflectors, (n, nproj))[:, :nrefl])
    tau = np.asarray(coeffs[:nrefl], CPLX)
    inter = np.asfortranarray(M.conj())
    lw = max(1, 64 * n)
    inter, _, info = lapack.zunmqr("R", "N", refl, tau, inter, lw)
    inter, _, info2 = lapack.zunmqr("L", "C", refl, tau, np.asfortranarray(inter), lw)
    if info or info2:
        raise RuntimeError("zunmqr failed")
    inter = np.array(inter, order="F")
    _swap_matrix_inplace(swaps[:nrefl], inter)
    return np.asfortranarray(inter[start:, start:].conj())


def _swap_matrix_inplace(swaps, M):
    """core.f95:506-530, literally: for each i with swaps(i) /= i, zswap the
    first n entries of columns i and swaps(i) (stride 1), then zswap n entries
    starting at rows i and swaps(i) with strides n and leading_dimension."""
    nsw = len(swaps)
    ld = M.shape[0]
    flat = M.reshape(-1, order="F")
    assert np.shares_memory(flat, M)
    for i in range(nsw):
        s = int(swaps[i]) - 1
        if s == i:
            continue
        for t in range(nsw):
            a, b = i * ld + t, s * ld + t
            flat[a], flat[b] = flat[b], flat[a]
        for t in range(nsw):
            a, b = i + t * nsw, s + t * ld
            flat[a], flat[b] = flat[b], flat[a]


def compute_opt_matrix_all_cases(bl, br, cl, cr, d, L, idx, mats, R, nproj, nrefl, orth, reflectors, coeffs, swaps):
    """core.f95:1764-1864."""
    n = d * br * bl
    if orth > n:
        raise ValueError("orthogonal_subspace_dimension too large")
    M = compute_optimization_matrix(bl, br, cl, cr, d, L, idx, mats, R)
    if nproj == 0:
        return M
    return project_matrix_into_orthog_space(n, nproj, nrefl, orth, reflectors, coeffs, swaps, M)


class _Arpack:
    """The ``run_arpack`` internal procedure of core.f95:2465-2527 (and the
    identical one at 2232-2291), driving ARPACK's znaupd/zneupd reverse
    communication with nev=1, bmat='I', exact shifts, mode 1 and info=1
    (resid = the projected guess)."""

    def __init__(self, op, which, tol):
        self.op = op
        self.which = which
        self.tol = tol
        self.matvecs = 0

    def run(self, n, nev, ncv, guess, max_restarts):
        resid = np.array(guess, CPLX).reshape(n)
        v = np.zeros((n, ncv), CPLX, order="F")
        iparam = np.zeros(11, np.int32)
        ipntr = np.zeros(14, np.int32)
        iparam[0] = 1
        iparam[2] = max_restarts
        iparam[6] = 1
        workd = np.zeros(3 * n, CPLX)
        lworkl = 3 * ncv * ncv + 5 * ncv
        workl = np.zeros(lworkl, CPLX)
        rwork = np.zeros(ncv, np.float64)
        ido = 0
        info = 1
        while True:
            ido, _tol, resid, v, iparam, ipntr, info = _arpack.znaupd(
                ido, "I", self.which, nev, self.tol, resid, v, iparam, ipntr, workd, workl, rwork, info
            )
            if ido in (-1, 1):
                i0 = int(ipntr[0]) - 1
                i1 = int(ipntr[1]) - 1
                workd[i1:i1 + n] = self.op(workd[i0:i0 + n].copy())
                self.matvecs += 1
            else:
                break
        iterations = int(iparam[2])
        nconv = int(iparam[4])
        select = np.zeros(ncv, np.int32)
        workev = np.zeros(3 * ncv, CPLX)
        d, z, info = _arpack.zneupd(
            1, b"A", select, 0j, workev, b"I", self.which, nev, self.tol, resid, v, iparam, ipntr, workd, workl, rwork, info
        )
        if nconv >= 1 and info != -14:
            k = int(np.argmin(np.real(d[:nconv])))
            return info, iterations, complex(d[k]), np.array(z[:, k], CPLX)
        return info, iterations, 0j, np.zeros(n, CPLX)


def _doit(n_full, orth, projected_guess, which, tol, iterations, op, unproject, strategy_1):
    """The ``doit`` procedure of core.f95:2565-2607 (and 2299-2340): ncv = 3;
    on info = -14 either fall back to strategy 1 (when n < iterations, where
    ``iterations`` is ARPACK's returned restart count) or grow ncv by 3/2.
    Returns (info, iterations, eigenvalue, result in full space, matvecs)."""
    ar = _Arpack(op, which, tol)
    nev = 1
    ncv = nev * 2 + 1
    info, iterations, ev, pres = ar.run(orth, nev, ncv, projected_guess, iterations)
    if info == -14:
        if n_full < iterations:
            info, iterations, ev, res = strategy_1()
            return info, iterations, ev, res, ar.matvecs
        while ncv <= iterations:
            ncv = ncv * 3 // 2
            info, iterations, ev, pres = ar.run(orth, nev, ncv, projected_guess, iterations)
            if info != -14:
                return info, iterations, ev, unproject(pres), ar.matvecs
        return info, iterations, ev, np.zeros(n_full, CPLX), ar.matvecs
    return info, iterations, ev, unproject(pres), ar.matvecs


def optimize(bl, br, cl, cr, d, L, idx, mats, R, nproj, nrefl, orth, reflectors, coeffs, swaps,
             which, tol, iterations, guess, stats=None):
    """core.f95:1911-2047 (dispatcher) with strategies 1 (2049-2129),
    2 (2131-2360) and 3 (2362-2609).

    Returns (info, iterations, result(br,bl,d), eigenvalue, normal)."""
    n = d * bl * br
    L = _f(L, (bl, bl, cl))
    R = _f(R, (br, br, cr))
    guess = np.asarray(guess, CPLX).reshape(n, order="F")
    if nproj:
        reflectors = _f(np.asarray(reflectors, CPLX), (n, nproj))
        swaps = np.asarray(swaps, np.int32)
        coeffs = np.asarray(coeffs, CPLX)
    else:
        reflectors = _fz((n, 0))
        swaps = np.zeros(0, np.int32)
        coeffs = np.zeros(0, CPLX)
    zero = np.zeros((br, bl, d), CPLX, order="F")
    if orth > n:
        return 5, iterations, zero, 0j, 0.0
    overlap = compute_overlap_with_projectors(nproj, nrefl, reflectors, coeffs, swaps, n, guess)
    if orth <= 0:
        return 10, iterations, zero, 0j, 0.0
    if orth < n and overlap > 1e-7:
        return 11, iterations, zero, 0j, 0.0

    def unproject(v):
        return unproject_from_orthogonal_space(n, nproj, nrefl, orth, reflectors, coeffs, swaps, v)

    def project(v):
        return project_into_orthogonal_space(n, nproj, nrefl, orth, reflectors, coeffs, swaps, v)

    def strategy_1():
        M = compute_opt_matrix_all_cases(bl, br, cl, cr, d, L, idx, mats, R, nproj, nrefl, orth, reflectors, coeffs, swaps)
        if which in ("SR", "LR"):
            ev, pv = lapack_eigenvalue_real_optimizer(orth, M, which)
        elif which == "LM":
            ev, pv = lapack_eigenvalue_mag_maximizer(orth, M)
        else:
            raise ValueError(which)
        return 0, 0, ev, unproject(pv)

    matvecs = 0
    if orth <= 4:
        info, iters, ev, res = strategy_1()
        used = 1
    elif bl * br < cl * cr:
        M = compute_opt_matrix_all_cases(bl, br, cl, cr, d, L, idx, mats, R, nproj, nrefl, orth, reflectors, coeffs, swaps)
        # zhemv('U') reads only the upper triangle of the Hermitian matrix
        Mu = np.triu(M) + np.triu(M, 1).conj().T

        def op2(v):
            return Mu @ v

        info, iters, ev, res, matvecs = _doit(n, orth, project(guess), which, tol, iterations, op2, unproject, strategy_1)
        used = 2
    else:
        stage1 = iteration_stage_1(bl, cl, cr, d, L, idx, mats)

        def op3(v):
            full = unproject(v)
            y = matvec(bl, br, cl, cr, d, L, idx, mats, R, full, stage1)
            return project(np.asarray(y).reshape(n, order="F"))

        info, iters, ev, res, matvecs = _doit(n, orth, project(guess), which, tol, iterations, op3, unproject, strategy_1)
        used = 3
    result = np.asfortranarray(np.asarray(res, CPLX).reshape((br, bl, d), order="F"))
    normal = float(np.linalg.norm(result))
    if stats is not None:
        stats["strategy"] = used
        stats["matvecs"] = matvecs
    return info, iters, result, ev, normal


# ---------------------------------------------------------------------------
# Overlap-site formation (core.f95:3315-3373)
# ---------------------------------------------------------------------------

def form_norm_overlap_tensors(bl, bm, br, dl, dr, unnorm_1, right_norm_2):
    """core.f95:3315-3373.  Returns (left_norm_overlap_1(bl,dl,bm),
    unnormalized_overlap_1(bl,dl,bm), unnormalized_state_2(br,bm,dr),
    right_norm_overlap_2(bm,dr,br))."""
    info, left_norm_1, unnorm_2 = norm_denorm_going_right(bl, bm, br, dl, dr, unnorm_1, right_norm_2)
    if info != 0:
        raise RuntimeError("Unable to normalize tensor.")
    unnorm_ov_1 = form_overlap_site_tensor(bm, bl, dl, unnorm_1)
    left_norm_ov_1 = form_overlap_site_tensor(bm, bl, dl, left_norm_1)
    right_norm_ov_2 = form_overlap_site_tensor(br, bm, dr, right_norm_2)
    return left_norm_ov_1, unnorm_ov_1, unnorm_2, right_norm_ov_2
