"""The benchmark configurations at full size on the GPU.

* Config 3 (XXZ Delta = 0.8, chi = 1024) bulk site: the FP64 matvec
  (core.f95:532-621) within 1e-12 of the oracle and the complex64 one within
  2e-6; contract_sos_left/right (623-848) within 1e-12; the polar site move
  (2708-2878) at 1024 x 2048 within 1e-10 of the zgesdd polar factor on
  well-conditioned input, and the two-site state kept on a site with decaying
  singular values.
* Config 2 (XXZ Delta = 1, N = 64, chi = 256, 6 sweeps): per sweep the energy
  does not rise, <H> is real and the centre site has unit norm; at two bulk
  sites the device eigenpair is checked against the oracle's H_eff built from
  the same environments (residual and Rayleigh quotient).
* Config 4 (64 TFIM chains, N = 32, h ~ U[0.5, 1.5], chi in {32, ..., 256}):
  the lockstep batch converges to the free-fermion closed form within 1e-8.
* Config 5 (spin-1 Heisenberg pattern, d = 3, chi = 2048): one bulk-site solve
  in the space orthogonal to three projector vectors, checked by its
  properties (orthogonality, Rayleigh quotient, projected residual), and the
  bytes of the resident set the full run would hold.
"""

import ctypes as C

import numpy as np
import pytest

from helpers import crand, rel, tfim_zz_x_ground_energy

import oracle.core as oc

pytestmark = pytest.mark.gpu


def _herm(rng, b, c):
    e = (rng.standard_normal((b, b, c)) + 1j * rng.standard_normal((b, b, c))) * np.sqrt(0.5)
    return np.asfortranarray(e + e.transpose(1, 0, 2).conj())


@pytest.fixture(scope="module")
def mods():
    from nutcracker_amd import chain, core, operators
    lib = core.lib()
    for name in ("nc_bench_heff", "nc_bench_heff_c64"):
        f = getattr(lib, name)
        f.restype = C.c_int
        f.argtypes = [C.c_uint32] * 6 + [C.POINTER(C.c_uint32)] + [C.c_void_p] * 5 + \
            [C.c_uint32, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return chain, core, operators, lib


def _heff(lib, name, bl, br, op, L, R, x):
    y = np.zeros_like(x)
    idx = np.ascontiguousarray(np.asarray(op.idx, np.uint32).reshape(-1, order="F"))
    mats = np.ascontiguousarray(np.asarray(op.mats).reshape(-1, order="F"))
    total = C.c_double(0)
    rc = getattr(lib, name)(bl, br, op.cl, op.cr, op.d, op.nmat, idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                            mats.ctypes.data, L.ctypes.data, R.ctypes.data, x.ctypes.data, y.ctypes.data, 0, 1,
                            C.byref(total), None)
    assert rc == 0
    return y


@pytest.fixture(scope="module")
def config3_site(mods):
    _, _, operators, _ = mods
    op = operators.xxz_operator(100, 2, 1.0, 0.8)[50]
    chi = 1024
    rng = np.random.default_rng(3)
    L = _herm(rng, chi, op.cl)
    R = _herm(rng, chi, op.cr)
    x = np.asfortranarray((rng.standard_normal((chi, chi, 2)) + 1j * rng.standard_normal((chi, chi, 2))) * 0.5 ** 0.5)
    return op, chi, L, R, x


def test_config3_matvec(mods, config3_site):
    _, _, _, lib = mods
    op, chi, L, R, x = config3_site
    print("config 3 matvec chi=%d" % chi, flush=True)
    ref = oc.matvec(chi, chi, op.cl, op.cr, op.d, L, op.idx, op.mats, R, x)
    y64 = _heff(lib, "nc_bench_heff", chi, chi, op, L, R, x)
    e64 = rel(y64, ref)
    y32 = _heff(lib, "nc_bench_heff_c64", chi, chi, op, L, R, x)
    e32 = rel(y32, ref)
    print("rel err fp64 %.2e complex64 %.2e" % (e64, e32), flush=True)
    assert e64 < 1e-12
    assert 1e-9 < e32 < 2e-6


def test_config3_environments(mods, config3_site):
    _, core, _, _ = mods
    op, chi, L, R, x = config3_site
    print("config 3 environments chi=%d" % chi, flush=True)
    nl = core.contract_sos_left(chi, chi, op.cl, op.cr, op.d, L, op.idx, op.mats, x)
    assert rel(nl, oc.contract_sos_left(chi, chi, op.cl, op.cr, op.d, L, op.idx, op.mats, x)) < 1e-12
    nr = core.contract_sos_right(chi, chi, op.cl, op.cr, op.d, R, op.idx, op.mats, x)
    assert rel(nr, oc.contract_sos_right(chi, chi, op.cl, op.cr, op.d, R, op.idx, op.mats, x)) < 1e-12


def test_config3_site_move(mods, config3_site):
    """norm_denorm_going_right (core.f95:2790-2878) at the config-3 bulk shape."""
    _, core, _, _ = mods
    op, chi, L, R, x = config3_site
    print("config 3 site move 1024 x 2048", flush=True)
    rng = np.random.default_rng(31)
    nb = crand(rng, chi, chi, 2)
    info, on, od = core.norm_denorm_going_right(chi, chi, chi, 2, 2, x, nb)
    oinfo, oon, ood = oc.norm_denorm_going_right(chi, chi, chi, 2, 2, x, nb)
    assert info == 0 and oinfo == 0
    assert rel(on, oon) < 1e-10
    assert rel(od, ood) < 1e-10
    # an MPS-like site: singular values decaying over 14 decades, the last ones zero
    U, _ = np.linalg.qr(crand(rng, chi, chi))
    V, _ = np.linalg.qr(crand(rng, 2 * chi, chi))
    s = 10.0 ** (-14.0 * np.arange(chi) / chi)
    s[-16:] = 0
    A = np.asfortranarray(((U * s) @ V.conj().T).reshape(chi, chi, 2, order="F"))
    info, on, od = core.norm_denorm_going_right(chi, chi, chi, 2, 2, A, nb)
    assert info == 0
    before = np.einsum("mla,rmb->labr", A, nb).reshape(-1)
    after = np.einsum("mla,rmb->labr", on, od).reshape(-1)
    assert rel(after, before) < 1e-12
    Bm = on.reshape(chi, 2 * chi, order="F")
    assert np.abs(Bm @ Bm.conj().T - np.eye(chi)).max() < 1e-10


def test_config2_chain(mods):
    """Config 2: six full sweeps, invariants per sweep, eigenpairs at bulk sites."""
    chain, core, operators, _ = mods
    ops = operators.xxz_operator(64, 2, 1.0, 1.0)
    ch = chain.Chain(ops, 256, seed=2, ncv=16, rethrow=True)
    energies = [ch.energy]
    for sweep in range(6):
        ch.perform_optimization_sweep()
        e = ch.energy
        ex = ch.expectation()
        i = ch.current_site
        nrm = np.linalg.norm(ch.site(i))
        print("sweep %d energy %.12f" % (sweep, e), flush=True)
        assert abs(ex.imag) <= 1e-10 * abs(ex.real)
        assert abs(nrm - 1) < 1e-12
        assert e <= energies[-1] + 1e-10 * abs(e)
        energies.append(e)
    assert energies[-1] < -27.0  # the chi = 256 ground energy is -28.1754...
    # eigenpairs at two bulk sites against the oracle's operator on the same environments
    target = [20, 40]
    for step in range(max(target) + 1):
        if step in target:
            ch.optimize_site()
            i = ch.current_site
            assert i == step
            lam = ch.energy
            x = ch.site(i)
            L, R = ch.environments(i)
            bl, br = ch.bonds[i], ch.bonds[i + 1]
            op = ops[i]
            y = oc.matvec(bl, br, op.cl, op.cr, op.d, L, op.idx, op.mats, R, x)
            xn = np.vdot(x, x).real
            rq = np.vdot(x, y) / xn
            res = np.linalg.norm((y - lam * x).ravel()) / np.sqrt(xn)
            print("site %d lambda %.12f rq %.12f residual %.2e" % (i, lam, rq.real, res), flush=True)
            assert abs(rq - lam) <= 1e-10 * abs(lam)
            assert res <= 1e-9 * abs(lam)
        ch.move(+1)
    ch.close()


def test_config4_batch_closed_form(mods):
    chain, core, operators, _ = mods
    from nutcracker_amd.chain import ChainBatch
    brng = np.random.default_rng(4)
    nb, nsites = 64, 32
    hs = brng.uniform(0.5, 1.5, nb)
    chis = brng.choice([32, 64, 128, 256], nb)
    print("config 4: %d chains" % nb, flush=True)
    cs = [chain.Chain(operators.tfim_zz_x_operator(nsites, float(hs[i])), int(chis[i]), seed=5000 + i, ncv=8,
                      mixed=True, reset=False) for i in range(nb)]
    b = ChainBatch(cs, reset=True)
    b.sweep_until_converged()
    b.close()
    got = np.array([c.energy for c in cs])
    for c in cs:
        c.close()
    exact = np.array([tfim_zz_x_ground_energy(nsites, float(h)) for h in hs])
    err = np.abs(got - exact) / np.abs(exact)
    print("max rel err vs closed form %.2e" % err.max(), flush=True)
    assert err.max() < 1e-8


def config5_resident_bytes(N=64, d=3, w=5, chi=2048, levels=4):
    """Bytes the config-5 run keeps in HBM: every site, every left and right
    environment, and per lower level three overlap-site copies
    (projectors.hpp:205-288) plus its left and right overlap boundaries."""
    from nutcracker_amd.operators import compute_bandwidth_dimension_sequence
    b = compute_bandwidth_dimension_sequence(chi, [d] * N)
    z = 16
    sites = sum(b[i + 1] * b[i] * d for i in range(N)) * z
    envs = 2 * sum(b[i] * b[i] * w for i in range(N + 1)) * z
    per_level = 3 * sites + 2 * sum(b[i] * b[i] for i in range(N + 1)) * z
    return {"sites": sites, "environments": envs, "per_lower_level": per_level,
            "total_4_levels": sites + envs + (levels - 1) * per_level}


def test_config5_projected_site_solve(mods):
    """One chi = 2048, d = 3 bulk site in the space orthogonal to three lower
    levels' overlap vectors.  H_eff is made Hermitian (Hermitian blocks on the
    spin-1 Heisenberg from/to pattern, Hermitian environments)."""
    chain, core, operators, lib = mods
    rb = config5_resident_bytes()
    print("config 5 resident set: %.1f GB (sites %.1f, environments %.1f, per lower level %.1f)" % (
        rb["total_4_levels"] / 1e9, rb["sites"] / 1e9, rb["environments"] / 1e9, rb["per_lower_level"] / 1e9),
        flush=True)
    assert rb["total_4_levels"] < 288e9
    op0 = operators.xxz_operator(64, 3, 1.0, 1.0)[32]
    mats_h = np.asfortranarray(op0.mats + op0.mats.transpose(1, 0, 2).conj())
    chi, d = 2048, 3
    n = chi * chi * d
    rng = np.random.default_rng(5)
    L = _herm(rng, chi, op0.cl)
    R = _herm(rng, chi, op0.cr)
    V = np.asfortranarray((rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))) * 0.5 ** 0.5)
    rank, refl, tau, swaps = core.convert_vectors_to_reflectors(n, 3, V)
    assert rank == 3
    orth = n - 3
    g = core.filter_components_outside_orthog(n, 3, 3, orth, refl, tau, swaps, crand(rng, n))
    prev = core.set_optimizer_ncv(16)
    try:
        print("config 5 solve n=%d" % n, flush=True)
        info, it, res, ev, nrm = core.optimize(chi, chi, op0.cl, op0.cr, d, L, op0.idx, mats_h, R, 3, 3, orth, refl,
                                               tau, swaps, "SR", 1e-10, 10000, g.reshape(chi, chi, d, order="F"))
    finally:
        core.set_optimizer_ncv(prev)
    assert info == 0
    r = res.reshape(n, order="F")
    rn = np.linalg.norm(r)
    assert abs(rn - nrm) < 1e-12 * rn
    # orthogonal to the lower levels (bilinear, tests/python/test.py:953)
    assert np.abs(V.T @ r).max() <= 1e-12 * rn * np.linalg.norm(V, axis=0).max()
    # Rayleigh quotient and projected residual with the device matvec
    class Op:
        pass
    op = Op()
    op.idx, op.mats, op.cl, op.cr, op.d, op.nmat = op0.idx, mats_h, op0.cl, op0.cr, d, op0.nmat
    y = _heff(lib, "nc_bench_heff", chi, chi, op, L, R, res).reshape(n, order="F")
    rq = np.vdot(r, y) / rn ** 2
    py = core.filter_components_outside_orthog(n, 3, 3, orth, refl, tau, swaps, y)
    resid = np.linalg.norm(py - ev * r) / rn
    print("eigenvalue %.10f rq %.10f projected residual %.2e" % (ev.real, rq.real, resid), flush=True)
    assert abs(rq - ev) <= 1e-9 * abs(ev)
    assert resid <= 1e-7 * abs(ev)
