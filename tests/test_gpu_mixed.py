"""The mixed-precision form: complex64 matvecs (and optionally environment
updates) on FP32 matrix cores, FP64 Krylov vectors, Rayleigh quotient and site
moves.

Matvec: against the oracle's FP64 strategy-3 matvec (core.f95:532-621) within a
complex64 tolerance of 2e-6 relative (FP32 unit roundoff 6e-8 times the growth
of a sum over K <= 4 * 96 terms).  Chains: against the plain FP64 GPU chain on
the same MPO and seeded initial state -- within 1e-10 relative (the drop-in's
own bar) when the environments stay FP64, within 1e-5 relative when they are
built in complex64 too."""

import ctypes as C

import numpy as np
import pytest

from helpers import crand, hermitian_env, random_sparse_matrices, rel

import oracle.core as oc

pytestmark = pytest.mark.gpu


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


def _heff(lib, name, bl, br, cl, cr, d, idx, mats, L, R, x):
    y = np.zeros_like(x)
    idx = np.ascontiguousarray(np.asarray(idx, np.uint32).reshape(-1, order="F"))
    mats = np.ascontiguousarray(np.asarray(mats).reshape(-1, order="F"))
    total = C.c_double(0)
    rc = getattr(lib, name)(bl, br, cl, cr, d, idx.shape[0] // 2, idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                            mats.ctypes.data, L.ctypes.data, R.ctypes.data, x.ctypes.data, y.ctypes.data, 0, 1,
                            C.byref(total), None)
    assert rc == 0
    return y


@pytest.mark.parametrize("shape", [(2, 16, 16, 3, 3), (3, 24, 40, 5, 4), (2, 96, 64, 5, 5), (4, 33, 17, 2, 6)])
def test_complex64_matvec_against_oracle(mods, shape):
    _, _, _, lib = mods
    d, bl, br, cl, cr = shape
    print("shape", shape, flush=True)
    rng = np.random.default_rng(sum(shape) * 7919)
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    x = crand(rng, br, bl, d)
    ref = oc.matvec(bl, br, cl, cr, d, L, idx, mats, R, x, oc.iteration_stage_1(bl, cl, cr, d, L, idx, mats))
    y64 = _heff(lib, "nc_bench_heff", bl, br, cl, cr, d, idx, mats, L, R, x)
    y32 = _heff(lib, "nc_bench_heff_c64", bl, br, cl, cr, d, idx, mats, L, R, x)
    assert rel(y64, ref) < 1e-12
    e32 = rel(y32, ref)
    print("complex64 rel err", e32, flush=True)
    assert 1e-9 < e32 < 2e-6  # really complex64 arithmetic, and within its tolerance


def _pair(chain, ops, chi, seed, **mixed):
    a = chain.Chain(ops, chi, seed=seed, rethrow=True)
    b = chain.Chain(ops, chi, seed=seed, rethrow=True, mixed=True, **mixed)
    return a, b


def _model(ops_mod, model):
    if model == "tfim":  # gapped (paramagnetic) side, so that ncv = 3 converges quickly
        return ops_mod.tfim_zz_x_operator(12, 1.5), 16
    if model == "xxz":
        return ops_mod.xxz_operator(12, 2, 1.0, 0.8), 24
    return ops_mod.xxz_operator(8, 3, 1.0, 1.0), 12


@pytest.mark.parametrize("model", ["tfim", "xxz", "spin1"])
def test_mixed_matvecs_meet_drop_in_bar(mods, model):
    """complex64 matvecs, FP64 environments: the FP64 Rayleigh quotient is
    quadratic in the eigenvector error, so the converged energy keeps the drop-in
    bar of 1e-10 relative to the plain FP64 chain."""
    chain, _, ops_mod, _ = mods
    ops, chi = _model(ops_mod, model)
    print("model", model, flush=True)
    a, b = _pair(chain, ops, chi, 5)
    assert abs(a.energy - b.energy) <= 1e-12 * abs(a.energy)
    a.sweep_until_converged()
    b.sweep_until_converged()
    de = abs(a.energy - b.energy)
    print("E64", a.energy, "Emixed", b.energy, "dE", de, "stats", b.stats(), flush=True)
    assert de <= 1e-10 * abs(a.energy)
    assert b.stats()["matvecs32"] > 0
    assert a.stats()["matvecs32"] == 0


@pytest.mark.parametrize("model", ["tfim", "xxz"])
def test_complex64_environments(mods, model):
    """complex64 environment updates as well: the energy carries their FP32
    round-off at first order (accumulated over the chain), so the bound is 1e-5
    relative -- this mode cannot reach 1e-8 per site."""
    chain, _, ops_mod, _ = mods
    ops, chi = _model(ops_mod, model)
    a, b = _pair(chain, ops, chi, 5, env32=True)
    a.sweep_until_converged()
    b.sweep_until_converged()
    de = abs(a.energy - b.energy)
    print("E64", a.energy, "Emixed-env32", b.energy, "dE", de, "dE/site", de / len(ops), flush=True)
    assert de <= 1e-5 * abs(a.energy)


def test_mixed_with_fp64_polish(mods):
    chain, _, ops_mod, _ = mods
    ops = ops_mod.xxz_operator(10, 2, 1.0, 0.8)
    a, b = _pair(chain, ops, 16, 9, polish=True)
    a.sweep_until_converged()
    b.sweep_until_converged()
    print("E64", a.energy, "Emixed+polish", b.energy, b.stats(), flush=True)
    assert abs(a.energy - b.energy) <= 1e-10 * abs(a.energy)


@pytest.mark.parametrize("model", ["xxz", "spin1"])
def test_mixed_chain_on_bf16_split_kernel(mods, model):
    """The complex64 contractions forced onto the bf16 matrix-core kernel (each
    FP32 operand split into three bf16 parts): the mixed chain still converges
    to the plain FP64 chain's energy within the drop-in bar of 1e-10."""
    chain, _, ops_mod, lib = mods
    f = lib.nc_gemm_force_choice
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int]
    ops, chi = _model(ops_mod, model)
    print("model", model, flush=True)
    a = chain.Chain(ops, chi, seed=7, rethrow=True)
    a.sweep_until_converged()
    assert f(6, 1, 1) == 0
    try:
        b = chain.Chain(ops, chi, seed=7, rethrow=True, mixed=True)
        b.sweep_until_converged()
    finally:
        assert f(0, 0, -1) == 0
    de = abs(a.energy - b.energy)
    print("E64", a.energy, "Emixed(bf16 split)", b.energy, "dE", de, flush=True)
    assert de <= 1e-10 * abs(a.energy)
    assert b.stats()["matvecs32"] > 0
