"""The mixed form's matvec on 16-bit planes: operands split once (L and R per
site solve, x per matvec, Q by the mix kernel) and a GEMM that only moves
planes into LDS and runs MFMAs (nutcracker-1_amd/csrc/planes.hip), in the
default fp16 form (each matrix scaled by a power of two, two fp16 parts, three
partial products) unless a test selects the bf16 form (three parts, six
products).

Matvec: against the oracle's FP64 strategy-3 matvec (core.f95:532-621) within
the complex64 tolerance of 2e-6 relative that the complex64 GEMM path meets
(both part forms give FP32 accuracy per product -- dropped terms below 2^-22 --
and accumulate in FP32 over K <= a few thousand terms); at the chi = 1024
bulk-site shape against the FP64 GPU matvec, which the oracle pins to 1e-12.
The ragged shapes cover K padding (K not a multiple of 32), row padding (rows
not a multiple of 128), K split over several workgroups (nc_planes_splits > 1)
and every branch of the mix kernel (d = 1, 2, 3 and d = 5 in passes of two).
Chains: mixed chains forced onto planes converge to the plain FP64 chain's
energy within the drop-in bar of 1e-10 relative, with and without projectors.
"""

import ctypes as C

import numpy as np
import pytest

from helpers import crand, hermitian_env, random_sparse_matrices, rel

import oracle.core as oc

pytestmark = pytest.mark.gpu

_SIG = [C.c_uint32] * 6 + [C.POINTER(C.c_uint32)] + [C.c_void_p] * 5 + \
    [C.c_uint32, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]


@pytest.fixture(scope="module")
def mods():
    from nutcracker_amd import chain, core, operators
    lib = core.lib()
    for name in ("nc_bench_heff", "nc_bench_heff_planes"):
        f = getattr(lib, name)
        f.restype = C.c_int
        f.argtypes = _SIG
    lib.nc_planes_splits.restype = C.c_int
    lib.nc_planes_splits.argtypes = [C.c_uint32] * 5 + [C.POINTER(C.c_int32)]
    return chain, core, operators, lib


def _heff(lib, name, bl, br, cl, cr, d, idx, mats, L, R, x, phases=None):
    y = np.zeros_like(x)
    idx = np.ascontiguousarray(np.asarray(idx, np.uint32).reshape(-1, order="F"))
    mats = np.ascontiguousarray(np.asarray(mats).reshape(-1, order="F"))
    total = C.c_double(0)
    ph = None if phases is None else (C.c_double * phases)()
    rc = getattr(lib, name)(bl, br, cl, cr, d, idx.shape[0] // 2, idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                            mats.ctypes.data, L.ctypes.data, R.ctypes.data, x.ctypes.data, y.ctypes.data, 1, 2,
                            C.byref(total), ph)
    assert rc == 0
    return y


def _splits(lib, bl, br, d, idx):
    idx = np.asarray(idx)
    out = (C.c_int32 * 2)()
    assert lib.nc_planes_splits(bl, br, d, len(set(idx[0])), len(set(idx[1])), out) == 0
    return out[0], out[1]


SHAPES = [(2, 16, 16, 3, 3), (3, 24, 40, 5, 4), (2, 96, 64, 5, 5), (4, 33, 17, 2, 6), (1, 40, 56, 3, 3),
          (5, 20, 36, 3, 3), (2, 160, 200, 4, 4), (3, 130, 129, 3, 3)]


@pytest.mark.parametrize("shape", SHAPES)
def test_planes_matvec_against_oracle(mods, shape):
    _, _, _, lib = mods
    d, bl, br, cl, cr = shape
    rng = np.random.default_rng(sum(shape) * 104729)
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    x = crand(rng, br, bl, d)
    print("shape", shape, "K slices (A, C)", _splits(lib, bl, br, d, idx), flush=True)
    ref = oc.matvec(bl, br, cl, cr, d, L, idx, mats, R, x, oc.iteration_stage_1(bl, cl, cr, d, L, idx, mats))
    y = _heff(lib, "nc_bench_heff_planes", bl, br, cl, cr, d, idx, mats, L, R, x, phases=5)
    e = rel(y, ref)
    print("planes rel err", e, flush=True)
    assert 1e-9 < e < 2e-6  # complex64 arithmetic, within its tolerance


def test_planes_cover_k_splits(mods):
    """At least one of the ragged shapes splits K over workgroups in each GEMM."""
    _, _, _, lib = mods
    seen_a = seen_c = False
    for d, bl, br, cl, cr in SHAPES:
        idx, _ = random_sparse_matrices(np.random.default_rng(sum((d, bl, br, cl, cr)) * 104729), cl, cr, d)
        sa, sc = _splits(lib, bl, br, d, idx)
        seen_a |= sa > 1
        seen_c |= sc > 1
    assert seen_a and seen_c


def test_planes_matvec_bulk_site_chi1024(mods):
    """The chi = 1024 bulk site of the XXZ chain (d = 2, 5 x 5 MPO bonds): the
    plane matvec against the FP64 GPU matvec within 2e-6 relative."""
    _, _, ops_mod, lib = mods
    op = ops_mod.xxz_operator(8, 2, 1.0, 0.8)[3]
    d, cl, cr = op.d, op.cl, op.cr
    bl = br = 1024
    rng = np.random.default_rng(1024)
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    x = crand(rng, br, bl, d)
    idx, mats = op.idx, op.mats
    print("K slices (A, C)", _splits(lib, bl, br, d, idx), flush=True)
    y64 = _heff(lib, "nc_bench_heff", bl, br, cl, cr, d, idx, mats, L, R, x)
    yp = _heff(lib, "nc_bench_heff_planes", bl, br, cl, cr, d, idx, mats, L, R, x, phases=5)
    e = rel(yp, y64)
    print("planes vs FP64 rel err", e, flush=True)
    assert 1e-9 < e < 2e-6


def _model(ops_mod, model):
    if model == "tfim":
        return ops_mod.tfim_zz_x_operator(12, 1.5), 16
    if model == "xxz":
        return ops_mod.xxz_operator(12, 2, 1.0, 0.8), 24
    return ops_mod.xxz_operator(8, 3, 1.0, 1.0), 12


@pytest.mark.parametrize("model", ["tfim", "xxz", "spin1"])
def test_mixed_chain_on_planes(mods, model):
    chain, core, ops_mod, _ = mods
    ops, chi = _model(ops_mod, model)
    print("model", model, flush=True)
    a = chain.Chain(ops, chi, seed=5, rethrow=True)
    a.sweep_until_converged()
    prev = core.set_mixed_planes(2)
    try:
        b = chain.Chain(ops, chi, seed=5, rethrow=True, mixed=True)
        b.sweep_until_converged()
    finally:
        core.set_mixed_planes(prev)
    de = abs(a.energy - b.energy)
    print("E64", a.energy, "Emixed(planes)", b.energy, "dE", de, b.stats(), flush=True)
    assert de <= 1e-10 * abs(a.energy)
    assert b.stats()["matvecs32"] > 0


def test_mixed_levels_on_planes(mods):
    """Two TFIM levels through projectors with the complex64 matvec on planes:
    the same levels as the FP64 chain within 1e-10 relative."""
    chain, core, ops_mod, _ = mods
    ops = ops_mod.tfim_zz_x_operator(10, 1.5)
    a = chain.Chain(ops, 16, seed=3, rethrow=True, ncv=8)
    ea = a.solve_for_multiple_levels(2, seed=3)
    prev = core.set_mixed_planes(2)
    try:
        b = chain.Chain(ops, 16, seed=3, rethrow=True, ncv=8, mixed=True)
        eb = b.solve_for_multiple_levels(2, seed=3)
    finally:
        core.set_mixed_planes(prev)
    print("E64", ea, "Emixed(planes)", eb, flush=True)
    for u, v in zip(ea, eb):
        assert abs(u - v) <= 1e-10 * abs(u)


def test_mixed_planes_mode_setter(mods):
    _, core, _, _ = mods
    prev = core.set_mixed_planes(0)
    try:
        assert core.set_mixed_planes(2) == 0
        assert core.set_mixed_planes(1) == 2
        with pytest.raises(ValueError):
            core.set_mixed_planes(3)
    finally:
        core.set_mixed_planes(prev)


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("shape", [SHAPES[1], SHAPES[3], SHAPES[6], SHAPES[7]])
def test_plane_forms_against_oracle(mods, form, shape):
    """Both part forms (0: three bf16 parts, six products; 1: power-of-two
    scaled matrices in two fp16 parts, three products) within the complex64
    tolerance of the oracle's FP64 matvec, on ragged shapes whose environments
    span six decades across the MPO components."""
    _, core, _, lib = mods
    d, bl, br, cl, cr = shape
    rng = np.random.default_rng(sum(shape) * 7919 + form)
    L = hermitian_env(rng, bl, cl) * np.logspace(-3, 3, cl)[None, None, :]
    R = hermitian_env(rng, br, cr) * np.logspace(2, -4, cr)[None, None, :]
    L, R = np.asfortranarray(L), np.asfortranarray(R)
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    x = crand(rng, br, bl, d) * 1e-5
    prev = core.set_plane_form(form)
    try:
        y = _heff(lib, "nc_bench_heff_planes", bl, br, cl, cr, d, idx, mats, L, R, x, phases=5)
    finally:
        core.set_plane_form(prev)
    ref = oc.matvec(bl, br, cl, cr, d, L, idx, mats, R, x, oc.iteration_stage_1(bl, cl, cr, d, L, idx, mats))
    e = rel(y, ref)
    print("form", form, "shape", shape, "rel err", e, flush=True)
    assert np.all(np.isfinite(y))
    assert 1e-9 < e < 2e-6


def test_fp16_planes_zero_inputs(mods):
    """fp16 planes: an environment component that is zero (a scale of 1 and a
    zero column norm in the bound of Q) and a zero x (every bound zero) give
    the oracle's result and zero, with no overflow."""
    _, core, _, lib = mods
    d, bl, br, cl, cr = 2, 96, 64, 5, 5
    rng = np.random.default_rng(4242)
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    L[:, :, 1] = 0
    R[:, :, 2] = 0
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    x = crand(rng, br, bl, d)
    prev = core.set_plane_form(1)
    try:
        y = _heff(lib, "nc_bench_heff_planes", bl, br, cl, cr, d, idx, mats, L, R, x)
        y0 = _heff(lib, "nc_bench_heff_planes", bl, br, cl, cr, d, idx, mats, L, R, np.zeros_like(x))
    finally:
        core.set_plane_form(prev)
    ref = oc.matvec(bl, br, cl, cr, d, L, idx, mats, R, x, oc.iteration_stage_1(bl, cl, cr, d, L, idx, mats))
    e = rel(y, ref)
    print("zero component rel err", e, flush=True)
    assert e < 2e-6
    assert np.all(y0 == 0)


def test_plane_form_setter(mods):
    _, core, _, _ = mods
    prev = core.set_plane_form(0)
    try:
        assert core.set_plane_form(1) == 0
        assert core.set_plane_form(1) == 1
        with pytest.raises(ValueError):
            core.set_plane_form(2)
    finally:
        core.set_plane_form(prev)
