"""The GEMM autotuner: contraction plans keyed on (chi_left, chi_right, d, MPO
bonds, step) are timed over their launch choices on first use.  Whatever it
picks, the matvec must keep the drop-in's accuracy against the oracle's
strategy-3 matvec (core.f95:532-621): 1e-12 relative in complex128, 2e-6 in
complex64 -- and agree with the untuned launch to round-off."""

import ctypes as C
import re

import numpy as np
import pytest

from helpers import crand, hermitian_env, random_sparse_matrices, rel

import oracle.core as oc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from nutcracker_amd import core
    h = core.lib()
    for name in ("nc_bench_heff", "nc_bench_heff_c64"):
        f = getattr(h, name)
        f.restype = C.c_int
        f.argtypes = [C.c_uint32] * 6 + [C.POINTER(C.c_uint32)] + [C.c_void_p] * 5 + \
            [C.c_uint32, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return core, h


def _heff(h, name, bl, br, cl, cr, d, idx, mats, L, R, x):
    y = np.zeros_like(x)
    idx = np.ascontiguousarray(np.asarray(idx, np.uint32).reshape(-1, order="F"))
    mats = np.ascontiguousarray(np.asarray(mats).reshape(-1, order="F"))
    total = C.c_double(0)
    rc = getattr(h, name)(bl, br, cl, cr, d, idx.shape[0] // 2, idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                          mats.ctypes.data, L.ctypes.data, R.ctypes.data, x.ctypes.data, y.ctypes.data, 1, 2,
                          C.byref(total), None)
    assert rc == 0
    return y


# (d, chi_left, chi_right, w_left, w_right): a K-split-prone shape (few output
# tiles, long K), a ragged one, and a spin-1 one -- each GEMM above the 2e8-flop
# tuning bound
SHAPES = [(2, 192, 192, 5, 5), (2, 200, 176, 5, 5), (3, 144, 128, 5, 5)]


@pytest.mark.parametrize("shape", SHAPES)
def test_tuned_matvec_against_oracle(lib, shape):
    core, h = lib
    d, bl, br, cl, cr = shape
    print("shape", shape, flush=True)
    rng = np.random.default_rng(sum(shape) * 104729)
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    x = crand(rng, br, bl, d)
    ref = oc.matvec(bl, br, cl, cr, d, L, idx, mats, R, x, oc.iteration_stage_1(bl, cl, cr, d, L, idx, mats))
    prev = core.set_autotune(True)
    try:
        y64 = _heff(h, "nc_bench_heff", bl, br, cl, cr, d, idx, mats, L, R, x)
        y32 = _heff(h, "nc_bench_heff_c64", bl, br, cl, cr, d, idx, mats, L, R, x)
        core.set_autotune(False)
        u64 = _heff(h, "nc_bench_heff", bl, br, cl, cr, d, idx, mats, L, R, x)
    finally:
        core.set_autotune(prev)
    assert rel(y64, ref) < 1e-12
    assert rel(y64, u64) < 1e-13
    e32 = rel(y32, ref)
    assert 1e-9 < e32 < 2e-6
    table = core.autotune_table()
    mine = [ln for ln in table if re.search(r"chi_l %d chi_r %d d %d w %d/%d chains 1:" % (bl, br, d, cl, cr), ln)]
    print("\n".join(mine), flush=True)
    steps = {(re.search(r"^fp(\d+)", ln).group(1), re.search(r"step (\S)", ln).group(1)) for ln in mine}
    assert {("64", "A"), ("64", "C"), ("32", "A"), ("32", "C")} <= steps
    for ln in mine:
        m = re.search(r"\(([\d.]+) us, default ([\d.]+) us, (\d+) candidates\)", ln)
        assert m and float(m.group(1)) <= float(m.group(2)) and int(m.group(3)) >= 4


VARIANTS = [(16, 1, 0), (16, 2, 1), (32, 4, 0), (32, 1, 1), (-16, 1, 0), (-16, 8, 1), (-16, 0, -1), (6, 1, 1),
            (6, 4, 1)]


@pytest.mark.parametrize("variant", VARIANTS)
def test_every_launch_variant_against_oracle(lib, variant):
    """Each kernel variant the tuner may pick (K step 16 / 32 / 16 with two LDS
    stages, K split, 3M / 4M, the bf16 split kernel for complex64), forced on
    every plan: the matvec and both environment contractors on ragged shapes
    keep the drop-in's accuracy, and the complex64 matvec FP32 accuracy."""
    core, h = lib
    f = h.nc_gemm_force_choice
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int]
    d, bl, br, cl, cr = 2, 100, 72, 5, 5
    rng = np.random.default_rng(4242)
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    x = crand(rng, br, bl, d)
    ref = oc.matvec(bl, br, cl, cr, d, L, idx, mats, R, x, oc.iteration_stage_1(bl, cl, cr, d, L, idx, mats))
    refL = oc.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, x)
    refR = oc.contract_sos_right(bl, br, cl, cr, d, R, idx, mats, x)
    assert f(*variant) == 0
    try:
        y64 = _heff(h, "nc_bench_heff", bl, br, cl, cr, d, idx, mats, L, R, x)
        y32 = _heff(h, "nc_bench_heff_c64", bl, br, cl, cr, d, idx, mats, L, R, x)
        newL = core.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, x)
        newR = core.contract_sos_right(bl, br, cl, cr, d, R, idx, mats, x)
    finally:
        assert f(0, 0, -1) == 0
    print("variant", variant, "fp64", rel(y64, ref), "c64", rel(y32, ref), flush=True)
    assert rel(y64, ref) < 1e-12
    assert 1e-9 < rel(y32, ref) < 2e-6
    assert rel(newL, refL) < 1e-12
    assert rel(newR, refR) < 1e-12
    assert f(24, 0, -1) == -1 and f(16, 3, 0) == -1 and f(7, 1, 1) == -1


# (d, chi_left, chi_right): bonds that are multiples of 64, so that the
# complex128 matvec's plans may run on the LDS-copy pipeline kernel
LDS_COPY_SHAPES = [(2, 128, 192), (3, 64, 128), (2, 256, 256)]
LDS_COPY_VARIANTS = [(8, 1, 1), (8, 1, 0), (9, 1, 1), (9, 1, 0), (8, 2, 1), (9, 4, 1), (8, 8, 0)]


@pytest.mark.parametrize("variant", LDS_COPY_VARIANTS)
@pytest.mark.parametrize("shape", LDS_COPY_SHAPES)
def test_lds_copy_kernel_against_oracle(lib, shape, variant):
    """The complex128 kernel fed by global -> LDS copies (3- or 4-stage ring,
    3M or 4M, K split 1-8), forced on every plan: on shapes it takes (M, N
    multiples of 64, K of 8) the matvec keeps the drop-in's 1e-12 against the
    oracle's strategy-3 matvec and agrees with the register-staged kernel to
    round-off; the environment contractors (transposed operands: the
    register-staged kernel, or this one where op(B) = B^H) stay exact."""
    core, h = lib
    f = h.nc_gemm_force_choice
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int]
    d, bl, br = shape
    cl = cr = 5
    print("shape", shape, "variant", variant, flush=True)
    rng = np.random.default_rng(bl * 7919 + br * 31 + d)
    L = hermitian_env(rng, bl, cl)
    R = hermitian_env(rng, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    x = crand(rng, br, bl, d)
    ref = oc.matvec(bl, br, cl, cr, d, L, idx, mats, R, x, oc.iteration_stage_1(bl, cl, cr, d, L, idx, mats))
    refL = oc.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, x)
    assert f(16, 1, variant[2]) == 0
    try:
        y16 = _heff(h, "nc_bench_heff", bl, br, cl, cr, d, idx, mats, L, R, x)
        assert f(*variant) == 0
        y = _heff(h, "nc_bench_heff", bl, br, cl, cr, d, idx, mats, L, R, x)
        newL = core.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, x)
    finally:
        assert f(0, 0, -1) == 0
    print("rel vs oracle %.3e, vs register-staged kernel %.3e" % (rel(y, ref), rel(y, y16)), flush=True)
    assert rel(y, ref) < 1e-12
    assert rel(y, y16) < 1e-13
    assert rel(newL, refL) < 1e-12


LDS_COPY_POLAR_DIMS = [(64, 128, 64, 2, 2), (128, 192, 96, 2, 3)]
LDS_COPY_POLAR_VARIANTS = [(8, 1, 1), (8, 2, 1), (9, 1, 0), (8, 4, 0)]


@pytest.mark.parametrize("variant", LDS_COPY_POLAR_VARIANTS)
@pytest.mark.parametrize("dims", LDS_COPY_POLAR_DIMS)
def test_lds_copy_conjugate_transposed_b(lib, dims, variant):
    """The LDS-copy kernel with op(B) = B^H (the Newton-Schulz Gram matrices
    X X^H on their upper tiles, and P = A U_p^H of the site move; the update
    X <- H X runs on its untransposed form), forced on every plan:
    norm_denorm_going_right (core.f95:2790-2878) keeps the oracle's zgesdd
    polar factor to 1e-10 and agrees with the register-staged kernels to
    round-off."""
    core, h = lib
    f = h.nc_gemm_force_choice
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int]
    bl, bm, br, dl, dr = dims
    print("dims", dims, "variant", variant, flush=True)
    rng = np.random.default_rng(sum(dims) + 17)
    tn = crand(rng, bm, bl, dl)
    td = crand(rng, br, bm, dr)
    oinfo, oon, ood = oc.norm_denorm_going_right(bl, bm, br, dl, dr, tn, td)
    assert f(16, 1, variant[2]) == 0
    try:
        info16, on16, od16 = core.norm_denorm_going_right(bl, bm, br, dl, dr, tn, td)
        assert f(*variant) == 0
        info, on, od = core.norm_denorm_going_right(bl, bm, br, dl, dr, tn, td)
    finally:
        assert f(0, 0, -1) == 0
    assert oinfo == 0 and info16 == 0 and info == 0
    print("rel vs oracle %.3e / %.3e, vs register-staged %.3e / %.3e"
          % (rel(on, oon), rel(od, ood), rel(on, on16), rel(od, od16)), flush=True)
    assert rel(on, oon) < 1e-10 and rel(od, ood) < 1e-10
    assert rel(on, on16) < 1e-11 and rel(od, od16) < 1e-11
    A = on.reshape(bm, bl * dl, order="F")
    assert np.abs(A @ A.conj().T - np.eye(bm)).max() < 1e-12
