"""nc_contract_sos_left_noalloc (the allocation-free variant of
contract_sos_left_, core.f95:623-677) against contract_sos_left_ on the same
inputs: two calls in a row, the second with a smaller input, both within 1e-13
relative; an input above the scratch maximum is refused with -90.  Device
buffers come from the HIP runtime directly (the test allocates, the entry point
does not)."""

import ctypes as C

import numpy as np
import pytest

from helpers import crand, hermitian_env, rel

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, "%s failed: %d" % (what, rc)

    def upload(self, a):
        a = np.asfortranarray(a)
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), max(16, a.nbytes)), "hipMalloc")
        self.bufs.append(p)
        self.check(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), max(16, nbytes)), "hipMalloc")
        self.bufs.append(p)
        return p

    def download(self, p, shape):
        out = np.zeros(shape, np.complex128, order="F")
        self.check(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self.check(self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


@pytest.fixture(scope="module")
def lib():
    from nutcracker_amd import core
    h = core.lib()
    f = h.nc_contract_sos_left_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_uint32] * 5 + [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]
    return core, h


def _call(h, hip, bl, br, op, L, psi):
    dL, dpsi = hip.upload(L), hip.upload(psi)
    dout = hip.empty(br * br * op.cr * 16)
    idx = np.ascontiguousarray(np.asarray(op.idx, np.uint32).reshape(-1, order="F"))
    mats = np.ascontiguousarray(np.asarray(op.mats, np.complex128).reshape(-1, order="F"))
    rc = h.nc_contract_sos_left_noalloc(bl, br, op.cl, op.cr, op.d, dL, op.nmat, idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                                        mats.ctypes.data, dpsi, dout, None)
    return rc, (hip.download(dout, (br, br, op.cr)) if rc == 0 else None)


def test_noalloc_matches_contract_sos_left(lib):
    core, h = lib
    from nutcracker_amd.operators import xxz_operator
    hip = Hip()
    try:
        op = xxz_operator(6, 2, 1.0, 0.8)[2]
        rng = np.random.default_rng(71)
        for bl, br in [(200, 192), (33, 17)]:
            print("noalloc bl %d br %d" % (bl, br), flush=True)
            L = hermitian_env(rng, bl, op.cl)
            psi = crand(rng, br, bl, op.d)
            rc, got = _call(h, hip, bl, br, op, L, psi)
            assert rc == 0, core.last_error()
            want = core.contract_sos_left(bl, br, op.cl, op.cr, op.d, L, op.idx, op.mats, psi)
            assert rel(got, want) < 1e-13
        # an MPO whose last right bond no block reaches: that slice is zero
        from nutcracker_amd.operators import construct_operator_site
        I2 = np.eye(2, dtype=complex)
        op2 = construct_operator_site(2, 2, 3, [(1, 1, I2), (2, 2, np.diag([1.0, -1.0]).astype(complex))])
        L = hermitian_env(rng, 9, 2)
        psi = crand(rng, 7, 9, 2)
        rc, got = _call(h, hip, 9, 7, op2, L, psi)
        assert rc == 0, core.last_error()
        assert np.abs(got[:, :, 2]).max() == 0
        assert rel(got, core.contract_sos_left(9, 7, 2, 3, 2, L, op2.idx, op2.mats, psi)) < 1e-13
    finally:
        hip.free()


def test_noalloc_refuses_inputs_above_the_scratch_maximum(lib):
    core, h = lib
    from nutcracker_amd.operators import xxz_operator
    op = xxz_operator(6, 2, 1.0, 0.8)[2]
    idx = np.ascontiguousarray(np.asarray(op.idx, np.uint32).reshape(-1, order="F"))
    mats = np.ascontiguousarray(np.asarray(op.mats, np.complex128).reshape(-1, order="F"))
    # br = bl = 1024 needs 1024^2 * 2 * 10 * 16 bytes = 320 MiB of slabs; the
    # size is refused on the host before any device pointer is touched
    rc = h.nc_contract_sos_left_noalloc(1024, 1024, op.cl, op.cr, op.d, None, op.nmat,
                                        idx.ctypes.data_as(C.POINTER(C.c_uint32)), mats.ctypes.data, None, None, None)
    assert rc == -90
    assert "scratch" in core.last_error()
