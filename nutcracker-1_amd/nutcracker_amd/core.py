"""Python mirror of the reference's ``Nutcracker::Core`` wrappers
(includes/nutcracker/core.hpp:34-302, sources/common/core.cpp) over the C-ABI of
libnutcracker_hip.so.

Each function takes and returns numpy complex128 arrays in Fortran order, with the
same dimension arguments, in the same order, as the C++ wrapper of the same name.
Every call runs on the GPU: if the shared library is missing or no device is
visible, the call raises -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Tuple

import numpy as np

CPLX = np.complex128
_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnutcracker_hip.so")

_u32p = C.POINTER(C.c_uint32)
_zp = C.c_void_p
_dp = C.POINTER(C.c_double)

_SIGNATURES = {
    "contract_sos_left_": (None, [_u32p] * 5 + [_zp, _u32p, _u32p, _zp, _zp, _zp]),
    "contract_sos_right_": (None, [_u32p] * 5 + [_zp, _u32p, _u32p, _zp, _zp, _zp]),
    "compute_expectation_": (None, [_u32p] * 5 + [_zp, _zp, _u32p, _u32p, _zp, _zp, _zp]),
    "contract_expectation_boundaries_": (None, [_u32p, _u32p, _zp, _zp, _zp]),
    "compute_optimization_matrix_": (None, [_u32p] * 5 + [_zp, _u32p, _u32p, _zp, _zp, _zp]),
    "iteration_stage_1_": (None, [_u32p] * 4 + [_zp, _u32p, _u32p, _zp, _zp]),
    "iteration_stage_2_": (None, [_u32p] * 4 + [_zp, _zp, _zp]),
    "iteration_stage_3_": (None, [_u32p] * 4 + [_zp, _zp, _zp]),
    "optimize_": (C.c_uint32, [_u32p] * 5 + [_zp, _u32p, _u32p, _zp, _zp, _u32p, _u32p, _u32p, _zp, _zp, _u32p,
                                            C.c_char_p, _dp, _u32p, _zp, _zp, _zp, _dp]),
    "norm_denorm_going_left_": (C.c_int32, [_u32p] * 5 + [_zp] * 4),
    "norm_denorm_going_right_": (C.c_int32, [_u32p] * 5 + [_zp] * 4),
    "compute_overlap_with_projectors_": (None, [_u32p, _u32p, _zp, _zp, _u32p, _u32p, _zp, _zp]),
    "filter_components_outside_orthog_": (None, [_u32p] * 4 + [_zp, _zp, _u32p, _zp, _zp]),
    "contract_vs_left_": (None, [_u32p] * 5 + [_zp] * 4),
    "contract_vs_right_": (None, [_u32p] * 5 + [_zp] * 4),
    "form_overlap_vector_": (None, [_u32p] * 5 + [_zp] * 4),
    "convert_vectors_to_reflectors_": (C.c_uint32, [_u32p, _u32p, _zp, _u32p, _zp, _u32p]),
    "rand_norm_state_site_tensor_": (None, [_u32p] * 3 + [_zp]),
    "rand_unnorm_state_site_tensor_": (None, [_u32p] * 3 + [_zp]),
    "random_projector_matrix_": (None, [_u32p, _u32p, _u32p, _zp, _zp, _u32p]),
    "norm_for_left_": (C.c_int32, [_u32p] * 3 + [_zp] * 2),
    "norm_for_right_": (C.c_int32, [_u32p] * 3 + [_zp] * 2),
    "increase_bandwidth_between_": (C.c_int32, [_u32p] * 6 + [_zp] * 4),
    "increase_bandwidth_with_environment_": (None, [_u32p] * 4 + [_zp] * 6),
    "construct_left_exp_boundary_": (None, [_u32p] * 2 + [_zp] * 3),
    "construct_right_exp_boundary_": (None, [_u32p] * 2 + [_zp] * 3),
    "contract_matrix_left_": (None, [_u32p] * 2 + [_zp] * 3),
    "extend_state_vector_fragment_": (None, [_u32p] * 4 + [_zp] * 3),
    "form_overlap_site_tensor_": (None, [_u32p] * 3 + [_zp] * 2),
    "form_norm_overlap_tensors_": (None, [_u32p] * 5 + [_zp] * 6),
    "nc_last_error": (C.c_char_p, []),
    "nc_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), _dp, C.POINTER(C.c_size_t)]),
    "nc_project_into_orthogonal_space": (None, [C.c_uint32] * 4 + [_zp, _zp, _u32p, _zp, _zp]),
    "nc_unproject_from_orthogonal_space": (None, [C.c_uint32] * 4 + [_zp, _zp, _u32p, _zp, _zp]),
}

_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load the HIP core (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libnutcracker_hip.so is not built (run python nutcracker-1_amd/build.py); "
                "the MI355X core has no CPU fallback")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def set_autotune(on: bool) -> bool:
    """Switch the GEMM autotuner (keyed on chi_left, chi_right, d, MPO bonds and
    contraction step) on or off; returns the previous setting."""
    f = lib().nc_set_autotune
    f.restype = C.c_int
    f.argtypes = [C.c_int]
    return bool(f(1 if on else 0))


def autotune_table() -> List[str]:
    """The autotuner's choices so far, one line per shape key."""
    f = lib().nc_autotune_table
    f.restype = C.c_int64
    f.argtypes = [C.c_char_p, C.c_int64]
    n = f(None, 0)
    buf = C.create_string_buffer(int(n))
    f(buf, n)
    return [ln for ln in buf.value.decode().splitlines() if ln]


def set_mixed_planes(mode: int) -> int:
    """How mixed-precision single-site solves apply H_eff: 0 on the complex64
    GEMM, 1 on operands split once into 16-bit planes when bl * br >= 512 * 512
    (the default), 2 always on planes.  Returns the previous mode."""
    f = lib().nc_set_mixed_planes
    f.restype = C.c_int
    f.argtypes = [C.c_int]
    prev = f(int(mode))
    if prev < 0:
        raise ValueError("mode must be 0, 1 or 2")
    return prev


def set_plane_form(form: int) -> int:
    """Part form of the mixed form's planes for site solves started afterwards:
    1 (default) two fp16 parts per value, each matrix scaled by a power of two;
    0 three bf16 parts.  Returns the previous form."""
    f = lib().nc_set_plane_form
    f.restype = C.c_int
    f.argtypes = [C.c_int]
    prev = f(int(form))
    if prev < 0:
        raise ValueError("form must be 0 or 1")
    return prev


def set_optimizer_ncv(ncv: int) -> int:
    """Krylov basis size of optimize_ (the reference's 3 by default); returns the
    previous value."""
    f = lib().nc_set_optimizer_ncv
    f.restype = C.c_int
    f.argtypes = [C.c_int]
    prev = f(int(ncv))
    if prev < 0:
        raise ValueError("ncv must be in 3..62")
    return prev


def optimizer_stats() -> dict:
    """Site-optimizer counters since load (see nc_optimizer_stats)."""
    f = lib().nc_optimizer_stats
    f.restype = C.c_int
    f.argtypes = [C.POINTER(C.c_int64)]
    out = (C.c_int64 * 5)()
    f(out)
    keys = ("solves", "first_run_failures", "ncv_retries", "dense_fallbacks", "unconverged")
    return {k: int(out[i]) for i, k in enumerate(keys)}


def last_error() -> str:
    v = lib().nc_last_error()
    return v.decode() if v else ""


def device_info() -> dict:
    name = C.create_string_buffer(128)
    cu = C.c_int(0)
    clk = C.c_double(0)
    mem = C.c_size_t(0)
    rc = lib().nc_device_info(name, 128, C.byref(cu), C.byref(clk), C.byref(mem))
    if rc != 0:
        raise RuntimeError(last_error())
    return {"arch": name.value.decode(), "compute_units": cu.value, "clock_ghz": clk.value, "memory": mem.value}


def _u(v) -> C.POINTER(C.c_uint32):
    return C.pointer(C.c_uint32(int(v)))


def _z(a: np.ndarray) -> int:
    return a.ctypes.data


def _zin(a, n: int) -> np.ndarray:
    a = np.asarray(a, dtype=CPLX)
    a = np.asfortranarray(a)
    if a.size != n:
        raise ValueError("array has %d elements, expected %d" % (a.size, n))
    return a


def _idx(idx) -> Tuple[np.ndarray, int]:
    i = np.asfortranarray(np.asarray(idx, dtype=np.uint32).reshape(2, -1, order="F"))
    return i, i.shape[1]


def _check_error(before: str = ""):
    msg = last_error()
    if msg and msg != before:
        raise RuntimeError(msg)


class _ErrorWatch:
    def __enter__(self):
        self.before = last_error()
        return self

    def __exit__(self, *exc):
        if exc[0] is None:
            _check_error(self.before)
        return False


def set_host_residency(max_bytes: int) -> int:
    """Host-pointer residency of the C-ABI (nc_set_host_residency): device copies
    of the host arrays the entry points read and write, up to max_bytes (0: off).
    Returns the previous budget."""
    f = lib().nc_set_host_residency
    f.restype = C.c_int64
    f.argtypes = [C.c_int64]
    return int(f(int(max_bytes)))


def host_residency_forget(a: np.ndarray) -> None:
    """Drop the device copy of a host array (before editing it in place)."""
    f = lib().nc_host_residency_forget
    f.restype = C.c_int
    f.argtypes = [C.c_void_p]
    f(a.ctypes.data)


def host_residency_stats() -> dict:
    f = lib().nc_host_residency_stats
    f.restype = C.c_int
    f.argtypes = [C.POINTER(C.c_int64)]
    out = (C.c_int64 * 7)()
    f(out)
    keys = ("hits", "misses", "bytes_uploaded", "bytes_reused", "outputs_kept", "entries", "bytes_held")
    return dict(zip(keys, (int(v) for v in out)))


def contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi, out=None) -> np.ndarray:
    """core.hpp contract_sos_left; returns the new left environment (br, br, cr),
    written into `out` when given (a Fortran-ordered complex128 array of that
    shape, whose address then stays the same from call to call)."""
    i, nmat = _idx(idx)
    L = _zin(L, bl * bl * cl)
    m = _zin(mats, d * d * nmat)
    p = _zin(psi, br * bl * d)
    if out is None:
        out = np.zeros((br, br, cr), CPLX, order="F")
    elif out.dtype != CPLX or out.size != br * br * cr or not out.flags.f_contiguous:
        raise ValueError("out must be a Fortran-ordered complex128 array of %d elements" % (br * br * cr))
    with _ErrorWatch():
        lib().contract_sos_left_(_u(bl), _u(br), _u(cl), _u(cr), _u(d), _z(L), _u(nmat),
                                 i.ctypes.data_as(_u32p), _z(m), _z(p), _z(out))
    return out


def contract_sos_right(bl, br, cl, cr, d, R, idx, mats, psi) -> np.ndarray:
    """core.hpp contract_sos_right; returns the new right environment (bl, bl, cl)."""
    i, nmat = _idx(idx)
    R = _zin(R, br * br * cr)
    m = _zin(mats, d * d * nmat)
    p = _zin(psi, br * bl * d)
    out = np.zeros((bl, bl, cl), CPLX, order="F")
    with _ErrorWatch():
        lib().contract_sos_right_(_u(bl), _u(br), _u(cl), _u(cr), _u(d), _z(R), _u(nmat),
                                  i.ctypes.data_as(_u32p), _z(m), _z(p), _z(out))
    return out


def compute_expectation(bl, br, cl, cr, d, L, psi, idx, mats, R) -> complex:
    i, nmat = _idx(idx)
    L = _zin(L, bl * bl * cl)
    R = _zin(R, br * br * cr)
    m = _zin(mats, d * d * nmat)
    p = _zin(psi, br * bl * d)
    out = np.zeros(1, CPLX)
    with _ErrorWatch():
        lib().compute_expectation_(_u(bl), _u(br), _u(cl), _u(cr), _u(d), _z(L), _z(p), _u(nmat),
                                   i.ctypes.data_as(_u32p), _z(m), _z(R), _z(out))
    return complex(out[0])


def contract_expectation_boundaries(b, c, L, R) -> complex:
    L = _zin(L, b * b * c)
    R = _zin(R, b * b * c)
    out = np.zeros(1, CPLX)
    with _ErrorWatch():
        lib().contract_expectation_boundaries_(_u(b), _u(c), _z(L), _z(R), _z(out))
    return complex(out[0])


def compute_optimization_matrix(bl, br, cl, cr, d, L, idx, mats, R) -> np.ndarray:
    i, nmat = _idx(idx)
    L = _zin(L, bl * bl * cl)
    R = _zin(R, br * br * cr)
    m = _zin(mats, d * d * nmat)
    n = br * bl * d
    out = np.zeros((n, n), CPLX, order="F")
    with _ErrorWatch():
        lib().compute_optimization_matrix_(_u(bl), _u(br), _u(cl), _u(cr), _u(d), _z(L), _u(nmat),
                                           i.ctypes.data_as(_u32p), _z(m), _z(R), _z(out))
    return out


def iteration_stage_1(bl, cl, cr, d, L, idx, mats) -> np.ndarray:
    i, nmat = _idx(idx)
    L = _zin(L, bl * bl * cl)
    m = _zin(mats, d * d * nmat)
    out = np.zeros((bl, d, cr, bl, d), CPLX, order="F")
    with _ErrorWatch():
        lib().iteration_stage_1_(_u(bl), _u(cl), _u(cr), _u(d), _z(L), _u(nmat), i.ctypes.data_as(_u32p),
                                 _z(m), _z(out))
    return out


def iteration_stage_2(bl, br, cr, d, I, x) -> np.ndarray:
    I = _zin(I, bl * d * cr * bl * d)
    x = _zin(x, br * bl * d)
    out = np.zeros((br, cr, bl, d), CPLX, order="F")
    with _ErrorWatch():
        lib().iteration_stage_2_(_u(bl), _u(br), _u(cr), _u(d), _z(I), _z(x), _z(out))
    return out


def iteration_stage_3(bl, br, cr, d, T, R) -> np.ndarray:
    T = _zin(T, br * cr * bl * d)
    R = _zin(R, br * br * cr)
    out = np.zeros((br, bl, d), CPLX, order="F")
    with _ErrorWatch():
        lib().iteration_stage_3_(_u(bl), _u(br), _u(cr), _u(d), _z(T), _z(R), _z(out))
    return out


def optimize(bl, br, cl, cr, d, L, idx, mats, R, nproj, nrefl, orth, reflectors, coeffs, swaps,
             which, tol, iterations, guess):
    """core.hpp optimize.  Returns (info, iterations, result(br,bl,d), eigenvalue, normal);
    info follows the reference: 0, -14, 5, 10, 11."""
    i, nmat = _idx(idx)
    L = _zin(L, bl * bl * cl)
    R = _zin(R, br * br * cr)
    m = _zin(mats, d * d * nmat)
    n = br * bl * d
    g = _zin(guess, n)
    if nproj:
        refl = _zin(reflectors, n * nproj)
        co = _zin(coeffs, max(1, nrefl)) if nrefl else np.zeros(1, CPLX)
        sw = np.ascontiguousarray(np.asarray(swaps, np.uint32).reshape(-1)[:max(1, nrefl)]) if nrefl else np.ones(1, np.uint32)
    else:
        refl = np.zeros(1, CPLX)
        co = np.zeros(1, CPLX)
        sw = np.ones(1, np.uint32)
    res = np.zeros((br, bl, d), CPLX, order="F")
    ev = np.zeros(1, CPLX)
    normal = C.c_double(0.0)
    it = C.c_uint32(int(iterations))
    tolv = C.c_double(float(tol))
    w = which.encode() if isinstance(which, str) else bytes(which)
    with _ErrorWatch():
        info = lib().optimize_(_u(bl), _u(br), _u(cl), _u(cr), _u(d), _z(L), _u(nmat), i.ctypes.data_as(_u32p),
                               _z(m), _z(R), _u(nproj), _u(nrefl), _u(orth), _z(refl), _z(co),
                               sw.ctypes.data_as(_u32p), w, C.byref(tolv), C.byref(it), _z(g), _z(res), _z(ev),
                               C.byref(normal))
    info = C.c_int32(info).value
    return info, int(it.value), res, complex(ev[0]), float(normal.value)


def norm_denorm_going_left(bll, bl, br, dl, d, to_denorm, to_norm):
    """core.hpp norm_denorm_going_left (C-side names): to_denorm (bl, bll, dl),
    to_norm (br, bl, d).  Returns (info, denormalized, normalized)."""
    td = _zin(to_denorm, bl * bll * dl)
    tn = _zin(to_norm, br * bl * d)
    od = np.zeros((bl, bll, dl), CPLX, order="F")
    on = np.zeros((br, bl, d), CPLX, order="F")
    with _ErrorWatch():
        info = lib().norm_denorm_going_left_(_u(bll), _u(bl), _u(br), _u(dl), _u(d), _z(td), _z(tn), _z(od), _z(on))
    return int(info), od, on


def norm_denorm_going_right(bl, br, brr, d, dr, to_norm, to_denorm):
    """core.hpp norm_denorm_going_right (C-side names): to_norm (br, bl, d),
    to_denorm (brr, br, dr).  Returns (info, normalized, denormalized)."""
    tn = _zin(to_norm, br * bl * d)
    td = _zin(to_denorm, brr * br * dr)
    on = np.zeros((br, bl, d), CPLX, order="F")
    od = np.zeros((brr, br, dr), CPLX, order="F")
    with _ErrorWatch():
        info = lib().norm_denorm_going_right_(_u(bl), _u(br), _u(brr), _u(d), _u(dr), _z(tn), _z(td), _z(on), _z(od))
    return int(info), on, od


def compute_overlap_with_projectors(nproj, nrefl, reflectors, coeffs, swaps, n, vector) -> float:
    refl = _zin(reflectors, n * nproj) if nproj else np.zeros(1, CPLX)
    co = _zin(coeffs, nrefl) if nrefl else np.zeros(1, CPLX)
    sw = np.ascontiguousarray(np.asarray(swaps, np.uint32).reshape(-1)) if nrefl else np.ones(1, np.uint32)
    v = _zin(vector, n)
    out = np.zeros(1, CPLX)
    with _ErrorWatch():
        lib().compute_overlap_with_projectors_(_u(nproj), _u(nrefl), _z(refl), _z(co), sw.ctypes.data_as(_u32p),
                                               _u(n), _z(v), _z(out))
    return float(out.view(np.float64)[0])


def project_into_orthogonal_space(n, nproj, nrefl, orth, reflectors, coeffs, swaps, vector) -> np.ndarray:
    refl = _zin(reflectors, n * nproj) if nproj else np.zeros(1, CPLX)
    co = _zin(coeffs, nrefl) if nrefl else np.zeros(1, CPLX)
    sw = np.ascontiguousarray(np.asarray(swaps, np.uint32).reshape(-1)) if nrefl else np.ones(1, np.uint32)
    v = _zin(vector, n)
    out = np.zeros(orth if nproj else n, CPLX)
    with _ErrorWatch():
        lib().nc_project_into_orthogonal_space(n, nproj, nrefl, orth, _z(refl), _z(co), sw.ctypes.data_as(_u32p),
                                               _z(v), _z(out))
    return out


def unproject_from_orthogonal_space(n, nproj, nrefl, orth, reflectors, coeffs, swaps, vector) -> np.ndarray:
    refl = _zin(reflectors, n * nproj) if nproj else np.zeros(1, CPLX)
    co = _zin(coeffs, nrefl) if nrefl else np.zeros(1, CPLX)
    sw = np.ascontiguousarray(np.asarray(swaps, np.uint32).reshape(-1)) if nrefl else np.ones(1, np.uint32)
    v = _zin(vector, orth if nproj else n)
    out = np.zeros(n, CPLX)
    with _ErrorWatch():
        lib().nc_unproject_from_orthogonal_space(n, nproj, nrefl, orth, _z(refl), _z(co), sw.ctypes.data_as(_u32p),
                                                 _z(v), _z(out))
    return out


def filter_components_outside_orthog(n, nproj, nrefl, orth, reflectors, coeffs, swaps, vector) -> np.ndarray:
    refl = _zin(reflectors, n * nproj) if nproj else np.zeros(1, CPLX)
    co = _zin(coeffs, nrefl) if nrefl else np.zeros(1, CPLX)
    sw = np.ascontiguousarray(np.asarray(swaps, np.uint32).reshape(-1)) if nrefl else np.ones(1, np.uint32)
    v = _zin(vector, n)
    out = np.zeros(n, CPLX)
    with _ErrorWatch():
        lib().filter_components_outside_orthog_(_u(n), _u(nproj), _u(nrefl), _u(orth), _z(refl), _z(co),
                                                sw.ctypes.data_as(_u32p), _z(v), _z(out))
    return out


def contract_vs_left(blo, bro, bln, brn, d, L, proj_site, state_site) -> np.ndarray:
    L = _zin(L, bln * blo)
    P = _zin(proj_site, blo * d * bro)
    S = _zin(state_site, brn * bln * d)
    out = np.zeros((brn, bro), CPLX, order="F")
    with _ErrorWatch():
        lib().contract_vs_left_(_u(blo), _u(bro), _u(bln), _u(brn), _u(d), _z(L), _z(P), _z(S), _z(out))
    return out


def contract_vs_right(blo, bro, bln, brn, d, R, proj_site, state_site) -> np.ndarray:
    R = _zin(R, bro * brn)
    P = _zin(proj_site, blo * d * bro)
    S = _zin(state_site, brn * bln * d)
    out = np.zeros((blo, bln), CPLX, order="F")
    with _ErrorWatch():
        lib().contract_vs_right_(_u(blo), _u(bro), _u(bln), _u(brn), _u(d), _z(R), _z(P), _z(S), _z(out))
    return out


def form_overlap_vector(blo, bro, bln, brn, d, L, R, proj_site) -> np.ndarray:
    """Core::form_overlap_vector (core.cpp:517-537): returns the vector as (brn, bln, d)."""
    L = _zin(L, bln * blo)
    R = _zin(R, bro * brn)
    P = _zin(proj_site, blo * d * bro)
    out = np.zeros((brn, bln, d), CPLX, order="F")
    with _ErrorWatch():
        lib().form_overlap_vector_(_u(blo), _u(bro), _u(bln), _u(brn), _u(d), _z(L), _z(R), _z(P), _z(out))
    return out


def convert_vectors_to_reflectors(n, m, vectors):
    """Core::convert_vectors_to_reflectors (core.cpp:371-397) for ``m`` vectors of
    length ``n``: returns (rank, reflectors(n, m), coefficients, swaps) -- the
    reference overwrites ``vectors`` with the reflectors."""
    V = np.array(_zin(vectors, n * m).reshape((n, m), order="F"), order="F", copy=True)
    k = min(n, m)
    coeffs = np.zeros(max(1, k), CPLX)
    swaps = np.zeros(max(1, k), np.uint32)
    rank = C.c_uint32(0)
    with _ErrorWatch():
        lib().convert_vectors_to_reflectors_(_u(n), _u(m), _z(V), C.byref(rank), _z(coeffs),
                                             swaps.ctypes.data_as(_u32p))
    return int(rank.value), V, coeffs[:k], swaps[:k]


def form_overlap_site_tensor(br, bl, d, state_site) -> np.ndarray:
    """Core::form_overlap_site_tensor (core.cpp:498-509): (bl, d, br)."""
    S = _zin(state_site, br * bl * d)
    out = np.zeros((bl, d, br), CPLX, order="F")
    with _ErrorWatch():
        lib().form_overlap_site_tensor_(_u(br), _u(bl), _u(d), _z(S), _z(out))
    return out


def form_norm_overlap_tensors(bl, bm, br, dl, dr, unnorm_1, right_norm_2):
    """Core::form_norm_overlap_tensors (core.cpp:463-491): returns
    (left_norm_overlap_1(bl,dl,bm), unnormalized_overlap_1(bl,dl,bm),
    unnormalized_state_2(br,bm,dr), right_norm_overlap_2(bm,dr,br))."""
    U1 = _zin(unnorm_1, bm * bl * dl)
    R2 = _zin(right_norm_2, br * bm * dr)
    lo1 = np.zeros((bl, dl, bm), CPLX, order="F")
    uo1 = np.zeros((bl, dl, bm), CPLX, order="F")
    u2 = np.zeros((br, bm, dr), CPLX, order="F")
    ro2 = np.zeros((bm, dr, br), CPLX, order="F")
    with _ErrorWatch():
        lib().form_norm_overlap_tensors_(_u(bl), _u(bm), _u(br), _u(dl), _u(dr), _z(U1), _z(R2), _z(lo1), _z(uo1),
                                         _z(u2), _z(ro2))
    return lo1, uo1, u2, ro2


def rand_norm_state_site_tensor(br, bl, d) -> np.ndarray:
    """Core::rand_norm_state_site_tensor (core.cpp:735-750): a random (br, bl, d)
    site with sum_{r,s} psi(r,l,s) conj(psi(r,l',s)) = delta(l,l')."""
    out = np.zeros((br, bl, d), CPLX, order="F")
    with _ErrorWatch():
        lib().rand_norm_state_site_tensor_(_u(br), _u(bl), _u(d), _z(out))
    return out


def rand_unnorm_state_site_tensor(br, bl, d) -> np.ndarray:
    """Core::rand_unnorm_state_site_tensor (core.cpp:752-767): unit Frobenius norm."""
    out = np.zeros((br, bl, d), CPLX, order="F")
    with _ErrorWatch():
        lib().rand_unnorm_state_site_tensor_(_u(br), _u(bl), _u(d), _z(out))
    return out


def random_projector_matrix(n, nproj):
    """Core::random_projector_matrix (core.cpp:769-783): returns
    (rank, reflectors(n, nproj), coefficients, swaps)."""
    k = min(n, nproj)
    refl = np.zeros((n, max(1, nproj)), CPLX, order="F")
    coeffs = np.zeros(max(1, k), CPLX)
    swaps = np.zeros(max(1, k), np.uint32)
    rank = C.c_uint32(0)
    with _ErrorWatch():
        lib().random_projector_matrix_(_u(n), _u(nproj), C.byref(rank), _z(refl), _z(coeffs),
                                       swaps.ctypes.data_as(_u32p))
    return int(rank.value), refl[:, :nproj], coeffs[:k], swaps[:k]


def norm_for_left(br, bl, d, site):
    """Core::norm_for_left (core.cpp:638-655): (info, normalized (br, bl, d))."""
    S = _zin(site, br * bl * d)
    out = np.zeros((br, bl, d), CPLX, order="F")
    with _ErrorWatch():
        info = lib().norm_for_left_(_u(br), _u(bl), _u(d), _z(S), _z(out))
    return int(info), out


def norm_for_right(br, bl, d, site):
    """Core::norm_for_right (core.cpp:657-674): (info, normalized (br, bl, d))."""
    S = _zin(site, br * bl * d)
    out = np.zeros((br, bl, d), CPLX, order="F")
    with _ErrorWatch():
        info = lib().norm_for_right_(_u(br), _u(bl), _u(d), _z(S), _z(out))
    return int(info), out


def increase_bandwidth_between(bl, bm, br, dl, dr, new_bm, left_site, right_site):
    """Core::increase_bandwidth_between (core.cpp:548-578) for a left site
    (bm, bl, dl) and a right site (br, bm, dr): (info, new left (new_bm, bl, dl),
    new right (br, new_bm, dr))."""
    A = _zin(left_site, bm * bl * dl)
    B = _zin(right_site, br * bm * dr)
    o1 = np.zeros((new_bm, bl, dl), CPLX, order="F")
    o2 = np.zeros((br, new_bm, dr), CPLX, order="F")
    with _ErrorWatch():
        info = lib().increase_bandwidth_between_(_u(bl), _u(bm), _u(br), _u(dl), _u(dr), _u(new_bm), _z(A), _z(B),
                                                 _z(o1), _z(o2))
    return int(info), o1, o2


def increase_bandwidth_with_environment(b, c, d, new_b, site, L, R):
    """Core::increase_bandwidth_with_environment (core.cpp:785-819):
    (new site (new_b, new_b, d), new left (new_b, new_b, c), new right)."""
    S = _zin(site, b * b * d)
    Lz = _zin(L, b * b * c)
    Rz = _zin(R, b * b * c)
    so = np.zeros((new_b, new_b, d), CPLX, order="F")
    lo = np.zeros((new_b, new_b, c), CPLX, order="F")
    ro = np.zeros((new_b, new_b, c), CPLX, order="F")
    with _ErrorWatch():
        lib().increase_bandwidth_with_environment_(_u(b), _u(c), _u(d), _u(new_b), _z(S), _z(Lz), _z(Rz), _z(so),
                                                   _z(lo), _z(ro))
    return so, lo, ro


def construct_left_exp_boundary(b, c, state_boundary, operator_boundary) -> np.ndarray:
    s = _zin(state_boundary, b)
    o = _zin(operator_boundary, c)
    out = np.zeros((b, b, c), CPLX, order="F")
    with _ErrorWatch():
        lib().construct_left_exp_boundary_(_u(b), _u(c), _z(s), _z(o), _z(out))
    return out


def construct_right_exp_boundary(b, c, state_boundary, operator_boundary) -> np.ndarray:
    s = _zin(state_boundary, b)
    o = _zin(operator_boundary, c)
    out = np.zeros((b, b, c), CPLX, order="F")
    with _ErrorWatch():
        lib().construct_right_exp_boundary_(_u(b), _u(c), _z(s), _z(o), _z(out))
    return out


def contract_matrix_left(bl, br, L, matrix) -> np.ndarray:
    Lz = _zin(L, bl * bl)
    X = _zin(matrix, br * bl)
    out = np.zeros((br, br), CPLX, order="F")
    with _ErrorWatch():
        lib().contract_matrix_left_(_u(bl), _u(br), _z(Lz), _z(X), _z(out))
    return out


def extend_state_vector_fragment(bm, br, dl, dr, old, site) -> np.ndarray:
    O = _zin(old, bm * dl)
    S = _zin(site, br * bm * dr)
    out = np.zeros((br, dr, dl), CPLX, order="F")
    with _ErrorWatch():
        lib().extend_state_vector_fragment_(_u(bm), _u(br), _u(dl), _u(dr), _z(O), _z(S), _z(out))
    return out
