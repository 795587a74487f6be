"""ctypes handle on the reference's C API front door as the library implements it
(include/nutcracker_c_api.h; c-interface/operator.cpp, state.cpp,
serialization.cpp, errors.cpp of the reference).  Operators and states cross
the boundary as protobuf bytes (nutcracker_amd.formats)."""

from __future__ import annotations

import ctypes as C
from typing import List, Sequence, Tuple

import numpy as np

from . import core, formats
from .operators import OperatorSite


class NutcrackerError(RuntimeError):
    pass


class _Z(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


_SIGS = {
    "Nutcracker_clearError": (None, []),
    "Nutcracker_getError": (C.c_char_p, []),
    "Nutcracker_Serialization_free": (None, [C.c_void_p]),
    "Nutcracker_Serialization_getSize": (C.c_uint32, [C.c_void_p]),
    "Nutcracker_Serialization_write": (None, [C.c_void_p, C.c_void_p]),
    "Nutcracker_Operator_deserialize": (C.c_void_p, [C.c_uint32, C.c_char_p]),
    "Nutcracker_Operator_free": (None, [C.c_void_p]),
    "Nutcracker_Operator_serialize": (C.c_void_p, [C.c_void_p]),
    "Nutcracker_Operator_simpleSolveForLeastEigenvalues": (None, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float)]),
    "Nutcracker_Operator_simpleSolveForLeastEigenvaluesWithEigenvectors":
        (None, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_void_p)]),
    "Nutcracker_State_deserialize": (C.c_void_p, [C.c_uint32, C.c_char_p]),
    "Nutcracker_State_free": (None, [C.c_void_p]),
    "Nutcracker_State_serialize": (C.c_void_p, [C.c_void_p]),
    "Nutcracker_State_computeExpectation": (None, [C.c_void_p, C.c_void_p, C.POINTER(_Z)]),
    "Nutcracker_State_computeOverlap": (None, [C.c_void_p, C.c_void_p, C.POINTER(_Z)]),
}


def _lib():
    h = core.lib()
    if not getattr(h, "_c_api_sigs", False):
        for name, (res, args) in _SIGS.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        h._c_api_sigs = True
    return h


def _check(ok=True):
    err = _lib().Nutcracker_getError()
    if err is not None or not ok:
        _lib().Nutcracker_clearError()
        raise NutcrackerError(err.decode() if err else "call failed")


def _bytes_of(ser) -> bytes:
    lib = _lib()
    try:
        n = lib.Nutcracker_Serialization_getSize(ser)
        buf = C.create_string_buffer(n)
        lib.Nutcracker_Serialization_write(ser, buf)
        return buf.raw[:n]
    finally:
        lib.Nutcracker_Serialization_free(ser)


class Operator:
    """NutcrackerOperator: built from protobuf bytes or from operator sites."""

    def __init__(self, data: bytes):
        lib = _lib()
        lib.Nutcracker_clearError()
        self._h = lib.Nutcracker_Operator_deserialize(len(data), data)
        _check(self._h is not None)

    @classmethod
    def from_sites(cls, ops: Sequence[OperatorSite]) -> "Operator":
        return cls(formats.encode_operator(ops))

    def serialize(self) -> bytes:
        s = _lib().Nutcracker_Operator_serialize(self._h)
        _check(s is not None)
        return _bytes_of(s)

    def simple_solve_for_least_eigenvalues(self, levels: int) -> np.ndarray:
        ev = (C.c_float * levels)()
        _lib().Nutcracker_clearError()
        _lib().Nutcracker_Operator_simpleSolveForLeastEigenvalues(self._h, levels, ev)
        _check()
        return np.array(ev[:], dtype=np.float32)

    def simple_solve_for_least_eigenvalues_with_eigenvectors(self, levels: int) -> Tuple[np.ndarray, List["State"]]:
        ev = (C.c_float * levels)()
        vecs = (C.c_void_p * levels)()
        _lib().Nutcracker_clearError()
        _lib().Nutcracker_Operator_simpleSolveForLeastEigenvaluesWithEigenvectors(self._h, levels, ev, vecs)
        _check()
        return np.array(ev[:], dtype=np.float32), [State._own(v) for v in vecs]

    def close(self):
        if getattr(self, "_h", None):
            _lib().Nutcracker_Operator_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class State:
    """NutcrackerState: built from protobuf bytes (StateBuffer)."""

    def __init__(self, data: bytes = None):
        self._h = None
        if data is not None:
            lib = _lib()
            lib.Nutcracker_clearError()
            self._h = lib.Nutcracker_State_deserialize(len(data), data)
            _check(self._h is not None)

    @classmethod
    def _own(cls, handle) -> "State":
        s = cls()
        s._h = handle
        return s

    def serialize(self) -> bytes:
        s = _lib().Nutcracker_State_serialize(self._h)
        _check(s is not None)
        return _bytes_of(s)

    def sites(self) -> List[np.ndarray]:
        return formats.decode_state(self.serialize())[0]

    def expectation(self, op: Operator) -> complex:
        z = _Z()
        _lib().Nutcracker_clearError()
        _lib().Nutcracker_State_computeExpectation(self._h, op._h, C.byref(z))
        _check()
        return complex(z.re, z.im)

    def overlap(self, other: "State") -> complex:
        z = _Z()
        _lib().Nutcracker_clearError()
        _lib().Nutcracker_State_computeOverlap(self._h, other._h, C.byref(z))
        _check()
        return complex(z.re, z.im)

    def close(self):
        if getattr(self, "_h", None):
            _lib().Nutcracker_State_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
