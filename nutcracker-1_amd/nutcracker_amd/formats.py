"""Reference file formats for operators and states, on the host.

* YAML operators (sources/common/yaml.cpp:169-273): a map with ``sequence``
  (1-based positions in ``sites``) and ``sites``, each with ``physical
  dimension``, ``left dimension``, ``right dimension`` and ``matrices`` (``from``,
  ``to``, ``data``: the d x d matrix row-major, each entry a number, a
  [real, imag] pair or a {real, imag} map -- includes/nutcracker/yaml.hpp:56-69).
  The data list fills the block's storage verbatim, as construct_operator_site
  does (operators.cpp:83-112).
* Protobuf messages OperatorBuffer / StateBuffer (nutcracker.proto:10-35) in the
  protobuf wire format, written by hand (protobuf.cpp:24-110: matrix_data and
  data are the complex elements in storage order, real then imaginary part;
  index_data the 1-based (from, to) pairs; sequence the 0-based positions of
  the distinct sites, io.cpp:26-50).

The library's C API (include/nutcracker_c_api.h) reads and writes the same
bytes."""

from __future__ import annotations

import struct
from typing import IO, List, Sequence, Tuple, Union

import numpy as np
import yaml

from .operators import OperatorSite

CPLX = np.complex128

# Normalization enum of the schema
MIDDLE, LEFT, RIGHT = 0, 1, 2


class YAMLInputError(ValueError):
    pass


class NoSuchOperatorSiteNumberError(YAMLInputError):
    """A sequence entry names no site; ``index`` is the 1-based entry value."""

    def __init__(self, index: int):
        super().__init__("no such operator site number: %d" % index)
        self.index = index


# ---------------------------------------------------------------------------
# YAML
# ---------------------------------------------------------------------------
def _complex(node) -> complex:
    if isinstance(node, (list, tuple)):
        return complex(float(node[0]), float(node[1]))
    if isinstance(node, dict):
        return complex(float(node["real"]), float(node["imag"]))
    return complex(float(node), 0.0)


def _distinct_sites(ops: Sequence[OperatorSite]) -> Tuple[List[OperatorSite], List[int]]:
    """io.cpp:26-50 (deconstructOperatorTo): distinct sites in order of first
    appearance and the 0-based sequence (sites are equal when their contents are)."""
    unique: List[OperatorSite] = []
    seq: List[int] = []
    for o in ops:
        for k, u in enumerate(unique):
            if (u is o) or (u.d == o.d and u.cl == o.cl and u.cr == o.cr and np.array_equal(u.idx, o.idx)
                            and np.array_equal(u.mats, o.mats)):
                seq.append(k)
                break
        else:
            seq.append(len(unique))
            unique.append(o)
    return unique, seq


def read_operator_yaml(source: Union[str, IO]) -> List[OperatorSite]:
    """yaml.cpp:169-183, 238-273: the chain's operator sites from a YAML document."""
    doc = yaml.safe_load(source)
    sites_node = doc.get("sites") or []
    unique = []
    for node in sites_node:
        d = int(node["physical dimension"])
        cl = int(node["left dimension"])
        cr = int(node["right dimension"])
        mats_node = node["matrices"]
        n = len(mats_node)
        idx = np.zeros((2, n), dtype=np.uint32)
        mats = np.zeros((d, d, n), dtype=CPLX, order="F")
        for k, m in enumerate(mats_node):
            f, t = int(m["from"]), int(m["to"])
            if f < 1 or t < 1:
                raise YAMLInputError("index too low")
            if f > cl or t > cr:
                raise YAMLInputError("index too high")
            data = m["data"]
            if len(data) != d * d:
                raise YAMLInputError("wrong data length %d (expected %d)" % (len(data), d * d))
            idx[0, k], idx[1, k] = f, t
            # the data list fills the block's storage (column-major) verbatim
            mats[:, :, k] = np.array([_complex(x) for x in data], dtype=CPLX).reshape((d, d), order="F")
        unique.append(OperatorSite(d, cl, cr, idx, mats))
    ops = []
    for v in doc.get("sequence") or []:
        k = int(v) - 1
        if k < 0 or k >= len(unique):
            raise NoSuchOperatorSiteNumberError(int(v))
        ops.append(unique[k])
    return ops


def write_operator_yaml(ops: Sequence[OperatorSite]) -> str:
    """yaml.cpp:184-205, 274-300: the document read_operator_yaml reads back
    exactly (numbers written with 17 significant digits)."""
    unique, seq = _distinct_sites(ops)

    def num(x: complex):
        if x.imag == 0:
            return float(repr(float(x.real)))
        return [float(x.real), float(x.imag)]

    doc = {"sequence": [k + 1 for k in seq], "sites": []}
    for u in unique:
        mats = []
        for k in range(u.nmat):
            data = np.asarray(u.mats)[:, :, k].reshape(-1, order="F")
            mats.append({"from": int(u.idx[0, k]), "to": int(u.idx[1, k]), "data": [num(complex(x)) for x in data]})
        doc["sites"].append({"physical dimension": u.d, "left dimension": u.cl, "right dimension": u.cr,
                             "matrices": mats})
    return yaml.safe_dump(doc, sort_keys=False)


# ---------------------------------------------------------------------------
# Protobuf wire format
# ---------------------------------------------------------------------------
def _varint(v: int) -> bytes:
    out = bytearray()
    while v >= 0x80:
        out.append((v & 0x7F) | 0x80)
        v >>= 7
    out.append(v)
    return bytes(out)


def _tag(field: int, wire: int) -> bytes:
    return _varint((field << 3) | wire)


def _len_field(field: int, payload: bytes) -> bytes:
    return _tag(field, 2) + _varint(len(payload)) + payload


def _u32_field(field: int, v: int) -> bytes:
    return _tag(field, 0) + _varint(int(v))


def _packed_doubles(field: int, values: np.ndarray) -> bytes:
    if len(values) == 0:
        return b""
    return _len_field(field, np.ascontiguousarray(values, dtype="<f8").tobytes())


def _packed_u32(field: int, values) -> bytes:
    if len(values) == 0:
        return b""
    return _len_field(field, b"".join(_varint(int(v)) for v in values))


def _read_varint(buf: bytes, pos: int) -> Tuple[int, int]:
    v = 0
    shift = 0
    while True:
        if pos >= len(buf):
            raise ValueError("protobuf: truncated varint")
        b = buf[pos]
        pos += 1
        v |= (b & 0x7F) << shift
        if not b & 0x80:
            return v, pos
        shift += 7
        if shift >= 70:
            raise ValueError("protobuf: varint too long")


def _fields(buf: bytes):
    """(field, wire, value) with value: int (varint), bytes (length-delimited or fixed)."""
    pos = 0
    while pos < len(buf):
        key, pos = _read_varint(buf, pos)
        field, wire = key >> 3, key & 7
        if wire == 0:
            v, pos = _read_varint(buf, pos)
        elif wire == 1:
            v, pos = buf[pos:pos + 8], pos + 8
        elif wire == 2:
            n, pos = _read_varint(buf, pos)
            v, pos = buf[pos:pos + n], pos + n
        elif wire == 5:
            v, pos = buf[pos:pos + 4], pos + 4
        else:
            raise ValueError("protobuf: unsupported wire type %d" % wire)
        if pos > len(buf):
            raise ValueError("protobuf: truncated field")
        yield field, wire, v


def _doubles(wire: int, v) -> List[float]:
    if wire == 2:
        return list(np.frombuffer(v, dtype="<f8"))
    return [struct.unpack("<d", v)[0]]


def _u32s(wire: int, v) -> List[int]:
    if wire == 0:
        return [v]
    out, pos = [], 0
    while pos < len(v):
        x, pos = _read_varint(v, pos)
        out.append(x)
    return out


def encode_operator(ops: Sequence[OperatorSite]) -> bytes:
    """An OperatorBuffer message (protobuf.cpp:24-37, 53-71)."""
    unique, seq = _distinct_sites(ops)
    out = b""
    for u in unique:
        z = np.asarray(u.mats, dtype=CPLX).reshape(-1, order="F")
        md = np.empty(2 * z.size)
        md[0::2], md[1::2] = z.real, z.imag
        site = (_u32_field(1, u.nmat) + _u32_field(2, u.d) + _u32_field(3, u.cl) + _u32_field(4, u.cr)
                + _packed_doubles(5, md) + _packed_u32(6, np.asarray(u.idx, dtype=np.uint32).reshape(-1, order="F")))
        out += _len_field(1, site)
    return out + _packed_u32(2, seq)


def decode_operator(buf: bytes) -> List[OperatorSite]:
    """OperatorBuffer -> the chain's operator sites (protobuf.cpp:38-52, 73-93)."""
    unique, seq = [], []
    for field, wire, v in _fields(buf):
        if field == 1 and wire == 2:
            f = {}
            md, ix = [], []
            for f2, w2, v2 in _fields(v):
                if f2 in (1, 2, 3, 4):
                    f[f2] = v2
                elif f2 == 5:
                    md += _doubles(w2, v2)
                elif f2 == 6:
                    ix += _u32s(w2, v2)
            if set(f) != {1, 2, 3, 4}:
                raise ValueError("protobuf: OperatorSiteBuffer without a required field")
            n, d, cl, cr = f[1], f[2], f[3], f[4]
            if len(md) != 2 * n * d * d or len(ix) != 2 * n:
                raise ValueError("protobuf: OperatorSiteBuffer data of the wrong length")
            z = np.array(md[0::2]) + 1j * np.array(md[1::2])
            mats = np.asfortranarray(z.reshape((d, d, n), order="F"))
            idx = np.array(ix, dtype=np.uint32).reshape((2, n), order="F")
            unique.append(OperatorSite(d, cl, cr, idx, mats))
        elif field == 2:
            seq += _u32s(wire, v)
    out = []
    for k in seq:
        if k >= len(unique):
            raise NoSuchOperatorSiteNumberError(k + 1)
        out.append(unique[k])
    return out


def encode_state(sites: Sequence[np.ndarray], normalizations: Sequence[int] = None) -> bytes:
    """A StateBuffer message from site tensors (br, bl, d); by default the first
    site MIDDLE and the rest RIGHT (setState, protobuf.hpp:87-93)."""
    if normalizations is None:
        normalizations = [MIDDLE] + [RIGHT] * (len(sites) - 1)
    out = b""
    for A, nz in zip(sites, normalizations):
        A = np.asarray(A, dtype=CPLX)
        br, bl, d = A.shape
        z = A.reshape(-1, order="F")
        data = np.empty(2 * z.size)
        data[0::2], data[1::2] = z.real, z.imag
        site = (_u32_field(1, nz) + _u32_field(2, d) + _u32_field(3, bl) + _u32_field(4, br)
                + _packed_doubles(5, data))
        out += _len_field(1, site)
    return out


def decode_state(buf: bytes) -> Tuple[List[np.ndarray], List[int]]:
    """StateBuffer -> site tensors (br, bl, d) and their normalizations."""
    sites, norms = [], []
    for field, wire, v in _fields(buf):
        if field != 1 or wire != 2:
            continue
        f = {1: MIDDLE}
        data = []
        for f2, w2, v2 in _fields(v):
            if f2 in (1, 2, 3, 4):
                f[f2] = v2
            elif f2 == 5:
                data += _doubles(w2, v2)
        if not {2, 3, 4} <= set(f):
            raise ValueError("protobuf: StateSiteBuffer without a required field")
        d, bl, br = f[2], f[3], f[4]
        if len(data) != 2 * d * bl * br:
            raise ValueError("protobuf: StateSiteBuffer data of the wrong length")
        z = np.array(data[0::2]) + 1j * np.array(data[1::2])
        sites.append(np.asfortranarray(z.reshape((br, bl, d), order="F")))
        norms.append(f[1])
    return sites, norms
