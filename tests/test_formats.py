"""Reference file formats on the host: the YAML operator format (yaml.cpp:169-273)
and the protobuf OperatorBuffer / StateBuffer messages (nutcracker.proto:10-35,
protobuf.cpp:24-110), in Python (nutcracker_amd.formats) and in the library's C
API codec (include/nutcracker_c_api.h; host-only calls, no device work).

Pins: the external-field operator of tests/yaml.cpp:135-200 (the same three
distinct sites and sequence, here in a YAML document of our own) equals the
operator the library builds for it; the NoSuchOperatorSiteNumberError case of
tests/yaml.cpp:206-220; the three complex-number forms of yaml.hpp:56-69; and a
protobuf message whose bytes are worked out by hand from the wire-format rules
(varint keys (field << 3) | wire type, packed little-endian doubles)."""

import ctypes as C
import struct

import numpy as np
import pytest

from nutcracker_amd import formats, operators

EXTERNAL_FIELD_YAML = """
sequence: [1, {middle}3]
paulis:
  - &I [1, 0, 0, 1]
  - &Z [1, 0, 0, -1]
sites:
  - {{physical dimension: 2, left dimension: 1, right dimension: 2,
      matrices: [{{from: 1, to: 1, data: *I}}, {{from: 1, to: 2, data: *Z}}]}}
  - {{physical dimension: 2, left dimension: 2, right dimension: 2,
      matrices: [{{from: 1, to: 1, data: *I}}, {{from: 1, to: 2, data: *Z}}, {{from: 2, to: 2, data: *I}}]}}
  - {{physical dimension: 2, left dimension: 2, right dimension: 1,
      matrices: [{{from: 1, to: 1, data: *Z}}, {{from: 2, to: 1, data: *I}}]}}
"""


def _same_operator(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert (x.d, x.cl, x.cr) == (y.d, y.cl, y.cr)
        # the same blocks, in any order
        bx = sorted((int(x.idx[0, k]), int(x.idx[1, k]), tuple(np.asarray(x.mats)[:, :, k].ravel())) for k in range(x.nmat))
        by = sorted((int(y.idx[0, k]), int(y.idx[1, k]), tuple(np.asarray(y.mats)[:, :, k].ravel())) for k in range(y.nmat))
        assert bx == by


@pytest.mark.parametrize("middle", [0, 1, 5])
def test_yaml_external_field(middle):
    text = EXTERNAL_FIELD_YAML.format(middle="2, " * middle)
    got = formats.read_operator_yaml(text)
    _same_operator(got, operators.external_field_operator(middle + 2, operators.Z))


def test_yaml_no_such_operator_site_number():
    with pytest.raises(formats.NoSuchOperatorSiteNumberError) as e:
        formats.read_operator_yaml("---\nsites:\nsequence: [1]\n")
    assert e.value.index == 1


def test_yaml_complex_forms_and_round_trip():
    text = """
sequence: [1]
sites:
  - physical dimension: 2
    left dimension: 1
    right dimension: 1
    matrices:
      - from: 1
        to: 1
        data: [0.5, [1.25, -2.0], {real: 3.0, imag: 0.75}, -1]
"""
    (site,) = formats.read_operator_yaml(text)
    # storage order of the block = the data list
    assert np.array_equal(np.asarray(site.mats)[:, :, 0].ravel(order="F"), [0.5, 1.25 - 2j, 3 + 0.75j, -1])
    rng = np.random.default_rng(3)
    ops = operators.xxz_operator(5, 2, 0.7, 1.3)
    ops[2] = operators.OperatorSite(ops[2].d, ops[2].cl, ops[2].cr, ops[2].idx,
                                    np.asfortranarray(ops[2].mats + 1j * rng.standard_normal(ops[2].mats.shape)))
    back = formats.read_operator_yaml(formats.write_operator_yaml(ops))
    _same_operator(back, ops)


def test_protobuf_bytes_worked_out_by_hand():
    site = operators.OperatorSite(1, 1, 1, np.array([[1], [1]], np.uint32), np.full((1, 1, 1), 2.0 + 0j, order="F"))
    payload = (b"\x08\x01" b"\x10\x01" b"\x18\x01" b"\x20\x01"          # fields 1-4, varint 1
               + b"\x2a\x10" + struct.pack("<dd", 2.0, 0.0)            # field 5, packed doubles
               + b"\x32\x02\x01\x01")                                  # field 6, packed uint32
    expect = b"\x0a" + bytes([len(payload)]) + payload + b"\x12\x01\x00"  # sites, then sequence [0]
    assert formats.encode_operator([site]) == expect
    (back,) = formats.decode_operator(expect)
    _same_operator([back], [site])
    # a multi-byte varint: 300 = 0xac 0x02
    assert formats._varint(300) == b"\xac\x02"


def test_protobuf_operator_and_state_round_trip():
    ops = operators.xxz_operator(6, 3, 1.0, 0.4)
    data = formats.encode_operator(ops)
    _same_operator(formats.decode_operator(data), ops)
    rng = np.random.default_rng(11)
    bonds = [1, 3, 5, 5, 3, 1]
    sites = [np.asfortranarray(rng.standard_normal((bonds[i + 1], bonds[i], 3))
                               + 1j * rng.standard_normal((bonds[i + 1], bonds[i], 3))) for i in range(5)]
    back, norms = formats.decode_state(formats.encode_state(sites))
    assert norms == [formats.MIDDLE] + [formats.RIGHT] * 4
    for a, b in zip(sites, back):
        assert np.array_equal(a, b)


def test_c_api_codec_matches_python():
    """The library's decoder and encoder (host only) give back the same bytes,
    and it rejects a sequence that names a missing site."""
    from nutcracker_amd import c_api
    ops = operators.tfim_zz_x_operator(7, 0.8)
    data = formats.encode_operator(ops)
    op = c_api.Operator(data)
    assert op.serialize() == data
    op.close()
    rng = np.random.default_rng(5)
    bonds = [1, 2, 4, 2, 1]
    sites = [np.asfortranarray(rng.standard_normal((bonds[i + 1], bonds[i], 2)) + 0j) for i in range(4)]
    sdata = formats.encode_state(sites)
    st = c_api.State(sdata)
    assert st.serialize() == sdata
    st.close()
    bad = data[:-1] + bytes([len(formats.decode_operator(data)) + 3])  # last sequence entry out of range
    with pytest.raises(c_api.NutcrackerError, match="no such operator site"):
        c_api.Operator(bad)
