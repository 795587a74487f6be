"""Parity of the HIP kernels (through the C-ABI) with the oracle and with the
reference's own golden vectors.  Integer-valued goldens must match exactly;
random inputs within 1e-12 relative (complex FP64, different summation order)."""

import numpy as np
import pytest

from helpers import crand, goldens, hermitian_env, random_sparse_matrices, rel, z

import oracle.core as oc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from nutcracker_amd import core as c
    info = c.device_info()
    print("device", info, flush=True)
    return c


def _idx(e):
    return np.array(e["idx"], dtype=np.uint32).reshape(2, -1, order="F")


def test_contract_sos_left_golden(core):
    e = goldens()["contract_sos_left_non_trivial"]
    D = e["dims"]
    out = core.contract_sos_left(D["bl"], D["br"], D["cl"], D["cr"], D["d"], z(e["L"]), _idx(e), z(e["mats"]), z(e["psi"]))
    assert np.array_equal(out.ravel(order="F"), z(e["expected"]))


def test_contract_sos_right_golden(core):
    e = goldens()["contract_sos_right_non_trivial"]
    D = e["dims"]
    out = core.contract_sos_right(D["bl"], D["br"], D["cl"], D["cr"], D["d"], z(e["R"]), _idx(e), z(e["mats"]), z(e["psi"]))
    assert np.array_equal(out.ravel(order="F"), z(e["expected"]))


def test_trivial_operator_dimension_2(core):
    g = goldens()
    e = g["contract_sos_left_trivial_operator_dimension_2"]
    D = e["dims"]
    out = core.contract_sos_left(D["bl"], D["br"], D["cl"], D["cr"], D["d"], z(e["L"]), _idx(e), z(e["mats"]), z(e["psi"]))
    assert np.array_equal(out.ravel(order="F"), z(e["expected"]))
    e = g["contract_sos_right_trivial_operator_dimension_2"]
    D = e["dims"]
    out = core.contract_sos_right(D["bl"], D["br"], D["cl"], D["cr"], D["d"], z(e["R"]), _idx(e), z(e["mats"]), z(e["psi"]))
    assert np.array_equal(out.ravel(order="F"), z(e["expected"]))


def test_expectation_goldens(core):
    g = goldens()
    e = g["contract_expectation_boundaries_non_trivial"]
    v = core.contract_expectation_boundaries(3, 2, z(e["L"]), z(e["R"]))
    assert v == complex(*e["expected"])
    for cse in g["compute_expectation_trivial"]["cases"]:
        psi = z(cse["psi"])
        d = len(psi)
        v = core.compute_expectation(1, 1, 1, 1, d, np.ones(1), psi, np.array(cse["idx"]).reshape(2, 1), z(cse["mats"]), np.ones(1))
        assert v == complex(*cse["expected"])


def test_contract_vs_goldens(core):
    g = goldens()
    e = g["contract_vs_left_non_trivial"]
    D = e["dims"]
    out = core.contract_vs_left(D["blo"], D["bro"], D["bln"], D["brn"], D["d"], z(e["L"]), z(e["proj"]), z(e["psi"]))
    assert np.array_equal(out.ravel(order="F"), z(e["expected"]))
    e = g["contract_vs_right_non_trivial"]
    D = e["dims"]
    out = core.contract_vs_right(D["blo"], D["bro"], D["bln"], D["brn"], D["d"], z(e["R"]), z(e["proj"]), z(e["psi"]))
    assert np.array_equal(out.ravel(order="F"), z(e["expected"]))


SHAPES = [  # (d, bl, br, cl, cr): edge cases and ragged sizes around the 64 x 64 x 16 tile
    (1, 1, 1, 1, 1), (2, 1, 2, 1, 3), (2, 3, 5, 2, 4), (3, 7, 9, 5, 5), (2, 20, 17, 10, 3),
    (4, 13, 6, 3, 7), (2, 65, 63, 5, 5), (3, 33, 70, 4, 2), (2, 128, 128, 5, 5),
]


@pytest.mark.parametrize("shape", SHAPES)
def test_environments_and_matvec_random(core, shape):
    d, bl, br, cl, cr = shape
    rng = np.random.default_rng(abs(hash(shape)) % (2 ** 32))
    print("shape", shape, flush=True)
    L = crand(rng, bl, bl, cl)
    R = crand(rng, br, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d)
    psi = crand(rng, br, bl, d)
    nl = core.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi)
    assert rel(nl, oc.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi)) < 1e-12
    nr = core.contract_sos_right(bl, br, cl, cr, d, R, idx, mats, psi)
    assert rel(nr, oc.contract_sos_right(bl, br, cl, cr, d, R, idx, mats, psi)) < 1e-12
    e = core.compute_expectation(bl, br, cl, cr, d, L, psi, idx, mats, R)
    eo = oc.compute_expectation(bl, br, cl, cr, d, L, psi, idx, mats, R)
    assert abs(e - eo) <= 1e-12 * max(1.0, abs(eo))
    # the three stages, exactly as the reference composes them
    I = core.iteration_stage_1(bl, cl, cr, d, L, idx, mats)
    Io = oc.iteration_stage_1(bl, cl, cr, d, L, idx, mats)
    assert rel(I, Io) < 1e-15
    T = core.iteration_stage_2(bl, br, cr, d, Io, psi)
    To = oc.iteration_stage_2(bl, br, cr, d, Io, psi)
    assert rel(T, To) < 1e-12
    y = core.iteration_stage_3(bl, br, cr, d, To, R)
    assert rel(y, oc.iteration_stage_3(bl, br, cr, d, To, R)) < 1e-12


@pytest.mark.parametrize("shape", [(2, 3, 4, 2, 3), (3, 4, 3, 3, 2), (2, 5, 6, 5, 5)])
def test_optimization_matrix(core, shape):
    d, bl, br, cl, cr = shape
    rng = np.random.default_rng(7 + sum(shape))
    L = crand(rng, bl, bl, cl)
    R = crand(rng, br, br, cr)
    idx, mats = random_sparse_matrices(rng, cl, cr, d)
    M = core.compute_optimization_matrix(bl, br, cl, cr, d, L, idx, mats, R)
    assert rel(M, oc.compute_optimization_matrix(bl, br, cl, cr, d, L, idx, mats, R)) < 1e-12


def test_environment_consistency(core):
    """tests/boundaries.cpp:468-548: <L'|R> = <L|R'> = expectation at the site."""
    rng = np.random.default_rng(11)
    for _ in range(5):
        d, bl, br, cl, cr = (int(rng.integers(1, 5)), int(rng.integers(1, 12)), int(rng.integers(1, 12)),
                             int(rng.integers(1, 6)), int(rng.integers(1, 6)))
        L = crand(rng, bl, bl, cl)
        R = crand(rng, br, br, cr)
        idx, mats = random_sparse_matrices(rng, cl, cr, d)
        psi = crand(rng, br, bl, d)
        e = core.compute_expectation(bl, br, cl, cr, d, L, psi, idx, mats, R)
        nl = core.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi)
        nr = core.contract_sos_right(bl, br, cl, cr, d, R, idx, mats, psi)
        a = core.contract_expectation_boundaries(br, cr, nl, R)
        b = core.contract_expectation_boundaries(bl, cl, L, nr)
        assert abs(a - e) <= 1e-10 * abs(e)
        assert abs(b - e) <= 1e-10 * abs(e)
