"""Host-pointer residency of the drop-in C-ABI (nc_set_host_residency): the
reference's C++ layer hands the Fortran-core entry points host arrays that keep
their address and are handed back unchanged (the environment a contraction
writes is the next contraction's input, core.cpp:235-282), so the library may
keep device copies of them.  Every result must stay what the oracle computes
(core.f95:623-677) whether an input was served from a copy or uploaded: reused
outputs, arrays rewritten at the same address, copies evicted by a small budget
and arrays forgotten after an in-place edit."""

import numpy as np
import pytest

from helpers import crand, hermitian_env, phase_overlap, random_sparse_matrices, rel

import oracle.core as oc

pytestmark = pytest.mark.gpu


@pytest.fixture
def core():
    from nutcracker_amd import core as c
    prev = c.set_host_residency(1 << 30)
    try:
        yield c
    finally:
        c.set_host_residency(prev)


def _chain_step(core, rng, bl, br, cl, cr, d):
    idx, mats = random_sparse_matrices(rng, cl, cr, d, symmetric=True)
    L = hermitian_env(rng, bl, cl)
    psi = crand(rng, br, bl, d)
    return idx, mats, L, psi


def test_reused_output_is_served_from_the_device(core):
    """newL written by one contraction is the next contraction's input: the
    second call finds its device copy (a hit) and both results match the
    oracle; a repeated call with unchanged inputs uploads nothing."""
    rng = np.random.default_rng(31)
    bl = br = 48
    cl = cr = 5
    d = 2
    idx, mats, L, psi = _chain_step(core, rng, bl, br, cl, cr, d)
    out1 = np.zeros((br, br, cr), complex, order="F")
    out2 = np.zeros((br, br, cr), complex, order="F")
    s0 = core.host_residency_stats()
    core.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi, out=out1)
    ref1 = oc.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi)
    assert rel(out1, ref1) < 1e-12
    s1 = core.host_residency_stats()
    assert s1["misses"] - s0["misses"] >= 2 and s1["outputs_kept"] - s0["outputs_kept"] == 1
    psi2 = crand(rng, br, br, d)
    core.contract_sos_left(br, br, cl, cr, d, out1, idx, mats, psi2, out=out2)
    ref2 = oc.contract_sos_left(br, br, cl, cr, d, ref1, idx, mats, psi2)
    assert rel(out2, ref2) < 1e-12
    s2 = core.host_residency_stats()
    assert s2["hits"] - s1["hits"] >= 1  # out1 came from its device copy
    up = s2["bytes_uploaded"]
    core.contract_sos_left(br, br, cl, cr, d, out1, idx, mats, psi2, out=out2)
    s3 = core.host_residency_stats()
    assert rel(out2, ref2) < 1e-12
    assert s3["bytes_uploaded"] == up and s3["hits"] - s2["hits"] >= 2


def test_array_rewritten_at_the_same_address_is_uploaded_again(core):
    """New contents in the same host buffer (a freed tensor's address reused, or
    an in-place renormalisation that rewrites every element) change the
    fingerprint: the input is uploaded again and the result is the oracle's."""
    rng = np.random.default_rng(32)
    bl = br = 40
    cl = cr = 4
    d = 3
    idx, mats, L, psi = _chain_step(core, rng, bl, br, cl, cr, d)
    out = np.zeros((br, br, cr), complex, order="F")
    core.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi, out=out)
    assert rel(out, oc.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi)) < 1e-12
    psi *= 0.5  # every element rewritten in place
    L[...] = hermitian_env(rng, bl, cl)
    s0 = core.host_residency_stats()
    core.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi, out=out)
    s1 = core.host_residency_stats()
    assert s1["misses"] - s0["misses"] >= 2
    assert rel(out, oc.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi)) < 1e-12


def test_small_budget_evicts_and_stays_exact(core):
    """A budget smaller than one call's inputs: copies are evicted while the call
    still uses them, and the results stay the oracle's."""
    rng = np.random.default_rng(33)
    bl = br = 64
    cl = cr = 5
    d = 2
    idx, mats, L, psi = _chain_step(core, rng, bl, br, cl, cr, d)
    R = hermitian_env(rng, br, cr)
    core.set_host_residency(bl * bl * cl * 16 + 1024)  # room for about one environment
    for _ in range(3):
        e = core.compute_expectation(bl, br, cl, cr, d, L, psi, idx, mats, R)
        want = oc.compute_expectation(bl, br, cl, cr, d, L, psi, idx, mats, R)
        assert abs(e - want) <= 1e-12 * abs(want)
    st = core.host_residency_stats()
    assert st["bytes_held"] <= bl * bl * cl * 16 + 1024


def test_forget_after_an_in_place_edit(core):
    """An edit of a single element of a large array may touch no sampled word;
    forgetting the array first makes the next call upload it."""
    rng = np.random.default_rng(34)
    bl = br = 96
    cl = cr = 5
    d = 2
    idx, mats, L, psi = _chain_step(core, rng, bl, br, cl, cr, d)
    out = np.zeros((br, br, cr), complex, order="F")
    core.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi, out=out)
    L[5, 7, 2] += 1.0 + 0.5j
    L[7, 5, 2] += 1.0 - 0.5j
    core.host_residency_forget(L)
    core.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi, out=out)
    assert rel(out, oc.contract_sos_left(bl, br, cl, cr, d, L, idx, mats, psi)) < 1e-12


def test_optimize_and_move_with_residency(core):
    """optimize_ and norm_denorm_going_right_ with residency on give the results
    they give with it off (eigenvalue to 1e-12, the two-site state up to its
    phase to 1e-10)."""
    rng = np.random.default_rng(35)
    bl = br = 32
    cl = cr = 5
    d = 2
    idx, mats, L, psi = _chain_step(core, rng, bl, br, cl, cr, d)
    R = hermitian_env(rng, br, cr)
    n = bl * br * d
    runs = []
    for budget in (1 << 30, 0):
        core.set_host_residency(budget)
        info, it, res, ev, nrm = core.optimize(bl, br, cl, cr, d, L, idx, mats, R, 0, 0, n, None, None, None,
                                               "SR", 1e-12, 10000, psi)
        assert info == 0
        nxt = crand(np.random.default_rng(36), br, br, d)
        i2, on, od = core.norm_denorm_going_right(bl, br, br, d, d, res, nxt)
        assert i2 == 0
        runs.append((ev, on, od))
    (e1, on1, od1), (e0, on0, od0) = runs
    assert abs(e1 - e0) <= 1e-12 * abs(e0)
    # the two-site states agree up to the eigenvector's phase
    two1 = np.einsum("mla,rmb->labr", on1, od1)
    two0 = np.einsum("mla,rmb->labr", on0, od0)
    assert abs(1 - phase_overlap(two1, two0)) < 1e-10
