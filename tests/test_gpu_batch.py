"""The batching form: independent chains of uneven bond dimension sweeping in
lockstep with grouped launches (nc_chain_batch_*).  Every chain must end where
it ends when it sweeps alone (same MPO, same initial state), and the TFIM
ground energies must match the free-fermion closed form."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from nutcracker_amd import chain, operators
    return chain, operators


def tfim_ground_energy(n, h):
    """H = -sum Z_i Z_{i+1} - h sum X_i, open chain: free fermions (the same
    construction as the oracle's closed-form check)."""
    A = np.zeros((n, n))
    B = np.zeros((n, n))
    for i in range(n):
        A[i, i] = 2 * h
    for i in range(n - 1):
        A[i, i + 1] = A[i + 1, i] = -1.0
        B[i, i + 1] = -1.0
        B[i + 1, i] = 1.0
    lam = np.sqrt(np.clip(np.linalg.eigvalsh((A - B) @ (A + B)), 0, None))
    return 0.5 * (np.trace(A) - lam.sum()) - n * h


def _chains(chain, ops_mod, n, spec, **kw):
    # ncv = 8: the ordered phase (h < 1) has a nearly degenerate ground doublet,
    # on which ncv = 3 needs thousands of restarts per site
    kw.setdefault("ncv", 8)
    return [chain.Chain(ops_mod.tfim_zz_x_operator(n, h), chi, seed=seed, rethrow=True, **kw)
            for (h, chi, seed) in spec]


SPEC = [(0.7, 8, 1), (1.3, 16, 2), (0.9, 32, 3), (1.1, 8, 4), (1.45, 16, 5), (0.55, 32, 6)]


def test_batched_sweeps_match_sequential(mods):
    chain, ops_mod = mods
    n = 12
    a = _chains(chain, ops_mod, n, SPEC)
    b = _chains(chain, ops_mod, n, SPEC)
    batch = chain.ChainBatch(a)
    batch.sweep_until_converged()
    for c in b:
        c.sweep_until_converged()
    for ca, cb, (h, chi, seed) in zip(a, b, SPEC):
        print("h", h, "chi", chi, "batched", ca.energy, "alone", cb.energy, "sweeps", ca.stats()["sweeps"],
              cb.stats()["sweeps"], flush=True)
        assert abs(ca.energy - cb.energy) <= 1e-10 * abs(cb.energy)
        assert ca.stats()["sweeps"] == cb.stats()["sweeps"]
    batch.close()


def test_batched_tfim_closed_form(mods):
    chain, ops_mod = mods
    n = 16
    spec = [(0.6, 32, 11), (1.4, 32, 12), (0.8, 16, 13), (1.2, 24, 14)]
    cs = _chains(chain, ops_mod, n, spec)
    batch = chain.ChainBatch(cs)
    batch.sweep_until_converged()
    for c, (h, chi, seed) in zip(cs, spec):
        exact = tfim_ground_energy(n, h)
        print("h", h, "chi", chi, "E", c.energy, "exact", exact, flush=True)
        assert abs(c.energy - exact) < 1e-8 * abs(exact)
    batch.close()


def test_batched_mixed_precision(mods):
    """complex64 contractions with FP64 eigen-steps, batched: the converged
    energies keep the FP64 bar (1e-10 relative)."""
    chain, ops_mod = mods
    n = 12
    a = _chains(chain, ops_mod, n, SPEC, mixed=True)
    b = _chains(chain, ops_mod, n, SPEC)
    batch = chain.ChainBatch(a)
    batch.sweep_until_converged()
    for c in b:
        c.sweep_until_converged()
    for ca, cb in zip(a, b):
        assert abs(ca.energy - cb.energy) <= 1e-10 * abs(cb.energy)
        assert ca.stats()["matvecs32"] > 0
    batch.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_graph_replay_matches_direct_launches(mods, mixed):
    """The lockstep Krylov extensions replayed from HIP graphs give bit-identical
    results to the same kernels launched one by one."""
    chain, ops_mod = mods
    n = 10
    spec = SPEC[:4]
    runs = []
    for graphs in (True, False):
        prev = chain.set_krylov_graphs(graphs)
        try:
            cs = _chains(chain, ops_mod, n, spec, mixed=mixed)
            batch = chain.ChainBatch(cs)
            batch.sweep_until_converged()
            runs.append([(c.energy, c.stats()["sweeps"], c.stats()["matvecs"] + c.stats()["matvecs32"]) for c in cs])
            batch.close()
        finally:
            chain.set_krylov_graphs(prev)
    print("graphs", runs[0], "direct", runs[1], flush=True)
    assert runs[0] == runs[1]


def test_batched_reset_matches_per_chain_reset(mods):
    """ChainBatch(..., reset=True): every right normalisation of every chain in one
    batched polar split gives the states, environments and energies of
    Chain.reset (chain.cpp:175-237) one chain at a time."""
    chain, ops_mod = mods
    n = 12
    a = _chains(chain, ops_mod, n, SPEC)
    b = _chains(chain, ops_mod, n, SPEC, reset=False)
    batch = chain.ChainBatch(b, reset=True)
    for ca, cb in zip(a, b):
        assert abs(ca.energy - cb.energy) <= 1e-12 * abs(ca.energy)
        for i in range(n):
            assert np.allclose(ca.site(i), cb.site(i), atol=1e-12, rtol=0)
    # and the batch sweeps on from there like the alone chains
    batch.sweep_until_converged()
    for c in a:
        c.sweep_until_converged()
    for ca, cb in zip(a, b):
        assert abs(ca.energy - cb.energy) <= 1e-10 * abs(ca.energy)
    batch.close()
