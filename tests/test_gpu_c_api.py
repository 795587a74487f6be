"""The C API front door on the device (include/nutcracker_c_api.h):
Nutcracker_Operator_simpleSolveForLeastEigenvalues[WithEigenvectors]
(c-interface/operator.cpp:23-52: initial bond dimension 2 grown until the
energy settles, thresholds 1e-12 / 1e-9 / 1e-8 / 1e-8, float eigenvalues) and
the state expectation values and overlaps (state.cpp:9-14), against dense
diagonalisation and dense contraction of the same operator and states."""

import numpy as np
import pytest

from helpers import dense_hamiltonian, dense_mps

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from nutcracker_amd import c_api, formats, operators
    return c_api, formats, operators


def test_simple_solve_least_eigenvalues(mods):
    c_api, formats, operators = mods
    ops = operators.tfim_zz_x_operator(8, 1.5)
    exact = np.sort(np.linalg.eigvalsh(dense_hamiltonian(ops)))[:3]
    op = c_api.Operator.from_sites(ops)
    got = op.simple_solve_for_least_eigenvalues(3)
    print("levels", got, "exact", exact, flush=True)
    assert got.dtype == np.float32
    # float output; chain convergence 1e-8
    assert np.allclose(got, exact, rtol=2e-6, atol=0)


def test_simple_solve_with_eigenvectors(mods):
    c_api, formats, operators = mods
    ops = operators.tfim_zz_x_operator(8, 1.5)
    H = dense_hamiltonian(ops)
    w = np.sort(np.linalg.eigvalsh(H))[:3]
    op = c_api.Operator.from_sites(ops)
    ev, states = op.simple_solve_for_least_eigenvalues_with_eigenvectors(3)
    assert np.allclose(ev, w, rtol=2e-6, atol=0)
    vecs = []
    for k, st in enumerate(states):
        sites, norms = formats.decode_state(st.serialize())
        assert norms == [formats.MIDDLE] + [formats.RIGHT] * (len(sites) - 1)
        for A in sites[1:]:  # right-normalised: sum_{r,s} A(r,l,s) conj(A(r,l',s)) = delta(l,l')
            M = np.transpose(A, (1, 0, 2)).reshape(A.shape[1], -1, order="F")
            assert np.allclose(M @ M.conj().T, np.eye(A.shape[1]), atol=1e-10)
        v = dense_mps(sites)
        vecs.append(v)
        e_dense = np.vdot(v, H @ v).real / np.vdot(v, v).real
        e_dev = st.expectation(op)
        print("level", k, "eigenvalue", ev[k], "dense <H>", e_dense, "device <H>", e_dev, flush=True)
        assert abs(e_dev - e_dense) < 1e-10 * abs(e_dense)
        assert abs(e_dense - w[k]) < 1e-6 * abs(w[k])
    for i in range(3):
        for j in range(3):
            o_dev = states[i].overlap(states[j])
            o_dense = np.vdot(vecs[i], vecs[j])
            assert abs(o_dev - o_dense) < 1e-12
            if i == j:
                assert abs(o_dev - 1) < 1e-10
            else:
                assert abs(o_dev) < 1e-4  # orthogonal to the chain convergence threshold
    for st in states:
        st.close()


def test_state_round_trip_through_device(mods):
    """A state built on the host, deserialized, contracted on the device."""
    c_api, formats, operators = mods
    rng = np.random.default_rng(17)
    ops = operators.xxz_operator(5, 2, 1.0, 0.6)
    bonds = [1, 2, 4, 4, 2, 1]
    sites = [rng.standard_normal((bonds[i + 1], bonds[i], 2)) + 1j * rng.standard_normal((bonds[i + 1], bonds[i], 2))
             for i in range(5)]
    sites = [np.asfortranarray(a) for a in sites]
    st = c_api.State(formats.encode_state(sites))
    op = c_api.Operator.from_sites(ops)
    v = dense_mps(sites)
    H = dense_hamiltonian(ops)
    assert np.allclose(H, H.T)  # real symmetric, so the dense basis orientation does not matter
    assert abs(st.expectation(op) - np.vdot(v, H @ v)) < 1e-11 * abs(np.vdot(v, H @ v))
    assert abs(st.overlap(st) - np.vdot(v, v)) < 1e-11 * np.vdot(v, v).real
