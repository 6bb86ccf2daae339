"""Host-side circuit construction (no GPU): generators, read-back, validation.

The QFT op sequence follows the reference's construction (tests/gen/qft.hpp:47-108:
for each qubit i, H on i, then the controlled phase pi/2^(j-i) on (i, j) for
j > i).  Random layers are Haar SU(2) gates (unitary, det 1) on every qubit and
CZ on a perfect matching of the qubits.
"""
import numpy as np
import pytest

from qtnh_amd import _lib, ops

H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)


def test_qft_op_sequence_matches_reference_construction():
    n = 6
    phases = np.exp(1j * np.pi / 2.0 ** np.arange(n))
    c = ops.Circuit(n).qft(phases)
    got = c.ops()
    expect = []
    for i in range(n):
        expect.append((0, [i], H.reshape(-1)))
        for j in range(i + 1, n):
            expect.append((1, [i, j], np.array([1, 1, 1, phases[j - i]])))
    assert len(got) == len(expect) == n * (n + 1) // 2
    for (k1, q1, c1), (k2, q2, c2) in zip(got, expect):
        assert k1 == k2 and q1 == q2
        assert np.allclose(c1, c2, atol=0, rtol=2 ** -52)


def test_qft_default_phases_are_the_controlled_phases():
    n = 5
    got = ops.Circuit(n).qft().ops()
    for kind, q, co in got:
        if kind == 1:
            m = q[1] - q[0]
            assert abs(co[3] - np.exp(1j * np.pi / 2 ** m)) < 1e-16


@pytest.mark.parametrize("n,depth,seed", [(1, 3, 0), (7, 4, 11), (12, 2, 2 ** 63 + 1)])
def test_random_layers_are_su2_and_cz_matchings(n, depth, seed):
    c = ops.Circuit(n).random(depth, seed)
    got = c.ops()
    per_layer = n + n // 2
    assert len(got) == depth * per_layer
    for layer in range(depth):
        lo = got[layer * per_layer:(layer + 1) * per_layer]
        for q in range(n):
            kind, qs, co = lo[q]
            assert kind == 0 and qs == [q]
            U = co.reshape(2, 2).T  # G[in][out] = U[out][in]
            assert np.abs(U.conj().T @ U - np.eye(2)).max() < 1e-14
            assert abs(np.linalg.det(U) - 1) < 1e-14
        seen = set()
        for kind, qs, co in lo[n:]:
            assert kind == 1 and len(qs) == 2
            assert np.array_equal(co, [1, 1, 1, -1])
            seen.update(qs)
        assert len(seen) == 2 * (n // 2)


def test_random_is_deterministic_and_seed_sensitive():
    a = ops.Circuit(6).random(3, 5).ops()
    b = ops.Circuit(6).random(3, 5).ops()
    c = ops.Circuit(6).random(3, 6).ops()
    assert all(x[1] == y[1] and np.array_equal(x[2], y[2]) for x, y in zip(a, b))
    assert any(not np.array_equal(x[2], y[2]) for x, y in zip(a, c))


def test_haar_first_moment():
    # E|U_00|^2 = 1/2 for Haar SU(2)
    c = ops.Circuit(1).random(4000, 3)
    v = np.array([abs(co[0]) ** 2 for kind, q, co in c.ops()])
    assert abs(v.mean() - 0.5) < 0.02


def test_runs_group_ops_by_tile_bits():
    c = ops.Circuit(20).random(2, 1)
    ops.tune_circuit(12)
    n, nops, runs12 = c.info()
    ops.tune_circuit(4)
    _, _, runs4 = c.info()
    ops.tune_circuit(12)
    assert nops == 2 * 30
    assert runs12 < runs4 <= nops
    small = ops.Circuit(10).random(5, 1)
    assert small.info()[2] == 1  # a state of <= tile bits is one tile: one run


@pytest.mark.parametrize("qubits", [[5], [0, 0], [0, 1, 2, 3], []])
def test_invalid_gates_rejected(qubits):
    c = ops.Circuit(5)
    with pytest.raises((_lib.QtnhError, ValueError)):
        c.gate(qubits, np.eye(1 << len(qubits)))


def test_invalid_sizes_rejected():
    with pytest.raises(_lib.QtnhError):
        ops.Circuit(0)
    with pytest.raises(_lib.QtnhError):
        ops.Circuit(63)
    c = ops.Circuit(3)
    with pytest.raises(ValueError):
        c.diag([0, 1], [1, 1, 1])
