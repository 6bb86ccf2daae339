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


def _fewest_runs(n, op_list, tb, low_bits):
    """Dynamic programme over contiguous segments: fewest runs whose gate
    qubits (with the low tile bits) fit tb bits; diagonals need no tile bit."""
    tbe = min(tb, n)
    low = (1 << min(low_bits, max(tbe - 3, 0), n)) - 1
    masks = []
    for kind, q, _ in op_list:
        m = 0
        if kind == 0:
            for x in q:
                m |= 1 << (n - 1 - x)
        masks.append(m)
    inf = float("inf")
    best = [0] + [inf] * len(masks)
    for end in range(1, len(masks) + 1):
        m = low
        for start in range(end - 1, -1, -1):
            m |= masks[start]
            if bin(m).count("1") > tbe:
                break
            best[end] = min(best[end], best[start] + 1)
    return best[-1]


@pytest.mark.parametrize("n,depth,tb", [(20, 3, 12), (24, 4, 12), (18, 3, 8), (22, 2, 13)])
def test_runs_are_fewest_contiguous_segments(n, depth, tb):
    """The planner's run count equals the optimum over contiguous partitions,
    with diagonals (CZ layers, extra 3-qubit diagonals) costing no tile bits."""
    c = ops.Circuit(n).random(depth, n + tb)
    rng = np.random.default_rng(n)
    for _ in range(10):
        qs = [int(v) for v in rng.choice(n, 3, replace=False)]
        c.diag(qs, np.exp(1j * rng.standard_normal(8)))
    ops.tune_circuit(tb, 4)
    try:
        assert c.info()[2] == _fewest_runs(n, c.ops(), tb, 4)
    finally:
        ops.tune_circuit(12, 4)
    only_diag = ops.Circuit(n)
    for q in range(n - 1):
        only_diag.diag([q, q + 1], [1, 1, 1, -1])
    assert only_diag.info()[2] == 1


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
