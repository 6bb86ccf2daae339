"""Circuits on resident states, through the C ABI, against the oracle.

Both executors (one launch per gate, and fused tiles of up to 13 qubits in
LDS) must equal the oracle's gate-by-gate application bit for bit: every op
does the reference's arithmetic (sum over inputs in row-major order from
zero, products without FMA) on the same values in the same order.
<Z_q> is checked against numpy brute force; large states through norm
preservation and fused == per-gate.
"""
import numpy as np
import pytest
import torch

import oracle
from qtnh_amd import _lib, ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield
    ops.tune_circuit(12)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def rnd_state(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return x / np.linalg.norm(x)


def oracle_run(circuit, x):
    x = np.array(x, dtype=np.complex128, copy=True)
    dims = [2] * circuit.n
    for kind, q, co in circuit.ops():
        if kind == 0:
            oracle.gate_apply(x, dims, q, co, inplace=True)
        else:
            oracle.diag_apply(x, dims, q, co, inplace=True)
    return x


@pytest.mark.parametrize("n", [1, 3, 8, 12, 14, 16])
@pytest.mark.parametrize("mode", [_lib.RUN_FUSED, _lib.RUN_PER_GATE])
def test_qft_bitexact(n, mode):
    phases = np.exp(1j * np.pi / 2.0 ** np.arange(n + 1))
    x = rnd_state(n, n)
    ref = oracle.qft(x, n, phases)
    c = ops.Circuit(n).qft(phases)
    s = dev(x)
    c.run(s, mode)
    assert same_bits(host(s), ref)


def test_qft_plus_state_known_answer():
    # QFT of |+>^n is |0...0> (reference tests/gen/qft.hpp, c2-network.cpp:155-198)
    n = 12
    s = dev(np.full(1 << n, 2.0 ** (-n / 2), dtype=np.complex128))
    ops.Circuit(n).qft().run(s)
    out = host(s)
    assert abs(out[0] - 1) < 1e-12
    assert np.abs(out[1:]).max() < 1e-12


@pytest.mark.parametrize("n,depth", [(5, 6), (13, 3), (17, 2), (20, 2)])
def test_random_circuit_bitexact_both_modes(n, depth):
    c = ops.Circuit(n).random(depth, 100 + n)
    x = rnd_state(n, 7)
    ref = oracle_run(c, x)
    for mode in (_lib.RUN_FUSED, _lib.RUN_PER_GATE):
        s = dev(x)
        c.run(s, mode)
        assert same_bits(host(s), ref), mode


@pytest.mark.parametrize("tb", [3, 5, 8, 10, 11, 12, 13])
def test_every_tile_size_bitexact(tb):
    n = 15
    c = ops.Circuit(n).random(2, 5)
    # also 3-qubit dense gates and 3-qubit diagonals
    rng = np.random.default_rng(tb)
    g3 = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    c.gate([14, 0, 7], g3).diag([3, 12, 1], np.exp(1j * rng.standard_normal(8))).gate([2, 9], rng.standard_normal(16))
    x = rnd_state(n, 9)
    ref = oracle_run(c, x)
    ops.tune_circuit(tb)
    try:
        s = dev(x)
        c.run(s)
        assert same_bits(host(s), ref)
    finally:
        ops.tune_circuit(12)


@pytest.mark.parametrize("mode", [3, 2, 1, 0])
def test_large_state_kernels_bitexact(mode):
    """22 qubits: 1024 tiles of 12 bits.  Register-blocked with the next tile
    in flight (3), register-blocked (2), pipelined (1) and plain (0) fused
    kernels all equal the oracle; the circuit mixes 1-, 2- and 3-qubit gates
    and diagonals on low and high qubits."""
    n = 22
    c = ops.Circuit(n).random(2, 2222)
    rng = np.random.default_rng(mode)
    g3 = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    c.gate([21, 20, 19], g3).diag([0, 21, 10], np.exp(1j * rng.standard_normal(8))).gate([20, 3], rng.standard_normal(16))
    c.gate([21], [0, 1, 1, 0]).diag([18, 21], [1, 1j, -1, -1j])
    x = rnd_state(n, 22)
    ref = oracle_run(c, x)
    ops.tune_circuit_pipeline(mode)
    try:
        s = dev(x)
        c.run(s)
        assert same_bits(host(s), ref), mode
    finally:
        ops.tune_circuit_pipeline(2)


@pytest.mark.parametrize("mode,tb", [(3, 12), (2, 12), (1, 12), (0, 12), (0, 6)])
def test_diagonals_outside_the_tile_bitexact(mode, tb):
    """Diagonals on qubits outside a run's tile take the coefficient block
    those qubits select per tile.  Random unequal coefficients (so a wrong
    index order or outer bit shows) on 1-3 qubits mixing tile and outer ones,
    between layers of dense gates; every large-state kernel and a 6-bit tile."""
    n = 22
    rng = np.random.default_rng(10 * mode + tb)
    c = ops.Circuit(n)
    for layer in range(2):
        for q in range(n):
            c.gate([q], rng.standard_normal(4) + 1j * rng.standard_normal(4))
        for _ in range(10):
            k = int(rng.integers(1, 4))
            qs = [int(v) for v in rng.choice(n, k, replace=False)]
            c.diag(qs, np.exp(1j * rng.standard_normal(1 << k)) * (1 + 0.25 * rng.standard_normal(1 << k)))
    x = rnd_state(n, 23)
    ref = oracle_run(c, x)
    ops.tune_circuit(tb)
    ops.tune_circuit_pipeline(mode)
    try:
        s_ = dev(x)
        c.run(s_)
        assert same_bits(host(s_), ref), (mode, tb)
    finally:
        ops.tune_circuit(12)
        ops.tune_circuit_pipeline(2)


def test_batch_equals_individual_runs():
    sizes = [3, 9, 12, 13, 14, 16, 12, 5]
    circuits = [ops.Circuit(n).random(3, 40 + i) for i, n in enumerate(sizes)]
    xs = [rnd_state(n, 60 + i) for i, n in enumerate(sizes)]
    states = [dev(x) for x in xs]
    ops.run_batch(circuits, states)
    for c, x, s in zip(circuits, xs, states):
        assert same_bits(host(s), oracle_run(c, x))


def test_batch_rejects_aliased_states():
    c = ops.Circuit(4).random(1, 1)
    s = dev(rnd_state(4, 1))
    with pytest.raises(_lib.QtnhError):
        ops.run_batch([c, c], [s, s])


@pytest.mark.parametrize("n", [1, 4, 11, 12, 13, 18])
def test_expect_z_matches_brute_force(n):
    x = rnd_state(n, 3 * n)
    qs = list(range(n))
    got = ops.expect_z(dev(x), n, qs)
    p = np.abs(x) ** 2
    idx = np.arange(1 << n)
    for q, v in zip(qs, got):
        bit = (idx >> (n - 1 - q)) & 1
        assert abs(v - np.sum(p * (1 - 2 * bit))) < 1e-12


def test_expect_z_basis_states_exact():
    n = 14
    for x0 in (0, 1, 5000, (1 << n) - 1):
        e = np.zeros(1 << n, dtype=np.complex128)
        e[x0] = 1
        got = ops.expect_z(dev(e), n, list(range(n)))
        want = [1.0 - 2.0 * ((x0 >> (n - 1 - q)) & 1) for q in range(n)]
        assert got == want


def test_large_random_circuit_keeps_norm_and_modes_agree():
    n = 24
    c = ops.Circuit(n).random(2, 77)
    a = ops.fill_(ops.empty(1 << n), 5)
    ops.scale_(a, 1 / np.sqrt(ops.norm2(a)))
    b = a.clone()
    c.run(a, _lib.RUN_FUSED)
    c.run(b, _lib.RUN_PER_GATE)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.float64), b.view(torch.float64))
    assert abs(ops.norm2(a) - 1) < 1e-12
    z = ops.expect_z(a, n, [0, n // 2, n - 1])
    assert all(abs(v) <= 1 for v in z)


@pytest.mark.parametrize("kind,n", [("qft", 12), ("random", 12), ("random", 16)])
def test_circuit_tune_keeps_results(kind, n):
    """A circuit tuned to its own fastest tile gives the per-gate executor's bits."""
    c = ops.Circuit(n).qft() if kind == "qft" else ops.Circuit(n).random(6, 3)
    x = rnd_state(n, 40 + n)
    want = dev(x)
    c.run(want, _lib.RUN_PER_GATE)
    tb, lb, us = c.tune()
    assert 7 <= tb <= 12 and lb in (2, 4) and us > 0
    got = dev(x)
    c.run(got, _lib.RUN_FUSED)
    assert same_bits(host(got), host(want))
