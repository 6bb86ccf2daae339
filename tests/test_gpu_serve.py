"""Serving loop: answers are independent of policy, batch and cache, and the
batched policy admits arrivals while other circuits are in flight.

Every answer must equal the same query computed alone (circuit applied to
|0...0> with one launch per gate, then the amplitude or <Z>), bit for bit:
the batched policy uses fused runs, which do the same arithmetic in the same
order; resident states are the same buffers.
"""
import numpy as np
import pytest
import torch

from qtnh_amd import _lib, ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def solo(q):
    n = q["n_qubits"]
    c = ops.Circuit(n).random(q["depth"], q["seed"])
    s = torch.zeros(1 << n, dtype=torch.complex128, device="cuda")
    s[0] = 1
    c.run(s, _lib.RUN_PER_GATE)
    torch.cuda.synchronize()
    if q["kind"] == _lib.QUERY_AMPLITUDE:
        v = s[q["basis"]].item()
        return (v.real, v.imag), None
    z = ops.expect_z(s, n, list(range(n)))
    return (z[0], z[-1] if n > 1 else 0.0), z


def check_answers(queries, results):
    for q, r in zip(queries, results):
        want, z = solo(q)
        assert np.array_equal(np.array(r["value"]).view(np.uint64), np.array(want).view(np.uint64)), (q, r)
        if z is not None:
            assert all(abs(v) <= 1 for v in z)
        assert r["start_s"] >= q["arrival_s"] - 1e-9
        assert r["done_s"] >= r["start_s"]


@pytest.mark.parametrize("policy", [_lib.SERVE_SEQUENTIAL, _lib.SERVE_BATCHED])
def test_answers_equal_solo_runs(policy):
    qs = ops.poisson_queries(24, rate=2000.0, seed=3, qubits=(3, 14), depth=(2, 8))
    rs = ops.serve(qs, budget_bytes=1 << 30, policy=policy)
    check_answers(qs, rs)


def test_resident_states_answer_repeats():
    qs = ops.poisson_queries(40, rate=0, seed=5, qubits=(4, 12), depth=(3, 6), pool=6)
    rs = ops.serve(qs, budget_bytes=1 << 30, policy=_lib.SERVE_BATCHED)
    check_answers(qs, rs)
    # all arrive at once and fit: all admitted at step 0, repeats of a spec share its state
    assert {r["batch"] for r in rs} == {0}


def test_resident_states_across_rounds():
    qs = ops.poisson_queries(30, rate=300.0, seed=9, qubits=(6, 12), depth=(3, 5), pool=3)
    rs = ops.serve(qs, budget_bytes=1 << 30, policy=_lib.SERVE_BATCHED)
    check_answers(qs, rs)
    assert any(r["cached"] for r in rs)


def test_small_budget_forces_rounds_and_evictions():
    qs = ops.poisson_queries(20, rate=0, seed=11, qubits=(12, 14), depth=(2, 4))
    budget = (16 << 14) * 2  # two of the largest states at a time
    rs = ops.serve(qs, budget_bytes=budget, policy=_lib.SERVE_BATCHED)
    check_answers(qs, rs)
    assert len({r["batch"] for r in rs}) >= 4


def test_continuous_batching_admits_while_a_large_circuit_runs():
    """A 26-qubit circuit (about 80 fused runs, one per step) arrives first;
    small queries arriving while it is on the GPU are admitted at later steps,
    run beside it and finish before it; every answer equals the query alone."""
    big = {"n_qubits": 26, "depth": 30, "seed": 77, "kind": _lib.QUERY_EXPECT_Z, "basis": 0, "arrival_s": 0.0}
    small = ops.poisson_queries(12, rate=0, seed=21, qubits=(6, 12), depth=(4, 10))
    for i, q in enumerate(small):
        q["arrival_s"] = 0.001 + 0.0005 * i
    qs = [big] + small
    rs = ops.serve(qs, budget_bytes=2 << 30, policy=_lib.SERVE_BATCHED)
    check_answers(qs, rs)
    b = rs[0]
    assert b["batch"] == 0
    for r in rs[1:]:
        assert r["batch"] > 0
        assert r["start_s"] < b["done_s"]
        assert r["done_s"] < b["done_s"]


def test_oversized_query_rejected():
    qs = [{"n_qubits": 20, "depth": 1, "seed": 1, "kind": 0, "basis": 0, "arrival_s": 0.0}]
    with pytest.raises(_lib.QtnhError):
        ops.serve(qs, budget_bytes=1 << 20)
