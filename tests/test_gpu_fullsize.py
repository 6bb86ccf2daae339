"""Parity at the benchmark sizes, not only at the sizes the oracle finishes in
seconds on one thread.

  * rank-28 (4 GiB) random full permutation, bitwise against the oracle's
    permutation split over host threads (the 32-bit byte-offset path of the
    power-of-two permute runs at its limit here), plus the inverse round trip;
  * rank-29 (8 GiB) random permutation, bitwise (the 64-bit offset path);
  * rank-28 2-qubit gate, 3-qubit gate and diagonals on the benchmark's axes
    and on low-bit placements, bitwise against the threaded oracle;
  * the rank-30 depth-20 random circuit (16 GiB): fused runs bitwise equal to
    one launch per gate, and the norm stays 1;
  * the 4096^3 complex contraction of two rank-24 tensors sharing 12 axes:
    FP64 within 1e-12 and split precision within 1e-10 (relative Frobenius
    norm) of an independent FP64 product (numpy transposes, torch matmul).

The threaded oracle functions are pinned bitwise to the one-thread oracle
(tests/test_oracle.py).  Inputs are filled on the device with the seeded
generator and copied to the host for the oracle.
"""
import math
import os
import random
import sys

import numpy as np
import pytest
import torch

import oracle
from qtnh_amd import _lib, ops

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from bench import workload  # noqa: E402  (the benchmark's seeded placements)

SEED = 2024


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield
    torch.cuda.empty_cache()


def threads():
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def dev_equal(a, b):
    return bool(torch.equal(torch.view_as_real(a), torch.view_as_real(b)))


def say(msg):
    print(msg, flush=True)


def inverse(p):
    inv = [0] * len(p)
    for i, v in enumerate(p):
        inv[v] = i
    return inv


@pytest.mark.parametrize("rank", [28, 29])
def test_full_size_permute_bitwise(rank):
    say("full-size permute rank %d" % rank)
    dims = [2] * rank
    if rank == 28:
        ptup = workload(rank, SEED)[0]  # the benchmark's permutation
    else:
        ptup = list(range(rank))
        random.Random(SEED + rank).shuffle(ptup)
    x = ops.fill_(ops.empty(1 << rank), SEED + rank)
    y = ops.permute(x, dims, ptup)  # first call: per-permutation tile tuning
    y2 = ops.permute(x, dims, ptup)  # tuned shape, cached
    assert dev_equal(y, y2)
    del y2
    xh = host(x)
    yh = host(y)
    ref = oracle.permute_mt(xh, dims, ptup, threads())
    assert same_bits(yh, ref), "rank-%d permute differs from the oracle" % rank
    del xh, yh, ref
    z = ops.permute(y, ops.permuted_dims(dims, ptup), inverse(ptup))
    assert dev_equal(z, x), "rank-%d permute round trip" % rank
    del x, y, z
    torch.cuda.empty_cache()


def _gate_cases():
    _, gate_axes, g, diag_axes, d = workload(28, SEED)
    rng = np.random.default_rng(SEED)
    g3 = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    d3 = np.exp(2j * np.pi * rng.random(8))
    return [
        ("gate2-bench-axes", "gate", gate_axes, np.asarray(g)),
        ("gate2-lowest-bits", "gate", [27, 26], np.asarray(g)),
        ("gate2-low-high", "gate", [0, 25], np.asarray(g)),
        ("gate3-low-bits", "gate", [27, 24, 22], g3),
        ("diag2-bench-axes", "diag", diag_axes, np.asarray(d)),
        ("diag3-mixed", "diag", [1, 14, 27], d3),
    ]


def test_full_size_gates_and_diagonals_bitwise():
    rank = 28
    dims = [2] * rank
    x = ops.fill_(ops.empty(1 << rank), SEED)
    state = host(x)
    for name, kind, axes, coef in _gate_cases():
        say("full-size %s on axes %s" % (name, axes))
        if kind == "gate":
            ops.gate_apply_(x, dims, axes, coef)
            oracle.gate_apply_mt(state, dims, axes, coef, threads(), inplace=True)
        else:
            ops.diag_apply_(x, dims, axes, coef)
            oracle.diag_apply_mt(state, dims, axes, coef, threads(), inplace=True)
        assert same_bits(host(x), state), name
    del x, state
    torch.cuda.empty_cache()


def test_full_size_dense_gate_default_placement_bitwise():
    """The benchmark's 2-qubit gate as a DenseTensor contraction with the
    default placement (gate outputs last): the row kernel keeps the
    reference's summation order, so it equals the oracle contraction bitwise;
    checked against the in-place gate followed by the permutation that moves
    the gate's axes last (both pinned to the oracle)."""
    rank = 28
    _, gate_axes, g, _, _ = workload(rank, SEED)
    say("full-size dense gate default placement on axes %s" % gate_axes)
    dims = [2] * rank
    x = ops.fill_(ops.empty(1 << rank), SEED + 1)
    gt = torch.tensor(np.asarray(g, dtype=np.complex128), device="cuda")
    wires = [(gate_axes[0], 0), (gate_axes[1], 1)]
    c, dc, _ = ops.contract(x, dims, gt, [2, 2, 2, 2], wires)
    assert dc == dims
    ops.gate_apply_(x, dims, gate_axes, g)
    # default placement: open state axes in order, then the gate's outputs
    rest = [a for a in range(rank) if a not in gate_axes]
    ptup = [0] * rank
    for pos, a in enumerate(rest + list(gate_axes)):
        ptup[a] = pos
    y = ops.permute(x, dims, ptup)
    assert dev_equal(c, y)
    del x, c, y
    torch.cuda.empty_cache()


def test_full_size_random_circuit_fused_equals_per_gate():
    n, depth = 30, 20
    say("rank-30 depth-20 random circuit: fused vs per gate")
    c = ops.Circuit(n).random(depth, SEED)
    a = ops.fill_(ops.empty(1 << n), SEED)
    ops.scale_(a, 1.0 / math.sqrt(ops.norm2(a)))
    b = a.clone()
    c.run(a, _lib.RUN_PER_GATE)
    c.run(b, _lib.RUN_FUSED)
    torch.cuda.synchronize()
    assert dev_equal(a, b), "fused run differs from per-gate launches"
    nrm = ops.norm2(b)
    assert abs(nrm - 1.0) < 1e-12, nrm
    del a, b
    c.close()
    torch.cuda.empty_cache()


def _reference_contraction(a, da, b, db, wires):
    """Independent FP64 product: numpy transposes both operands to matrices on
    the host, torch multiplies them on the device (library complex GEMM); the
    result is in the default placement (open axes of a in order, then open
    axes of b)."""
    wa = [w[0] for w in wires]
    wb = [w[1] for w in wires]
    oa = [i for i in range(len(da)) if i not in wa]
    ob = [i for i in range(len(db)) if i not in wb]
    m = int(np.prod([da[i] for i in oa]))
    k = int(np.prod([da[i] for i in wa]))
    n = int(np.prod([db[i] for i in ob]))
    ta = np.ascontiguousarray(host(a).reshape(da).transpose(oa + wa)).reshape(m, k)
    tb = np.ascontiguousarray(host(b).reshape(db).transpose(wb + ob)).reshape(k, n)
    return torch.matmul(torch.from_numpy(ta).cuda(), torch.from_numpy(tb).cuda()).reshape(-1)


def rel_err(x, ref):
    return float(((x - ref).norm() / ref.norm()).item())


@pytest.mark.parametrize("layout", ["matrix", "seeded-axes"])
def test_full_size_gemm_fp64_and_split(layout):
    r = 12
    say("4096^3 contraction, %s layout" % layout)
    dims = [2] * (2 * r)
    a = ops.fill_(ops.empty(1 << (2 * r)), SEED)
    b = ops.fill_(ops.empty(1 << (2 * r)), SEED + 1)
    if layout == "matrix":
        wires = [(r + i, i) for i in range(r)]
    else:
        rng = random.Random(SEED + 12)
        wa = rng.sample(range(2 * r), r)
        wb = rng.sample(range(2 * r), r)
        wires = list(zip(wa, wb))
    ref = _reference_contraction(a, dims, b, dims, wires)
    c64, dc, eta64 = ops.contract(a, dims, b, dims, wires, precision=_lib.PREC_FP64)
    assert dc == dims and eta64 == 0.0
    e64 = rel_err(c64, ref)
    assert e64 <= 1e-12, e64
    cs, _, eta = ops.contract(a, dims, b, dims, wires, precision=_lib.PREC_SPLIT)
    es = rel_err(cs, ref)
    say("  fp64 rel err %.3e, split rel err %.3e (eta %.2e)" % (e64, es, eta))
    assert es <= 1e-10, es
    del a, b, ref, c64, cs
    torch.cuda.empty_cache()
