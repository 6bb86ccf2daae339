"""Parity of the gfx950 kernels with the oracle, through the C ABI.

Bit-exact for data movement (permute, swap), seeded fill, gates and diagonal
gates (the kernels keep the reference's summation order and separate
rounding), and for small contractions (direct route).  The GEMM route is
checked at relative error <= 1e-12 (the FP64 target).
"""
import math
import random

import numpy as np
import pytest
import torch

import oracle
from _fixtures import contraction_cases, iden_tensor, swap_tensor
from qtnh_amd import _lib, ops

pytestmark = pytest.mark.gpu

X = _lib.X


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rnd(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


# ------------------------------------------------------------------ fill
@pytest.mark.parametrize("dist", [0, 1])
def test_fill_matches_oracle_bitwise(dist):
    n = 100003
    for seed in (0, 1, 2 ** 63 + 5):
        x = ops.fill_(ops.empty(n), seed, dist)
        assert same_bits(host(x), oracle.fill(n, seed, dist)), (seed, dist)


# ------------------------------------------------------------------ permute
def _perm_cases():
    pyr = random.Random(5)
    cases = []
    for rank in list(range(1, 21)) + [20, 21]:
        dims = [2] * rank
        p = list(range(rank))
        pyr.shuffle(p)
        cases.append((dims, p))
    for _ in range(12):  # power-of-two dims of several widths
        rank = pyr.randint(2, 8)
        dims = [pyr.choice([1, 2, 4, 8, 16]) for _ in range(rank)]
        p = list(range(rank))
        pyr.shuffle(p)
        cases.append((dims, p))
    for _ in range(12):  # general dims
        rank = pyr.randint(1, 7)
        dims = [pyr.choice([1, 2, 3, 5, 6]) for _ in range(rank)]
        p = list(range(rank))
        pyr.shuffle(p)
        cases.append((dims, p))
    # fixed structures: reversal, rotation, identity, swap of the two lowest / highest
    for rank in (12, 18):
        cases.append(([2] * rank, list(range(rank))[::-1]))
        cases.append(([2] * rank, [(i + 1) % rank for i in range(rank)]))
        cases.append(([2] * rank, list(range(rank))))
        cases.append(([2] * rank, list(range(rank - 2)) + [rank - 1, rank - 2]))
        cases.append(([2] * rank, [1, 0] + list(range(2, rank))))
        # low source bits kept in place (the tile grows past the kept bits)
        hi = list(range(rank - 8))
        pyr.shuffle(hi)
        cases.append(([2] * rank, hi + list(range(rank - 8, rank))))
        # lowest bit exchanged with the highest
        sw = list(range(rank))
        sw[0], sw[rank - 1] = rank - 1, 0
        cases.append(([2] * rank, sw))
    return cases


def test_permute_every_tile_shape():
    """Results do not depend on the tile parameters: a random permutation, and
    ones whose natural tile is small (a rotation, the low 8 axes kept), where
    the tile grows or is filled with outer bits."""
    dims = [2] * 16
    rng = random.Random(16)
    p = list(range(16))
    rng.shuffle(p)
    hi = list(range(8))
    rng.shuffle(hi)
    x = rnd(1 << 16, 16)
    perms = [p, [(i + 1) % 16 for i in range(16)], hi + list(range(8, 16))]
    # order + 2: the tile is filled with outer bits instead of longer runs;
    # order + 4: tiles are walked by the outer bits' distance from bit 10
    shapes = [(6, 6, 12, 0, 0, 16), (4, 4, 8, 1, 1, 2), (6, 3, 9, 1, 0, 4), (3, 6, 12, 0, 1, 8), (2, 2, 4, 0, 0, 1),
              (8, 8, 12, 1, 1, 16), (6, 6, 12, 3, 1, 2), (5, 6, 11, 2, 0, 8), (4, 6, 12, 3, 1, 4), (6, 6, 10, 3, 1, 16),
              (6, 6, 12, 4, 1, 2), (5, 5, 11, 6, 1, 8)]
    try:
        for p in perms:
            ref = oracle.permute(x, dims, p)
            for a, b, u, order, nt, bpc in shapes:
                ops.tune_permute(a, b, u, order, nt, bpc)
                assert same_bits(host(ops.permute(dev(x), dims, p)), ref), (p, a, b, u, order, nt, bpc)
    finally:
        ops.tune_permute_auto(True)


def test_permute_tuning_keeps_bits():
    """Per-permutation tuning gives the default tile's bits: on the tuning call
    (candidate shapes timed into dst), on the cached call after it, and in a
    graph captured before the signature was tuned (tuning skipped), replayed
    with eager work and synchronisation between replays."""
    r = 24
    dims = [2] * r
    rng = random.Random(2400)
    rand = list(range(r))
    rng.shuffle(rand)
    low = list(range(r - 8))
    rng.shuffle(low)
    swap = list(range(r))
    swap[0], swap[r - 1] = r - 1, 0
    perms = [rand, list(range(r))[::-1], low + list(range(r - 8, r)), swap]
    x = ops.fill_(ops.empty(1 << r), 31)
    try:
        ops.tune_permute_auto(False)
        refs = [ops.permute(x, dims, p) for p in perms]
        ops.tune_permute_auto(True)
        for p, ref in zip(perms, refs):
            for rep in range(2):
                y = ops.permute(x, dims, p)
                torch.cuda.synchronize()
                assert torch.equal(torch.view_as_real(y), torch.view_as_real(ref)), (p, rep)
        ops.tune_permute_auto(True)  # forget the tuned shapes
        y = ops.empty(1 << r)
        s = torch.cuda.Stream()
        ops.context(stream=s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            ops.permute(x, dims, perms[0], out=y)
        for rep in range(3):
            ops.fill_(y, 5)
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(torch.view_as_real(y), torch.view_as_real(refs[0])), rep
    finally:
        ops.tune_permute_auto(True)


@pytest.mark.parametrize("case", _perm_cases(), ids=lambda c: "%s-%s" % (c[0][:3], c[1][:3]))
def test_permute_bitexact(case):
    dims, p = case
    n = int(np.prod(dims))
    x = rnd(n, n + len(dims))
    y = ops.permute(dev(x), dims, p)
    assert same_bits(host(y), oracle.permute(x, dims, p))


def test_permute_roundtrip_rank24():
    dims = [2] * 24
    p = list(range(24))
    random.Random(24).shuffle(p)
    inv = list(np.argsort(p))
    x = ops.fill_(ops.empty(1 << 24), 77)
    y = ops.permute(x, dims, p)
    z = ops.permute(y, ops.permuted_dims(dims, p), inv)
    torch.cuda.synchronize()
    assert torch.equal(torch.view_as_real(x), torch.view_as_real(z))


def test_permute_rejects_bad_tuple():
    """Refused by the Python front end, and by the C ABI itself."""
    x = ops.empty(8)
    y = ops.empty(8)
    with pytest.raises(ValueError):
        ops.permute(x, [2, 2, 2], [0, 1, 1])
    d, dp = ops._u64([2, 2, 2])
    p, pp = ops._u16([0, 1, 1])
    ctx = ops.context()
    st = _lib.lib().qtnh_permute_c128(ctx, ops._ptr(x), ops._ptr(y), 3, dp, pp)
    assert st == 1, st
    assert b"permutation" in _lib.lib().qtnh_ctx_last_error(ctx)


# ------------------------------------------------------------------ swap
@pytest.mark.parametrize("dims,i,j", [([2] * 10, 0, 9), ([2] * 10, 3, 4), ([2] * 12, 11, 0), ([2] * 3, 1, 2),
                                      ([4, 2, 4, 2], 0, 2), ([3, 2, 3], 0, 2), ([2, 5, 5, 2], 1, 2),
                                      ([2, 2], 1, 1)])
def test_swap_bitexact(dims, i, j):
    n = int(np.prod(dims))
    x = rnd(n, 3)
    y = ops.swap_(dev(x), dims, i, j)
    assert same_bits(host(y), oracle.swap(x, dims, i, j))


def test_swap_rejects_unequal_dims():
    with pytest.raises(_lib.QtnhError):
        ops.swap_(ops.empty(6), [2, 3], 0, 1)


# ------------------------------------------------------------------ gates
# Placements that reach every gate kernel shape: all gate bits among the lowest
# 6 (lane exchange, registers; 3-qubit: LDS-staged tiles, in any bit order),
# some low and some high (lane exchange, LDS coefficients, one or two high
# bits), all high (one thread per group).
_GATE_PLACEMENTS = {1: [[13], [0]], 2: [[13, 12], [13, 0], [0, 13], [1, 2]],
                    3: [[13, 12, 11], [13, 12, 0], [13, 0, 5], [0, 13, 11], [0, 1, 2], [7, 13, 3],
                        [13, 9, 11], [12, 13, 8]]}


@pytest.mark.parametrize("k", [1, 2, 3])
def test_gate_bitexact_binary(k):
    pyr = random.Random(k)
    rank = 14
    dims = [2] * rank
    fixed = _GATE_PLACEMENTS[k]
    for trial in range(6 + len(fixed)):
        axes = pyr.sample(range(rank), k)
        if trial == 0:
            axes = list(range(rank - k, rank))  # lowest bits
        if trial == 1:
            axes = list(range(k))               # highest bits
        if trial >= 6:
            axes = fixed[trial - 6]
        x = rnd(1 << rank, trial)
        g = rnd(4 ** k, 100 + trial)
        y = ops.gate_apply_(dev(x), dims, axes, g)
        assert same_bits(host(y), oracle.gate_apply(x, dims, axes, g)), (k, axes)


@pytest.mark.parametrize("dims,axes", [([4, 2, 4], [2, 0]), ([2, 3, 3], [1, 2]), ([3, 3, 3, 2], [0]),
                                       ([2] * 10, [1, 5, 7, 8]), ([2] * 8, [0, 2, 4, 6, 7])])
def test_gate_bitexact_general(dims, axes):
    n = int(np.prod(dims))
    K = int(np.prod([dims[a] for a in axes]))
    x = rnd(n, 9)
    g = rnd(K * K, 10)
    y = ops.gate_apply_(dev(x), dims, axes, g)
    assert same_bits(host(y), oracle.gate_apply(x, dims, axes, g))


@pytest.mark.parametrize("dims,axes", [([2] * 12, [3]), ([2] * 12, [11, 0]), ([2] * 12, [5, 6, 1]),
                                       ([2] * 12, [0, 1, 2, 3, 4, 5]), ([3, 2, 3], [0, 2]), ([4, 2, 2], [0, 2])])
def test_diag_bitexact(dims, axes):
    n = int(np.prod(dims))
    K = int(np.prod([dims[a] for a in axes]))
    x = rnd(n, 4)
    d = np.exp(1j * np.random.default_rng(1).uniform(0, 2 * np.pi, K))
    y = ops.diag_apply_(dev(x), dims, axes, d)
    assert same_bits(host(y), oracle.diag_apply(x, dims, axes, d))


def test_gate_unitary_roundtrip_rank22():
    dims = [2] * 22
    x = ops.fill_(ops.empty(1 << 22), 5)
    x0 = x.clone()
    q, _ = np.linalg.qr(rnd(16, 8).reshape(4, 4))
    g = q.T.reshape(-1)                   # [in, out] = q^T: applies q
    ops.gate_apply_(x, dims, [3, 17], g)
    ops.gate_apply_(x, dims, [3, 17], q.conj().reshape(-1))  # applies q^dagger
    torch.cuda.synchronize()
    err = (x - x0).norm().item() / x0.norm().item()
    assert err < 1e-14


# ------------------------------------------------------------------ contraction: golden vectors
def _check_case(t1, d1, t2, d2, t3, d3, wires, repls=None):
    out, dc, _ = ops.contract(dev(t1), d1, dev(t2), d2, wires, *(repls or (None, None)))
    got = host(out)
    assert dc == d3
    if repls is None:
        repls = oracle.default_repls(len(d1), 0, len(d2), 0, wires)
    ref, _ = oracle.contract(t1, d1, t2, d2, wires, *repls)
    assert same_bits(got, ref)
    np.testing.assert_allclose(got, t3, rtol=0, atol=1e-12)


@pytest.mark.parametrize("case", list(contraction_cases("dense_vals")), ids=lambda c: c[0])
def test_contract_dense_vals(case):
    name, t1, d1, t2, d2, t3, d3, wires = case
    _check_case(t1, d1, t2, d2, t3, d3, wires)


@pytest.mark.parametrize("case", list(contraction_cases("swap_vals")), ids=lambda c: c[0])
def test_contract_swap_vals(case):
    name, t1, d1, _, d2, t3, d3, wires = case
    n = d2[0]
    _check_case(t1, d1, swap_tensor(n), [n] * 4, t3, d3, wires, ops.symm_repls(len(d1), 0, 4, 0, wires))


@pytest.mark.parametrize("case", list(contraction_cases("id_vals")), ids=lambda c: c[0])
def test_contract_id_vals(case):
    name, t1, d1, _, d2, t3, d3, wires = case
    k = len(d2) // 2
    _check_case(t1, d1, iden_tensor(d2[:k]), d2, t3, d3, wires, ops.symm_repls(len(d1), 0, len(d2), 0, wires))


@pytest.mark.parametrize("group,nd1,nd2", [("mpi2r_vals", 1, 0), ("mpi3r_vals", 1, 0), ("mpi4r_vals", 1, 1),
                                           ("mpi6r_vals", 1, 1), ("mpi8r_vals", 2, 1)])
def test_contract_distributed_goldens(group, nd1, nd2):
    for name, t1, d1, t2, d2, t3, d3, wires in contraction_cases(group):
        r = ops.default_repls(len(d1), nd1, len(d2), nd2, wires)
        _check_case(t1, d1, t2, d2, t3, d3, wires, r)


# ------------------------------------------------------------------ contraction: small dense gate, outputs last
@pytest.mark.parametrize("da,db,wires", [
    ([2] * 16, [2, 2, 2, 2], [(15, 0), (3, 1)]),             # 2-qubit gate: 4 terms, 4 results per row
    ([2] * 15, [2, 2], [(0, 0)]),                             # 1-qubit gate on the top axis
    ([2] * 16, [2] * 5, [(2, 0), (9, 1), (11, 2), (4, 3)]),   # 16 terms, 2 results per row
    ([2] * 16, [2] * 5, [(14, 0)]),                           # 2 terms, 16 results per row
    ([2] * 16, [2] * 5, [(1, 0), (12, 1), (6, 2)]),           # 8 terms, 4 results per row
    ([4] * 8, [4, 4], [(5, 0)]),                              # dim-4 axes: 4 terms, 4 results
    ([4] * 7 + [2], [2, 4, 2, 2], [(7, 0), (2, 1)]),           # mixed widths: 8 terms, 4 results
])
def test_contract_small_gate_outputs_last_bitexact(da, db, wires):
    """Default placement puts the gate's open axes last: the row kernel (one
    thread per row of gate results) must equal the oracle bit for bit."""
    a = rnd(int(np.prod(da)), 11)
    b = rnd(int(np.prod(db)), 12)
    out, dc, _ = ops.contract(dev(a), da, dev(b), db, wires)
    ref, rdc = oracle.contract(a, da, b, db, wires)
    assert dc == rdc
    assert same_bits(host(out), ref)


_ROW_SHAPES = [(nb, kb) for nb in range(1, 5) for kb in range(1, 6 - nb)]


@pytest.mark.parametrize("nb,kb", _ROW_SHAPES)
def test_contract_row_kernel_every_shape_bitexact(nb, kb):
    """Every (result bits, term bits) shape of the row kernel: staged reads
    (up to 8 terms) or direct ones (16), one or two rows per thread, one or two
    result passes; contracted axes scattered over a rank-16 state (>= 12 row bits)."""
    rng = np.random.default_rng(100 * nb + kb)
    da = [2] * 16
    db = [2] * (kb + nb)
    axes = rng.choice(16, size=kb, replace=False)
    wires = [(int(a), i) for i, a in enumerate(axes)]
    a = rnd(1 << 16, 31 + nb)
    b = rnd(1 << (kb + nb), 41 + kb)
    out, dc, _ = ops.contract(dev(a), da, dev(b), db, wires)
    ref, rdc = oracle.contract(a, da, b, db, wires)
    assert dc == rdc
    assert same_bits(host(out), ref)


@pytest.mark.parametrize("rank,axes,flip", [
    (16, [15, 14], False),  # the state's two lowest bits
    (17, [0, 1], False),    # its two highest
    (18, [3, 12], True),    # mixed, the wires given in the other order (other term order)
    (20, [19, 0], True),    # lowest and highest
    (19, [6, 7], False),
    (16, [7], False),       # 1-qubit gates
    (19, [18], False),
    (19, [0], False),
    (16, [15, 14, 13], False),  # 3-qubit gates
    (18, [0, 9, 17], True),
    (17, [2, 11, 5], False),
])
def test_contract_dense_gate_fused_permute_bitexact(rank, axes, flip):
    """A square dense gate (1 to 3 term bits) contracted into a state of at
    least 2^16 elements with the default placement runs as the bit permutation
    that moves the gate's axes last, each row of results formed in the
    permute's tiles: equal to the oracle bit for bit."""
    k = len(axes)
    da, db = [2] * rank, [2] * (2 * k)
    wires = [(a, i) for i, a in enumerate(axes)]
    if flip:
        wires = wires[::-1]
    a = rnd(1 << rank, 60 + rank)
    b = rnd(1 << (2 * k), 70 + k)
    out, dc, _ = ops.contract(dev(a), da, dev(b), db, wires)
    ref, rdc = oracle.contract(a, da, b, db, wires)
    assert dc == rdc
    assert same_bits(host(out), ref)


def test_contract_small_gate_custom_placement_bitexact():
    """Gate outputs placed below the state's open axes too (not only last)."""
    da, db = [2] * 15, [2, 2, 2, 2]
    wires = [(4, 0), (10, 1)]
    # state open axes keep their order at 0..12; gate outputs go to 14 and 13 (swapped)
    open1 = [i for i in range(15) if i not in (4, 10)]
    r1 = [X] * 15
    for k, i in enumerate(open1):
        r1[i] = k
    r2 = [X, X, 14, 13]
    a = rnd(1 << 15, 21)
    b = rnd(16, 22)
    out, dc, _ = ops.contract(dev(a), da, dev(b), db, wires, r1, r2)
    ref, _ = oracle.contract(a, da, b, db, wires, r1, r2)
    assert same_bits(host(out), ref)


# ------------------------------------------------------------------ contraction: GEMM route
@pytest.mark.parametrize("da,db,wires", [
    ([2] * 16, [2] * 16, [(i + 8, i) for i in range(8)]),                          # 256x256x256, no permutes
    ([2] * 16, [2] * 16, [(0, 15), (3, 2), (9, 7), (12, 0), (5, 9), (14, 11), (7, 4), (1, 13)]),
    ([3, 5, 7, 4, 9], [4, 9, 6, 11], [(3, 0), (4, 1)]),                             # odd sizes, tile edges
    ([8, 16, 32], [32, 16, 40], [(2, 0), (1, 1)]),
    ([2] * 20, [2] * 21, [(7 + i, i) for i in range(13)]),                        # 128x256x8192: 16 depth chunks
    ([2] * 20, [2] * 20, [(8 + i, 11 - i) for i in range(12)]),                   # 256x256x4096: 8 chunks, permuted B
])
def test_contract_gemm_route(da, db, wires):
    a = rnd(int(np.prod(da)), 1)
    b = rnd(int(np.prod(db)), 2)
    out, dc, _ = ops.contract(dev(a), da, dev(b), db, wires)
    ref, rdc = oracle.contract(a, da, b, db, wires)
    got = host(out)
    assert dc == rdc
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-12, err


def test_contract_gemm_route_placement():
    """GEMM route with a non-natural placement: the output permute follows repl1/repl2."""
    da, db = [2] * 14, [2] * 14
    wires = [(i, 13 - i) for i in range(7)]
    rng = random.Random(3)
    perm = list(range(14))
    rng.shuffle(perm)
    # wires use t1 axes 0..6 and t2 axes 13..7; open t1 axes 7..13, open t2 axes 0..6
    r1 = [X] * 7 + perm[:7]
    r2 = perm[7:] + [X] * 7
    a = rnd(1 << 14, 5)
    b = rnd(1 << 14, 6)
    out, dc, _ = ops.contract(dev(a), da, dev(b), db, wires, r1, r2)
    ref, _ = oracle.contract(a, da, b, db, wires, r1, r2)
    err = np.linalg.norm(host(out) - ref) / np.linalg.norm(ref)
    assert err < 1e-12, err


@pytest.mark.parametrize("da,db,wires", [
    ([2] * 16, [2] * 16, [(i, i) for i in range(16)]),                   # inner product to a scalar, K = 65536
    ([2] * 15, [2] * 15, [(i, 14 - i) for i in range(15)]),              # reversed wires: permutes before the sums
    ([4, 2] + [2] * 12, [2] * 12 + [3, 2], [(2 + i, i) for i in range(12)]),  # 8 x 6 outputs, K = 4096
    ([5, 7, 9, 11, 13], [9, 13, 11, 2], [(2, 0), (3, 2), (4, 1)]),        # ragged dims, K = 1287
])
def test_contract_few_outputs_long_sums(da, db, wires):
    """Few outputs over long sums (split into ordered chunks): within 1e-12 of the oracle."""
    a = rnd(int(np.prod(da)), 11)
    b = rnd(int(np.prod(db)), 12)
    out, dc, _ = ops.contract(dev(a), da, dev(b), db, wires)
    ref, rdc = oracle.contract(a, da, b, db, wires)
    assert dc == rdc
    got = host(out)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-12, err
    # deterministic: the same call gives the same bits
    out2, _, _ = ops.contract(dev(a), da, dev(b), db, wires)
    assert same_bits(host(out2), got)


def test_kernels_are_ordered_on_the_torch_stream():
    """Results are visible to torch work queued after them on the same stream,
    with no explicit synchronisation (the context runs on torch's stream)."""
    dims = [2] * 20
    p = list(range(20))[::-1]
    for _ in range(3):
        x = ops.fill_(ops.empty(1 << 20), 9)
        y = ops.permute(x, dims, p)
        z = ops.permute(y, dims, p)
        assert torch.equal(torch.view_as_real(z), torch.view_as_real(x))  # no synchronize() before this
        ops.gate_apply_(x, dims, [19], [0, 1, 1, 0])  # X gate on the last axis
        w = x.view(-1, 2).flip(1).reshape(-1)
        assert torch.equal(torch.view_as_real(w), torch.view_as_real(z))


def test_norm2_and_scale():
    x = rnd(12345, 3)
    t = dev(x)
    assert abs(ops.norm2(t) - np.vdot(x, x).real) < 1e-9
    ops.scale_(t, 0.5 - 2j)
    y = host(t)
    exp = (x.real * 0.5 - x.imag * -2.0) + 1j * (x.real * -2.0 + x.imag * 0.5)
    assert same_bits(y, exp)
