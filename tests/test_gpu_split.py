"""Split-precision contractions (INT8 matrix-core slices, FP64 recombination).

Checked against the oracle (small sizes) and against FP64 complex GEMM
(torch.matmul on complex128, larger sizes):
  * normwise relative error <= 1e-10 at the default 5 slices;
  * the componentwise a-priori bound returned in err_out holds,
    |C - C~|_ij <= eta max_k|A_ik| max_k|B_kj| (plus FP64 rounding slack);
  * the error falls by about 2^-8 per extra slice;
  * rows and columns of very different magnitude, zero rows, ragged
    (non-power-of-two) dims and placements all keep the bound;
  * problems small enough for the direct route, and few-output long sums,
    stay exact FP64 (err 0).
"""
import numpy as np
import pytest
import torch

import oracle
from qtnh_amd import _lib, ops

pytestmark = pytest.mark.gpu

SPLIT = _lib.PREC_SPLIT
FP64 = _lib.PREC_FP64


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield
    ops.tune_split(5)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rnd(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def matrix_wires(ra, rb, nk):
    """Wire the last nk axes of A with the first nk axes of B (a matrix product)."""
    return [(ra - nk + i, i) for i in range(nk)]


def rowmax(a):
    return np.maximum(np.abs(a.real), np.abs(a.imag)).max(axis=1)


def check_bound(A, B, C, Cref, eta, K):
    """|C - Cref| <= eta max|A_i.| max|B_.j| + FP64 slack of the reference
    (relative rounding, and two subnormal spacings for results that underflow:
    the recombination rounds twice there)."""
    bound = eta * np.outer(rowmax(A), rowmax(B.T))
    slack = 8 * K * 2.0 ** -52 * (np.abs(A) @ np.abs(B)) + 2.0 ** -1073
    excess = np.abs(C - Cref) - (bound + slack)
    assert excess.max() <= 0, excess.max()


def split_gemm(A, B, slices=5):
    """C = A @ B through qtnh_contract_c128 with power-of-two dims (M, K, N powers of 2)."""
    M, K = A.shape
    N = B.shape[1]
    dA = [2] * int(np.log2(M)) + [2] * int(np.log2(K))
    dB = [2] * int(np.log2(K)) + [2] * int(np.log2(N))
    nk = int(np.log2(K))
    ops.tune_split(slices)
    c, dc, err = ops.contract(dev(A), dA, dev(B), dB, matrix_wires(len(dA), len(dB), nk), precision=SPLIT)
    return host(c).reshape(M, N), err


@pytest.mark.parametrize("m,k,n", [(64, 256, 512), (128, 256, 256), (256, 128, 512), (64, 1024, 512)])
def test_split_matches_oracle(m, k, n):
    A = rnd((m, k), 1)
    B = rnd((k, n), 2)
    C, eta = split_gemm(A, B)
    ref, _ = oracle.contract(A.reshape(-1), [m, k], B.reshape(-1), [k, n], [(1, 0)])
    ref = ref.reshape(m, n)
    rel = np.linalg.norm(C - ref) / np.linalg.norm(ref)
    assert rel <= 1e-10, rel
    assert eta > 0
    check_bound(A, B, C, ref, eta, k)


@pytest.mark.parametrize("n", [1024, 2048])
def test_split_large_vs_torch_fp64(n):
    a = ops.fill_(ops.empty(n * n), 11)
    b = ops.fill_(ops.empty(n * n), 12)
    r = int(np.log2(n))
    wires = matrix_wires(2 * r, 2 * r, r)
    c, _, eta = ops.contract(a, [2] * (2 * r), b, [2] * (2 * r), wires, precision=SPLIT)
    ref = torch.matmul(a.view(n, n), b.view(n, n)).reshape(-1)
    rel = ((c - ref).norm() / ref.norm()).item()
    assert rel <= 1e-10, rel
    # componentwise bound, with maxima over re/im
    am = torch.maximum(a.view(n, n).real.abs(), a.view(n, n).imag.abs()).amax(dim=1)
    bm = torch.maximum(b.view(n, n).real.abs(), b.view(n, n).imag.abs()).amax(dim=0)
    diff = (c - ref).view(n, n).abs()
    bound = (eta + 2.0 * n * n * 2.0 ** -52) * am[:, None] * bm[None, :]
    assert bool((diff <= bound).all())


def test_split_error_falls_with_slices():
    A = rnd((256, 512), 3)
    B = rnd((512, 128), 4)
    ref = A @ B
    errs = []
    for s in range(1, 9):
        C, eta = split_gemm(A, B, s)
        errs.append(np.linalg.norm(C - ref) / np.linalg.norm(ref))
        check_bound(A, B, C, ref, eta, 512)
    for s in range(1, 6):
        # each slice adds 8 bits; allow slack for the random constant
        assert errs[s] < errs[s - 1] * 2.0 ** -5, errs
    assert errs[7] < 1e-14, errs


def test_split_wide_dynamic_range_and_zero_rows():
    rng = np.random.default_rng(9)
    A = rnd((256, 256), 5) * np.exp2(rng.integers(-300, 300, 256))[:, None]
    B = rnd((256, 128), 6) * np.exp2(rng.integers(-300, 300, 128))[None, :]
    A[7] = 0
    B[:, 3] = 0
    A[9, 5] = 1e-310  # a subnormal in an otherwise zero row
    A[9, :5] = 0
    A[9, 6:] = 0
    C, eta = split_gemm(A, B)
    ref = A @ B
    assert np.all(C[7] == 0) and np.all(C[:, 3] == 0)
    check_bound(A, B, C, ref, eta, 256)
    # row-relative error: each row has its own scale (rows whose result
    # cancels below max|A_i| max|B_j| can lose more; the bound above holds)
    rel = np.abs(C - ref).max(axis=1) / np.maximum(np.abs(ref).max(axis=1), 1e-300)
    assert np.all(rel[np.abs(ref).max(axis=1) > 0] <= 1e-8), rel.max()


@pytest.mark.parametrize("dA,dB,wires", [
    ([60, 40, 7], [7, 40, 5, 59], [(1, 1), (2, 0)]),     # ragged K = 280, not a multiple of 32
    ([25, 6, 7, 8], [8, 6, 133], [(3, 0), (1, 1)]),      # non-adjacent wired axes
    ([2] * 14, [2] * 13, [(0, 8), (3, 2), (13, 0), (5, 4), (7, 6)]),
])
def test_split_ragged_and_scattered_wires(dA, dB, wires):
    A = rnd(int(np.prod(dA)), 7)
    B = rnd(int(np.prod(dB)), 8)
    c, dc, eta = ops.contract(dev(A), dA, dev(B), dB, wires, precision=SPLIT)
    ref, dref = oracle.contract(A, dA, B, dB, wires)
    assert list(dc) == list(dref)
    got = host(c)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert rel <= 1e-10, rel


def test_split_placement_matches_fp64():
    dA, dB = [64, 16, 32], [32, 16, 8, 64]
    wires = [(1, 1), (2, 0)]
    A = rnd(int(np.prod(dA)), 21)
    B = rnd(int(np.prod(dB)), 22)
    r1, r2 = [1, _lib.X, _lib.X], [_lib.X, _lib.X, 2, 0]
    cs, dcs, eta = ops.contract(dev(A), dA, dev(B), dB, wires, r1, r2, precision=SPLIT)
    cf, dcf, e0 = ops.contract(dev(A), dA, dev(B), dB, wires, r1, r2, precision=FP64)
    assert dcs == dcf == [64, 64, 8]
    assert e0 == 0.0
    rel = np.linalg.norm(host(cs) - host(cf)) / np.linalg.norm(host(cf))
    assert rel <= 1e-10, rel


def test_split_few_outputs_is_exact_fp64():
    A = rnd((1, 8192), 41)
    B = rnd((8192, 4), 42)
    dA = [2] * 13
    dB = [2] * 13 + [4]
    cs, _, eta = ops.contract(dev(A), dA, dev(B), dB, [(i, i) for i in range(13)], precision=SPLIT)
    cf, _, _ = ops.contract(dev(A), dA, dev(B), dB, [(i, i) for i in range(13)], precision=FP64)
    assert eta == 0.0
    assert np.array_equal(host(cs).view(np.uint64), host(cf).view(np.uint64))


def test_split_small_problem_is_exact_fp64():
    dA, dB = [4, 4], [4, 8]
    A = rnd(16, 31)
    B = rnd(32, 32)
    cs, _, eta = ops.contract(dev(A), dA, dev(B), dB, [(1, 0)], precision=SPLIT)
    cf, _, _ = ops.contract(dev(A), dA, dev(B), dB, [(1, 0)], precision=FP64)
    assert eta == 0.0
    assert np.array_equal(host(cs).view(np.uint64), host(cf).view(np.uint64))


@pytest.mark.parametrize("m,k,n", [(384, 512, 320), (400, 520, 330), (200, 5000, 100)])
def test_split_block_tile_does_not_change_bits(m, k, n):
    """Every block tile of the split GEMM sums the same exact integer products:
    whole tiles, ragged edges with a zero-padded depth, and two depth chunks."""
    rng = np.random.default_rng(77)
    A = rng.standard_normal(m * k) + 1j * rng.standard_normal(m * k)
    B = rng.standard_normal(k * n) + 1j * rng.standard_normal(k * n)
    da, db = [m, k], [k, n]
    a = torch.from_numpy(A).cuda()
    b = torch.from_numpy(B).cuda()
    outs = []
    try:
        for shape in (0, 1, 2, 3):
            ops.tune_split_tile(shape)
            c, _, _ = ops.contract(a, da, b, db, [(1, 0)], precision=_lib.PREC_SPLIT)
            torch.cuda.synchronize()
            outs.append(c.cpu().numpy())
    finally:
        ops.tune_split_tile(0)
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint64), outs[0].view(np.uint64))


@pytest.mark.parametrize("m,k,n", [(256, 1024, 512), (96, 300, 200)])
def test_split_b_by_rows_and_by_columns_bitwise(m, k, n):
    """B stored [closed, open] (read by columns in place) and B stored
    [open, closed] (read by rows) give the same digit planes: the results
    are bit-identical, and within the bound of the oracle product."""
    A = rnd((m, k), 21)
    B = rnd((k, n), 22) * np.exp2(np.random.default_rng(3).integers(-40, 40, n))[None, :]
    B[:, 5] = 0  # a zero column: exponent 0 on both paths
    c1, d1, eta = ops.contract(dev(A), [m, k], dev(B), [k, n], [(1, 0)], precision=SPLIT)
    c2, d2, _ = ops.contract(dev(A), [m, k], dev(B.T), [n, k], [(1, 1)], precision=SPLIT)
    g1, g2 = host(c1), host(c2)
    assert list(d1) == list(d2) == [m, n]
    assert np.array_equal(g1.view(np.uint64), g2.view(np.uint64))
    C = g1.reshape(m, n)
    assert np.all(C[:, 5] == 0)
    check_bound(A, B, C, A @ B, eta, k)
