"""qtnh_contract_c128_noalloc (static device scratch, no allocation) against
qtnh_contract_c128, through the C ABI.

Both entry points run the same routes and kernels, so results agree bitwise:
the direct route (reference summation order), the FP64 GEMM route with depth
chunks, split-K and split precision.  Each route is called twice in a row, the
second call smaller than the first (the scratch keeps the first call's
contents).  A contraction whose workspace exceeds the scratch is refused on
the host before any launch.
"""
import ctypes
import random

import numpy as np
import pytest
import torch

from qtnh_amd import _lib, ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _u64(v):
    a = np.ascontiguousarray(np.asarray(list(v) or [0], dtype=np.uint64))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _u16(v):
    a = np.ascontiguousarray(np.asarray(list(v) or [0], dtype=np.uint16))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))


def call(name, A, dA, B, dB, wires, precision, out):
    """One call of the named C entry point; returns (status, eta)."""
    ctx = ops.context()
    repl1, repl2 = ops.default_repls(len(dA), 0, len(dB), 0, wires)
    a, ap = _u64(dA)
    b, bp = _u64(dB)
    w, wp = _u16([v for pair in wires for v in pair])
    r1, r1p = _u16(repl1)
    r2, r2p = _u16(repl2)
    eta = ctypes.c_double(-1.0)
    fn = getattr(_lib.lib(), name)
    st = fn(ctx, ctypes.c_void_p(A.data_ptr()), len(dA), ap, ctypes.c_void_p(B.data_ptr()), len(dB), bp,
            len(wires), wp, r1p, r2p, ctypes.c_void_p(out.data_ptr()), precision, ctypes.byref(eta))
    return st, eta.value


def case(seed, m_bits, n_bits, k_bits):
    """A (rank m+k) and B (rank k+n), all dims 2, contracted over k seeded axis pairs."""
    rng = random.Random(seed)
    ra, rb = m_bits + k_bits, k_bits + n_bits
    wa = rng.sample(range(ra), k_bits)
    wb = rng.sample(range(rb), k_bits)
    A = ops.fill_(ops.empty(1 << ra), seed)
    B = ops.fill_(ops.empty(1 << rb), seed + 1)
    return A, [2] * ra, B, [2] * rb, list(zip(wa, wb)), 1 << (m_bits + n_bits)


def same_bits(x, y):
    torch.cuda.synchronize()
    return bool(torch.equal(torch.view_as_real(x), torch.view_as_real(y)))


# (route, precision, first call (m, n, k bits), second smaller call)
ROUTES = [
    ("gemm-fp64", _lib.PREC_FP64, (9, 9, 9), (8, 8, 6)),
    ("split", _lib.PREC_SPLIT, (9, 9, 9), (8, 8, 7)),
    ("split-k", _lib.PREC_FP64, (3, 3, 14), (2, 2, 13)),
    ("direct", _lib.PREC_FP64, (14, 2, 2), (8, 2, 2)),
]


@pytest.mark.parametrize("route,precision,first,second", ROUTES, ids=[r[0] for r in ROUTES])
def test_noalloc_equals_contract(route, precision, first, second):
    for call_no, (mb, nb, kb) in enumerate((first, second)):
        print("noalloc %s call %d: M=2^%d N=2^%d K=2^%d" % (route, call_no, mb, nb, kb), flush=True)
        A, dA, B, dB, wires, nC = case(100 * call_no + mb + nb + kb, mb, nb, kb)
        want = ops.empty(nC)
        got = ops.empty(nC)
        got.fill_(float("nan"))
        st, eta_want = call("qtnh_contract_c128", A, dA, B, dB, wires, precision, want)
        assert st == 0, _lib.lib().qtnh_ctx_last_error(ops.context())
        st, eta_got = call("qtnh_contract_c128_noalloc", A, dA, B, dB, wires, precision, got)
        assert st == 0, _lib.lib().qtnh_ctx_last_error(ops.context())
        assert same_bits(got, want), (route, call_no)
        assert eta_got == eta_want


def test_noalloc_dense_gate_on_a_large_state():
    """A 2-qubit dense gate contracted into a rank-24 state needs no workspace
    and runs as the fused permute.  Called first through the allocation-free
    entry point (a permutation no call has tuned yet), it keeps the default
    tile shape without timing shapes; the regular entry point afterwards may
    time them.  Tile shapes do not change bits: both agree, on two calls in a
    row (the second smaller)."""
    for call_no, mb in enumerate((22, 20)):
        print("noalloc dense gate call %d: rank %d" % (call_no, mb + 2), flush=True)
        A, dA, B, dB, wires, nC = case(900 + mb, mb, 2, 2)
        got = ops.empty(nC)
        want = ops.empty(nC)
        got.fill_(float("nan"))
        st, _ = call("qtnh_contract_c128_noalloc", A, dA, B, dB, wires, _lib.PREC_FP64, got)
        assert st == 0, _lib.lib().qtnh_ctx_last_error(ops.context())
        st, _ = call("qtnh_contract_c128", A, dA, B, dB, wires, _lib.PREC_FP64, want)
        assert st == 0, _lib.lib().qtnh_ctx_last_error(ops.context())
        assert same_bits(got, want), call_no
        del A, B, got, want
        torch.cuda.empty_cache()


def test_noalloc_refuses_a_workspace_beyond_the_scratch():
    """M = N = K = 1024 with permuted operands needs 16 MiB per operand copy,
    16 MiB for the natural result and 64 MiB of depth-chunk partials: more
    than the 64 MiB scratch, so the call returns EINVAL before launching."""
    A, dA, B, dB, wires, nC = case(7, 10, 10, 10)
    out = ops.empty(nC)
    st, _ = call("qtnh_contract_c128_noalloc", A, dA, B, dB, wires, _lib.PREC_FP64, out)
    assert st == 1
    msg = _lib.lib().qtnh_ctx_last_error(ops.context()).decode()
    assert "scratch" in msg and str(_lib.NOALLOC_SCRATCH_BYTES) in msg, msg
    # the context is still usable, and the allocating entry point takes it
    st, _ = call("qtnh_contract_c128", A, dA, B, dB, wires, _lib.PREC_FP64, out)
    assert st == 0
    torch.cuda.synchronize()


def test_noalloc_split_captured_in_a_graph():
    """The split-precision route forks its operand splitting onto the
    context's second stream and joins it back: captured into a HIP graph
    through the entry point that allocates nothing, replays equal the eager
    result bitwise, with other work between replays."""
    cs = torch.cuda.Stream()
    with torch.cuda.stream(cs):
        A, dA, B, dB, wires, nC = case(41, 8, 8, 8)
        want = ops.empty(nC)
        got = ops.empty(nC)
        st, _ = call("qtnh_contract_c128_noalloc", A, dA, B, dB, wires, _lib.PREC_SPLIT, want)
        assert st == 0
    cs.synchronize()
    g = torch.cuda.CUDAGraph()
    print("noalloc split: capture", flush=True)
    with torch.cuda.graph(g, stream=cs):
        st, _ = call("qtnh_contract_c128_noalloc", A, dA, B, dB, wires, _lib.PREC_SPLIT, got)
    assert st == 0
    other = ops.fill_(ops.empty(1 << 16), 3)
    for rep in range(3):
        got.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        ops.scale_(other, 0.5)
        assert same_bits(got, want), rep
