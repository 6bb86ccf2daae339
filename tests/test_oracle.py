"""Pins the C oracle (oracle/oracle.c) before anything is checked against it.

Each oracle function is checked against (a) the golden vectors the reference's
own tests hold (tests/gen/random-tensors.hpp, random-tn.hpp, qft known answer),
(b) closed forms (numpy einsum / transpose / FFT), and (c) invariants chosen so
that a dropped term, a wrong sign, a transposed operand or a wrong placement
fails at least one check.  CPU only.
"""
import itertools
import math
import random

import numpy as np
import pytest

import oracle
from _fixtures import contraction_cases, iden_tensor, network_cases, swap_tensor

X = oracle.X


# ---------------------------------------------------------------- golden vectors
@pytest.mark.parametrize("case", list(contraction_cases("dense_vals")), ids=lambda c: c[0])
def test_dense_vals(case):
    """reference tests/c2-tensor.cpp:142-159: default placement, dense MAC."""
    name, t1, d1, t2, d2, t3, d3, wires = case
    out, dc = oracle.contract(t1, d1, t2, d2, wires)
    assert dc == d3
    # Inputs carry one decimal, so the exact result has two: the printed golden is exact.
    np.testing.assert_allclose(out, t3, rtol=0, atol=1e-12)


def test_dense_invalid_dims():
    """reference tests/c2-tensor.cpp:161-164: unequal wired dims are rejected."""
    with pytest.raises(ValueError):
        oracle.contract(np.ones(4), [2, 2], np.ones(3), [3], [(0, 0)])


@pytest.mark.parametrize("case", list(contraction_cases("swap_vals")), ids=lambda c: c[0])
def test_swap_vals(case):
    """reference tests/c2-tensor.cpp:167-184: SwapTensor through the symmetric placement."""
    name, t1, d1, _, d2, t3, d3, wires = case
    n = d2[0]
    r = oracle.symm_repls(len(d1), 0, 4, 0, wires)
    assert r is not None
    out, dc = oracle.contract(t1, d1, swap_tensor(n), [n] * 4, wires, *r)
    assert dc == d3
    np.testing.assert_allclose(out, t3, rtol=0, atol=1e-12)
    # The dedicated swap restates the same result as pure data movement, bit-exact.
    i1, i2 = wires[0][0], wires[1][0]
    sw = oracle.swap(t1, d1, i1, i2)
    assert np.array_equal(sw, out)


@pytest.mark.parametrize("case", list(contraction_cases("invalid_swaps")), ids=lambda c: c[0])
def test_invalid_swaps(case):
    """reference tests/c2-tensor.cpp:186-193: swaps of unequal dims are rejected."""
    name, t1, d1, _, d2, _, _, wires = case
    n = d2[0]
    r = oracle.symm_repls(len(d1), 0, 4, 0, wires)
    with pytest.raises(ValueError):
        oracle.contract(t1, d1, swap_tensor(n), [n] * 4, wires, *r)
    with pytest.raises(ValueError):
        oracle.swap(t1, d1, wires[0][0], wires[1][0])


@pytest.mark.parametrize("case", list(contraction_cases("id_vals")), ids=lambda c: c[0])
def test_id_vals(case):
    """reference tests/c2-tensor.cpp:196-214: multi-axis IdenTensor keeps the input."""
    name, t1, d1, _, d2, t3, d3, wires = case
    k = len(d2) // 2
    r = oracle.symm_repls(len(d1), 0, len(d2), 0, wires)
    assert r is not None
    out, dc = oracle.contract(t1, d1, iden_tensor(d2[:k]), d2, wires, *r)
    assert dc == d3
    np.testing.assert_allclose(out, t3, rtol=0, atol=1e-12)
    assert np.array_equal(out, t1)


# distributed-axis cases: (group, leading distributed axes of t1, of t2), reference c2-mpi.cpp:49-214
MPI_GROUPS = [("mpi2r_vals", 1, 0), ("mpi3r_vals", 1, 0), ("mpi4r_vals", 1, 1),
              ("mpi6r_vals", 1, 1), ("mpi8r_vals", 2, 1)]


@pytest.mark.parametrize("group,nd1,nd2", MPI_GROUPS)
def test_mpi_vals_global_view(group, nd1, nd2):
    """The distributed-axis goldens read as one global tensor: the default
    placement puts open distributed axes first (contract.cpp:224-247)."""
    for name, t1, d1, t2, d2, t3, d3, wires in contraction_cases(group):
        r1, r2 = oracle.default_repls(len(d1), nd1, len(d2), nd2, wires)
        out, dc = oracle.contract(t1, d1, t2, d2, wires, r1, r2)
        assert dc == d3, name
        np.testing.assert_allclose(out, t3, rtol=0, atol=1e-12, err_msg=name)


# ---------------------------------------------------------------- closed forms
def _einsum_natural(a, da, b, db, wires):
    letters = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"
    ia = list(letters[:len(da)])
    ib = list(letters[len(da):len(da) + len(db)])
    for x, y in wires:
        ib[y] = ia[x]
    wa = {x for x, _ in wires}
    wb = {y for _, y in wires}
    out = [ia[i] for i in range(len(da)) if i not in wa] + [ib[j] for j in range(len(db)) if j not in wb]
    return np.einsum("".join(ia) + "," + "".join(ib) + "->" + "".join(out),
                     a.reshape(da), b.reshape(db))


def test_contract_bruteforce_random_placement():
    """or_contract equals einsum followed by the requested placement, on random
    shapes, wires and output permutations (dims 1..4, up to 3 wires)."""
    rng = np.random.default_rng(7)
    pyr = random.Random(7)
    for trial in range(300):
        ra, rb = pyr.randint(0, 5), pyr.randint(0, 5)
        da = [pyr.choice([1, 2, 3, 4]) for _ in range(ra)]
        db = [pyr.choice([1, 2, 3, 4]) for _ in range(rb)]
        nw = pyr.randint(0, min(ra, rb, 3))
        wa = pyr.sample(range(ra), nw)
        wb = pyr.sample(range(rb), nw)
        for x, y in zip(wa, wb):
            db[y] = da[x]
        wires = list(zip(wa, wb))
        a = rng.standard_normal(int(np.prod(da))) + 1j * rng.standard_normal(int(np.prod(da)))
        b = rng.standard_normal(int(np.prod(db))) + 1j * rng.standard_normal(int(np.prod(db)))
        nat = _einsum_natural(a, da, b, db, wires)
        rc = ra + rb - 2 * nw
        perm = list(range(rc))
        pyr.shuffle(perm)
        r1 = [X] * ra
        r2 = [X] * rb
        q = 0
        for i in range(ra):
            if i not in wa:
                r1[i] = perm[q]; q += 1
        for j in range(rb):
            if j not in wb:
                r2[j] = perm[q]; q += 1
        out, dc = oracle.contract(a, da, b, db, wires, r1, r2)
        # natural axis q goes to position perm[q]
        expect = np.transpose(nat, np.argsort(perm)) if rc else nat
        np.testing.assert_allclose(out, expect.reshape(-1), rtol=0, atol=1e-12)


def test_contract_sum_order_is_reference_order():
    """The MAC loop sums closed tuples with the last wire (by t2 axis) fastest and
    starts from 0 (contract.cpp:131-156): check the exact bits of a sum whose
    value depends on the order."""
    big, tiny = 1e16, 1.0
    # A = [big, tiny, -big, tiny] over 2 closed axes (2,2); B = ones.
    a = np.array([big, tiny, -big, tiny], dtype=np.complex128)
    b = np.ones(4, dtype=np.complex128)
    # wires given in reverse t2 order: the oracle must still sort by the t2 axis.
    out, _ = oracle.contract(a, [2, 2], b, [2, 2], [(1, 1), (0, 0)])
    seq = ((0.0 + big) + tiny) + (-big) + tiny  # row-major over (w0, w1)
    assert out[0].real == seq
    out2, _ = oracle.contract(a, [2, 2], b, [2, 2], [(1, 0), (0, 1)])
    seq2 = ((0.0 + big) + (-big)) + tiny + tiny  # t2 axis 0 is wired to a's axis 1
    assert out2[0].real == seq2


def test_complex_product_has_no_fma():
    """Products are rounded separately (std::complex<double> without FMA):
    re = fl(fl(ar*br) - fl(ai*bi))."""
    ar, ai = 1.0 + 2.0 ** -30, 1.0
    br, bi = 1.0 - 2.0 ** -30, 1.0
    a = np.array([complex(ar, ai)])
    b = np.array([complex(br, bi)])
    out, _ = oracle.contract(a, [1], b, [1], [(0, 0)])
    assert out[0].real == (ar * br) - (ai * bi)
    assert out[0].imag == (ar * bi) + (ai * br)


def test_default_repls_closed_form():
    """Open distributed axes of t1, then of t2, then open local axes of t1, then of t2."""
    pyr = random.Random(3)
    for _ in range(200):
        r1, r2 = pyr.randint(1, 7), pyr.randint(1, 7)
        nd1, nd2 = pyr.randint(0, r1), pyr.randint(0, r2)
        nwd = pyr.randint(0, min(nd1, nd2))
        nwl = pyr.randint(0, min(r1 - nd1, r2 - nd2))
        wires = list(zip(pyr.sample(range(nd1), nwd), pyr.sample(range(nd2), nwd)))
        wires += list(zip(pyr.sample(range(nd1, r1), nwl), pyr.sample(range(nd2, r2), nwl)))
        pyr.shuffle(wires)
        w1 = {x for x, _ in wires}
        w2 = {y for _, y in wires}
        order = ([("a", i) for i in range(nd1) if i not in w1] + [("b", j) for j in range(nd2) if j not in w2]
                 + [("a", i) for i in range(nd1, r1) if i not in w1] + [("b", j) for j in range(nd2, r2) if j not in w2])
        exp1 = [X] * r1
        exp2 = [X] * r2
        for p, (s, i) in enumerate(order):
            (exp1 if s == "a" else exp2)[i] = p
        assert oracle.default_repls(r1, nd1, r2, nd2, wires) == (exp1, exp2)


def test_symm_repls_policy_conditions():
    """The symmetric placement applies only when every input axis of t2 is wired
    and its distributed in/out counts balance (contract.cpp:260-262)."""
    # 2-qubit gate, both inputs wired -> applies; outputs take the wired positions.
    assert oracle.symm_repls(5, 0, 4, 0, [(3, 0), (1, 1)]) == ([0, X, 2, X, 4], [X, X, 3, 1])
    # Wires given in another order give the same placement (sorted by t2 axis).
    assert oracle.symm_repls(5, 0, 4, 0, [(1, 1), (3, 0)]) == ([0, X, 2, X, 4], [X, X, 3, 1])
    # Only one input wired -> default placement instead.
    assert oracle.symm_repls(5, 0, 4, 0, [(3, 0)]) is None
    # An output axis wired instead of an input -> does not apply.
    assert oracle.symm_repls(5, 0, 4, 0, [(3, 0), (1, 2)]) is None
    # Imbalanced distributed split (IdenTensor {2}|{2}, n_dis_in 0: dis axis is an output).
    assert oracle.symm_repls(3, 1, 1, 0, [(0, 1)]) is None
    # Balanced distributed gate {2,2}|{}, n_dis_in 1.
    assert oracle.symm_repls(3, 2, 0, 1, [(0, 0)]) == ([X, 1, 2], [X, 0])


def test_gate_apply_closed_form():
    """In-place gate: psi'[..o@axes..] = sum_i psi[..i@axes..] G[i, o] (G rows are inputs)."""
    rng = np.random.default_rng(11)
    for k, rank in [(1, 5), (2, 6), (3, 7), (2, 3)]:
        for _ in range(5):
            axes = list(rng.permutation(rank)[:k])
            dims = [2] * rank
            x = rng.standard_normal(2 ** rank) + 1j * rng.standard_normal(2 ** rank)
            g = rng.standard_normal(4 ** k) + 1j * rng.standard_normal(4 ** k)
            got = oracle.gate_apply(x, dims, axes, g)
            gm = g.reshape(2 ** k, 2 ** k)  # [in, out]
            xt = np.moveaxis(x.reshape(dims), axes, list(range(k))).reshape(2 ** k, -1)
            yt = gm.T @ xt
            exp = np.moveaxis(yt.reshape([2] * rank), list(range(k)), axes).reshape(-1)
            np.testing.assert_allclose(got, exp, rtol=0, atol=1e-12)


def test_diag_equals_dense_diagonal_gate():
    """DiagTensor(d) acts like SymmTensor(diag(d)) (value-equal, signed zeros aside)."""
    rng = np.random.default_rng(5)
    for k in (1, 2, 3):
        rank = 6
        axes = list(rng.permutation(rank)[:k])
        x = rng.standard_normal(2 ** rank) + 1j * rng.standard_normal(2 ** rank)
        d = np.exp(1j * rng.uniform(0, 2 * np.pi, 2 ** k))
        a = oracle.diag_apply(x, [2] * rank, axes, d)
        b = oracle.gate_apply(x, [2] * rank, axes, np.diag(d).reshape(-1))
        assert np.array_equal(a, b)
        # closed form
        xt = x.reshape([2] * rank)
        idx = np.indices([2] * rank)
        di = np.zeros([2] * rank, dtype=int)
        for m, ax in enumerate(axes):
            di = di * 2 + idx[ax]
        dd = d[di]
        # separately rounded products (numpy's complex multiply may fuse)
        exp = (xt.real * dd.real - xt.imag * dd.imag) + 1j * (xt.real * dd.imag + xt.imag * dd.real)
        np.testing.assert_array_equal(a.reshape([2] * rank), exp)


def test_permute_closed_form_and_inverse():
    """new_dims[ptup[i]] = old_dims[i] (dense.cpp:155-158) == np.transpose(x, argsort(ptup)); bit-exact."""
    rng = np.random.default_rng(1)
    pyr = random.Random(1)
    for _ in range(100):
        rank = pyr.randint(0, 9)
        dims = [pyr.choice([1, 2, 3, 4]) for _ in range(rank)]
        ptup = list(range(rank))
        pyr.shuffle(ptup)
        n = int(np.prod(dims)) if dims else 1
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        y = oracle.permute(x, dims, ptup)
        exp = np.transpose(x.reshape(dims), np.argsort(ptup)).reshape(-1) if rank else x
        assert np.array_equal(y, exp)
        new_dims = [0] * rank
        for i in range(rank):
            new_dims[ptup[i]] = dims[i]
        inv = list(np.argsort(ptup))
        assert np.array_equal(oracle.permute(y, new_dims, inv), x)


def test_permute_rejects_bad_tuple():
    with pytest.raises(ValueError):
        oracle.permute(np.ones(8), [2, 2, 2], [0, 0, 1])


def test_swap_closed_form():
    rng = np.random.default_rng(2)
    for dims, i, j in [([2, 3, 3, 2], 1, 2), ([2] * 8, 0, 7), ([4, 4], 1, 0), ([3, 2, 3], 0, 2), ([2, 2], 1, 1)]:
        x = rng.standard_normal(int(np.prod(dims))) + 0j
        y = oracle.swap(x, dims, i, j)
        assert np.array_equal(y, np.swapaxes(x.reshape(dims), i, j).reshape(-1))


def test_swap_via_swaptensor_equals_transposition_rank10():
    """Swap contraction (SwapTensor, symmetric placement) == axis transposition, bitwise."""
    rng = np.random.default_rng(4)
    dims = [2] * 10
    x = rng.standard_normal(1024) + 1j * rng.standard_normal(1024)
    for i, j in [(0, 9), (3, 4), (7, 2)]:
        w = [(min(i, j), 0), (max(i, j), 1)]
        r = oracle.symm_repls(10, 0, 4, 0, w)
        out, _ = oracle.contract(x, dims, swap_tensor(2), [2] * 4, w, *r)
        assert np.array_equal(out, oracle.swap(x, dims, i, j))


# ---------------------------------------------------------------- QFT
def _phases(n):
    # controlled phase exp(i*pi/2^(j-i)) as the reference builds it (qft.hpp:71, std::exp(1i*p))
    ph = np.zeros(n, dtype=np.complex128)
    for m in range(1, n):
        p = math.pi / 2.0 ** m
        ph[m] = complex(math.cos(p), math.sin(p))
    return ph


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_qft_known_answer_plus_state(n):
    """reference tests/c2-network.cpp:155-198: QFT(|+>^n) has amplitude 0 equal to 1."""
    plus = np.full(2 ** n, 2.0 ** (-n / 2), dtype=np.complex128)
    out = oracle.qft(plus, n, _phases(n))
    assert abs(out[0] - 1) < 1e-12
    assert np.allclose(out[1:], 0, atol=1e-12)


@pytest.mark.parametrize("n", [3, 6, 12])
def test_qft_dft_closed_form(n):
    """out[bitrev(k)] = (1/sqrt N) sum_x exp(+2 pi i x k / N) psi[x] (qubit 0 most significant, no final swaps)."""
    N = 2 ** n
    psi = oracle.fill(N, seed=1234 + n)
    psi = psi / math.sqrt(oracle.norm2(psi))
    out = oracle.qft(psi, n, _phases(n))
    dft = np.fft.ifft(psi) * math.sqrt(N)
    rev = [int(format(k, "0%db" % n)[::-1], 2) for k in range(N)]
    exp = np.empty(N, dtype=np.complex128)
    exp[rev] = dft
    np.testing.assert_allclose(out, exp, rtol=0, atol=1e-13)


# ---------------------------------------------------------------- RNG
def test_splitmix64_known_answer():
    """SplitMix64 outputs for seed 1234567 (the published test sequence of Vigna's splitmix64.c)."""
    got = [oracle.splitmix64(1234567, k) for k in range(5)]
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def test_log_and_sincos_accuracy():
    rng = np.random.default_rng(0)
    us = np.concatenate([rng.uniform(0, 1, 20000), [1.0, 0.5, 2.0 ** -53, 0.7071067811865476, 0.99999999]])
    for u in us:
        if u == 0:
            continue
        ref = math.log(u)
        got = oracle.log(float(u))
        assert abs(got - ref) <= 4 * math.ulp(abs(ref)) + 1e-300, (u, got, ref)
    for v in np.concatenate([rng.uniform(0, 1, 20000), np.arange(16) / 16.0]):
        s, c = oracle.sincos2pi(float(v))
        assert abs(s - math.sin(2 * math.pi * v)) < 1e-15
        assert abs(c - math.cos(2 * math.pi * v)) < 1e-15


def test_fill_distribution_and_offsets():
    x = oracle.fill(200000, seed=99)
    for part in (x.real, x.imag):
        assert abs(part.mean()) < 0.01
        assert abs(part.var() - 1) < 0.01
        # fraction within one sigma of a standard normal: 0.6827
        assert abs(np.mean(np.abs(part) < 1) - 0.6827) < 0.005
    # counter-based: any window equals the same slice of the whole stream
    assert np.array_equal(oracle.fill(100, seed=99, first=1000), x[1000:1100])
    u = oracle.fill(100000, seed=5, dist=1)
    assert u.real.min() >= -1 and u.real.max() < 1 and abs(u.real.mean()) < 0.01


def test_multiset_hash_is_permutation_invariant():
    x = oracle.fill(4096, seed=3)
    p = list(range(12))
    random.Random(0).shuffle(p)
    y = oracle.permute(x, [2] * 12, p)
    assert oracle.multiset_hash(x) == oracle.multiset_hash(y)
    y[5] += 1e-9
    assert oracle.multiset_hash(x) != oracle.multiset_hash(y)


# ------------------------------------------------------------------ threaded CPU baseline
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_baseline_equals_oracle_bitwise(threads):
    """The multi-threaded CPU baseline does the oracle's arithmetic element for
    element: permute, gate (1-3 axes, mixed dims) and diagonal agree bitwise."""
    rng = np.random.default_rng(threads)
    for dims in ([2] * 12, [2, 3, 2, 4, 5, 2]):
        n = int(np.prod(dims))
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        p = list(rng.permutation(len(dims)))
        assert np.array_equal(oracle.permute_mt(x, dims, p, threads).view(np.uint64),
                              oracle.permute(x, dims, p).view(np.uint64))
        for k in (1, 2, 3):
            axes = [int(a) for a in rng.choice(len(dims), k, replace=False)]
            kk = int(np.prod([dims[a] for a in axes]))
            g = rng.standard_normal(kk * kk) + 1j * rng.standard_normal(kk * kk)
            assert np.array_equal(oracle.gate_apply_mt(x, dims, axes, g, threads).view(np.uint64),
                                  oracle.gate_apply(x, dims, axes, g).view(np.uint64)), (dims, axes)
            dv = rng.standard_normal(kk) + 1j * rng.standard_normal(kk)
            assert np.array_equal(oracle.diag_apply_mt(x, dims, axes, dv, threads).view(np.uint64),
                                  oracle.diag_apply(x, dims, axes, dv).view(np.uint64)), (dims, axes)
