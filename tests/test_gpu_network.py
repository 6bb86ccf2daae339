"""Networks contracted on the device, against independent answers.

  * amplitude networks of grid circuits equal the state-vector amplitude of
    the same circuit (fused circuit executor) within 1e-12 relative;
  * two plans (different seeds, hence orders) agree within 1e-12 relative;
  * general networks equal numpy einsum over the same bond labels;
  * the batched executor (grouped launches, admission waves) equals the
    one-at-a-time executor bit for bit.
"""
import numpy as np
import pytest
import torch

from qtnh_amd import ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def statevector_amplitude(c, bits):
    n = c.n
    s = torch.zeros(1 << n, dtype=torch.complex128, device="cuda")
    s[0] = 1
    c.run(s)
    idx = sum(int(b) << (n - 1 - q) for q, b in enumerate(bits))
    torch.cuda.synchronize()
    return complex(s[idx].item()), s


@pytest.mark.parametrize("rows,cols,depth", [(2, 2, 4), (3, 3, 6), (3, 4, 8), (4, 4, 8)])
def test_amplitude_matches_state_vector(rows, cols, depth):
    n = rows * cols
    c = ops.Circuit(n).grid(rows, cols, depth, 1000 + n)
    rng = np.random.default_rng(n)
    bits = rng.integers(0, 2, n)
    want, psi = statevector_amplitude(c, bits)
    got = ops.Network.amplitude(c, bits).plan(3).run()
    assert got.shape == (1,)
    # amplitudes are O(2^-n/2): relative to the larger of |want| and that scale
    assert abs(got[0] - want) <= 1e-12 * max(abs(want), 2.0 ** (-n / 2))


def test_best_of_sampled_orders_matches_state_vector():
    # plan_best keeps a sampled greedy order (not the seed's own) on these
    # networks; its amplitude is the state vector's all the same
    for rows, cols, depth, seed in ((4, 4, 10, 7), (3, 5, 12, 8)):
        n = rows * cols
        c = ops.Circuit(n).grid(rows, cols, depth, 2000 + seed)
        bits = np.random.default_rng(seed).integers(0, 2, n)
        want, _ = statevector_amplitude(c, bits)
        net = ops.Network.amplitude(c, bits)
        chosen = net.plan_best(seed, 26, 32)
        got = net.run()
        assert abs(got[0] - want) <= 1e-12 * max(abs(want), 2.0 ** (-n / 2)), chosen


def test_order_independence():
    c = ops.Circuit(12).grid(3, 4, 10, 5)
    bits = [1, 0] * 6
    vals = [ops.Network.amplitude(c, bits).plan(seed).run()[0] for seed in (0, 1, 2, 99)]
    for v in vals[1:]:
        assert abs(v - vals[0]) <= 1e-12 * max(abs(vals[0]), 2.0 ** -6)


def test_general_network_matches_einsum():
    rng = np.random.default_rng(4)
    dims = {0: 3, 1: 4, 2: 5, 3: 2, 4: 6, 5: 3}
    tensors = [[0, 1], [1, 2, 3], [3, 4], [2, 4, 5], [0, 5]]
    letters = "abcdef"
    net = ops.Network()
    arrays = []
    for bonds in tensors:
        shape = [dims[b] for b in bonds]
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        arrays.append(x)
        net.add(shape, bonds, x.reshape(-1))
    got = net.plan(1).run()
    want = np.einsum(",".join("".join(letters[b] for b in t) for t in tensors) + "->", *arrays)
    assert abs(got[0] - want) <= 1e-12 * abs(want)


def test_open_network_matches_einsum():
    rng = np.random.default_rng(8)
    net = ops.Network()
    a = rng.standard_normal((2, 3, 4)) + 1j * rng.standard_normal((2, 3, 4))
    b = rng.standard_normal((4, 5)) + 1j * rng.standard_normal((4, 5))
    net.add([2, 3, 4], [10, 11, 12], a.reshape(-1))
    net.add([4, 5], [12, 13], b.reshape(-1))
    got = net.plan().run()
    assert net.result_dims() == [2, 3, 5]
    want = np.einsum("abc,cd->abd", a, b).reshape(-1)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def mixed_networks():
    nets = []
    for i, (r, c, d) in enumerate([(2, 2, 3), (3, 3, 8), (4, 4, 8), (3, 4, 12), (4, 5, 6), (2, 3, 5), (4, 4, 12)]):
        circ = ops.Circuit(r * c).grid(r, c, d, 500 + i)
        bits = np.random.default_rng(i).integers(0, 2, r * c)
        nets.append(ops.Network.amplitude(circ, bits).plan(i))
    return nets


def test_batch_equals_one_at_a_time_bitwise():
    nets = mixed_networks()
    solo = [n.run() for n in nets]
    batch, waves = ops.run_networks(nets)
    assert waves == 1
    for s, b in zip(solo, batch):
        assert same_bits(s, b)


def test_batch_waves_under_a_small_budget():
    nets = mixed_networks()
    solo = [n.run() for n in nets]
    biggest = max(n.info()["device_bytes"] for n in nets)
    batch, waves = ops.run_networks(nets, budget_bytes=biggest)
    assert waves > 1
    for s, b in zip(solo, batch):
        assert same_bits(s, b)


def test_batch_over_many_step_ranges_and_after_trim():
    """Enough networks that the eager pass builds, uploads and launches its
    descriptors over several step ranges (2048 descriptors each); results equal
    one-at-a-time runs bit for bit, also after the context's cached buffers
    are released (qtnh_ctx_trim) and allocated again."""
    nets = []
    for i in range(40):
        circ = ops.Circuit(16).grid(4, 4, 8 + i % 5, 900 + i)
        bits = np.random.default_rng(100 + i).integers(0, 2, 16)
        nets.append(ops.Network.amplitude(circ, bits).plan(i))
    assert sum(n.info()["steps"] for n in nets) > 3 * 2048
    solo = [n.run() for n in nets]
    batch, waves = ops.run_networks(nets)
    assert waves == 1
    ops.trim()
    again, _ = ops.run_networks(nets)
    for s, b, a in zip(solo, batch, again):
        assert same_bits(s, b)
        assert same_bits(s, a)


def test_split_precision_networks():
    """GEMM-route steps in split precision: one-at-a-time and batched runs agree
    bit for bit (the same deterministic kernels), the amplitudes stay within
    1e-9 relative of the FP64 ones (each split GEMM is within about 1e-10
    normwise), and setting FP64 back restores the FP64 results bit for bit.
    Only steps whose result fills the GPU take the split GEMM: the two smaller
    networks have none and stay exact, the 5x5 depth-14 one has some."""
    from qtnh_amd import _lib
    nets = []
    for i, (r, c, d) in enumerate([(4, 4, 12), (5, 5, 10), (5, 5, 14)]):
        circ = ops.Circuit(r * c).grid(r, c, d, 506 + i)
        bits = np.random.default_rng(6 + i).integers(0, 2, r * c)
        nets.append(ops.Network.amplitude(circ, bits).plan(6 + i))
    fp64 = [n.run() for n in nets]
    for n in nets:
        n.set_precision(_lib.PREC_SPLIT)
    solo = [n.run() for n in nets]
    batch, _ = ops.run_networks(nets)
    errs = []
    for f, s, b in zip(fp64, solo, batch):
        assert same_bits(s, b)
        errs.append(np.abs(s - f).max() / np.abs(f).max())
    print("split relative errors", errs)
    assert max(errs) <= 1e-9
    assert errs[0] == 0 and errs[1] == 0
    assert errs[2] > 0  # the split GEMM ran
    for n in nets:
        n.set_precision(_lib.PREC_FP64)
    back = [n.run() for n in nets]
    for f, b in zip(fp64, back):
        assert same_bits(f, b)


def test_prepared_batch_keeps_its_precision():
    """A batch prepared in FP64 (eager and graph) still gives the FP64 results
    after its networks are switched to split precision."""
    from qtnh_amd import _lib
    circ = ops.Circuit(25).grid(5, 5, 14, 508)
    net = ops.Network.amplitude(circ, np.random.default_rng(8).integers(0, 2, 25)).plan(8)
    want = net.run()
    eager = ops.NetworkBatch([net], use_graph=False)
    graph = ops.NetworkBatch([net], use_graph=True)
    net.set_precision(_lib.PREC_SPLIT)
    try:
        assert not same_bits(net.run(), want)  # this network has split-precision steps
        assert same_bits(eager.run()[0], want)
        assert same_bits(graph.run()[0], want)
    finally:
        net.set_precision(_lib.PREC_FP64)
        eager.close()
        graph.close()


def _two_leaf_net(m_bits, n_bits, k_bits, seed, a_wired_first=True, b_wired_first=False):
    """A[m, k] x B[k, n] as a two-leaf network of binary axes with the wired
    axes first or last in each leaf; returns (net, A, B, a_labels, b_labels)."""
    rng = np.random.default_rng(seed)
    opena = list(range(100, 100 + m_bits))
    openb = list(range(200, 200 + n_bits))
    wired = list(range(k_bits))
    la = wired + opena if a_wired_first else opena + wired
    lb = wired[::-1] + openb if b_wired_first else openb + wired[::-1]
    A = rng.standard_normal(1 << (m_bits + k_bits)) + 1j * rng.standard_normal(1 << (m_bits + k_bits))
    B = rng.standard_normal(1 << (n_bits + k_bits)) + 1j * rng.standard_normal(1 << (n_bits + k_bits))
    net = ops.Network()
    net.add([2] * len(la), la, A)
    net.add([2] * len(lb), lb, B)
    net.plan(0, 64)
    return net, A, B, la, lb


def _einsum(A, la, B, lb, out_labels):
    letters = {}
    def sym(v):
        if v not in letters:
            letters[v] = chr(ord("a") + len(letters)) if len(letters) < 26 else chr(ord("A") + len(letters) - 26)
        return letters[v]
    sa = "".join(sym(v) for v in la)
    sb = "".join(sym(v) for v in lb)
    so = "".join(sym(v) for v in out_labels)
    return np.einsum("%s,%s->%s" % (sa, sb, so), A.reshape([2] * len(la)), B.reshape([2] * len(lb)),
                     optimize=True).reshape(-1)


@pytest.mark.parametrize("use_graph", [False, True])
def test_grouped_gemm_route_steps_bitwise(use_graph):
    """Mid-size GEMM-route steps of many networks run in grouped launches
    (operand permutes, GEMM tiles with and without depth chunks, split-K):
    bitwise equal to one network at a time, and equal to einsum within
    1e-12 relative."""
    nets, refs = [], []
    shapes = [(9, 6, 11, True, False),   # 512 x 64 x 2048: depth chunks, both operands permuted
              (8, 8, 6, True, False),    # 256 x 256 x 64: plain tiles
              (3, 3, 14, False, True),   # 8 x 8 x 16384: split-K
              (4, 4, 14, False, True),   # 16 x 16 x 16384: depth chunks
              (9, 6, 11, False, True),
              (7, 9, 8, True, True)]
    for i, (mb, nb, kb, af, bf) in enumerate(shapes * 2):
        net, A, B, la, lb = _two_leaf_net(mb, nb, kb, 50 + i, af, bf)
        out = [v for v in la if v >= 100] + [v for v in lb if v >= 100]
        nets.append(net)
        refs.append(_einsum(A, la, B, lb, out))
    nets += mixed_networks()
    solo = [n.run() for n in nets]
    for s, r in zip(solo, refs):
        assert np.linalg.norm(s - r) <= 1e-12 * np.linalg.norm(r)
    batch, waves = ops.run_networks(nets)
    assert waves == 1
    prepared = ops.NetworkBatch(nets, use_graph=use_graph)
    stats = prepared.stats()
    assert stats["grouped_gemm"] >= len(shapes) * 2, stats
    for rep in range(3):
        got = prepared.run()
        for s, b, g in zip(solo, batch, got):
            assert same_bits(s, b)
            assert same_bits(s, g), rep
        torch.cuda.synchronize()
    prepared.close()
