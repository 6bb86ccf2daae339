"""Prepared network batches (optionally replayed as one HIP graph) and the
circuit tile autotuner: results must not depend on either.

  * a prepared batch, eager or graph-replayed, equals the one-at-a-time
    executor bitwise, on every replay, with unrelated eager GPU work and
    synchronisation between replays;
  * new leaf values (set_leaf) give the same bits as a network built with
    those values from the start;
  * after autotuning, fused circuits give the same bits as before.
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


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


SHAPES = [(2, 3, 4), (3, 3, 8), (4, 4, 8), (3, 4, 10), (4, 5, 6)]


def grid_nets(seed_base, bits_seed=0):
    nets, circuits = [], []
    for i, (r, c, d) in enumerate(SHAPES):
        circ = ops.Circuit(r * c).grid(r, c, d, seed_base + i)
        bits = np.random.default_rng(bits_seed + i).integers(0, 2, r * c)
        nets.append(ops.Network.amplitude(circ, bits).plan(i))
        circuits.append(circ)
    return nets, circuits


def leaf_values(circuit):
    """The op leaves of an amplitude network, in network order (after the n |0> leaves)."""
    out = []
    for kind, q, co in circuit.ops():
        if kind == 0:
            out.append(co)
        else:
            out.append(np.diag(co).reshape(-1))
    return out


@pytest.mark.parametrize("use_graph", [False, True])
def test_prepared_batch_replays_equal_solo(use_graph):
    nets, _ = grid_nets(100)
    solo = [n.run() for n in nets]
    batch = ops.NetworkBatch(nets, use_graph=use_graph)
    x = ops.fill_(ops.empty(1 << 20), 3)
    for rep in range(4):
        got = batch.run()
        for s, g in zip(solo, got):
            assert same_bits(s, g), (use_graph, rep)
        # unrelated eager work and a synchronisation between replays
        ops.scale_(x, 0.5)
        ops.permute(x, [2] * 20, list(range(20))[::-1])
        torch.cuda.synchronize()
    batch.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_set_leaf_equals_network_built_with_the_values(use_graph):
    nets_a, circ_a = grid_nets(200)
    nets_b, circ_b = grid_nets(300)  # same structures (and plans), other gate values
    want = [n.run() for n in nets_b]
    batch = ops.NetworkBatch(nets_a, use_graph=use_graph)
    first = batch.run()
    for s, g in zip([n.run() for n in nets_a], first):
        assert same_bits(s, g)
    for i, c in enumerate(circ_b):
        for j, v in enumerate(leaf_values(c)):
            batch.set_leaf(i, c.n + j, v)
    for rep in range(3):
        got = batch.run()
        for w, g in zip(want, got):
            assert same_bits(w, g), (use_graph, rep)
        torch.cuda.synchronize()
    batch.close()


def test_autotune_keeps_bits():
    n = 16
    c = ops.Circuit(n).random(3, 8)
    x0 = ops.fill_(ops.empty(1 << n), 4)
    a = x0.clone()
    c.run(a)
    tb, lb, ms = ops.autotune_circuit(n)
    assert (tb, lb) in {(10, 3), (11, 3), (11, 4), (12, 4), (12, 5), (13, 4), (13, 5)}
    assert ms > 0
    c2 = ops.Circuit(n).random(3, 8)
    b = x0.clone()
    c2.run(b)
    torch.cuda.synchronize()
    assert torch.equal(torch.view_as_real(a), torch.view_as_real(b))
