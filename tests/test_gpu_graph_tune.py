"""Prepared network batches (optionally replayed as one HIP graph), fused
circuit runs captured into graphs, and the circuit tile autotuner: results
must not depend on any of them.

  * a prepared batch, eager or graph-replayed, equals the one-at-a-time
    executor bitwise, on every replay, with unrelated eager GPU work and
    synchronisation between replays;
  * new leaf values (set_leaf) give the same bits as a network built with
    those values from the start;
  * a fused circuit run captured into a HIP graph replays bitwise like
    eager runs; a run that would need a descriptor upload is refused during
    capture;
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


@pytest.mark.parametrize("n", [12, 20])
def test_fused_circuit_run_captured_in_a_graph(n):
    """After one eager run on a state, a fused run of the same circuit on that
    state launches kernels only: captured into a HIP graph and replayed (with
    eager work in between) it equals the eager runs bitwise, replay after
    replay."""
    c = ops.Circuit(n).random(4, 31 + n)
    x0 = ops.fill_(ops.empty(1 << n), 8)
    ref = x0.clone()
    snaps = []
    for _ in range(4):
        c.run(ref)
        snaps.append(ref.clone())
    torch.cuda.synchronize()
    s = x0.clone()
    cs = torch.cuda.Stream()
    with torch.cuda.stream(cs):
        c.run(s)  # eager: uploads the run descriptors for this state
    cs.synchronize()
    assert torch.equal(torch.view_as_real(s), torch.view_as_real(snaps[0]))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=cs):
        c.run(s)
    other = ops.fill_(ops.empty(1 << 16), 3)
    for k in range(1, 4):
        g.replay()
        ops.scale_(other, 0.5)
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(s), torch.view_as_real(snaps[k])), k


@pytest.mark.filterwarnings("ignore:The CUDA Graph is empty")
def test_first_fused_run_inside_capture_is_refused():
    """A fused run on a state the circuit has not run on would have to upload
    descriptors: during capture it is refused before touching the stream."""
    c = ops.Circuit(10).random(2, 3)
    t = ops.fill_(ops.empty(1 << 10), 2)
    c.run(t)  # compiled, descriptors cached for t (kept alive: s gets another address)
    s = ops.fill_(ops.empty(1 << 10), 1)
    assert s.data_ptr() != t.data_ptr()
    cs = torch.cuda.Stream()
    ops.context(stream=cs)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with pytest.raises(_lib.QtnhError):
        with torch.cuda.graph(g, stream=cs):
            c.run(s)


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


def test_captured_run_survives_recompiles():
    """A graph captured from a fused run keeps replaying the same bits after
    the circuit is recompiled for other tile shapes (a process-wide tile
    change, the per-circuit tuner) and run on another context: the earlier
    compiled form's device memory is kept until the circuit is destroyed."""
    n = 16
    c = ops.Circuit(n).random(4, 77)
    x0 = ops.fill_(ops.empty(1 << n), 21)
    ref = x0.clone()
    snaps = []
    for _ in range(4):
        c.run(ref)
        snaps.append(ref.clone())
    s = x0.clone()
    cs = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(cs):
        c.run(s)
    cs.synchronize()
    assert torch.equal(torch.view_as_real(s), torch.view_as_real(snaps[0]))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=cs):
        c.run(s)
    other = ops.fill_(ops.empty(1 << n), 5)
    try:
        ops.tune_circuit(10, 2)
        c.run(other)  # recompiled for 10-bit tiles
        tb, lb, us = c.tune(ops.empty(1 << n))  # six more compiles, the winner left compiled
        assert us > 0
        cs2 = torch.cuda.Stream()
        with torch.cuda.stream(cs2):
            c.run(other)  # another context (stream) on the same device
        cs2.synchronize()
    finally:
        ops.tune_circuit(12, 4)
    for k in range(1, 4):
        g.replay()
        ops.scale_(other, 0.5)
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(s), torch.view_as_real(snaps[k])), k
    c.close()
