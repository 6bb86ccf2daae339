"""Network construction and the greedy order planner (host only, no GPU).

The planner is pinned by: determinism from the seed; every step contracting
two live tensors that share a bond; the flop and byte counts of hand-sized
contractions; the rank cap; and the amplitude network of a circuit holding
one tensor per op plus the |0> and <bits| vectors.
"""
import numpy as np
import pytest

from qtnh_amd import _lib, ops


def chain(n, d=3):
    """A matrix chain v0 - M1 - ... - M(n-2) - v(n-1), closed to a scalar."""
    net = ops.Network()
    rng = np.random.default_rng(n)
    net.add([d], [0], rng.standard_normal(d))
    for i in range(1, n - 1):
        net.add([d, d], [i - 1, i], rng.standard_normal(d * d))
    net.add([d], [n - 2], rng.standard_normal(d))
    return net


def test_plan_is_deterministic_and_complete():
    c = ops.Circuit(9).grid(3, 3, 6, 1)
    a = ops.Network.amplitude(c).plan(7)
    b = ops.Network.amplitude(c).plan(7)
    ia, ib = a.info(), b.info()
    assert ia == ib
    n_ops = c.info()[1]
    assert ia["tensors"] == n_ops + 2 * 9
    assert ia["steps"] == ia["tensors"] - 1  # connected: one tensor left
    assert a.result_dims() == []


def test_plan_steps_consume_each_tensor_once():
    import ctypes
    net = ops.Network.amplitude(ops.Circuit(8).grid(2, 4, 5, 3)).plan(11)
    info = net.info()
    used = set()
    L = info["tensors"]
    for s in range(info["steps"]):
        a, b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().qtnh_net_step(net.handle, s, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                             ctypes.byref(d)))
        assert a.value not in used and b.value not in used
        assert a.value < c.value and b.value < c.value == L + s
        used.update((a.value, b.value))
    assert len(used) == L + info["steps"] - 1


def test_flop_and_byte_counts_by_hand():
    net = ops.Network()
    net.add([4, 8], [0, 1], np.zeros(32))
    net.add([8, 16], [1, 2], np.zeros(128))
    net.plan()
    i = net.info()
    assert i["flops"] == 8 * 4 * 16 * 8
    assert i["bytes"] == 16 * (32 + 128 + 64)
    assert net.result_dims() == [4, 16]


def test_chain_ranks_stay_small():
    net = chain(40).plan(0)
    i = net.info()
    assert i["max_rank"] <= 2
    assert net.result_dims() == []


def test_rank_cap_rejects():
    c = ops.Circuit(16).grid(4, 4, 12, 2)
    net = ops.Network.amplitude(c)
    with pytest.raises(_lib.QtnhError) as e:
        net.plan(0, 6)
    assert e.value.status == _lib.QTNH_EUNSUPPORTED
    net.plan(0, 26)
    assert 6 < net.info()["max_rank"] <= 26


def test_invalid_networks_rejected():
    net = ops.Network()
    with pytest.raises(_lib.QtnhError):
        net.add([2, 2], [5, 5], np.zeros(4))  # repeated bond on one tensor
    net.add([2], [1], np.zeros(2))
    net.add([2], [1], np.zeros(2))
    net.add([2], [1], np.zeros(2))
    with pytest.raises(_lib.QtnhError):
        net.plan()  # a bond on three axes
    with pytest.raises(ValueError):
        ops.Network().add([2, 3], [0, 1], np.zeros(5))


def test_disconnected_parts_end_in_an_outer_product():
    net = ops.Network()
    net.add([2, 3], [0, 1], np.zeros(6))
    net.add([3], [1], np.zeros(3))
    net.add([5], [7], np.zeros(5))
    net.plan()
    assert sorted(net.result_dims()) == [2, 5]


def _step_nodes(net):
    import ctypes
    out = []
    for s in range(net.info()["steps"]):
        a, b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().qtnh_net_step(net.handle, s, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                             ctypes.byref(d)))
        out.append((a.value, b.value, c.value))
    return out


def test_plan_best_keeps_the_fewest_flops_of_its_tries():
    c = ops.Circuit(25).grid(5, 5, 12, 4)
    bits = np.random.default_rng(4).integers(0, 2, 25)
    first = ops.Network.amplitude(c, bits).plan(9, 26).info()
    prev = None
    for k in (1, 4, 16, 64):
        net = ops.Network.amplitude(c, bits)
        seed = net.plan_best(9, 26, k)
        i = net.info()
        assert 9 <= seed < 9 + k
        if k == 1:
            # try 0 is the deterministic greedy order of the seed
            assert seed == 9 and i == first and _step_nodes(net) == _step_nodes(ops.Network.amplitude(c, bits).plan(9, 26))
        assert i["flops"] <= first["flops"] and i["max_rank"] <= 26
        if prev is not None:
            assert i["flops"] <= prev  # the tries of a larger K include those of a smaller one
        prev = i["flops"]
        # a complete order: every tensor consumed once, one result left
        used = set()
        for a, b, cc in _step_nodes(net):
            assert a not in used and b not in used and a < cc and b < cc
            used.update((a, b))
        assert len(used) == i["tensors"] + i["steps"] - 1
    assert prev < first["flops"]  # sampled orders find a cheaper plan on this network


def test_plan_best_is_deterministic():
    c = ops.Circuit(24).grid(4, 6, 12, 11)
    bits = np.random.default_rng(11).integers(0, 2, 24)
    runs = []
    for _ in range(3):
        net = ops.Network.amplitude(c, bits)
        seed = net.plan_best(100, 26, 48)
        runs.append((seed, net.info(), _step_nodes(net)))
    assert runs[0] == runs[1] == runs[2]


def test_plan_best_rank_cap():
    c = ops.Circuit(16).grid(4, 4, 12, 2)
    with pytest.raises(_lib.QtnhError) as e:
        ops.Network.amplitude(c).plan_best(0, 6, 16)
    assert e.value.status == _lib.QTNH_EUNSUPPORTED
