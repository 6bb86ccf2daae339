"""CPU checks of the C ABI library: it loads, exports every declared symbol, and
its host-side placement logic (reference contract.cpp:219-289) agrees with the
oracle.  No compute call is made (there is no GPU here)."""
import ctypes
import os
import random
import re

import numpy as np
import pytest

import oracle
from qtnh_amd import _lib, ops

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
X = _lib.X


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qtnh_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions("qtnh_hip.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the binding covers them all
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_status_strings():
    L = _lib.lib()
    assert L.qtnh_status_string(0) == b"ok"
    assert L.qtnh_status_string(1) == b"invalid argument"


def test_device_count_never_fails():
    assert _lib.device_count() >= 0


def test_ctx_create_without_gpu_fails_loudly():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    st = _lib.lib().qtnh_ctx_create(0, None, ctypes.byref(h))
    assert st == _lib.QTNH_EHIP and not h.value


def test_null_context_is_rejected():
    L = _lib.lib()
    assert L.qtnh_sync(None) == _lib.QTNH_EINVAL
    assert L.qtnh_permute_c128(None, None, None, 0, None, None) == _lib.QTNH_EINVAL


def _random_wires(pyr, r1, nd1, r2, nd2):
    nwd = pyr.randint(0, min(nd1, nd2))
    nwl = pyr.randint(0, min(r1 - nd1, r2 - nd2))
    w = list(zip(pyr.sample(range(nd1), nwd), pyr.sample(range(nd2), nwd)))
    w += list(zip(pyr.sample(range(nd1, r1), nwl), pyr.sample(range(nd2, r2), nwl)))
    pyr.shuffle(w)
    return w


def test_default_repls_match_oracle():
    pyr = random.Random(21)
    for _ in range(500):
        r1, r2 = pyr.randint(0, 9), pyr.randint(0, 9)
        nd1, nd2 = pyr.randint(0, r1), pyr.randint(0, r2)
        w = _random_wires(pyr, r1, nd1, r2, nd2)
        assert ops.default_repls(r1, nd1, r2, nd2, w) == oracle.default_repls(r1, nd1, r2, nd2, w)


def test_symm_repls_match_oracle():
    pyr = random.Random(22)
    for _ in range(500):
        r1 = pyr.randint(1, 9)
        nd2 = pyr.choice([0, 0, 0, 1, 2, 2])
        nl2 = pyr.choice([0, 2, 2, 4, 4, 6])
        n_dis_in = pyr.randint(0, nd2)
        r2 = nd2 + nl2
        if r2 == 0:
            continue
        nw = pyr.randint(0, min(r1, r2))
        w = list(zip(pyr.sample(range(r1), nw), pyr.sample(range(r2), nw)))
        assert ops.symm_repls(r1, nd2, nl2, n_dis_in, w) == oracle.symm_repls(r1, nd2, nl2, n_dis_in, w)


def test_contract_dims_and_validation():
    # gate placement keeps the state's dims
    r = ops.symm_repls(4, 0, 4, 0, [(2, 0), (0, 1)])
    assert ops.contract_dims([2, 3, 2, 5], [2, 2, 2, 2], [(2, 0), (0, 1)], *r) == [2, 3, 2, 5]
    # default placement: open t1 axes then open t2 axes
    r1, r2 = ops.default_repls(3, 0, 2, 0, [(1, 0)])
    assert ops.contract_dims([2, 3, 4], [3, 5], [(1, 0)], r1, r2) == [2, 4, 5]
    # unequal wired dims are rejected (reference contract.cpp:209-211)
    with pytest.raises(_lib.QtnhError):
        ops.contract_dims([2, 2], [3], [(0, 0)], [X, 0], [X])
    # a placement that is not a permutation is rejected
    with pytest.raises(_lib.QtnhError):
        ops.contract_dims([2, 2], [2], [], [0, 0], [1])
    # duplicate wire endpoints are rejected
    with pytest.raises(_lib.QtnhError):
        ops.contract_dims([2, 2], [2, 2], [(0, 0), (0, 1)], [X, 0], [X, X])


def test_result_buffers_are_checked_before_any_launch():
    """An undersized, mistyped or non-contiguous out= is refused on the host
    before the library is called (the kernels write every element of out)."""
    import torch
    dims = [2] * 6
    x = torch.zeros(64, dtype=torch.complex128)
    ptup = [5, 4, 3, 2, 1, 0]
    with pytest.raises(ValueError):
        ops.permute(x, dims, ptup, out=torch.zeros(63, dtype=torch.complex128))
    with pytest.raises(TypeError):
        ops.permute(x, dims, ptup, out=torch.zeros(64, dtype=torch.float64))
    with pytest.raises(ValueError):
        ops.permute(x, dims, ptup, out=torch.zeros(128, dtype=torch.complex128)[::2])
    with pytest.raises(ValueError):
        ops.permute(x, dims, [0, 1, 2, 3, 4, 4])
    a = torch.zeros(16, dtype=torch.complex128)
    b = torch.zeros(16, dtype=torch.complex128)
    wires = [(3, 0), (2, 1)]
    with pytest.raises(ValueError):
        ops.contract(a, [2] * 4, b, [2] * 4, wires, out=torch.zeros(15, dtype=torch.complex128))
    with pytest.raises(ValueError):
        ops.contract(a, [2] * 4, b, [2] * 4, wires, out=torch.zeros(17, dtype=torch.complex128))
