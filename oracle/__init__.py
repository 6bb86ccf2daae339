"""numpy front end of the C oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package (qtnh_amd) never imports it.

Arrays are numpy complex128 (interleaved re/im, the std::complex<double>
layout).  Every wrapper raises on a non-zero oracle status.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
X = 0xFFFF

_u16p = ctypes.POINTER(ctypes.c_uint16)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_dp = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so from oracle.c with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.or_size.restype = ctypes.c_uint64
    lib.or_splitmix64.restype = ctypes.c_uint64
    lib.or_splitmix64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    lib.or_log.restype = ctypes.c_double
    lib.or_log.argtypes = [ctypes.c_double]
    lib.or_sincos2pi.argtypes = [ctypes.c_double, _dp, _dp]
    lib.or_fill_c128.argtypes = [_dp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
    lib.or_permute.argtypes = [_dp, _dp, ctypes.c_int, _u64p, _u16p]
    lib.or_swap.argtypes = [_dp, ctypes.c_int, _u64p, ctypes.c_int, ctypes.c_int]
    lib.or_default_repls.argtypes = [ctypes.c_int] * 5 + [_u16p, _u16p, _u16p]
    lib.or_symm_repls.argtypes = [ctypes.c_int] * 5 + [_u16p, _u16p, _u16p]
    lib.or_contract.argtypes = [_dp, ctypes.c_int, _u64p, _dp, ctypes.c_int, _u64p,
                                ctypes.c_int, _u16p, _u16p, _u16p, _dp]
    lib.or_gate_apply.argtypes = [_dp, ctypes.c_int, _u64p, ctypes.c_int, _u16p, _dp]
    lib.or_diag_apply.argtypes = [_dp, ctypes.c_int, _u64p, ctypes.c_int, _u16p, _dp]
    lib.or_qft.argtypes = [_dp, ctypes.c_int, _dp]
    lib.or_permute_mt.argtypes = [_dp, _dp, ctypes.c_int, _u64p, _u16p, ctypes.c_int]
    lib.or_gate_apply_mt.argtypes = [_dp, ctypes.c_int, _u64p, ctypes.c_int, _u16p, _dp, ctypes.c_int]
    lib.or_diag_apply_mt.argtypes = [_dp, ctypes.c_int, _u64p, ctypes.c_int, _u16p, _dp, ctypes.c_int]
    lib.or_norm2.restype = ctypes.c_double
    lib.or_norm2.argtypes = [_dp, ctypes.c_uint64]
    lib.or_multiset_hash.restype = ctypes.c_uint64
    lib.or_multiset_hash.argtypes = [_dp, ctypes.c_uint64]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a, a.ctypes.data_as(_dp)


def _u64(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.uint64))
    return a, a.ctypes.data_as(_u64p)


def _u16(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.uint16).reshape(-1))
    if a.size == 0:
        a = np.zeros(1, dtype=np.uint16)
    return a, a.ctypes.data_as(_u16p)


def _check(rc, what):
    if rc != 0:
        raise ValueError("oracle %s failed with status %d" % (what, rc))


def splitmix64(seed, k):
    return lib().or_splitmix64(seed, k)


def log(u):
    return lib().or_log(u)


def sincos2pi(v):
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().or_sincos2pi(v, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def fill(n, seed, dist=0, first=0):
    out = np.empty(n, dtype=np.complex128)
    lib().or_fill_c128(out.ctypes.data_as(_dp), first, n, seed, dist)
    return out


def permute(x, dims, ptup):
    x, xp = _c128(x)
    out = np.empty_like(x)
    d, dp = _u64(dims)
    p, pp = _u16(ptup)
    _check(lib().or_permute(xp, out.ctypes.data_as(_dp), len(dims), dp, pp), "permute")
    return out


def swap(x, dims, i, j):
    x = np.array(x, dtype=np.complex128, copy=True)
    d, dp = _u64(dims)
    _check(lib().or_swap(x.ctypes.data_as(_dp), len(dims), dp, i, j), "swap")
    return x


def default_repls(r1, ndis1, r2, ndis2, wires):
    w, wp = _u16(wires)
    r1a = np.zeros(max(r1, 1), dtype=np.uint16)
    r2a = np.zeros(max(r2, 1), dtype=np.uint16)
    _check(lib().or_default_repls(r1, ndis1, r2, ndis2, len(wires), wp,
                                  r1a.ctypes.data_as(_u16p), r2a.ctypes.data_as(_u16p)), "default_repls")
    return [int(v) for v in r1a[:r1]], [int(v) for v in r2a[:r2]]


def symm_repls(r1, nd2, nl2, n_dis_in, wires):
    """Returns (repl1, repl2) or None when the symmetric policy does not apply."""
    w, wp = _u16(wires)
    r2 = nd2 + nl2
    r1a = np.zeros(max(r1, 1), dtype=np.uint16)
    r2a = np.zeros(max(r2, 1), dtype=np.uint16)
    ok = lib().or_symm_repls(r1, nd2, nl2, n_dis_in, len(wires), wp,
                             r1a.ctypes.data_as(_u16p), r2a.ctypes.data_as(_u16p))
    if not ok:
        return None
    return [int(v) for v in r1a[:r1]], [int(v) for v in r2a[:r2]]


def contract(a, da, b, db, wires, repl1=None, repl2=None):
    """Contract a (dims da) with b (dims db) over wires; default placement if repls omitted.

    Returns (c, dc): the flat result and its dims."""
    if repl1 is None or repl2 is None:
        repl1, repl2 = default_repls(len(da), 0, len(db), 0, wires)
    a, ap = _c128(a)
    b, bp = _c128(b)
    dA, dAp = _u64(da)
    dB, dBp = _u64(db)
    w, wp = _u16(wires)
    r1, r1p = _u16(repl1)
    r2, r2p = _u16(repl2)
    rc_dims = len(da) + len(db) - 2 * len(wires)
    dc = [0] * rc_dims
    for i, d in enumerate(da):
        if repl1[i] != X:
            dc[repl1[i]] = d
    for i, d in enumerate(db):
        if repl2[i] != X:
            dc[repl2[i]] = d
    n = int(np.prod(dc)) if dc else 1
    out = np.empty(n, dtype=np.complex128)
    _check(lib().or_contract(ap, len(da), dAp, bp, len(db), dBp, len(wires), wp, r1p, r2p,
                             out.ctypes.data_as(_dp)), "contract")
    return out, dc


def gate_apply(x, dims, axes, g, inplace=False):
    """In-place k-axis gate (x'[..o..] = sum_i x[..i..] g[i, o]); a copy unless inplace
    (then x must be a contiguous complex128 array)."""
    if not inplace:
        x = np.array(x, dtype=np.complex128, copy=True)
    d, dp = _u64(dims)
    a, ap = _u16(axes)
    g, gp = _c128(g)
    _check(lib().or_gate_apply(x.ctypes.data_as(_dp), len(dims), dp, len(axes), ap, gp), "gate_apply")
    return x


def diag_apply(x, dims, axes, dvals, inplace=False):
    if not inplace:
        x = np.array(x, dtype=np.complex128, copy=True)
    d, dp = _u64(dims)
    a, ap = _u16(axes)
    dv, dvp = _c128(dvals)
    _check(lib().or_diag_apply(x.ctypes.data_as(_dp), len(dims), dp, len(axes), ap, dvp), "diag_apply")
    return x


# ---- CPU timing baseline: the same arithmetic split over threads (not checkers).
def permute_mt(x, dims, ptup, threads):
    x, xp = _c128(x)
    out = np.empty_like(x)
    d, dp = _u64(dims)
    p, pp = _u16(ptup)
    _check(lib().or_permute_mt(xp, out.ctypes.data_as(_dp), len(dims), dp, pp, int(threads)), "permute_mt")
    return out


def gate_apply_mt(x, dims, axes, g, threads, inplace=False):
    if not inplace:
        x = np.array(x, dtype=np.complex128, copy=True)
    d, dp = _u64(dims)
    a, ap = _u16(axes)
    g, gp = _c128(g)
    _check(lib().or_gate_apply_mt(x.ctypes.data_as(_dp), len(dims), dp, len(axes), ap, gp, int(threads)),
           "gate_apply_mt")
    return x


def diag_apply_mt(x, dims, axes, dvals, threads, inplace=False):
    if not inplace:
        x = np.array(x, dtype=np.complex128, copy=True)
    d, dp = _u64(dims)
    a, ap = _u16(axes)
    dv, dvp = _c128(dvals)
    _check(lib().or_diag_apply_mt(x.ctypes.data_as(_dp), len(dims), dp, len(axes), ap, dvp, int(threads)),
           "diag_apply_mt")
    return x


def qft(x, n, phases):
    x = np.array(x, dtype=np.complex128, copy=True)
    ph, php = _c128(phases)
    _check(lib().or_qft(x.ctypes.data_as(_dp), n, php), "qft")
    return x


def norm2(x):
    x, xp = _c128(x)
    return lib().or_norm2(xp, x.size)


def multiset_hash(x):
    x, xp = _c128(x)
    return lib().or_multiset_hash(xp, x.size)
