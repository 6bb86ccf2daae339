"""Python front end of the gfx950 kernels, over torch device memory.

Tensors are torch complex128 CUDA (HIP) tensors holding row-major data; shapes
travel separately as `dims` lists (as in the reference, where a tensor is a
flat element vector plus its dims).  Every call goes through the C ABI of
libqtnh_hip.so on the caller's current torch stream; nothing here computes on
the CPU.

Placement arguments follow reference src/tensor/contract.cpp: repl1/repl2
give each open axis its position in the result (X for wired axes).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import PREC_FP64, PREC_SPLIT, RUN_FUSED, RUN_PER_GATE, X, check, lib

_CTX = {}


def context(device=None, stream=None):
    """The qtnh_ctx_t bound to (device, stream); created on first use."""
    if device is None:
        device = torch.cuda.current_device()
    if stream is None:
        stream = torch.cuda.current_stream(device)
    key = (device, int(stream.cuda_stream))
    ctx = _CTX.get(key)
    if ctx is None:
        h = ctypes.c_void_p()
        check(lib().qtnh_ctx_create(device, ctypes.c_void_p(int(stream.cuda_stream)), ctypes.byref(h)))
        ctx = h
        _CTX[key] = ctx
    return ctx


def trim(device=None, stream=None):
    """Release the context's cached device and pinned buffers (they are
    allocated again when needed)."""
    ctx = context(device, stream)
    check(lib().qtnh_ctx_trim(ctx), ctx)


def _u64(v):
    a = np.ascontiguousarray(np.asarray(list(v), dtype=np.uint64))
    if a.size == 0:
        a = np.zeros(1, dtype=np.uint64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _u16(v):
    a = np.ascontiguousarray(np.asarray(list(v), dtype=np.uint16).reshape(-1))
    if a.size == 0:
        a = np.zeros(1, dtype=np.uint16)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))


def _c128_host(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.complex128).reshape(-1))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ptr(t):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.complex128 or not t.is_cuda:
        raise TypeError("expected a complex128 device tensor")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _size(dims):
    n = 1
    for d in dims:
        n *= int(d)
    return n


def _check_len(t, dims):
    if t.numel() != _size(dims):
        raise ValueError("tensor has %d elements, dims %s need %d" % (t.numel(), list(dims), _size(dims)))


def _check_out(out, x, dims):
    """A caller-supplied result buffer must hold exactly prod(dims) complex128
    elements, contiguously, on x's device: the kernels write all of it."""
    if not isinstance(out, torch.Tensor) or out.dtype != torch.complex128:
        raise TypeError("out must be a complex128 tensor")
    _check_len(out, dims)
    if not out.is_contiguous():
        raise ValueError("out must be contiguous")
    if out.device != x.device:
        raise ValueError("out is on %s, the operands on %s" % (out.device, x.device))


def empty(n, device=None):
    return torch.empty(int(n), dtype=torch.complex128, device=device or "cuda")


def fill_(x, seed, dist=_lib.DIST_NORMAL):
    ctx = context(x.device.index)
    check(lib().qtnh_fill_c128(ctx, _ptr(x), x.numel(), seed, dist), ctx)
    return x


def permute(x, dims, ptup, out=None):
    """Result axis ptup[i] = axis i of x (reference dense.cpp:152-169)."""
    _check_len(x, dims)
    if len(ptup) != len(dims):
        raise ValueError("ptup length differs from rank")
    rdims = permuted_dims(dims, ptup)
    if out is None:
        out = empty(x.numel(), x.device)
    else:
        _check_out(out, x, rdims)
    ctx = context(x.device.index)
    d, dp = _u64(dims)
    p, pp = _u16(ptup)
    check(lib().qtnh_permute_c128(ctx, _ptr(x), _ptr(out), len(dims), dp, pp), ctx)
    return out


def permuted_dims(dims, ptup):
    if sorted(int(v) for v in ptup) != list(range(len(dims))):
        raise ValueError("ptup %s is not a permutation of %d axes" % (list(ptup), len(dims)))
    nd = [0] * len(dims)
    for i, d in enumerate(dims):
        nd[ptup[i]] = d
    return nd


def swap_(x, dims, i, j):
    """In-place exchange of equal-dim axes i and j (reference dense.cpp:171-218)."""
    _check_len(x, dims)
    ctx = context(x.device.index)
    d, dp = _u64(dims)
    check(lib().qtnh_swap_c128(ctx, _ptr(x), len(dims), dp, i, j), ctx)
    return x


def gate_apply_(x, dims, axes, g):
    """In place: x'[.. o@axes ..] = sum_i x[.. i@axes ..] g[i, o]."""
    _check_len(x, dims)
    ctx = context(x.device.index)
    d, dp = _u64(dims)
    a, ap = _u16(axes)
    gh, gp = _c128_host(g)
    k = _size(dims[ax] for ax in axes)
    if gh.size != k * k:
        raise ValueError("gate has %d elements, expected %d" % (gh.size, k * k))
    check(lib().qtnh_gate_apply_c128(ctx, _ptr(x), len(dims), dp, len(axes), ap, gp, 0), ctx)
    return x


def diag_apply_(x, dims, axes, dvals):
    """In place: x[i] *= dvals[i@axes]."""
    _check_len(x, dims)
    ctx = context(x.device.index)
    d, dp = _u64(dims)
    a, ap = _u16(axes)
    dh, dhp = _c128_host(dvals)
    if dh.size != _size(dims[ax] for ax in axes):
        raise ValueError("diagonal length mismatch")
    check(lib().qtnh_diag_apply_c128(ctx, _ptr(x), len(dims), dp, len(axes), ap, dhp), ctx)
    return x


def default_repls(r1, ndis1, r2, ndis2, wires):
    w, wp = _u16([v for pair in wires for v in pair])
    o1 = np.zeros(max(r1, 1), dtype=np.uint16)
    o2 = np.zeros(max(r2, 1), dtype=np.uint16)
    check(lib().qtnh_default_repls(r1, ndis1, r2, ndis2, len(wires), wp,
                                   o1.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                   o2.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))))
    return [int(v) for v in o1[:r1]], [int(v) for v in o2[:r2]]


def symm_repls(r1, nd2, nl2, n_dis_in, wires):
    w, wp = _u16([v for pair in wires for v in pair])
    r2 = nd2 + nl2
    o1 = np.zeros(max(r1, 1), dtype=np.uint16)
    o2 = np.zeros(max(r2, 1), dtype=np.uint16)
    ap = ctypes.c_int(0)
    check(lib().qtnh_symm_repls(r1, nd2, nl2, n_dis_in, len(wires), wp,
                                o1.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                o2.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), ctypes.byref(ap)))
    if not ap.value:
        return None
    return [int(v) for v in o1[:r1]], [int(v) for v in o2[:r2]]


def contract_dims(dA, dB, wires, repl1, repl2):
    a, ap = _u64(dA)
    b, bp = _u64(dB)
    w, wp = _u16([v for pair in wires for v in pair])
    r1, r1p = _u16(repl1)
    r2, r2p = _u16(repl2)
    rc = ctypes.c_int(0)
    out = np.zeros(64, dtype=np.uint64)
    check(lib().qtnh_contract_dims(len(dA), ap, len(dB), bp, len(wires), wp, r1p, r2p, ctypes.byref(rc),
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
    return [int(v) for v in out[:rc.value]]


def contract(A, dA, B, dB, wires, repl1=None, repl2=None, precision=PREC_FP64, out=None):
    """Pairwise contraction over `wires`; returns (C, dC, err).  Default placement
    when repl1/repl2 are omitted (reference contract.cpp:219-248)."""
    _check_len(A, dA)
    _check_len(B, dB)
    if repl1 is None or repl2 is None:
        repl1, repl2 = default_repls(len(dA), 0, len(dB), 0, wires)
    dC = contract_dims(dA, dB, wires, repl1, repl2)
    if B.device != A.device:
        raise ValueError("A is on %s, B on %s" % (A.device, B.device))
    if out is None:
        out = empty(_size(dC), A.device)
    else:
        _check_out(out, A, dC)
    ctx = context(A.device.index)
    a, ap = _u64(dA)
    b, bp = _u64(dB)
    w, wp = _u16([v for pair in wires for v in pair])
    r1, r1p = _u16(repl1)
    r2, r2p = _u16(repl2)
    err = ctypes.c_double(0.0)
    check(lib().qtnh_contract_c128(ctx, _ptr(A), len(dA), ap, _ptr(B), len(dB), bp, len(wires), wp, r1p, r2p,
                                   _ptr(out), precision, ctypes.byref(err)), ctx)
    return out, dC, err.value


def norm2(x):
    ctx = context(x.device.index)
    v = ctypes.c_double(0.0)
    check(lib().qtnh_norm2_c128(ctx, _ptr(x), x.numel(), ctypes.byref(v)), ctx)
    return v.value


def scale_(x, c):
    ctx = context(x.device.index)
    c = complex(c)
    check(lib().qtnh_scale_c128(ctx, _ptr(x), x.numel(), c.real, c.imag), ctx)
    return x


def tune_permute(a=6, b=6, u_target=12, order=1, nt=1, blocks_per_cu=2):
    """Tile parameters of the power-of-two permute kernel (process-wide);
    turns per-permutation tuning off."""
    lib().qtnh_tune_permute(a, b, u_target, order, nt, blocks_per_cu)


def tune_permute_auto(on=True):
    """Per-permutation tile tuning of the power-of-two permute (on by default;
    turning it on restores the default tile and forgets tuned shapes)."""
    lib().qtnh_tune_permute_auto(1 if on else 0)


def tune_zgemm(bn=0):
    """Block width (64 or 128 complex columns; 0 chooses by size) of the FP64 complex GEMM (process-wide)."""
    lib().qtnh_tune_zgemm(bn)


def tune_split(slices=5):
    """Number of 8-bit slices of split-precision contractions (process-wide)."""
    lib().qtnh_tune_split(slices)


def tune_split_tile(shape=0):
    """Block tile of the split-precision GEMM (0..3, process-wide); results do not depend on it."""
    lib().qtnh_tune_split_tile(int(shape))


def tune_stream(max_blocks=1 << 20):
    """Most workgroups of the streaming kernel for gates on high bits (process-wide)."""
    lib().qtnh_tune_stream(max_blocks)


def tune_circuit(tile_bits=12, low_bits=4):
    """Tile bits (3..13) of fused circuit runs and the lowest state bits every
    tile keeps (0..8), process-wide; results do not depend on them."""
    lib().qtnh_tune_circuit(tile_bits, low_bits)


def autotune_circuit(n_qubits, device=None):
    """Pick the fused-run tile shape for n-qubit circuits on this context by timing;
    returns (tile_bits, low_bits, ms)."""
    ctx = context(device)
    tb, lb, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0)
    check(lib().qtnh_autotune_circuit(ctx, n_qubits, ctypes.byref(tb), ctypes.byref(lb), ctypes.byref(ms)), ctx)
    return tb.value, lb.value, ms.value


def tune_circuit_pipeline(mode=2):
    """Fused kernel for large states: 3 register-blocked with the next tile in
    flight, 2 register-blocked (default), 1 pipelined, 0 plain; results do not
    depend on it."""
    lib().qtnh_tune_circuit_pipeline(int(mode))


class Circuit:
    """A qubit circuit (gate list) applied in place to resident 2^n states.

    Qubit q is axis q of the rank-n state (axis 0 most significant); gates are
    G[in][out] contracted over `in`, as in the reference's SymmTensor gates."""

    def __init__(self, n_qubits):
        h = ctypes.c_void_p()
        check(lib().qtnh_circuit_create(n_qubits, ctypes.byref(h)))
        self._h = h
        self.n = n_qubits

    def close(self):
        if self._h is not None and self._h.value:
            lib().qtnh_circuit_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def gate(self, qubits, g):
        q, qp = _u16(qubits)
        gh, gp = _c128_host(g)
        if gh.size != (1 << (2 * len(qubits))):
            raise ValueError("gate needs %d elements" % (1 << (2 * len(qubits))))
        check(lib().qtnh_circuit_gate(self._h, len(qubits), qp, gp))
        return self

    def diag(self, qubits, d):
        q, qp = _u16(qubits)
        dh, dp = _c128_host(d)
        if dh.size != (1 << len(qubits)):
            raise ValueError("diagonal needs %d elements" % (1 << len(qubits)))
        check(lib().qtnh_circuit_diag(self._h, len(qubits), qp, dp))
        return self

    def qft(self, phases=None):
        if phases is None:
            check(lib().qtnh_circuit_qft(self._h, None))
        else:
            ph, pp = _c128_host(phases)
            check(lib().qtnh_circuit_qft(self._h, pp))
        return self

    def random(self, depth, seed):
        check(lib().qtnh_circuit_random(self._h, depth, seed))
        return self

    def grid(self, rows, cols, depth, seed):
        check(lib().qtnh_circuit_grid(self._h, rows, cols, depth, seed))
        return self

    def info(self):
        n = ctypes.c_int(0)
        nops = ctypes.c_uint64(0)
        nruns = ctypes.c_uint64(0)
        check(lib().qtnh_circuit_info(self._h, ctypes.byref(n), ctypes.byref(nops), ctypes.byref(nruns)))
        return n.value, nops.value, nruns.value

    def ops(self):
        """[(kind, qubits, coef)] with kind 0 for dense gates, 1 for diagonals."""
        out = []
        _, nops, _ = self.info()
        for i in range(nops):
            kind = ctypes.c_int(0)
            k = ctypes.c_int(0)
            q = np.zeros(3, dtype=np.uint16)
            co = np.zeros(64, dtype=np.complex128)
            check(lib().qtnh_circuit_op(self._h, i, ctypes.byref(kind), ctypes.byref(k),
                                        q.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                        co.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
            size = (1 << (2 * k.value)) if kind.value == 0 else (1 << k.value)
            out.append((kind.value, [int(v) for v in q[:k.value]], co[:size].copy()))
        return out

    def tune(self, scratch=None):
        """Pick this circuit's fused tile by timing its own runs on `scratch`
        (a 2^n complex128 device tensor, overwritten; allocated if None).
        Returns (tile_bits, low_bits, us_per_run); results do not depend on it."""
        if scratch is None:
            scratch = empty(1 << self.n)
        if scratch.numel() != (1 << self.n):
            raise ValueError("scratch has %d amplitudes, circuit needs %d" % (scratch.numel(), 1 << self.n))
        ctx = context(scratch.device.index)
        tb, lb, us = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0)
        check(lib().qtnh_circuit_tune(ctx, self._h, _ptr(scratch), ctypes.byref(tb), ctypes.byref(lb),
                                      ctypes.byref(us)), ctx)
        return tb.value, lb.value, us.value

    def run(self, state, mode=RUN_FUSED):
        if state.numel() != (1 << self.n):
            raise ValueError("state has %d amplitudes, circuit needs %d" % (state.numel(), 1 << self.n))
        ctx = context(state.device.index)
        check(lib().qtnh_circuit_run(ctx, self._h, _ptr(state), mode), ctx)
        return state


def run_batch(circuits, states):
    """Apply circuits[i] to states[i] in place, one launch per run index."""
    if len(circuits) != len(states):
        raise ValueError("one state per circuit")
    for c, s in zip(circuits, states):
        if s.numel() != (1 << c.n):
            raise ValueError("state size does not match its circuit")
    if not circuits:
        return states
    ctx = context(states[0].device.index)
    hs = (ctypes.c_void_p * len(circuits))(*[c.handle.value for c in circuits])
    ps = (ctypes.c_void_p * len(states))(*[s.data_ptr() for s in states])
    for s in states:
        _ptr(s)
    check(lib().qtnh_circuit_run_batch(ctx, len(circuits), hs, ps), ctx)
    return states


def expect_z(state, n_qubits, qubits):
    """[<psi|Z_q|psi> for q in qubits] (host floats; blocks)."""
    if state.numel() != (1 << n_qubits):
        raise ValueError("state size does not match n_qubits")
    ctx = context(state.device.index)
    q, qp = _u16(qubits)
    out = np.zeros(max(len(qubits), 1), dtype=np.float64)
    check(lib().qtnh_expect_z_c128(ctx, _ptr(state), n_qubits, len(qubits), qp,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), ctx)
    return [float(v) for v in out[:len(qubits)]]


class Network:
    """A tensor network of leaf tensors with labelled bonds, planned and
    contracted pairwise on the device (reference TensorNetwork::contractAll)."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().qtnh_net_create(ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h is not None and self._h.value:
            lib().qtnh_net_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def add(self, dims, bonds, data):
        d, dp = _u64(dims)
        b = np.ascontiguousarray(np.asarray(list(bonds) or [0], dtype=np.int64))
        x, xp = _c128_host(data)
        if x.size != _size(dims):
            raise ValueError("tensor data has %d elements, dims need %d" % (x.size, _size(dims)))
        i = ctypes.c_int(0)
        check(lib().qtnh_net_add_tensor(self._h, len(dims), dp, b.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        xp, ctypes.byref(i)))
        return i.value

    @classmethod
    def amplitude(cls, circuit, bits=None):
        """The network of <bits| circuit |0...0>."""
        net = cls()
        if bits is None:
            check(lib().qtnh_net_from_circuit(net._h, circuit.handle, None))
        else:
            bb = np.ascontiguousarray(np.asarray(bits, dtype=np.uint8))
            if bb.size != circuit.n:
                raise ValueError("one bit per qubit")
            check(lib().qtnh_net_from_circuit(net._h, circuit.handle, bb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return net

    def plan(self, seed=0, max_rank=26):
        check(lib().qtnh_net_plan(self._h, seed, max_rank))
        return self

    def plan_best(self, seed=0, max_rank=26, tries=4):
        """The fewest-flop plan among `tries` seeded greedy orders within max_rank;
        returns the chosen seed."""
        chosen = ctypes.c_uint64(0)
        check(lib().qtnh_net_plan_best(self._h, seed, max_rank, tries, ctypes.byref(chosen)))
        return chosen.value

    def set_precision(self, precision):
        """_lib.PREC_FP64 (default) or _lib.PREC_SPLIT for the GEMM-route steps."""
        check(lib().qtnh_net_set_precision(self._h, precision))
        return self

    def info(self):
        nt = ctypes.c_uint64(0)
        ns = ctypes.c_uint64(0)
        mr = ctypes.c_int(0)
        fl = ctypes.c_double(0)
        by = ctypes.c_double(0)
        db = ctypes.c_uint64(0)
        check(lib().qtnh_net_info(self._h, ctypes.byref(nt), ctypes.byref(ns), ctypes.byref(mr), ctypes.byref(fl),
                                  ctypes.byref(by), ctypes.byref(db)))
        return {"tensors": nt.value, "steps": ns.value, "max_rank": mr.value, "flops": fl.value,
                "bytes": by.value, "device_bytes": db.value}

    def result_dims(self):
        r = ctypes.c_int(0)
        d = np.zeros(64, dtype=np.uint64)
        check(lib().qtnh_net_result_dims(self._h, ctypes.byref(r), d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return [int(v) for v in d[:r.value]]

    def run(self, device=None):
        n = _size(self.result_dims())
        out = np.zeros(max(n, 1), dtype=np.complex128)
        ctx = context(device)
        check(lib().qtnh_net_run(ctx, self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n), ctx)
        return out[:n]


class NetworkBatch:
    """Networks prepared once for repeated contraction (optionally replayed as one
    HIP graph); new leaf values go in with set_leaf."""

    def __init__(self, nets, use_graph=True, device=None):
        self.nets = list(nets)
        self._ctx = context(device)
        hs = (ctypes.c_void_p * len(self.nets))(*[n.handle.value for n in self.nets])
        h = ctypes.c_void_p()
        check(lib().qtnh_net_batch_create(self._ctx, len(self.nets), hs, int(bool(use_graph)), ctypes.byref(h)),
              self._ctx)
        self._h = h

    def close(self):
        if self._h is not None and self._h.value:
            lib().qtnh_net_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_leaf(self, net, leaf, data):
        x, xp = _c128_host(data)
        check(lib().qtnh_net_batch_set_leaf(self._ctx, self._h, net, leaf, xp), self._ctx)

    def stats(self):
        """Steps by execution: grouped direct, grouped GEMM-route, one at a time."""
        v = [ctypes.c_uint64(0) for _ in range(3)]
        check(lib().qtnh_net_batch_stats(self._h, *[ctypes.byref(x) for x in v]))
        return {"grouped_direct": v[0].value, "grouped_gemm": v[1].value, "single": v[2].value}

    def run(self):
        outs = [np.zeros(max(_size(n.result_dims()), 1), dtype=np.complex128) for n in self.nets]
        ps = (ctypes.POINTER(ctypes.c_double) * len(outs))(*[o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                                                              for o in outs])
        check(lib().qtnh_net_batch_run(self._ctx, self._h, ps), self._ctx)
        return [o[:_size(n.result_dims())] for o, n in zip(outs, self.nets)]


def run_networks(nets, budget_bytes=1 << 40, device=None):
    """Contract planned networks as one batch; returns (results, waves)."""
    outs = [np.zeros(max(_size(n.result_dims()), 1), dtype=np.complex128) for n in nets]
    if not nets:
        return outs, 0
    ctx = context(device)
    hs = (ctypes.c_void_p * len(nets))(*[n.handle.value for n in nets])
    ps = (ctypes.POINTER(ctypes.c_double) * len(nets))(*[o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                                                          for o in outs])
    waves = ctypes.c_int(0)
    check(lib().qtnh_net_run_batch(ctx, len(nets), hs, ps, budget_bytes, ctypes.byref(waves)), ctx)
    return [o[:_size(n.result_dims())] for o, n in zip(outs, nets)], waves.value


def serve(queries, budget_bytes, policy=_lib.SERVE_BATCHED, device=None):
    """Run the serving loop on a list of dicts (n_qubits, depth, seed, kind,
    basis, arrival_s); returns one dict per query with start/done times,
    latency (done - arrival), value, zsum, batch, cached and route.  policy:
    one of _lib.SERVE_*, with _lib.SERVE_NETWORK optionally or-ed with
    _lib.SERVE_SPLIT / _lib.SERVE_ALL_NETWORK."""
    n = len(queries)
    qs = (_lib.Query * max(n, 1))()
    for i, q in enumerate(queries):
        qs[i] = _lib.Query(q["n_qubits"], q["depth"], q["seed"], q["kind"], q.get("basis", 0), q["arrival_s"])
    rs = (_lib.QueryResult * max(n, 1))()
    ctx = context(device)
    check(lib().qtnh_serve(ctx, n, ctypes.cast(qs, ctypes.c_void_p), ctypes.cast(rs, ctypes.c_void_p),
                           budget_bytes, policy), ctx)
    out = []
    for i in range(n):
        r = rs[i]
        out.append({"start_s": r.start_s, "done_s": r.done_s, "latency_s": r.done_s - queries[i]["arrival_s"],
                    "value": (r.value[0], r.value[1]), "zsum": r.zsum, "batch": r.batch, "cached": r.cached,
                    "route": "network" if r.route else "state"})
    return out


def poisson_queries(count, rate, seed, qubits=(10, 28), depth=(10, 30), pool=None, amp_fraction=0.5):
    """Seeded open-loop workload: Poisson arrivals at `rate` per second, circuit
    sizes uniform in the given ranges; with `pool`, circuits are drawn from that
    many distinct specs (repeats hit resident states)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    specs = None
    if pool:
        specs = [(int(rng.integers(qubits[0], qubits[1] + 1)), int(rng.integers(depth[0], depth[1] + 1)),
                  int(rng.integers(0, 2 ** 62))) for _ in range(pool)]
    out = []
    for _ in range(count):
        t += float(rng.exponential(1.0 / rate)) if rate else 0.0
        if specs:
            n, d, s = specs[int(rng.integers(0, len(specs)))]
        else:
            n, d, s = (int(rng.integers(qubits[0], qubits[1] + 1)), int(rng.integers(depth[0], depth[1] + 1)),
                       int(rng.integers(0, 2 ** 62)))
        kind = _lib.QUERY_AMPLITUDE if rng.random() < amp_fraction else _lib.QUERY_EXPECT_Z
        out.append({"n_qubits": n, "depth": d, "seed": s, "kind": kind,
                    "basis": int(rng.integers(0, 2 ** n)), "arrival_s": t})
    return out
