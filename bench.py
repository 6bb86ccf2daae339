"""Kernel microbenchmark of the qtnh local contraction path on one MI355X.

Workload (rank-28 complex128 state, 2^28 elements = 4 GiB, seeded N(0,1)):
  (a) a random full permutation of the 28 axes           (permute, out of place)
  (b) a random 2-qubit dense gate on 2 random axes        (SymmTensor in place)
  (c) a random unit-modulus diagonal phase on 2 axes      (DiagTensor in place)
One step = (a) + (b) + (c); the next step permutes the previous result.

Prints one JSON line.  value = complex GFLOP/s of the step (algorithmic flops
of (b) and (c): 8 N 2^k and 6 N); hbm_pct_peak = the step's algorithmic bytes
(3 x 32 N) per second against 8 TB/s.  Per-kernel times come from HIP events
recorded on the stream the kernels run on.

    python bench.py [--steps K] [--warmup W] [--rank R] [--no-cpu]
"""
import argparse
import json
import math
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "complex GFLOP/s + HBM %peak per contraction; p50/p99 query latency vs FP64"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--rank", type=int, default=28)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--cpu-rank", type=int, default=24, help="rank of the bounded CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--calib-copy", action="store_true",
                   help="run one plain device copy of the state first (HBM counter calibration)")
    p.add_argument("--no-extra", action="store_true", help="skip the other configurations (C1, C3, C4, C5)")
    p.add_argument("--only", default="", help="comma list of extra configurations to run (c2b,c2p,c1,c3c,c3g,c4,c5)")
    p.add_argument("--circuit-rank", type=int, default=30)
    p.add_argument("--circuit-depth", type=int, default=20)
    p.add_argument("--gemm-n", type=int, default=4096)
    p.add_argument("--networks", type=int, default=256)
    p.add_argument("--network-max-rank", type=int, default=26)
    p.add_argument("--serve-queries", type=int, default=48)
    p.add_argument("--serve-rate", type=float, default=8.0, help="Poisson arrival rate, queries per second")
    return p.parse_args()


def workload(rank, seed):
    rng = random.Random(seed)
    ptup = list(range(rank))
    rng.shuffle(ptup)
    gate_axes = rng.sample(range(rank), 2)
    diag_axes = rng.sample(range(rank), 2)
    g = [complex(rng.gauss(0, 1), rng.gauss(0, 1)) for _ in range(16)]
    d = [complex(math.cos(t), math.sin(t)) for t in (rng.uniform(0, 2 * math.pi) for _ in range(4))]
    return ptup, gate_axes, g, diag_axes, d


def _host_threads():
    """Threads for the threaded CPU baseline: the CPUs this process may use,
    at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(args, ptup_full, threads=1, rank=None):
    """The oracle's restatement of the reference path (one thread: the C
    oracle itself; several: the same per-element arithmetic split over POSIX
    threads) on a rank-`rank` sample of the same three kernels, scaled per
    element to the full rank."""
    import numpy as np

    import oracle

    r = rank or args.cpu_rank
    n = 1 << r
    rng = random.Random(args.seed + 1)
    ptup = list(range(r))
    rng.shuffle(ptup)
    ptup_gate, gate_axes, g, diag_axes, d = workload(r, args.seed)
    x = oracle.fill(n, args.seed, 1)
    dims = [2] * r
    t0 = time.perf_counter()
    if threads == 1:
        y = oracle.permute(x, dims, ptup)
        t1 = time.perf_counter()
        y = oracle.gate_apply(y, dims, gate_axes, np.asarray(g), inplace=True)
        t2 = time.perf_counter()
        y = oracle.diag_apply(y, dims, diag_axes, np.asarray(d), inplace=True)
    else:
        y = oracle.permute_mt(x, dims, ptup, threads)
        t1 = time.perf_counter()
        y = oracle.gate_apply_mt(y, dims, gate_axes, np.asarray(g), threads, inplace=True)
        t2 = time.perf_counter()
        y = oracle.diag_apply_mt(y, dims, diag_axes, np.asarray(d), threads, inplace=True)
    t3 = time.perf_counter()
    scale = float(1 << args.rank) / n
    t_step = (t3 - t0) * scale
    N = float(1 << args.rank)
    flops = 8 * N * 4 + 6 * N
    how = "the C oracle, one thread" if threads == 1 else "the oracle's arithmetic on %d POSIX threads" % threads
    return {
        "value": flops / t_step / 1e9,
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": "rank-%d state (2^%d elements), one pass each of permute, 2-qubit gate, diagonal phase "
                  "through %s; per-element times scaled x%d to rank %d" % (r, r, how, int(scale), args.rank),
        "ms_per_step": t_step * 1e3,
        "ms_sample": {"permute": (t1 - t0) * 1e3, "gate": (t2 - t1) * 1e3, "diag": (t3 - t2) * 1e3},
        "host_cpu": _cpu_model(),
    }


_KERNEL_PREFIX = {"permute": "permute_bits_kernel", "gate": "gate_", "diag": "diag_bits_kernel"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel`, read from the committed rocprofv3
    counter summary profiles/hbm_traffic.json (FETCH_SIZE and WRITE_SIZE, each
    scaled by what it reads for a plain 4 GiB copy), not measured in this run;
    and where it comes from.  Every instantiation of the kernel's family that
    the profile holds is listed; the bytes are given only when they agree
    within 1% (the bench's tuner may launch any of them), else None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None, None
    hits = {name: v["bytes"] for name, v in data.get("kernels", {}).items()
            if name.startswith(_KERNEL_PREFIX[kernel])}
    source = {"file": "profiles/hbm_traffic.json", "profiled_kernels": sorted(hits),
              "workload": data.get("workload"), "command": data.get("command")}
    if not hits or max(hits.values()) > 1.01 * min(hits.values()):
        return None, source
    return max(hits.values()), source


def _pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v), q))


def bench_c2b(torch, ops, rank, seed):
    """The same random 2-qubit gate as a DenseTensor contraction with the
    default placement (the gate's output axes last in the result, an output
    permute in the reference): one out-of-place pass, 32 N bytes."""
    import numpy as np
    _, gate_axes, g, _, _ = workload(rank, seed)
    n = 1 << rank
    x = ops.fill_(ops.empty(n), seed)
    gt = torch.tensor(np.asarray(g, dtype=np.complex128), device="cuda")
    out = ops.empty(n)
    dims = [2] * rank
    wires = [(gate_axes[0], 0), (gate_axes[1], 1)]
    ops.contract(x, dims, gt, [2, 2, 2, 2], wires, out=out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 5
    e0.record()
    for _ in range(reps):
        ops.contract(x, dims, gt, [2, 2, 2, 2], wires, out=out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = 32.0 * n / ms / 1e6
    del x, out
    torch.cuda.empty_cache()
    return {"rank": rank, "gate_axes": gate_axes, "placement": "default (gate outputs last)", "ms": round(ms, 4),
            "GB/s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "GFLOP/s": round(8.0 * n * 4 / ms / 1e6, 1)}


def bench_c2_placements(torch, ops, rank, seed):
    """C2 kernels over fixed sets of placements instead of one seeded draw:
    twelve permutations (eight seeded random ones, reversal, rotation,
    lowest <-> highest axis, low 8 axes kept), gates on 1, 2 and 3 axes at
    low, high and mixed positions, diagonals likewise.  GB/s = 32 N bytes per
    launch over its time (HIP events: the median of 7 launches after one untimed)."""
    import numpy as np
    n = 1 << rank
    dims = [2] * rank
    x = ops.fill_(ops.empty(n), seed)
    y = ops.empty(n)
    rng = random.Random(seed + 100)
    perms = []
    for _ in range(8):
        p = list(range(rank))
        rng.shuffle(p)
        perms.append(("random", p))
    perms.append(("reversal", list(range(rank))[::-1]))
    perms.append(("rotation", [(i + 1) % rank for i in range(rank)]))
    sw = list(range(rank))
    sw[0], sw[rank - 1] = rank - 1, 0
    perms.append(("lowest<->highest", sw))
    hi = list(range(rank - 8))
    rng.shuffle(hi)
    perms.append(("low 8 kept", hi + list(range(rank - 8, rank))))
    grng = np.random.default_rng(seed)

    def gate(k):
        m = 1 << k
        return grng.standard_normal(m * m) + 1j * grng.standard_normal(m * m)

    def phases(k):
        return np.exp(2j * np.pi * grng.random(1 << k))

    L = rank - 1
    gates = [([L], gate(1)), ([0], gate(1)), ([rank // 2], gate(1)), ([L, L - 1], gate(2)), ([0, 1], gate(2)),
             ([L, 0], gate(2)), ([rank // 2, rank // 2 + 1], gate(2)), ([L, L - 1, L - 2], gate(3)),
             ([0, 1, 2], gate(3)), ([L, rank // 2, 0], gate(3))]
    diags = [([L], phases(1)), ([0], phases(1)), ([L, L - 1], phases(2)), ([0, 1], phases(2)), ([L, 0], phases(2)),
             ([L, rank // 2, 0], phases(3))]

    def timed(fn, reps=7):
        # Median of per-launch event pairs: one launch delayed by something
        # outside the kernel does not set the placement's figure.
        fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for e0, e1 in ev:
            e0.record()
            fn()
            e1.record()
        torch.cuda.synchronize()
        ms = float(np.median([e0.elapsed_time(e1) for e0, e1 in ev]))
        return round(32.0 * n / ms / 1e6, 1)

    out = {"rank": rank, "permute": [], "gate": [], "diag": []}
    for name, p in perms:
        out["permute"].append({"kind": name, "GB/s": timed(lambda: ops.permute(x, dims, p, out=y))})
    for axes, g in gates:
        out["gate"].append({"axes": axes, "GB/s": timed(lambda: ops.gate_apply_(x, dims, axes, g))})
    for axes, d in diags:
        out["diag"].append({"axes": axes, "GB/s": timed(lambda: ops.diag_apply_(x, dims, axes, d))})
    for k in ("permute", "gate", "diag"):
        fr = sorted(e["GB/s"] / HBM_PEAK_GBS for e in out[k])
        out[k + "_frac"] = {"min": round(fr[0], 3), "median": round(float(np.median(fr)), 3), "max": round(fr[-1], 3)}
    del x, y
    torch.cuda.empty_cache()
    return out


def bench_c1(torch, ops, _lib):
    """QFT-12 (reference examples/qft.cpp, tests/gen/qft.hpp): latency of one
    circuit application to a resident 4096-amplitude state, one launch per
    gate vs fused (default tile, then the tile qtnh_circuit_tune picks for
    this circuit); the results are compared bitwise."""
    n = 12
    c = ops.Circuit(n).qft()
    out = {"qubits": n, "gates": c.info()[1]}
    res = {}
    for name, mode in (("per_gate", _lib.RUN_PER_GATE), ("fused", _lib.RUN_FUSED), ("fused_tuned", _lib.RUN_FUSED)):
        if name == "fused_tuned":
            tb, lb, _ = c.tune()
            out["tuned_tile"] = {"tile_bits": tb, "low_bits": lb, "runs": c.info()[2]}
        s = ops.fill_(ops.empty(1 << n), 7)
        ops.scale_(s, 1 / math.sqrt(ops.norm2(s)))
        x0 = s.clone()
        for _ in range(10):
            s.copy_(x0)
            c.run(s, mode)
        torch.cuda.synchronize()
        lat = []
        for _ in range(200):
            s.copy_(x0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            c.run(s, mode)
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - t0) * 1e6)
        res[name] = s
        out[name] = {"p50_us": round(_pct(lat, 50), 1), "p99_us": round(_pct(lat, 99), 1)}
    pg = torch.view_as_real(res["per_gate"])
    out["bitwise_equal"] = bool(torch.equal(pg, torch.view_as_real(res["fused"])) and
                                torch.equal(pg, torch.view_as_real(res["fused_tuned"])))
    out["cpu"] = cpu_qft(n, res["per_gate"].cpu().numpy(), x0.cpu().numpy())
    return out


def cpu_qft(n, gpu_out, psi0):
    """QFT-n on the host: the C oracle (one thread, the reference's gate
    order and arithmetic) and the same gates through the oracle's threaded
    gate and diagonal functions; per-circuit latency percentiles.  The host
    results are compared bitwise with the GPU's."""
    import numpy as np

    import oracle
    phases = np.zeros(n, dtype=np.complex128)
    for m in range(1, n):
        p = math.pi / 2.0 ** m
        phases[m] = complex(math.cos(p), math.sin(p))
    h = 2.0 ** -0.5
    H = np.array([h, h, h, -h], dtype=np.complex128)
    dims = [2] * n
    threads = _host_threads()
    out = {"threads": threads, "host_cpu": _cpu_model()}
    one, mt = [], []
    for _ in range(20):
        oracle.qft(psi0, n, phases)
    for _ in range(200):
        t0 = time.perf_counter()
        ref = oracle.qft(psi0, n, phases)
        one.append((time.perf_counter() - t0) * 1e6)
    for rep in range(60):
        x = psi0.copy()
        t0 = time.perf_counter()
        for i in range(n):
            oracle.gate_apply_mt(x, dims, [i], H, threads, inplace=True)
            for j in range(i + 1, n):
                oracle.diag_apply_mt(x, dims, [i, j], np.array([1, 1, 1, phases[j - i]]), threads, inplace=True)
        if rep >= 10:
            mt.append((time.perf_counter() - t0) * 1e6)
    out["one_core"] = {"p50_us": round(_pct(one, 50), 1), "p99_us": round(_pct(one, 99), 1)}
    out["threaded"] = {"p50_us": round(_pct(mt, 50), 1), "p99_us": round(_pct(mt, 99), 1)}
    out["bitwise_equal_to_gpu"] = bool(np.array_equal(ref.view(np.uint64), gpu_out.view(np.uint64)) and
                                       np.array_equal(x.view(np.uint64), gpu_out.view(np.uint64)))
    return out


def bench_c3_circuit(torch, ops, _lib, n, depth, seed):
    """Random circuit: every layer a Haar SU(2) on each qubit + CZ on a seeded
    random pairing; one launch per gate vs fused LDS-tiled runs (bit-identical)."""
    tb, lb, _ = ops.autotune_circuit(n)
    c = ops.Circuit(n).random(depth, seed)
    _, nops, _ = c.info()
    N = 1 << n
    a = ops.fill_(ops.empty(N), seed)
    ops.scale_(a, 1 / math.sqrt(ops.norm2(a)))
    b = a.clone()
    out = {"qubits": n, "depth": depth, "gates": nops, "state_GiB": round(N * 16 / 2 ** 30, 2),
           "autotuned_tile": {"tile_bits": tb, "low_bits": lb}}
    ops.tune_circuit(tb, lb)  # the run count below reflects the tuned tile
    nruns = c.info()[2]
    ops.tune_circuit()
    out["fused_runs"] = nruns
    for name, mode, st, passes in (("per_gate", _lib.RUN_PER_GATE, a, nops), ("fused", _lib.RUN_FUSED, b, nruns)):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        c.run(st, mode)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        out[name] = {"ms": round(ms, 2), "passes": passes, "ms_per_pass": round(ms / passes, 3),
                     "GB_per_s_per_pass": round(32.0 * N * passes / ms / 1e6, 1),
                     "gate_GB_per_s": round(32.0 * N * nops / ms / 1e6, 1)}
    out["speedup_fused"] = round(out["per_gate"]["ms"] / out["fused"]["ms"], 2)
    out["bitwise_equal"] = bool(torch.equal(torch.view_as_real(a), torch.view_as_real(b)))
    out["norm_after"] = ops.norm2(b)
    del a, b
    torch.cuda.empty_cache()
    return out


def bench_c3_gemm(torch, ops, _lib, n, seed):
    """Two rank-2r tensors sharing r axes (M = N = K = n): FP64 complex GEMM vs
    split precision (INT8 matrix cores, 5 slices); error vs the FP64 result."""
    r = int(round(math.log2(n)))
    a = ops.fill_(ops.empty(n * n), seed)
    b = ops.fill_(ops.empty(n * n), seed + 1)
    dims = [2] * (2 * r)
    wires = [(r + i, i) for i in range(r)]
    c64 = ops.empty(n * n)
    cs = ops.empty(n * n)
    flops = 8.0 * n ** 3
    out = {"M": n, "N": n, "K": n, "GFLOP": flops / 1e9}
    # Steady state: 3 untimed runs (the clocks settle over the first few
    # milliseconds of matrix-core work), then 10 timed, for both precisions.
    for name, prec, dst in (("fp64", _lib.PREC_FP64, c64), ("split", _lib.PREC_SPLIT, cs)):
        for _ in range(3):
            ops.contract(a, dims, b, dims, wires, precision=prec, out=dst)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 10
        e0.record()
        for _ in range(reps):
            _, _, eta = ops.contract(a, dims, b, dims, wires, precision=prec, out=dst)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"ms": round(ms, 3), "TFLOP/s": round(flops / ms / 1e9, 2)}
        if name == "split":
            out[name]["eta_bound"] = eta
    out["split_speedup"] = round(out["fp64"]["ms"] / out["split"]["ms"], 2)
    out["split_rel_err"] = float(((cs - c64).norm() / c64.norm()).item())
    out["fp64_frac_of_peak"] = round(out["fp64"]["TFLOP/s"] / 78.6, 3)
    # The split path's matrix-core work: three real products (Gauss), each the
    # S (S + 1) / 2 = 15 digit products of 5 slices, 2 n^3 int8 ops apiece, over
    # the whole contraction's time (operand splitting included), against the
    # 5.0 POP/s dense int8 peak.
    int8_ops = 3 * 15 * 2.0 * n ** 3
    out["split"]["int8_POP/s"] = round(int8_ops / out["split"]["ms"] / 1e12, 3)
    out["split"]["int8_frac_of_peak"] = round(out["split"]["int8_POP/s"] / 5.0, 3)
    del a, b, c64, cs
    torch.cuda.empty_cache()
    return out


PLAN_TRIES = 128


def bench_c4(torch, ops, _lib, count, max_rank, seed):
    """Amplitude networks of grid circuits (L x L, L in {4,5,6}, depth 8..16,
    SU(2) + CZ), greedy seeded order; one network at a time vs one batch."""
    import numpy as np
    rng = np.random.default_rng(seed)
    nets, redrawn, s, single = [], 0, 0, []
    tp = time.perf_counter()
    while len(nets) < count:
        L = int(rng.integers(4, 7))
        d = int(rng.integers(8, 17))
        s += 1
        c = ops.Circuit(L * L).grid(L, L, d, seed * 1000 + s)
        net = ops.Network.amplitude(c, rng.integers(0, 2, L * L))
        try:
            # The workload: networks whose seed-s greedy plan fits the rank cap.
            net.plan(s, max_rank)
        except _lib.QtnhError:
            redrawn += 1
            continue
        single.append(net.info()["flops"])
        # Time-aware: the least-cost plan (flops + 1024 per element each step
        # reads and writes) of 128 greedy orders (seed s's own and 127 sampled
        # at low temperature), each with its last tensors in their cheapest
        # order by exhaustive search.
        net.plan_best(s, max_rank, PLAN_TRIES)
        nets.append(net)
    plan_s = time.perf_counter() - tp
    info = [n.info() for n in nets]
    # Every network once, untimed: first-use costs (permute tile tuning for new
    # large permutations, context buffers) stay out of both timed passes.
    for n in nets:
        n.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solo = [n.run() for n in nets]
    t1 = time.perf_counter()
    # The batch once, untimed, as the networks were run once above: its
    # staging buffers are grown to size before the timed pass.
    ops.run_networks(nets)
    t2 = time.perf_counter()
    batch, waves = ops.run_networks(nets)
    t3 = time.perf_counter()
    # The prepared batch below holds all networks at once; the eager pass's
    # workspace (the same size) is released first.
    ops.trim()
    # Prepared once, replayed as one HIP graph (repeated evaluation of the same structures).
    prepared = ops.NetworkBatch(nets, use_graph=True)
    stats = prepared.stats()
    prepared.run()
    reps = 3
    t4 = time.perf_counter()
    for _ in range(reps):
        replay = prepared.run()
    t5 = time.perf_counter()
    prepared.close()
    # The same batch with the GEMM-route steps in split precision (small steps
    # stay exact FP64); one untimed pass first (permute tuning for the split
    # route's operand layouts), error per network against the FP64 results.
    for n in nets:
        n.set_precision(_lib.PREC_SPLIT)
    ops.run_networks(nets)
    torch.cuda.synchronize()
    t6 = time.perf_counter()
    split, _ = ops.run_networks(nets)
    t7 = time.perf_counter()
    for n in nets:
        n.set_precision(_lib.PREC_FP64)
    ops.trim()
    rel = [float(np.abs(s - b).max() / max(np.abs(b).max(), 1e-300)) for s, b in zip(split, batch)]
    flops = sum(i["flops"] for i in info)
    return {"networks": count, "redrawn_over_rank_cap": redrawn, "rank_cap": max_rank,
            "planner": ("least cost (flops + 1024 per element moved) of %d greedy orders within the rank cap "
                        "(the seed's own + sampled), last 10 tensors (12 in the kept order) by exhaustive search"
                        % PLAN_TRIES),
            "plan_s": round(plan_s, 1),
            "GFLOP_first_seed_plans": round(sum(single) / 1e9, 2),
            "steps_by_execution": stats,
            "max_rank": max(i["max_rank"] for i in info), "steps": sum(i["steps"] for i in info),
            "GFLOP": round(flops / 1e9, 2), "GB_moved": round(sum(i["bytes"] for i in info) / 1e9, 2),
            "one_at_a_time_ms": round((t1 - t0) * 1e3, 1), "batched_ms": round((t3 - t2) * 1e3, 1),
            "speedup": round((t1 - t0) / (t3 - t2), 2), "waves": waves,
            "batched_GFLOP_per_s": round(flops / (t3 - t2) / 1e9, 1),
            "graph_replay_ms": round((t5 - t4) / reps * 1e3, 1),
            "graph_replay_speedup": round((t1 - t0) / ((t5 - t4) / reps), 2),
            "split": {"batched_ms": round((t7 - t6) * 1e3, 1),
                      "speedup_vs_one_at_a_time_fp64": round((t1 - t0) / (t7 - t6), 2),
                      "speedup_vs_batched_fp64": round((t3 - t2) / (t7 - t6), 2),
                      "max_rel_err": max(rel), "median_rel_err": float(np.median(rel))},
            "bitwise_equal": all(np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(solo, batch))
            and all(np.array_equal(a.view(np.uint64), b.view(np.uint64)) for a, b in zip(solo, replay))}


def _serve_stats(count, r):
    lat = [x["latency_s"] * 1e3 for x in r]
    return {"p50_ms": round(_pct(lat, 50), 2), "p99_ms": round(_pct(lat, 99), 2),
            "throughput_qps": round(count / max(x["done_s"] for x in r), 2),
            "network_route": sum(1 for x in r if x["route"] == "network")}


def bench_c5(torch, ops, _lib, count, rate, seed):
    """Serving: Poisson arrivals, circuits of 10..28 qubits and depth 10..30
    (random SU(2) + CZ), half amplitude and half <Z> queries, at the arrival
    rate and at saturation (every query arriving at once).  Policies:
    sequential (one launch per gate) and sequential fused runs, one query at a
    time; batched (continuous batching of fused runs, resident states); network
    (batched, each amplitude query on the cheaper route: state vector or its
    tensor network), in FP64 and with split-precision GEMM steps.  Then every
    amplitude query whose network plan fits rank 28 forced onto the network
    route, FP64 vs split, with the per-result error of split against FP64."""
    import numpy as np
    qs = ops.poisson_queries(count, rate, seed, qubits=(10, 28), depth=(10, 30))
    out = {"queries": count, "rate_per_s": rate, "qubits": [10, 28], "depth": [10, 30]}
    pols = (("sequential", _lib.SERVE_SEQUENTIAL), ("sequential_fused", _lib.SERVE_SEQUENTIAL_FUSED),
            ("batched", _lib.SERVE_BATCHED), ("network", _lib.SERVE_NETWORK),
            ("network_split", _lib.SERVE_NETWORK | _lib.SERVE_SPLIT))
    same = lambda a, b: np.array_equal(np.array(a["value"]).view(np.uint64),  # noqa: E731
                                       np.array(b["value"]).view(np.uint64)) and a["zsum"] == b["zsum"]
    sat = [dict(q, arrival_s=0.0) for q in qs]
    res = {}
    out["saturation"] = {}
    for label, qlist, dst in (("rate", qs, out), ("saturation", sat, out["saturation"])):
        for name, pol in pols:
            t0 = time.perf_counter()
            r = ops.serve(qlist, budget_bytes=200 << 30, policy=pol)
            dst[name] = _serve_stats(count, r)
            dst[name]["wall_s"] = round(time.perf_counter() - t0, 2)
            res[(label, name)] = r
        for base in ("sequential", "sequential_fused"):
            dst["throughput_speedup_vs_" + base] = round(
                dst["batched"]["throughput_qps"] / dst[base]["throughput_qps"], 2)
            dst["p99_speedup_vs_" + base] = round(dst[base]["p99_ms"] / dst["batched"]["p99_ms"], 2)
        # State-route answers are identical under every policy.
        dst["answers_bitwise_equal"] = all(
            same(a, b) for name in ("sequential_fused", "batched") for a, b in
            zip(res[(label, "sequential")], res[(label, name)])) and all(
            same(a, b) for name in ("network", "network_split") for a, b in
            zip(res[(label, "batched")], res[(label, name)]) if b["route"] == "state")
    out["answers_equal_at_rate_and_saturation"] = all(
        same(a, b) for a, b in zip(res[("rate", "batched")], res[("saturation", "batched")]))
    # Forced network route at saturation: FP64 vs split precision.
    allnet = _lib.SERVE_NETWORK | _lib.SERVE_ALL_NETWORK
    forced = {}
    for name, pol in (("fp64", allnet), ("split", allnet | _lib.SERVE_SPLIT)):
        r = ops.serve(sat, budget_bytes=200 << 30, policy=pol)
        forced[name] = r
        out["saturation"]["all_network_" + name] = _serve_stats(count, r)
    sv = res[("saturation", "batched")]
    errs, rels, errs64 = [], [], []
    for a, b, c in zip(forced["split"], forced["fp64"], sv):
        if a["route"] != "network":
            continue
        va, vb, vc = complex(*a["value"]), complex(*b["value"]), complex(*c["value"])
        errs.append(abs(va - vb))
        rels.append(abs(va - vb) / max(abs(vb), 1e-300))
        errs64.append(abs(vb - vc))
    out["saturation"]["all_network_errors"] = {
        "network_answers": len(errs),
        "split_vs_fp64_network_abs": {"max": max(errs, default=0.0), "median": float(np.median(errs)) if errs else 0.0},
        "split_vs_fp64_network_rel": {"max": max(rels, default=0.0), "median": float(np.median(rels)) if rels else 0.0},
        "fp64_network_vs_state_vector_abs_max": max(errs64, default=0.0),
        "per_result_abs": [float("%.3e" % e) for e in errs]}
    return out


def extra_configs(args):
    import torch

    from qtnh_amd import _lib, ops

    only = set(x for x in args.only.split(",") if x)
    want = lambda k: not only or k in only  # noqa: E731
    out = {}
    if want("c2b"):
        out["C2b_dense_gate_default_placement"] = bench_c2b(torch, ops, args.rank, args.seed)
    if want("c2p"):
        out["C2_placements"] = bench_c2_placements(torch, ops, args.rank, args.seed)
    if want("c1"):
        out["C1_qft12_latency"] = bench_c1(torch, ops, _lib)
    if want("c3c"):
        out["C3_random_circuit"] = bench_c3_circuit(torch, ops, _lib, args.circuit_rank, args.circuit_depth, args.seed)
    if want("c3g"):
        out["C3_gemm"] = bench_c3_gemm(torch, ops, _lib, args.gemm_n, args.seed)
    if want("c4"):
        out["C4_batched_networks"] = bench_c4(torch, ops, _lib, args.networks, args.network_max_rank, args.seed)
    if want("c5"):
        out["C5_serving"] = bench_c5(torch, ops, _lib, args.serve_queries, args.serve_rate, args.seed)
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    from qtnh_amd import ops

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    rank = args.rank
    N = 1 << rank
    dims = [2] * rank
    ptup, gate_axes, g, diag_axes, d = workload(rank, args.seed)
    x = ops.fill_(ops.empty(N), args.seed)
    y = ops.empty(N)
    stream = torch.cuda.current_stream()
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def step(src, dst, record=None):
        if record is not None:
            record[0].record(stream)
        ops.permute(src, dims, ptup, out=dst)
        if record is not None:
            record[1].record(stream)
        ops.gate_apply_(dst, dims, gate_axes, g)
        if record is not None:
            record[2].record(stream)
        ops.diag_apply_(dst, dims, diag_axes, d)
        if record is not None:
            record[3].record(stream)
        return dst, src

    if args.calib_copy:
        y.copy_(x)
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        x, y = step(x, y)
    torch.cuda.synchronize()
    evs = [[ev() for _ in range(4)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    start, stop = ev(), ev()
    start.record(stream)
    for s in range(args.steps):
        x, y = step(x, y, evs[s])
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = start.elapsed_time(stop)
    ms = {"permute": 0.0, "gate": 0.0, "diag": 0.0}
    for e in evs:
        ms["permute"] += e[0].elapsed_time(e[1])
        ms["gate"] += e[1].elapsed_time(e[2])
        ms["diag"] += e[2].elapsed_time(e[3])
    for k in ms:
        ms[k] /= args.steps
    ms_step = total_ms / args.steps
    bytes_per = 32.0 * N  # read + write of every element
    flops = {"permute": 0.0, "gate": 8.0 * N * 4, "diag": 6.0 * N}
    kernels = {}
    for k in ms:
        kernels[k] = {"ms": ms[k], "GB/s": bytes_per / ms[k] / 1e6, "hbm_frac": bytes_per / ms[k] / 1e6 / HBM_PEAK_GBS,
                      "GFLOP/s": flops[k] / ms[k] / 1e6}
    dominant = max(ms, key=ms.get)
    step_flops = flops["gate"] + flops["diag"]
    value = step_flops / ms_step / 1e6
    achieved = bytes_per / ms[dominant] / 1e6
    traffic, traffic_source = measured_traffic(dominant)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic: seeded N(0,1) complex128 state (on-device SplitMix64 + Box-Muller), seeded gate and phases",
        "config": {"workload": "rank-%d complex128 state (2^%d el, %.1f GiB): random full-axis permute + random "
                               "2-qubit dense gate (SymmTensor, in place) + random 2-axis diagonal phase "
                               "(DiagTensor, in place)" % (rank, rank, N * 16 / 2 ** 30),
                   "rank": rank, "ptup_seed": args.seed, "gate_axes": gate_axes, "diag_axes": diag_axes},
        "hbm_pct_peak": round(100.0 * 3 * bytes_per / ms_step / 1e6 / HBM_PEAK_GBS, 2),
        "kernels": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in kernels.items()},
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_source, "algorithmic_bytes_per_launch": bytes_per},
        "wall_s_timed": round(wall, 3),
    }
    del x, y
    torch.cuda.empty_cache()
    if not args.no_extra:
        out["configs"] = extra_configs(args)
    if not args.no_cpu:
        threads = _host_threads()
        out["cpu_baseline"] = cpu_baseline(args, ptup, threads=threads, rank=args.cpu_rank + 2 if threads > 1 else None)
        out["cpu_baseline_1core"] = cpu_baseline(args, ptup)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
