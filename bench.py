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


def cpu_baseline(args, ptup_full):
    """The oracle (single-threaded C restatement of the reference path) on a
    rank-`cpu_rank` sample of the same three kernels, scaled per element to
    the full rank."""
    import numpy as np

    import oracle

    r = args.cpu_rank
    n = 1 << r
    rng = random.Random(args.seed + 1)
    ptup = list(range(r))
    rng.shuffle(ptup)
    ptup_gate, gate_axes, g, diag_axes, d = workload(r, args.seed)
    x = oracle.fill(n, args.seed, 1)
    dims = [2] * r
    t0 = time.perf_counter()
    y = oracle.permute(x, dims, ptup)
    t1 = time.perf_counter()
    y = oracle.gate_apply(y, dims, gate_axes, np.asarray(g))
    t2 = time.perf_counter()
    y = oracle.diag_apply(y, dims, diag_axes, np.asarray(d))
    t3 = time.perf_counter()
    scale = float(1 << args.rank) / n
    t_step = (t3 - t0) * scale
    N = float(1 << args.rank)
    flops = 8 * N * 4 + 6 * N
    return {
        "value": flops / t_step / 1e9,
        "unit": "GFLOP/s",
        "cores": 1,
        "kind": "port",
        "sample": "rank-%d state (2^%d elements), one pass each of permute, 2-qubit gate, diagonal phase "
                  "through the oracle; per-element times scaled x%d to rank %d" % (r, r, int(scale), args.rank),
        "ms_per_step": t_step * 1e3,
        "ms_sample": {"permute": (t1 - t0) * 1e3, "gate": (t2 - t1) * 1e3, "diag": (t3 - t2) * 1e3},
        "host_cpu": _cpu_model(),
    }


_KERNEL_PREFIX = {"permute": "permute_bits_kernel", "gate": "gate_", "diag": "diag_bits_kernel"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for a
    plain 4 GiB copy); None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    for name, v in data.get("kernels", {}).items():
        if name.startswith(_KERNEL_PREFIX[kernel]):
            return v["bytes"]
    return None


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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dominant),
                     "algorithmic_bytes_per_launch": bytes_per},
        "wall_s_timed": round(wall, 3),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, ptup)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
