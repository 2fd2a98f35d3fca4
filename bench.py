"""Benchmark of the extended-precision hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one double-double GEMM C = A*B at m = n = k = 4096 (hi limbs N(0,1),
lo = hi*U(-2^-53, 2^-53), seed 1) on operands already resident in HBM.  Prints ONE JSON
line.  `value` is dd-GEMM GFLOP/s = 2*m*n*k / (seconds per step), counting one dd
multiply-add as 2 flops.

The `roofline` object describes the dominant kernel, timed with HIP events recorded on
the stream the kernel runs on.  The `cpu_baseline` object times the oracle (the C++
op-for-op restatement of the reference's gemm, crates/quad/src/linalg.rs:150-165) on a
bounded sample of rows of the same product on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

M = N = K = 4096
FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 vector (= FP64 matrix) dense peak
DD_MAC_OPS = 29                  # FP64 ops per reference dd MAC (dd add 20 + dd mul 9)


def cpu_baseline(a, b, target_s=12.0):
    """Oracle gemm on a contiguous block of rows; each row is the same work, so the
    rate scales exactly to the full product."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = O.lib().hpo_num_threads()
    t0 = time.perf_counter()
    O.gemm_rows(a, b, M, N, K, 0, threads)  # one row per thread to size the sample
    t1 = time.perf_counter() - t0
    rows = max(threads, min(M, int(threads * target_s / max(t1, 1e-3))))
    rows = (rows // threads) * threads
    t0 = time.perf_counter()
    O.gemm_rows(a, b, M, N, K, threads, rows)
    t = time.perf_counter() - t0
    gflops = 2.0 * rows * N * K / t / 1e9
    return {"value": round(gflops, 4), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"{rows} of {M} rows of the 4096^3 dd gemm ({t:.1f} s), oracle C++ restatement, OpenMP rows"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU layout is described in DESIGN.md and is not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import quadprec_amd as qp

    dev = torch.device("cuda:0")
    a = qp.gen_dd_normal(1, M * K)
    b = qp.gen_dd_normal(1 + 0x5EED, K * N)
    da, db = qp.dev.to_planes(a, 2, dev), qp.dev.to_planes(b, 2, dev)
    dc = torch.empty((2, M * N), dtype=torch.float64, device=dev)

    def step():
        qp.dev.gemm(da, db, dc, M, N, K, mode=qp.HP_GEMM_PLAIN)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        step()
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = sum(s.elapsed_time(e) for s, e in ev) / len(ev)
    ms_per_step = wall * 1e3 / args.steps
    macs = M * N * K
    value = 2.0 * macs / (ms_per_step * 1e-3) / 1e9
    achieved = 2.0 * DD_MAC_OPS * macs / (kern_ms * 1e-3) / 1e12
    out = {
        "metric": "dd-GEMM GFLOP/s @4096^3 (plain dd, reference op order)",
        "value": round(value, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (double-double)",
        "data": "synthetic: seeded N(0,1) dd operands, generated on the host and resident in HBM",
        "config": {"workload": "dd gemm C=A*B, m=n=k=4096, row-major, seed 1", "parallelism": "1 GPU"},
        "roofline": {"bound": "valu", "kernel": "k_gemm_dd_plain", "achieved": round(achieved, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": None, "kernel_ms": round(kern_ms, 4),
                     "note": "29 FP64 VALU ops per dd MAC counted as FMA-equivalent flops (2 each)"},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a, b)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
