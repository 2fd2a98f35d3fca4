"""Benchmark of the extended-precision hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one double-double GEMM C = A*B at m = n = k = 4096 (hi limbs N(0,1),
lo = hi*U(-2^-53, 2^-53), seed 1) on operands already resident in HBM, computed by the
split-precision path (INT8 matrix cores, 30-modulus CRT; its accuracy against the f256
oracle is established by tests/test_gpu_split.py).  The plain dd kernel (reference op
order, bit-identical) is timed on the same operands for the split-vs-plain ratio.

Prints ONE JSON line.  `value` is dd-GEMM GFLOP/s = 2*m*n*k / (seconds per step),
counting one dd multiply-add as 2 flops.  The `roofline` object describes the dominant
kernel of the step (the int8 GEMMs), timed with HIP events recorded on the stream it runs
on.  The `cpu_baseline` object times the oracle (the C++ op-for-op restatement of the
reference's gemm, crates/quad/src/linalg.rs:150-165) on a bounded sample of rows.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

M = N = K = 4096
DD_BITS = 104                    # accuracy the headline asks the autotuner for (dd)
SLICES = 30                      # CRT moduli the autotuner picks for DD_BITS at k = 4096
FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 vector (= FP64 matrix) dense peak
INT8_PEAK_TOPS = 5000.0          # MI355X dense INT8 matrix peak
DD_MAC_OPS = 29                  # FP64 ops per reference dd MAC (dd add 20 + dd mul 9)
# int8 GEMM kernel per autotuner variant code
I8_KERNELS = {32: "k_gemm_i8_mod<32>", 64: "k_gemm_i8_mod<64>", 128: "k_gemm_i8_mod<128>", 3064: "k_gemm_i8_mod_dl",
              4064: "k_gemm_i8_mod_dlp8", 5064: "k_gemm_i8_mod_dlp4", 6064: "k_gemm_i8_mod_rs",
              7064: "k_gemm_i8_mod_dlpp"}
# HBM bytes one launch at 4096^3 x 30 moduli moves, from rocprofv3 FETCH_SIZE / WRITE_SIZE
# counters scaled by a device copy of known size (k_gemm_i8_mod_dlp8 and its persistent form
# on the 64-deep K-slab residue layout: read 3.020e9 + write 0.503e9); null for a kernel
# these were not measured on
I8_TRAFFIC_BYTES = {4064: 3.523e9, 7064: 3.524e9}


def oracle():
    """The oracle (tests/oracle_lib.py): only the cpu_baseline legs and the error checks of
    the legs below use it, never inside a timed GPU region."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    return O


def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "threads": oracle().lib().hpo_num_threads(), "nproc": len(os.sched_getaffinity(0))}


def best_of(fn, reps, torch):
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return min(ts) * 1e3


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


SOLVE_N = 4096
SOLVE_CASES = [  # (name, alpha_lo, spacing, target rel err, split-path factor, seed)
    ("a_kappa_small", 0.5, 1.4, 1e-30, "f64", 2),
    ("b_kappa_1e16", 0.01, 1.4, 1e-28, "split", 2),
    ("b_extreme_kappa_1e25", 0.001, 0.7, 1e-28, "split", 2),
]


def dd_rel_err(O, x, xs):
    import numpy as np
    d = O.dd_binop("sub", x, xs)
    return float(np.max(np.abs(d[:, 0] + d[:, 1])) / np.max(np.abs(xs[:, 0])))


def oracle_refined_solution(O, s, b, n, l, x0):
    """x*: corrections of x0 through the oracle's factor l with MultiFloat<4> residuals b - S x
    (crates/quad/src/multifloat/linalg.rs:17-36) until they stop at x's dd rounding level."""
    import numpy as np

    def mf4(v):
        out = np.zeros((v.shape[0], 4))
        out[:, :2] = v
        return out
    x, s4, b4, prev = x0, mf4(s), mf4(b), np.inf
    for _ in range(12):
        r4 = O.mf4_binop("sub", b4, O.mf4_gemv(s4, mf4(x), n, n))
        dx = O.solve_cholesky(l, np.ascontiguousarray(r4[:, :2]), n)
        x = O.dd_binop("add", x, dx)
        rel = np.max(np.abs(dx[:, 0])) / np.max(np.abs(x[:, 0]))
        if rel <= 1e-34 or (rel <= 1e-31 and rel > 0.25 * prev):
            break
        prev = rel
    return x


def oracle_iterative_refinement(O, s, b, n):
    """The reference's CPU path for a dd system: cholesky_blocked(nb = 64)
    (crates/quad/src/linalg.rs:238-293), solve_cholesky, then iterative_refinement's loop
    (crates/solver/src/lib.rs:92-116): r = b - S x with the dd matvec, stop when
    ||r|| < 1e-28 ||b|| or, after 3 steps, when ||r|| > 0.5 best; dx = solve_cholesky(L, r)."""
    l = O.cholesky_blocked(s, n, 64)
    x = O.solve_cholesky(l, b, n)
    bn = O.norm_f128(b)
    best, iters = float("inf"), 0
    for it in range(30):
        r = O.dd_binop("sub", b, O.matvec(s, x, n))
        rn = O.norm_f128(r)
        if rn < 1e-28 * bn or (it > 3 and rn > 0.5 * best):
            break
        best = min(best, rn)
        x = O.dd_binop("add", x, O.solve_cholesky(l, r, n))
        iters += 1
    return l, x, iters


def split_factor_work_ms(n, nb=256, slices=SLICES):
    """Lower bound of the split-dd factor's time from its work at peak rates: the panels'
    dd operations on FP64 VALU (the panel's own columns ~ n^2 nb / 4 MACs, the next panel's
    diagonal blocks ~ n nb^2 / 2) at 29 ops per MAC, and the remaining n^3/6 MACs as split
    products, 2 slices int8 ops each, on the INT8 matrix cores."""
    valu_macs = n * n * nb / 4.0 + n * nb * nb / 2.0
    i8_macs = n ** 3 / 6.0 - valu_macs
    return (DD_MAC_OPS * valu_macs / (FP64_PEAK_TFLOPS / 2 * 1e12) + 2 * slices * i8_macs / (INT8_PEAK_TOPS * 1e12)) * 1e3


def bench_solves(qp, torch, dev, cpu):
    """N=4096 overlap SPD solves to the target error: split path vs the plain dd factor,
    both refined with quad-double residuals (time = factor + refinement, device-synced).
    Each solution's error is measured against x* (the oracle's dd factor corrected with
    MultiFloat<4> residuals, computed on the host after the timed runs)."""
    import numpy as np
    kinds = {"f64": qp.HP_FACTOR_F64, "dd": qp.HP_FACTOR_DD, "split": qp.HP_FACTOR_DD_SPLIT}
    out = {}
    n = SOLVE_N
    O = oracle()
    for name, lo, sp, target, split_kind, seed in SOLVE_CASES:
        c, a = qp.gen_overlap_basis(seed, n, lo, 10.0, sp)
        dc = torch.from_numpy(c.ravel()).to(dev)
        da = torch.from_numpy(a).to(dev)
        S = torch.empty((2, n * n), dtype=torch.float64, device=dev)
        qp.dev.overlap_s(dc, da, S, n)
        b_host = qp.gen_dd_normal(seed + 100, n)
        b = qp.dev.to_planes(b_host, 2, dev)
        x = torch.empty((2, n), dtype=torch.float64, device=dev)
        res, xs_gpu = {}, {}
        for label, kind in (("split", split_kind), ("plain", "dd")):
            best = None
            for _ in range(2):  # the first run also warms the autotuner and allocations
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                st = qp.dev.solve_spd_refined(S, b, x, n, kinds[kind], tol=target * 1e-2, max_iter=30)
                torch.cuda.synchronize()
                t = (time.perf_counter() - t0) * 1e3
                if best is None or t < best[0]:
                    best = (t, st.factor_ms, st.refine_ms, st.iterations, st.last_correction, st.converged)
            xs_gpu[label] = qp.dev.from_planes(x)
            res[label] = {"factor": kind, "ms": round(best[0], 2), "factor_ms": round(best[1], 2),
                          "refine_ms": round(best[2], 2), "iterations": best[3],
                          "last_correction": float(f"{best[4]:.3e}"),
                          "converged": {0: "no", 1: "measured", 2: "predicted"}[best[5]]}
            if kind == "dd":
                ach = 2.0 * DD_MAC_OPS * n ** 3 / 6.0 / (best[1] * 1e-3) / 1e12
                res[label]["roofline"] = {
                    "bound": "valu", "kernel": "reference-order factor (k_diag_factor, k_row_recur, k_rank_update)",
                    "achieved": round(ach, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(ach / FP64_PEAK_TFLOPS, 4),
                    "note": "n^3/6 dd MACs x 29 FP64 ops (FMA-equivalent) over factor_ms"}
            elif kind == "f64":
                ach = 2.0 * n ** 3 / 6.0 / (best[1] * 1e-3) / 1e12
                res[label]["roofline"] = {"bound": "valu", "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS,
                                          "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4),
                                          "note": "n^3/6 binary64 MACs (2 flops) over factor_ms: the 64-wide "
                                                  "diagonal blocks' pivot chain bounds it, not the FLOPs"}
            else:
                wb = split_factor_work_ms(n)
                res[label]["roofline"] = {
                    "bound": "latency", "work_bound_ms": round(wb, 3), "frac": round(wb / best[1], 4),
                    "note": "frac = the factor's work at the FP64 VALU and INT8 peaks over factor_ms; the "
                            "factor is held by the serial chain of 64 diagonal 64x64 sub-block factorisations "
                            "(k_diag_factor_split, ~35-55 us each)"}
        # x* and the errors (host, after the timed runs)
        s_host = O.overlap_s_dd(c, a)
        t0 = time.perf_counter()
        l_cpu, x_cpu, it_cpu = oracle_iterative_refinement(O, s_host, b_host, n)
        t_cpu = time.perf_counter() - t0
        xstar = oracle_refined_solution(O, s_host, b_host, n, l_cpu, x_cpu)
        for label in ("split", "plain"):
            res[label]["rel_err"] = float(f"{dd_rel_err(O, xs_gpu[label], xstar):.3e}")
        res["alpha"] = [lo, 10.0]
        res["spacing_bohr"] = sp
        res["target_rel_err"] = target
        res["split_over_plain"] = round(res["plain"]["ms"] / res["split"]["ms"], 2)
        if cpu:
            res["cpu_baseline"] = {
                "value": round(t_cpu * 1e3, 1), "unit": "ms", "cores": O.lib().hpo_num_threads(), "kind": "port",
                "rel_err": float(f"{dd_rel_err(O, x_cpu, xstar):.3e}"), "refinement_iters": it_cpu,
                "sample": "the full case: oracle cholesky_blocked(nb=64) + solve_cholesky + iterative_refinement's "
                          "dd-residual loop (lib.rs:92-116), OpenMP rows; its error is what dd residuals reach"}
        out[name] = res
    return out


def bench_batch(qp, torch, dev, cpu, count=512, nrhs=4, seed=3):
    """Config 3: 512 overlap systems with n ~ U[64, 1024] (seeded), each from its own seed,
    exponent floor log-uniform in [1e-3, 0.5] and lattice spacing U[0.7, 1.4] bohr so the
    conditioning spans several decades; 4 dd N(0,1) right-hand sides each.  Factored and
    solved in the schedule the autotuner picks; timed as a replayed hipGraph (both schedules
    are timed and reported)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = rng.integers(64, 1025, count).astype(np.int32)
    lo = 10 ** rng.uniform(np.log10(1e-3), np.log10(0.5), count)
    sp = rng.uniform(0.7, 1.4, count)
    tot = int((sizes.astype(np.int64) ** 2).sum())
    tr = int(sizes.astype(np.int64).sum()) * nrhs
    S = torch.empty((2, tot), dtype=torch.float64, device=dev)
    B = qp.dev.to_planes(qp.gen_dd_normal(seed + 7, tr), 2, dev)
    L = torch.empty_like(S)
    X = torch.empty_like(B)
    st = torch.empty(count, dtype=torch.int32, device=dev)
    off = 0
    for i, n in enumerate(sizes):
        n = int(n)
        c, a = qp.gen_overlap_basis(seed * 100000 + i, n, float(lo[i]), 10.0, float(sp[i]))
        dc = torch.from_numpy(c.ravel()).to(dev)
        da = torch.from_numpy(a).to(dev)
        qp.dev.overlap_s(dc, da, S[:, off:off + n * n], n)
        off += n * n
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    qp.dev.batch_cholesky_solve(sizes, nrhs, S, L, B, X, st)  # first call: the autotuner times both schedules
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3
    stream = torch.cuda.current_stream()

    def replay_ms(g, reps=3):
        g.launch()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            g.launch()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    per_schedule = {}
    for sched, label in ((qp.HP_BATCH_PER_INSTANCE, "per_instance"), (qp.HP_BATCH_PANEL_STEPS, "panel_steps")):
        qp.dev.set_batch_schedule(sched)
        gp = qp.dev.batch_cholesky_solve(sizes, nrhs, S, L, B, X, st, graph=True)
        per_schedule[label] = round(replay_ms(gp), 2)
        del gp
    qp.dev.set_batch_schedule(0)
    g = qp.dev.batch_cholesky_solve(sizes, nrhs, S, L, B, X, st, graph=True)  # the autotuned schedule
    tuned = {qp.HP_BATCH_PER_INSTANCE: "per_instance", qp.HP_BATCH_PANEL_STEPS: "panel_steps"}[g.schedule]
    ms = replay_ms(g)
    macs = float((sizes.astype(np.float64) ** 3).sum() / 6.0)
    solve_macs = float(nrhs * (sizes.astype(np.float64) ** 2).sum())
    failed = int((st.cpu().numpy() != 0).sum())
    ach = 2.0 * DD_MAC_OPS * (macs + solve_macs) / (ms * 1e-3) / 1e12
    out = {"instances": count, "n_range": [int(sizes.min()), int(sizes.max())], "nrhs": nrhs,
           "ms_graph_replay": round(ms, 2), "schedule": tuned, "ms_graph_replay_by_schedule": per_schedule,
           "ms_eager_first_call_incl_tuning": round(eager_ms, 2),
           "factor_macs": macs, "dd_gflops": round(2.0 * macs / (ms * 1e-3) / 1e9, 1),
           "breakdowns": failed, "matrix_bytes": 2 * 16 * tot,
           "roofline": {"bound": "valu", "kernel": "k_batch_chol_solve", "achieved": round(ach, 3),
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4),
                        "note": "(sum n^3/6 factor + nrhs sum n^2 solve) dd MACs x 29 FP64 ops (FMA-equivalent) "
                                "over the replay of the autotuned schedule (per_instance: one workgroup per "
                                "instance; panel_steps: each panel step of all instances across the chip)"}}
    if cpu:
        O = oracle()
        pick = list(range(0, count, count // 16))[:16]
        mats = []
        for i in pick:
            c, a = qp.gen_overlap_basis(seed * 100000 + i, int(sizes[i]), float(lo[i]), 10.0, float(sp[i]))
            mats.append(O.overlap_s_dd(c, a))
        rhs = [qp.gen_dd_normal(99 + i, int(sizes[i])) for i in pick]
        t0 = time.perf_counter()
        for j, i in enumerate(pick):
            n = int(sizes[i])
            l = O.cholesky_blocked(mats[j], n, 64)
            for _ in range(nrhs):
                O.solve_cholesky(l, rhs[j], n)
        t = time.perf_counter() - t0
        smacs = float(sum(float(sizes[i]) ** 3 / 6.0 for i in pick))
        rate = 2.0 * smacs / t / 1e9
        out["cpu_baseline"] = {"value": round(rate, 3), "unit": "dd-GFLOP/s (factor MACs)",
                               "est_ms_full_batch": round(2.0 * macs / (rate * 1e9) * 1e3, 1),
                               "cores": O.lib().hpo_num_threads(), "kind": "port",
                               "sample": f"16 of the {count} instances (every {count // 16}th, n "
                                         f"{min(int(sizes[i]) for i in pick)}..{max(int(sizes[i]) for i in pick)}): "
                                         f"oracle cholesky_blocked(nb=64) + {nrhs} solve_cholesky each ({t:.1f} s), "
                                         "OpenMP rows"}
    return out


MF4_N = 1024
# mean FP64 operations of one MF4 MAC s = s + a*b along the accumulation chains of these
# inputs, from the oracle's op-counting restatement (re-derived by tests/test_oracle_mf4.py)
MF4_OPS_PER_MAC_MEAN = 531


def bench_mf4(qp, torch, dev, steps, cpu):
    """Config 1, second part: MultiFloat<4> mf_gemm C = A*B at 1024^3 (limbs generated as
    the dd operands, seed 1), the reference's per-element order (bit-identical)."""
    n = MF4_N
    a = qp.gen_mf4_normal(1, n * n)
    b = qp.gen_mf4_normal(1 + 0x5EED, n * n)
    da, db = qp.dev.to_planes(a, 4, dev), qp.dev.to_planes(b, 4, dev)
    dc = torch.empty((4, n * n), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    ms, kern = time_steps(lambda: qp.dev.mf_gemm(da, db, dc, n, n, n), max(2, min(steps, 5)), 1, stream, torch)
    macs = float(n) ** 3
    ops = MF4_OPS_PER_MAC_MEAN * macs / (kern * 1e-3) / 1e12
    out = {"workload": "MF4 mf_gemm C=A*B, m=n=k=1024, seed 1", "ms_per_step": round(ms, 3),
           "value": round(2.0 * macs / (ms * 1e-3) / 1e9, 3), "unit": "MF4 GFLOP/s (2 per MAC)",
           "roofline": {"bound": "valu", "kernel": "k_mf4_gemm", "kernel_ms": round(kern, 3),
                        "ops_per_mac": MF4_OPS_PER_MAC_MEAN, "achieved": round(2.0 * ops, 3),
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(2.0 * ops / FP64_PEAK_TFLOPS, 4),
                        "note": "FP64 VALU ops per MAC from the op-counting restatement on these inputs (mean), "
                                "FMA-equivalent (2 flops each)"}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        threads = O.lib().hpo_num_threads()
        rows = threads * 2
        t0 = time.perf_counter()
        O.mf4_gemm_rows(a, b, n, n, n, 0, rows)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(2.0 * rows * n * n / t / 1e9, 4), "unit": "MF4 GFLOP/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{rows} of {n} rows ({t:.1f} s), oracle C++ restatement, OpenMP rows"}
    return out


MFN_NL, MFN_N = 8, 512
# FP64 operations per f512 MAC, the op-counting restatement's mean over four accumulation
# chains of these inputs (tests/test_oracle_mfn.py re-derives it)
MFN8_OPS_PER_MAC_MEAN = 1966


def bench_mfn(qp, torch, dev, steps, cpu):
    """f512 = MultiFloat<8> mf_gemm C = A*B at 512^3 (multifloat/linalg.rs:39-61; operands
    from gen_mfn_normal(8, 1 / 2, .)), the reference's per-element order (bit-identical,
    tests/test_gpu_mfn.py)."""
    nl, n = MFN_NL, MFN_N
    a = qp.gen_mfn_normal(nl, 1, n * n)
    b = qp.gen_mfn_normal(nl, 2, n * n)
    da, db = qp.dev.to_planes(a, nl, dev), qp.dev.to_planes(b, nl, dev)
    dc = torch.empty((nl, n * n), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    ms, kern = time_steps(lambda: qp.dev.mfn_gemm(nl, da, db, dc, n, n, n), max(2, min(steps, 5)), 1, stream, torch)
    macs = float(n) ** 3
    ops = MFN8_OPS_PER_MAC_MEAN * macs / (kern * 1e-3) / 1e12
    out = {"workload": f"f512 (MultiFloat<8>) mf_gemm C=A*B, m=n=k={n}, gen_mfn_normal seeds 1, 2",
           "ms_per_step": round(ms, 3), "value": round(2.0 * macs / (ms * 1e-3) / 1e9, 3),
           "unit": "f512 GFLOP/s (2 per MAC)",
           "roofline": {"bound": "valu", "kernel": "k_mfn_gemm<8>", "kernel_ms": round(kern, 3),
                        "ops_per_mac": MFN8_OPS_PER_MAC_MEAN, "achieved": round(2.0 * ops, 3),
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(2.0 * ops / FP64_PEAK_TFLOPS, 4),
                        "note": "FP64 VALU ops per MAC from the op-counting restatement on these inputs (mean), "
                                "FMA-equivalent (2 flops each); the sorts and compactions add selects and compares "
                                "beyond the counted arithmetic"}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        threads = O.lib().hpo_num_threads()
        rows = threads
        t0 = time.perf_counter()
        O.mfn_gemm(nl, a, b, n, n, n, row0=0, nrows=rows)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(2.0 * rows * n * n / t / 1e9, 4), "unit": "f512 GFLOP/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{rows} of {n} rows ({t:.1f} s), oracle C++ restatement of the generic "
                                         "MultiFloat<N> code, OpenMP rows"}
    return out


def bench_eigen(qp, torch, dev, cpu):
    """The consumer of S: solve_gen_eigen F C = S C eps (crates/solver/src/ortho.rs:201-248)
    with S the config-0 overlap matrix (lattice, seed 0, exponents in [0.01, 10]) and F a
    seeded symmetric dd matrix, threshold 1e-8; both Jacobi eigenproblems on the GPU
    (parallel round-robin sweeps to max |a_pq| < 1e-60).  Accuracy against the oracle:
    tests/test_gpu_eigen.py."""
    import ctypes as C
    import numpy as np
    L = qp.lib()

    def inputs(n):
        c, a = qp.gen_overlap_basis(0, n, 0.01, 10.0)
        s_dd = qp.overlap_s(c, a)
        rng = np.random.default_rng(1)
        b = rng.standard_normal((n, n)) / np.sqrt(n)
        f_dd = np.zeros((n * n, 2))
        f_dd[:, 0] = ((b + b.T) / 2).ravel()
        return f_dd, s_dd

    def gpu(n, reps=2):
        f_dd, s_dd = inputs(n)
        fd, sd = qp.dev.to_planes(f_dd, 2, dev), qp.dev.to_planes(s_dd, 2, dev)
        ev = torch.zeros((2, n), dtype=torch.float64, device=dev)
        cc = torch.zeros((2, n * n), dtype=torch.float64, device=dev)
        m, sw = C.c_size_t(0), C.c_size_t(0)
        best = None
        for _ in range(reps + 1):  # the first run also builds the sweep graphs' allocations
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rc = L.hp_dev_solve_gen_eigen(fd[0].data_ptr(), fd[1].data_ptr(), sd[0].data_ptr(), sd[1].data_ptr(), n,
                                          1e-8, ev[0].data_ptr(), ev[1].data_ptr(), cc[0].data_ptr(),
                                          cc[1].data_ptr(), C.byref(m), C.byref(sw), None)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) * 1e3
            if rc:
                raise RuntimeError(f"hp_dev_solve_gen_eigen: status {rc}")
            best = dt if best is None else min(best, dt)
        return best, m.value, sw.value, (f_dd, s_dd)

    n = EIG_N
    ms, m, sweeps, _ = gpu(n)
    n2 = n + (n & 1)
    steps = sweeps * (n2 - 1)
    # A is read and written once per step (32 N^2 bytes of dd); V once per batch of up to
    # 128 stored rotation steps, the batches restarting at each sweep
    vbatches = sweeps * -(-(n2 - 1) // 128)
    nbytes = 32.0 * n2 * n2 * (steps + vbatches)
    out = {"workload": f"solve_gen_eigen, S = overlap N={n} (lattice, seed 0, alpha in [0.01, 10]), F seeded "
                       "symmetric dd, threshold 1e-8",
           "ms": round(ms, 2), "n_basis": m, "jacobi_sweeps": sweeps,
           "roofline": {"bound": "hbm", "kernel": "k_jac_step",
                        "achieved": round(nbytes / (ms * 1e-3) / 1e9, 1), "peak": 8000.0,
                        "unit": "GB/s", "frac": round(nbytes / (ms * 1e-3) / 8e12, 4),
                        "note": "32 N^2 bytes per round-robin step (A read and written, dd) plus 32 N^2 per batch of "
                                "up to 128 steps applied to V, over the whole call's time; the steps are latency "
                                "bound (each step's rotations need the previous step's diagonal: loads, a dependent "
                                "dd divide / sqrt chain, the update, one launch per step)"}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        nc = EIG_CPU_N
        ms_g, _, _, (f_dd, s_dd) = gpu(nc)
        t0 = time.perf_counter()
        O.solve_gen_eigen(f_dd, s_dd, nc, 1e-8)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(t * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                               "sample": f"the full solve at N={nc} (max-pivot Jacobi as ortho.rs:61-152, "
                                         f"oracle C++; one core: the reference's Jacobi sweeps are serial), "
                                         f"against {ms_g:.1f} ms on the GPU at the same N"}
    return out


def bench_text(qp, torch, dev, cpu, reps=5):
    """The serde string form of f128 (dd.rs:446-450 and 496-511) over TEXT_N dd values
    resident in HBM: format to Arrow-style text + offsets in device memory, and parse that
    text back (bit-identical round trip; parity in tests/test_gpu_text.py).  Bytes per value:
    16 in, 53 of text + 8 of offset out (format); the reverse for parse."""
    import ctypes as C
    import numpy as np
    L = qp.lib()
    n = TEXT_N
    rng = np.random.default_rng(2)
    hi = rng.standard_normal(n) * 10.0 ** rng.integers(-300, 300, n)
    lo = hi * 2.0 ** -53 * rng.uniform(-0.5, 0.5, n)
    dhi, dlo = torch.tensor(hi, device=dev), torch.tensor(lo, device=dev)
    planes = (C.c_void_p * 4)(dhi.data_ptr(), dlo.data_ptr(), 0, 0)
    text = torch.zeros(53 * n, dtype=torch.uint8, device=dev)
    off = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    ohi, olo = torch.empty_like(dhi), torch.empty_like(dlo)
    oplanes = (C.c_void_p * 4)(ohi.data_ptr(), olo.data_ptr(), 0, 0)
    status = torch.empty(n, dtype=torch.uint8, device=dev)
    need, bad = C.c_size_t(0), C.c_size_t(0)

    def fmt():
        rc = L.hp_dev_format(3, C.cast(planes, C.c_void_p), n, 18, 1, C.c_void_p(text.data_ptr()), 53 * n,
                             C.c_void_p(off.data_ptr()), C.byref(need), None)
        if rc:
            raise RuntimeError(f"hp_dev_format: status {rc}")

    def parse():
        rc = L.hp_dev_parse(3, C.c_void_p(text.data_ptr()), C.c_void_p(off.data_ptr()), n,
                            C.cast(oplanes, C.c_void_p), C.c_void_p(status.data_ptr()), C.byref(bad), None)
        if rc or bad.value:
            raise RuntimeError(f"hp_dev_parse: status {rc}, {bad.value} rejected")

    def best(fn):
        fn()
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    tf, tp = best(fmt), best(parse)
    ok = torch.equal(ohi, dhi) and torch.equal(olo, dlo)
    nbytes = int(need.value)
    out = {"workload": f"serde string form of f128, {n} dd values (N(0,1) x 10^U[-300,300), seed 2) in HBM",
           "text_bytes": nbytes, "round_trip_bit_exact": bool(ok),
           "format": {"ms": round(tf * 1e3, 3), "Mvalues_per_s": round(n / tf / 1e6, 1),
                      "GB_per_s": round((16 * n + nbytes + 8 * n) / tf / 1e9, 1)},
           "parse": {"ms": round(tp * 1e3, 3), "Mvalues_per_s": round(n / tp / 1e6, 1),
                     "GB_per_s": round((16 * n + nbytes + 8 * n) / tp / 1e9, 1)},
           "roofline": {"bound": "valu", "note": "per f64 one 64 x 128-bit multiply against a table of 10^s "
                                                 "(the exact 1024-bit path only near rounding boundaries), byte-wise "
                                                 "string work per lane: integer VALU bound; ~77 algorithmic bytes "
                                                 "per value (+128 through the format slots)"}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        m = TEXT_CPU_N
        v = np.stack([hi[:m], lo[:m]], axis=1)
        t0 = time.perf_counter()
        txt, offs = O.dd_serialize_n(v)
        t1 = time.perf_counter()
        back, nbad = O.dd_deserialize_n(txt, offs)
        t2 = time.perf_counter()
        assert nbad == 0 and np.array_equal(back, v)
        out["cpu_baseline"] = {"value": round(m / (t1 - t0) / 1e6, 3), "unit": "Mvalues/s (format)",
                               "parse_Mvalues_per_s": round(m / (t2 - t1) / 1e6, 3),
                               "cores": O.lib().hpo_num_threads(), "kind": "port",
                               "sample": f"the first {m} values, oracle C++ (glibc printf / strtod behind the "
                                         "reference's grammar and split rules), OpenMP over values"}
    return out


TEXT_N, TEXT_CPU_N = 1 << 22, 1 << 19


def molecule(seed, n_atoms, diffuse=True):
    """a jittered planar lattice of atoms (1.6 bohr), each with s, s, p, d, f shells of 1-3
    normalised primitives, exponents 10^U(-1.6 or -0.5, 1.8)"""
    import numpy as np
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(n_atoms)))
    shells = []
    for a in range(n_atoms):
        c = (1.6 * (a % side) + rng.uniform(-0.2, 0.2), 1.6 * (a // side) + rng.uniform(-0.2, 0.2),
             rng.uniform(-0.3, 0.3))
        for l in (0, 0, 1, 2, 3):
            npr = int(rng.integers(1, 4))
            alphas = np.sort(10 ** rng.uniform(-1.6 if diffuse else -0.5, 1.8, npr))[::-1]
            shells.append((c, l, [(float(x), float(rng.uniform(0.1, 1.0) * (2 * x / np.pi) ** 0.75))
                                  for x in alphas]))
    return shells


def bench_overlap_mixed(qp, torch, dev, cpu, reps=3):
    """assemble_overlap_mixed (crates/overlap/src/assemble.rs:35-85) over OVL_ATOMS atoms of
    s, s, p, d, f contracted shells (n_ao = 21 per atom), the default screening; S written
    to HBM (hp_dev_assemble_overlap_mixed).  Parity: tests/test_gpu_overlap_shells.py."""
    import ctypes as C
    import numpy as np
    L = qp.lib()

    def flat(shells):
        centers = np.ascontiguousarray([c for c, _, _ in shells], dtype=np.float64).ravel()
        ls = np.ascontiguousarray([l for _, l, _ in shells], dtype=np.int32)
        off = np.zeros(len(shells) + 1, dtype=np.int64)
        for i, (_, _, prims) in enumerate(shells):
            off[i + 1] = off[i] + len(prims)
        alpha = np.ascontiguousarray([a for _, _, pr in shells for a, _ in pr], dtype=np.float64)
        coeff = np.ascontiguousarray([c for _, _, pr in shells for _, c in pr], dtype=np.float64)
        return centers, ls, off, alpha, coeff, int(sum((l + 1) * (l + 2) // 2 for l in ls))

    shells = molecule(5, OVL_ATOMS)
    centers, ls, off, alpha, coeff, n_ao = flat(shells)
    prm = np.ascontiguousarray(qp.SCREEN_DEFAULT, dtype=np.float64)
    s = torch.empty(n_ao * n_ao, dtype=torch.float64, device=dev)
    counts = (C.c_size_t * 3)()
    ts = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc = L.hp_dev_assemble_overlap_mixed(centers, ls, off, alpha, coeff, len(shells), prm, s.data_ptr(), n_ao,
                                             counts, None)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
        if rc:
            raise RuntimeError(f"hp_dev_assemble_overlap_mixed: status {rc}")
    t = min(ts[1:])
    elems = n_ao * (n_ao + 1) / 2  # the computed half; the mirror is a second store
    out = {"workload": f"assemble_overlap_mixed, {OVL_ATOMS} atoms x (s, s, p, d, f) shells of 1-3 primitives "
                       f"(seed 5), n_ao {n_ao}, default screening",
           "ms": round(t * 1e3, 3), "shells": len(shells), "n_ao": n_ao,
           "tiers": {"f64_pairs": counts[0], "f128_pairs": counts[1], "skipped_pairs": counts[2]},
           "Melements_per_s": round(elems / t / 1e6, 1),
           "roofline": {"bound": "valu", "hbm_write_GB_per_s": round(8.0 * n_ao * n_ao / t / 1e9, 1),
                        "note": "FP64 VALU: the contracted Obara-Saika element (1-9 primitive pairs, f64 or dd) "
                                "per thread; the S write is 8 bytes per element"}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        small = molecule(5, OVL_CPU_ATOMS)
        n_small = sum((l + 1) * (l + 2) // 2 for _, l, _ in small)
        t0 = time.perf_counter()
        O.assemble_overlap_mixed(small, O.DEFAULT_SCREEN)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n_small * (n_small + 1) / 2 / tc / 1e6, 3), "unit": "Melements/s",
                               "cores": O.lib().hpo_num_threads(), "kind": "port",
                               "sample": f"{OVL_CPU_ATOMS} atoms (n_ao {n_small}) of the same kind, oracle C++, "
                                         f"OpenMP over shell pairs as the reference's par_iter ({tc:.2f} s)"}
    return out


def bench_reduce(qp, torch, dev, cpu, reps=5):
    """dot (linalg.rs:7-18) of RED_N dd values and gemv (linalg.rs:29-40) of a RED_ROWS^2 dd
    matrix, both resident in HBM; the reference's summation order (bit-identical to the
    oracle, tests/test_gpu_reduce.py)."""
    import ctypes as C
    import numpy as np
    L = qp.lib()
    n, r = RED_N, RED_ROWS
    a = qp.dev.to_planes(qp.gen_dd_normal(7, n), 2, dev)
    b = qp.dev.to_planes(qp.gen_dd_normal(8, n), 2, dev)
    o = torch.zeros(2, dtype=torch.float64, device=dev)
    A = qp.dev.to_planes(qp.gen_dd_normal(9, r * r), 2, dev)
    x = qp.dev.to_planes(qp.gen_dd_normal(10, r), 2, dev)
    y = torch.zeros((2, r), dtype=torch.float64, device=dev)

    def best(fn):
        fn()
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    def dot():
        if L.hp_dev_dot(a[0].data_ptr(), a[1].data_ptr(), b[0].data_ptr(), b[1].data_ptr(), n, o.data_ptr(),
                        o.data_ptr() + 8, None):
            raise RuntimeError("hp_dev_dot")

    def gemv():
        if L.hp_dev_gemv(A[0].data_ptr(), A[1].data_ptr(), r, r, x[0].data_ptr(), x[1].data_ptr(), y[0].data_ptr(),
                         y[1].data_ptr(), None):
            raise RuntimeError("hp_dev_gemv")

    td, tg = best(dot), best(gemv)
    out = {"workload": f"dd dot of {n} values, dd gemv {r} x {r} (gen_dd_normal seeds 7-10), in HBM",
           "dot": {"ms": round(td * 1e3, 3), "GB_per_s": round(32.0 * n / td / 1e9, 1),
                   "note": "1024-element chunks, one thread each (the reference's order), then the 16384 "
                           "chunk sums combined by rayon's halving tree, pairwise with DPP / ds_swizzle in each "
                           "wavefront: bound by the chunks' serial 1024-step dd chains, not HBM"},
           "gemv": {"ms": round(tg * 1e3, 3), "GB_per_s": round(16.0 * r * r / tg / 1e9, 1),
                    "roofline_frac_hbm": round(16.0 * r * r / tg / 8e12, 4),
                    "note": "one thread per row (the reference's order); 64 x 32 tiles through double-buffered "
                            "LDS, staged by three loader wavefronts: bound by the 8192-step dependent dd chain "
                            "of each row (~200 cycles a step), not HBM"}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        ah = np.stack([a[0].cpu().numpy(), a[1].cpu().numpy()], axis=1)
        bh = np.stack([b[0].cpu().numpy(), b[1].cpu().numpy()], axis=1)
        t0 = time.perf_counter()
        ref = O.dot(ah, bh)
        t1 = time.perf_counter()
        assert np.array_equal(ref, o.cpu().numpy())
        out["cpu_baseline"] = {"value": round(32.0 * n / (t1 - t0) / 1e9, 2), "unit": "GB/s (dot)",
                               "cores": O.lib().hpo_num_threads(), "kind": "port",
                               "sample": f"the same {n}-value dot, oracle C++ (chunk sums over OpenMP threads as "
                                         f"rayon's par_chunks, then the tree) ({t1 - t0:.2f} s), result bit-identical "
                                         "to the GPU's"}
    return out


def bench_mf4_chol(qp, torch, dev, cpu, reps=3):
    """mf_cholesky::<4> and mf_cholesky_solve::<4> (crates/quad/src/multifloat/linalg.rs:89-172)
    at n = MFC_N on B B^T + n I (B seeded MF4 N(0,1)), host arrays in and out (the drop-in
    entry points); bit-identical to the oracle (tests/test_gpu_mf4_linalg.py)."""
    import numpy as np
    n = MFC_N
    rng = np.random.default_rng(6)
    b = rng.standard_normal((n, n))
    a = b @ b.T + n * np.eye(n)  # exact f64 entries, promoted to MF4
    a4 = np.zeros((n * n, 4))
    a4[:, 0] = a.ravel()
    rhs = np.zeros((n, 4))
    rhs[:, 0] = rng.standard_normal(n)

    def best(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    L = qp.mf_cholesky(a4, n)
    tc = best(lambda: qp.mf_cholesky(a4, n))
    ts = best(lambda: qp.mf_cholesky_solve(L, rhs, n))
    macs = n ** 3 / 6.0
    out = {"workload": f"mf_cholesky + mf_cholesky_solve, MF4, n = {n}, A = B B^T + n I (seed 6)",
           "cholesky_ms": round(tc * 1e3, 2), "solve_ms": round(ts * 1e3, 2),
           "cholesky_MF4_GMAC_per_s": round(macs / tc / 1e9, 3),
           "note": "column-serial right-looking factor (one launch for the column's MF sqrt / recip chain, "
                   "one for the trailing one-MAC update); the reference's per-element order; host arrays "
                   "in and out"}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        t0 = time.perf_counter()
        ref = O.mf4_cholesky(a4, n)
        t1 = time.perf_counter()
        O.mf4_cholesky_solve(ref, rhs, n)
        t2 = time.perf_counter()
        assert np.array_equal(ref, L)
        out["cpu_baseline"] = {"value": round((t1 - t0) * 1e3, 1), "unit": "ms (cholesky)",
                               "solve_ms": round((t2 - t1) * 1e3, 1), "cores": 1, "kind": "port",
                               "sample": f"the same n = {n} factor and solve, oracle C++, one core (the reference's "
                                         "mf_cholesky and its solves are serial loops); factor bit-identical to the "
                                         "GPU's"}
    return out


MFC_N = 256
OVL_ATOMS, OVL_CPU_ATOMS = 400, 400
RED_N, RED_ROWS = 1 << 24, 8192
EIG_N, EIG_CPU_N = 512, 192
ADJ_N, ADJ_SEED, ADJ_B_SEED = 2048, 4, 44


def bench_adjoint(qp, torch, dev, cpu):
    """Config 4: gradients of log det S and b^T S^-1 b over 3N centre coordinates and N
    exponents, N = 2048 Gaussians (lattice, seed 4, exponents in [0.01, 10]), dd.  Forward =
    overlap build + factor + refined solve + log det; backward = L^-1, S^-1, contraction.
    Accuracy against f256 central differences: tests/test_gpu_adjoint.py."""
    n = ADJ_N
    c, a = qp.gen_overlap_basis(ADJ_SEED, n, 0.01, 10.0)
    b = qp.gen_dd_normal(ADJ_B_SEED, n)
    dc = torch.from_numpy(c.ravel()).to(dev)
    da = torch.from_numpy(a).to(dev)
    db = qp.dev.to_planes(b, 2, dev)
    out = {"workload": "d/d(R, alpha) of log det S and b^T S^-1 b, N=2048 Gaussians, seed 4, dd",
           "parameters": 4 * n}
    qp.set_phase_timing(True)
    for label, fk, gm in (("dd_factor_plain_gemm", qp.HP_FACTOR_DD, qp.HP_GEMM_PLAIN),
                          ("split_factor_split_gemm", qp.HP_FACTOR_DD_SPLIT, qp.HP_GEMM_SPLIT_I8)):
        best = None
        for _ in range(3):  # the first pass also warms allocations and the split autotuner
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = qp.dev.OverlapAdjoint(dc, da, db, fk)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            st.gradients(gm)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            inf = st.refresh_info()
            rec = ((t2 - t0) * 1e3, (t1 - t0) * 1e3, (t2 - t1) * 1e3, inf.inverse_ms, inf.gram_ms, inf.contract_ms)
            st.close()
            if best is None or rec[0] < best[0]:
                best = rec
        bw_macs = n ** 3 / 3.0  # algorithmic: L^-1 (n^3/6 dd MACs) + lower triangle of L^-T L^-1 (n^3/6)
        out[label] = {"ms": round(best[0], 2), "forward_ms": round(best[1], 2), "backward_ms": round(best[2], 2),
                      "inverse_ms": round(best[3], 3), "gram_ms": round(best[4], 3),
                      "contract_ms": round(best[5], 3),
                      "backward_dd_gflops": round(2.0 * bw_macs / (best[2] * 1e-3) / 1e9, 1)}
        if gm == qp.HP_GEMM_PLAIN:
            ach = 2.0 * DD_MAC_OPS * (n ** 3 / 6.0 + bw_macs) / (best[0] * 1e-3) / 1e12
            out[label]["roofline"] = {"bound": "valu", "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS,
                                      "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4),
                                      "note": "factor n^3/6 + L^-1 n^3/6 + S^-1 n^3/6 dd MACs x 29 FP64 ops "
                                              "(FMA-equivalent) over the whole forward + backward"}
        else:
            wb = split_factor_work_ms(n) + (DD_MAC_OPS * n ** 3 / 6.0 / (FP64_PEAK_TFLOPS / 2 * 1e12)
                                            + 2 * SLICES * n ** 3 / 6.0 / (INT8_PEAK_TOPS * 1e12)) * 1e3
            out[label]["roofline"] = {"bound": "latency", "work_bound_ms": round(wb, 3),
                                      "frac": round(wb / best[0], 4),
                                      "note": "frac = the work at peak rates (split factor as in spd_solve, L^-1 "
                                              "on FP64 VALU, S^-1 on the INT8 matrix cores) over the total"}
    qp.set_phase_timing(False)
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        t0 = time.perf_counter()
        O.overlap_adjoint_dd(c, a, b)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(t * 1e3, 1), "unit": "ms", "cores": O.lib().hpo_num_threads(),
                               "kind": "port", "sample": "the full N=2048 gradient pair: dd S build, "
                               "cholesky_blocked, S^-1 by solve_cholesky per column, contraction (oracle C++, "
                               "OpenMP over columns / rows)"}
    return out


def bench_config0(qp, torch, cpu, reps=5):
    """Config 0, the reference's CPU-runnable case, through the drop-in host entry points (host
    arrays in and out, PCIe included): N=512 lattice Gaussians (seed 0, exponents in [0.01, 10])
    -> dd overlap S (hp_overlap_s) -> cholesky_blocked(nb=64) -> solve_cholesky with b = ones;
    the solution must be the oracle's bit for bit."""
    import numpy as np
    n = 512
    c, a = qp.gen_overlap_basis(0, n, 0.01, 10.0, 1.4)
    ones = np.zeros((n, 2))
    ones[:, 0] = 1.0
    st = {}

    def run():
        t0 = time.perf_counter()
        s_ = qp.overlap_s(c, a)
        t1 = time.perf_counter()
        l_ = qp.cholesky_blocked(s_, n, 64)
        t2 = time.perf_counter()
        x_ = qp.solve_cholesky(l_, ones, n)
        t3 = time.perf_counter()
        return (t1 - t0, t2 - t1, t3 - t2), x_, l_

    run()
    best, x, l = None, None, None
    for _ in range(reps):
        tt, x, l = run()
        best = tt if best is None or sum(tt) < sum(best) else best
    kappa = qp.cond_estimate(l, n)
    ach = 2.0 * DD_MAC_OPS * n ** 3 / 6.0 / best[1] / 1e12
    out = {"workload": "config 0: N=512 s Gaussians (8^3 lattice, 1.4 bohr, seed 0, alpha in [0.01, 10]) -> dd S "
                       "-> cholesky_blocked(nb=64) -> solve_cholesky(L, ones), host entry points",
           "ms": round(sum(best) * 1e3, 3), "overlap_ms": round(best[0] * 1e3, 3),
           "cholesky_blocked_ms": round(best[1] * 1e3, 3), "solve_cholesky_ms": round(best[2] * 1e3, 3),
           "kappa_est": float(f"{kappa:.3e}"),
           "roofline": {"bound": "latency", "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(ach / FP64_PEAK_TFLOPS, 4),
                        "note": "the factor's n^3/6 dd MACs x 29 FP64 ops over cholesky_blocked_ms (PCIe staging "
                                "of 4 MB included); at N=512 the 8 diagonal blocks' pivot chains dominate"}}
    if cpu:
        O = oracle()
        t0 = time.perf_counter()
        s_o = O.overlap_s_dd(c, a)
        t1 = time.perf_counter()
        l_o = O.cholesky_blocked(s_o, n, 64)
        t2 = time.perf_counter()
        x_o = O.solve_cholesky(l_o, ones, n)
        t3 = time.perf_counter()
        out["bit_identical_to_oracle"] = bool(np.array_equal(x.view(np.uint64), x_o.view(np.uint64)))
        out["cpu_baseline"] = {"value": round((t3 - t0) * 1e3, 2), "unit": "ms",
                               "overlap_ms": round((t1 - t0) * 1e3, 2), "cholesky_blocked_ms": round((t2 - t1) * 1e3, 2),
                               "solve_cholesky_ms": round((t3 - t2) * 1e3, 2), "cores": O.lib().hpo_num_threads(),
                               "kind": "port", "sample": "the whole config, oracle C++ (OpenMP rows)"}
    return out


OVL_S_NS = (512, 2048, 4096)


def bench_overlap_s(qp, torch, dev, cpu, reps=5):
    """The dd overlap build (k_overlap_s): S_uv = (4 a_u a_v / (a_u + a_v)^2)^{3/4}
    exp(-a_u a_v / (a_u + a_v) |R_u - R_v|^2) for lattice Gaussians (seed 0, exponents in
    [0.01, 10]) on centres / exponents resident in HBM; bit-identical to the oracle
    (tests/test_gpu_linalg.py, tests/test_gpu_config_sizes.py)."""
    import numpy as np
    O = oracle()
    out = {"workload": "dd Gaussian overlap S, N in (512, 2048, 4096), lattice 1.4 bohr, seed 0, alpha in [0.01, 10]"}
    for n in OVL_S_NS:
        c, a = qp.gen_overlap_basis(0, n, 0.01, 10.0, 1.4)
        dc = torch.from_numpy(c.ravel()).to(dev)
        da = torch.from_numpy(a).to(dev)
        S = torch.empty((2, n * n), dtype=torch.float64, device=dev)
        ms = best_of(lambda: qp.dev.overlap_s(dc, da, S, n), reps, torch)
        out[f"n{n}"] = {"ms": round(ms, 4), "Mentries_per_s": round(n * (n + 1) / 2 / (ms * 1e-3) / 1e6, 1)}
    # FP64 operations per entry from the oracle's op-counting restatement, mean over a sample
    n = OVL_S_NS[-1]
    c, a = qp.gen_overlap_basis(0, n, 0.01, 10.0, 1.4)
    rng = np.random.default_rng(5)
    ii, jj = rng.integers(0, n, 4000), rng.integers(0, n, 4000)
    w = float(np.mean([O.count_overlap_entry(a[i], c[i], a[j], c[j]) for i, j in zip(ii, jj) if i != j]))
    ms = out[f"n{n}"]["ms"]
    ach = 2.0 * w * n * (n - 1) / 2 / (ms * 1e-3) / 1e12
    out["ops_per_entry"] = round(w, 1)
    out["roofline"] = {"bound": "valu", "kernel": "k_overlap_s", "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4),
                       "note": f"N=4096: N(N-1)/2 off-diagonal entries x {w:.0f} FP64 ops (the op-counting "
                               "restatement's mean; a dd division counts its IEEE divisions as one op each), "
                               "FMA-equivalent; HBM writes 16 B per stored entry"}
    if cpu:
        t0 = time.perf_counter()
        O.overlap_s_dd(c, a)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n * (n + 1) / 2 / t / 1e6, 2), "unit": "Mentries/s",
                               "ms": round(t * 1e3, 1), "cores": O.lib().hpo_num_threads(), "kind": "port",
                               "sample": "the full N=4096 build, oracle C++ (OpenMP rows)"}
    return out


def bench_gemm_atb(qp, torch, dev, steps, cpu):
    """Config 1: gemm_atb C = A^T B at m=n=k=4096 (A stored k x m; hi N(0,1), seed 1),
    plain (the reference's order, bit-identical) and split (HP_GEMM_AUTO at 104 bits)."""
    m = n = k = 4096
    a = qp.gen_dd_normal(1, k * m)
    b = qp.gen_dd_normal(1 + 0x5EED, k * n)
    da, db = qp.dev.to_planes(a, 2, dev), qp.dev.to_planes(b, 2, dev)
    dc = torch.empty((2, m * n), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    macs = float(m) * n * k
    ms_p, kern_p = time_steps(lambda: qp.dev.gemm(da, db, dc, m, n, k, trans_a=True, mode=qp.HP_GEMM_PLAIN), 2, 1,
                              stream, torch)
    ms_s, kern_s = time_steps(lambda: qp.dev.gemm(da, db, dc, m, n, k, trans_a=True, mode=qp.HP_GEMM_AUTO,
                                                  slices=DD_BITS), max(3, min(steps, 10)), 2, stream, torch)
    plan = qp.gemm_plan(m, n, k, DD_BITS, trans_a=True)
    ach_p = 2.0 * DD_MAC_OPS * macs / (kern_p * 1e-3) / 1e12
    out = {"workload": "dd gemm_atb C=A^T B, m=n=k=4096, A stored k x m, seed 1",
           "plain": {"ms": round(ms_p, 3), "GFLOP_s": round(2.0 * macs / (ms_p * 1e-3) / 1e9, 1),
                     "roofline": {"bound": "valu", "kernel": "k_gemm_dd_plain<true>", "achieved": round(ach_p, 3),
                                  "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                  "frac": round(ach_p / FP64_PEAK_TFLOPS, 4)}},
           "split": {"ms": round(ms_s, 3), "GFLOP_s": round(2.0 * macs / (ms_s * 1e-3) / 1e9, 1),
                     "bits": DD_BITS, "slices": plan.slices if plan else None, "w_bits": plan.w_bits if plan else None,
                     "variant": plan.variant if plan else None,
                     "note": "whole step (scales, residues, int8 GEMMs, CRT); the int8 kernel's roofline is the "
                             "headline's"}}
    out["split_over_plain"] = round(ms_p / ms_s, 2)
    if cpu:
        O = oracle()
        threads = O.lib().hpo_num_threads()
        rows = threads * 8
        t0 = time.perf_counter()
        O.gemm_atb_rows(a, b, m, n, k, 0, rows)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(2.0 * rows * n * k / t / 1e9, 4), "unit": "GFLOP/s", "cores": threads,
                               "kind": "port", "sample": f"{rows} of {m} rows ({t:.1f} s), oracle C++ restatement of "
                                                         "gemm_atb (both operands strided), OpenMP rows"}
    return out


def bench_autotune(qp, torch, dev, da, db, dc, steps):
    """Config 1, the autotuner: HP_GEMM_AUTO at 4096^3 for several requested accuracies --
    the moduli (int8 GEMMs) it takes for each and the time of a whole step."""
    stream = torch.cuda.current_stream()
    out = {}
    for bits in (53, 80, 96, 104):
        ms, _ = time_steps(lambda: qp.dev.gemm(da, db, dc, M, N, K, mode=qp.HP_GEMM_AUTO, slices=bits),
                           max(3, min(steps, 10)), 2, stream, torch)
        plan = qp.gemm_plan(M, N, K, bits)
        out[f"bits{bits}"] = {"mode": "split" if plan.mode == qp.HP_GEMM_SPLIT_I8 else "plain",
                              "slices": plan.slices, "w_bits": plan.w_bits, "variant": plan.variant,
                              "ms": round(ms, 3), "GFLOP_s": round(2.0 * M * N * K / (ms * 1e-3) / 1e9, 1),
                              "bound_rel_to_norm": f"2^-{bits}"}
    return out


def time_steps(step, steps, warmup, stream, torch):
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        step()
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall * 1e3 / steps, sum(s.elapsed_time(e) for s, e in ev) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-solve", action="store_true", help="skip the N=4096 SPD solve timings")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven solves")
    ap.add_argument("--no-mf4", action="store_true", help="skip the MF4 1024^3 gemm")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the N=2048 adjoint")
    ap.add_argument("--no-eigen", action="store_true", help="skip the N=512 generalized eigensolve")
    ap.add_argument("--no-text", action="store_true", help="skip the dd serde text round trip")
    ap.add_argument("--no-overlap", action="store_true", help="skip the mixed-tier overlap assembly")
    ap.add_argument("--no-reduce", action="store_true", help="skip dot / gemv")
    ap.add_argument("--no-mfchol", action="store_true", help="skip the MF4 Cholesky")
    ap.add_argument("--no-mfn", action="store_true", help="skip the f512 (MultiFloat<8>) gemm")
    ap.add_argument("--no-config0", action="store_true", help="skip config 0 (N=512 overlap, factor, solve)")
    ap.add_argument("--no-overlap-s", action="store_true", help="skip the dd overlap build")
    ap.add_argument("--no-atb", action="store_true", help="skip gemm_atb at 4096^3")
    ap.add_argument("--no-autotune", action="store_true", help="skip the autotuner's accuracy sweep")
    ap.add_argument("--i8-variant", type=int, default=0, metavar="V",
                    help="pin the headline's int8 GEMM kernel (variant code, e.g. 4064 or 7064; 0: the tuner's pick)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device three times (a plain copy kernel of known size, "
                         "to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU layout is described in DESIGN.md and is not built.", file=sys.stderr)
        sys.exit(2)

    import torch
    import quadprec_amd as qp

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        count = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(count, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read_per_copy": count * 8, "bytes_written_per_copy": count * 8,
                          "copies": 3}), flush=True)
        return
    a = qp.gen_dd_normal(1, M * K)
    b = qp.gen_dd_normal(1 + 0x5EED, K * N)
    da, db = qp.dev.to_planes(a, 2, dev), qp.dev.to_planes(b, 2, dev)
    dc = torch.empty((2, M * N), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    macs = M * N * K

    # split path (headline); per-phase HIP events on the launch stream, accumulated over the
    # timed steps without a host wait per call
    qp.set_phase_timing(2)

    if args.i8_variant:  # pin the int8 kernel instead of the tuner's choice
        qp.split_gemm_set_variant(M, N, K, args.i8_variant, slices=SLICES)

    def split_step():
        qp.dev.gemm(da, db, dc, M, N, K, mode=qp.HP_GEMM_AUTO, slices=DD_BITS)

    for _ in range(args.warmup):
        split_step()
    torch.cuda.synchronize()
    qp.collect_phase_times()  # the warmup calls' phases are dropped
    ms_split, kern_split = time_steps(split_step, args.steps, 0, stream, torch)
    sums, calls = qp.collect_phase_times()
    qp.set_phase_timing(False)
    ph = [x / calls for x in sums]  # prep, int8 gemms, crt
    gemm_ms = ph[1]
    plan = qp.gemm_plan(M, N, K, DD_BITS)
    assert plan is not None and plan.mode == qp.HP_GEMM_SPLIT_I8 and plan.slices == SLICES, "autotuner plan"
    variant = plan.variant
    achieved = 2.0 * SLICES * macs / (gemm_ms * 1e-3) / 1e12

    # plain path (reference op order) on the same operands
    def plain_step():
        qp.dev.gemm(da, db, dc, M, N, K, mode=qp.HP_GEMM_PLAIN)

    ms_plain, kern_plain = time_steps(plain_step, max(2, min(args.steps, 3)), 1, stream, torch)
    plain_ach = 2.0 * DD_MAC_OPS * macs / (kern_plain * 1e-3) / 1e12

    value = 2.0 * macs / (ms_split * 1e-3) / 1e9
    out = {
        "metric": "dd-GEMM GFLOP/s @4096^3 (split precision)",
        "value": round(value, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_split, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 residues (30-modulus CRT) -> double-double",
        "data": "synthetic: seeded N(0,1) dd operands, generated on the host and resident in HBM",
        "config": {"workload": "dd gemm C=A*B, m=n=k=4096, row-major, seed 1; split at dd accuracy (104 bits, "
                               "autotuned: 30 moduli = 30 int8 GEMMs + CRT)",
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "mfma", "kernel": I8_KERNELS.get(variant, str(variant)), "achieved": round(achieved, 2),
                     "peak": INT8_PEAK_TOPS, "unit": "TOP/s (int8)", "frac": round(achieved / INT8_PEAK_TOPS, 4),
                     "traffic": I8_TRAFFIC_BYTES.get(variant), "traffic_algorithmic": 3 * SLICES * M * N,
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000) per launch, "
                                       "calibrated on a 256 MiB device copy; profiles/k_gemm_i8_mod_dlp8_4096.md",
                     "kernel_ms": round(gemm_ms, 4),
                     "variant": variant, "slices": plan.slices, "w_bits": plan.w_bits,
                     "phases_ms": {"scale_and_residues": round(ph[0], 4), "int8_gemms": round(ph[1], 4),
                                   "crt": round(ph[2], 4)}},
        "plain": {"value": round(2.0 * macs / (ms_plain * 1e-3) / 1e9, 3), "unit": "GFLOP/s",
                  "ms_per_step": round(ms_plain, 4), "kernel": "k_gemm_dd_plain",
                  "roofline": {"bound": "valu", "achieved": round(plain_ach, 3), "peak": FP64_PEAK_TFLOPS,
                               "unit": "TFLOP/s", "frac": round(plain_ach / FP64_PEAK_TFLOPS, 4),
                               "note": "29 FP64 VALU ops per dd MAC, FMA-equivalent (2 flops each)"}},
        "split_over_plain": round(ms_plain / ms_split, 2),
    }
    if not args.no_autotune:
        out["autotune_4096"] = bench_autotune(qp, torch, dev, da, db, dc, args.steps)
    if not args.no_atb:
        out["gemm_atb_4096"] = bench_gemm_atb(qp, torch, dev, args.steps, not args.no_cpu)
    if not args.no_config0:
        out["config0_n512"] = bench_config0(qp, torch, not args.no_cpu)
    if not args.no_overlap_s:
        out["overlap_s_dd"] = bench_overlap_s(qp, torch, dev, not args.no_cpu)
    if not args.no_solve:
        out["spd_solve_n4096"] = bench_solves(qp, torch, dev, not args.no_cpu)
    if not args.no_batch:
        out["batched_solves"] = bench_batch(qp, torch, dev, not args.no_cpu)
    if not args.no_mf4:
        out["mf4_gemm_1024"] = bench_mf4(qp, torch, dev, args.steps, not args.no_cpu)
    if not args.no_mfn:
        out["f512_gemm_512"] = bench_mfn(qp, torch, dev, args.steps, not args.no_cpu)
    if not args.no_adjoint:
        out["adjoint_n2048"] = bench_adjoint(qp, torch, dev, not args.no_cpu)
    if not args.no_eigen:
        out["gen_eigen_n512"] = bench_eigen(qp, torch, dev, not args.no_cpu)
    if not args.no_text:
        out["dd_serde_text_4m"] = bench_text(qp, torch, dev, not args.no_cpu)
    if not args.no_overlap:
        out["overlap_mixed_400_atoms"] = bench_overlap_mixed(qp, torch, dev, not args.no_cpu)
    if not args.no_reduce:
        out["dot_gemv"] = bench_reduce(qp, torch, dev, not args.no_cpu)
    if not args.no_mfchol:
        out["mf4_cholesky_256"] = bench_mf4_chol(qp, torch, dev, not args.no_cpu)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a, b)
        out["host_cpu"] = host_cpu()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
