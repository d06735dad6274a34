#!/usr/bin/env python3
"""Benchmark of the batched hot path on one MI355X.

Workload (one "step"): 10,000 uniform random 3-SAT instances (n = 200, m = 852) with
64 partial assignments each (every var assigned with p = 0.3, uniform polarity),
packed as an int32 CSR literal arena plus literal-bitmap rows, all resident in HBM.
A step is one launch of the fused kernel that writes exact #SAT/#UNSAT/#UNIT clause
counts for every (instance, assignment) and the exact BCP fixpoint (closure row,
conflict flag, implied-literal count) of every assignment.

Prints ONE JSON line (rank 0).  value = instances per second of that step.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "verified instances/sec + BCP props/sec (HBM GB/s); learned split vs VSIDS"


def algorithmic_bytes_per_instance(n: int, m: int, L: int, A: int, eval_: bool, bcp: bool) -> int:
    """Compulsory HBM bytes of one instance: arena (4L literals + 4(m+1) offsets + 32 B
    metadata), the input rows (A * 4W), and the outputs."""
    W = (2 * n + 31) // 32
    b = 4 * L + 4 * (m + 1) + 32 + A * 4 * W
    if eval_:
        b += A * 12
    if bcp:
        b += A * 4 * W + A * 8
    return b


def cpu_baseline_eval_bcp(batch, rows, A, seconds: float):
    import numpy as np
    import oracle
    base = batch.row_base(A)
    done = 0
    t0 = time.perf_counter()
    i = 0
    while time.perf_counter() - t0 < seconds and i < batch.n_inst:
        idx = list(range(i, min(i + 4, batch.n_inst)))
        sb = batch.subset(idx)
        srows = np.concatenate([rows[base[j]:base[j + 1]] for j in idx])
        oracle.eval_batch(sb, srows, A)
        oracle.bcp_batch(sb, srows, A)
        done += len(idx)
        i += len(idx)
    dt = time.perf_counter() - t0
    return done / dt, done, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--instances", type=int, default=10_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the instance batch would shard across GPUs with no "
              "exchange (see DESIGN.md); multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import gpupsat_amd as G
    from gpupsat_amd import workloads as W

    N, n, m, A, p = args.instances, 200, 852, 64, 0.3
    b = W.random_3sat_batch(N, n, m, seed=0)
    rows = W.random_assignments(b, A, p, seed=0)
    L = int(b.clause_off[-1]) // N
    gb = G.GpuBatch(b)
    d_rows = G.DevBuf.from_host(rows)
    d_counts = G.DevBuf(N * A * 12)
    d_out = G.DevBuf(rows.nbytes)
    d_conf = G.DevBuf(N * A * 4)
    d_props = G.DevBuf(N * A * 4)

    def step():
        gb.eval_bcp_dev(d_rows, A, d_counts, d_out, d_conf, d_props)

    for _ in range(args.warmup):
        step()
    gb.sync()
    G.synchronize()
    e0, e1 = G.Event(), G.Event()
    t0 = time.perf_counter()
    e0.record(gb)
    for _ in range(args.steps):
        step()
    e1.record(gb)
    gb.sync()
    G.synchronize()
    wall = time.perf_counter() - t0
    ms_fused = e0.elapsed_ms(e1) / args.steps
    ms_per_step = max(ms_fused, 1e3 * wall / args.steps)

    # per-kernel breakdown on the same stream
    def timed(fn, k):
        a, z = G.Event(), G.Event()
        a.record(gb)
        for _ in range(k):
            fn()
        z.record(gb)
        gb.sync()
        return a.elapsed_ms(z) / k

    ms_eval = timed(lambda: gb.clause_eval_dev(d_rows, A, d_counts), args.steps)
    ms_bcp = timed(lambda: gb.bcp_fixpoint_dev(d_rows, A, d_out, d_conf, d_props), args.steps)

    conflict = d_conf.to_host(np.int32, N * A).reshape(N, A)
    props = d_props.to_host(np.int32, N * A).reshape(N, A)

    # low-density trails, where BCP actually implies literals (props/s)
    rows_lo = W.random_assignments(b, A, 0.05, seed=1)
    d_rows_lo = G.DevBuf.from_host(rows_lo)
    gb.bcp_fixpoint_dev(d_rows_lo, A, d_out, d_conf, d_props)
    ms_bcp_lo = timed(lambda: gb.bcp_fixpoint_dev(d_rows_lo, A, d_out, d_conf, d_props), args.steps)
    props_lo = d_props.to_host(np.int32, N * A)
    conf_lo = d_conf.to_host(np.int32, N * A)

    # practical HBM peak: streaming copy of 2 GiB
    cb = 1 << 31
    src, dst = G.DevBuf(cb), G.DevBuf(cb)
    cp = lambda: G.lib().gpsat_copy_dev(gb.h, dst.ptr, src.ptr, G.ctypes.c_int64(cb))
    cp()
    ms_copy = timed(cp, 5)
    copy_gbs = 2 * cb / (ms_copy * 1e-3) / 1e9
    src.close(); dst.close()

    # spot parity check against the oracle on 16 instances
    parity = None
    try:
        import oracle
        idx = list(range(0, N, max(1, N // 16)))[:16]
        gb.eval_bcp_dev(d_rows, A, d_counts, d_out, d_conf, d_props)
        gb.sync()
        counts = d_counts.to_host(np.int32, N * A * 3).reshape(N, A, 3)
        outr = d_out.to_host(np.uint32, rows.size)
        conf = d_conf.to_host(np.int32, N * A).reshape(N, A)
        base = b.row_base(A)
        sb = b.subset(idx)
        srows = np.concatenate([rows[base[j]:base[j + 1]] for j in idx])
        e_out, e_conf, _ = oracle.bcp_batch(sb, srows, A)
        parity = bool((counts[idx] == oracle.eval_batch(sb, srows, A)).all()
                      and (conf[idx] == e_conf).all()
                      and (np.concatenate([outr[base[j]:base[j + 1]] for j in idx]) == e_out).all())
    except Exception as ex:  # the oracle is a checker; its absence must not hide the number
        parity = f"unchecked: {ex}"

    bytes_fused = N * algorithmic_bytes_per_instance(n, m, L, A, True, True)
    bytes_eval = N * algorithmic_bytes_per_instance(n, m, L, A, True, False)
    bytes_bcp = N * algorithmic_bytes_per_instance(n, m, L, A, False, True)
    achieved = bytes_fused / (ms_fused * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "eval_bcp_config2_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu:
        rate, done, dt = cpu_baseline_eval_bcp(b, rows, A, args.cpu_seconds)
        cpu = {"value": rate, "unit": "instances/s", "cores": 1, "kind": "port",
               "sample": f"oracle clause-eval + UP-closure of the first {done} instances x 64 rows "
                         f"(same workload), {dt:.1f} s on one host core"}

    out = {
        "metric": METRIC,
        "value": N / (ms_per_step * 1e-3),
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "batched clause-eval + BCP fixpoints: 10,000 random 3-SAT (n=200, m=852), "
                               "64 partial assignments each (p=0.3)",
                   "instances": N, "n": n, "m": m, "assignments": A, "p": p},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "kernels": {
            "eval_bcp_fused_ms": ms_fused, "eval_ms": ms_eval, "bcp_ms": ms_bcp,
            "eval_gbs": bytes_eval / (ms_eval * 1e-3) / 1e9, "bcp_gbs": bytes_bcp / (ms_bcp * 1e-3) / 1e9,
            "algorithmic_bytes_per_launch": {"fused": bytes_fused, "eval": bytes_eval, "bcp": bytes_bcp},
            "copy_peak_gbs": copy_gbs,
        },
        "bcp": {
            "p0.3_conflict_frac": float(conflict.mean()),
            "p0.05_ms": ms_bcp_lo,
            "p0.05_conflict_frac": float(conf_lo.mean()),
            "p0.05_props_per_s": float(props_lo.sum()) / (ms_bcp_lo * 1e-3),
        },
        "parity_sample_ok": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
