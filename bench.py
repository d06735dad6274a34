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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "verified instances/sec + BCP props/sec (HBM GB/s); learned split vs VSIDS"


def survey_bytes_per_instance(n: int, m: int, L: int, A: int, eval_: bool, bcp: bool) -> int:
    """Compulsory HBM bytes of one instance in the survey's accounting (the roofline's
    denominator): the CSR arena 4(L + m + 1), the A input assignments at 2 bits per var
    A*ceil(n/4), #SAT/#UNSAT/#UNIT out 12A, and the closure rows plus a conflict/props word
    out A*(ceil(n/4) + 4).  n=200, m=852, L=2556, A=64: eval 17,604 B, eval+BCP 21,060 B."""
    q = (n + 3) // 4
    b = 4 * (L + m + 1) + A * q
    if eval_:
        b += 12 * A
    if bcp:
        b += A * (q + 4)
    return b


def packed_bytes_per_instance(n: int, m: int, A: int, eval_: bool, bcp: bool) -> int:
    """What the 3-literal kernel itself must move: one 4-byte packed word per clause, 32 B of
    metadata, the rows as stored (A * 4W in, A * 4W out) and the counts/flags."""
    W = (2 * n + 31) // 32
    b = 4 * m + 32 + A * 4 * W
    if eval_:
        b += 12 * A
    if bcp:
        b += A * (4 * W + 8)
    return b


def _eval_bcp_chunk(args):
    """Worker: oracle clause counts + UP closure of a slice of instances (64 rows each)."""
    batch, rows, A = args
    import oracle
    oracle.eval_batch(batch, rows, A)
    oracle.bcp_batch(batch, rows, A)
    return batch.n_inst


def _pin_worker(cpus):
    """Pool initializer: pin this worker process to the next host CPU of the shared list."""
    try:
        os.sched_setaffinity(0, {cpus.get_nowait()})
    except Exception:  # pinning is best effort (a queue shorter than the pool, no permission)
        pass


def _pinned_pool(ctx, cores: int):
    """P forked worker processes, each pinned to its own CPU of this process's affinity set."""
    q = ctx.Queue()
    for c in sorted(os.sched_getaffinity(0))[:cores]:
        q.put(c)
    return ctx.Pool(cores, initializer=_pin_worker, initargs=(q,))


def cpu_baseline_eval_bcp(batch, rows, A, seconds: float, cores: int):
    """All-core CPU baseline of the headline step: P forked worker processes pinned to one
    CPU each, running the oracle's clause evaluation and UP closure on 32-instance slices of
    the same instances and rows (cycling through them).  The sample is sized from a one-core
    probe to about 1.5 x `seconds` of CPU time over all workers."""
    import multiprocessing as mp
    import oracle
    base = batch.row_base(A)

    def piece(lo, hi):
        idx = list(range(lo, hi))
        return batch.subset(idx), np.ascontiguousarray(rows[base[lo]:base[hi]]), A

    probe = piece(0, min(32, batch.n_inst))
    t0 = time.perf_counter()
    _eval_bcp_chunk(probe)
    per = (time.perf_counter() - t0) / probe[0].n_inst
    n_sample = int(max(cores * 8, 1.5 * seconds / max(per, 1e-7)))
    # the distinct slices once; the sample walks through them again until it is long enough
    slices = [piece(lo, min(batch.n_inst, lo + 32)) for lo in range(0, batch.n_inst, 32)]
    chunks, k = [], 0
    while sum(c[0].n_inst for c in chunks) < n_sample:
        chunks.append(slices[k % len(slices)])
        k += 1
    ctx = mp.get_context("fork")
    with _pinned_pool(ctx, cores) as pool:
        t0 = time.perf_counter()
        done = sum(pool.map(_eval_bcp_chunk, chunks, chunksize=1))
        dt = time.perf_counter() - t0
    return done / dt, done, dt, per


def _ref_solve_chunk(sub):
    """Worker: the oracle's C restatement of the reference's sequential path
    (preprocessing, then decision = last free var, decision-scheme learning, 20-clause
    learned DB, restarts 100/x1.3) over a slice of instances, one at a time."""
    import oracle
    return oracle.solve_ref_batch(sub)


def cpu_baseline_solve(batch, seconds: float, cores: int):
    """All-core CPU baseline of the solve leg: P worker processes pinned to one CPU each,
    each solving its share sequentially (-b 1 -t 1 semantics).  Returns instances/s over
    the sample (the instances in order, as many as about `seconds` of wall time allows)."""
    import multiprocessing as mp
    import oracle
    # size the sample from a short single-core probe so the leg takes about `seconds`
    probe = batch.subset(range(min(64, batch.n_inst)))
    t0 = time.perf_counter()
    oracle.solve_ref_batch(probe)
    per = (time.perf_counter() - t0) / probe.n_inst
    n_sample = int(min(batch.n_inst, max(cores, seconds * cores / max(per, 1e-7))))
    chunks = [batch.subset(range(s, min(n_sample, s + 64))) for s in range(0, n_sample, 64)]
    ctx = mp.get_context("fork")
    with _pinned_pool(ctx, cores) as pool:
        t0 = time.perf_counter()
        res = pool.map(_ref_solve_chunk, chunks, chunksize=1)
        dt = time.perf_counter() - t0
    return n_sample / dt, n_sample, dt, np.concatenate(res)


def solve_leg(G, W, b, cpu):
    """Verified random 3-SAT instances/s: N uniform random 3-SAT instances (n = 50,
    m = 213), one wavefront per instance; every SAT model re-verified on the GPU."""
    import oracle
    N = b.n_inst
    gb = G.GpuBatch(b)
    o = G.SolveOpts.default(max_ms=30000)
    # one untimed solve of the same batch: the timed call then reuses the device blocks the
    # library keeps from it, as any repeated solve of one shape does
    gb.solve(o, want_models=False)
    t0 = time.perf_counter()
    res = gb.solve(o, want_models=False)
    wall = time.perf_counter() - t0
    inf = res["info"]
    st = res["status"]
    decided = int((st != 2).sum())
    verified_sat = int(res["verified"].sum())
    ms = inf["kernel_ms"] + inf["verify_ms"]
    # verdict agreement on a sample, against the oracle's textbook CDCL
    idx = list(range(0, N, max(1, N // 200)))[:200]
    agree = sum(int(oracle.solve_cdcl(b.instance(i))[0] == st[i]) for i in idx)
    js = res["job_stats"]
    leg = {
        "workload": f"{N} uniform random 3-SAT (n=50, m=213), one wavefront per instance",
        "verified_instances_per_s": (verified_sat + int((st == 1).sum())) / (ms * 1e-3),
        # the same count over the host's wall clock around gpsat_solve (job table set-up,
        # copies, solve, model re-check and read-back included): the like-for-like figure
        # beside the CPU pool's wall clock
        "verified_instances_per_s_wall": (verified_sat + int((st == 1).sum())) / wall,
        "kernel_ms": inf["kernel_ms"], "verify_ms": inf["verify_ms"], "wall_s": wall,
        "sat": int((st == 0).sum()), "unsat": int((st == 1).sum()), "undef": int((st == 2).sum()),
        "sat_models_verified_on_gpu": verified_sat, "decided": decided,
        "verdict_agreement_sample": f"{agree}/{len(idx)}",
        "mean_conflicts": float(js[:, 1].mean()), "mean_props": float(js[:, 3].mean()),
        "props_per_s": float(js[:, 3].sum()) / (inf["kernel_ms"] * 1e-3),
        "waves": inf["waves"], "clause_mem": {1: "LDS", 2: "HBM", 3: "split", 4: "HBM-all"}[inf["clause_mem"]],
    }
    # the reference's own sequential search on the GPU (trajectory mode), same instances
    gb.solve_ref(max_iterations=0, learned_cap=2048, learned_lits=16384)
    t0 = time.perf_counter()
    tr = gb.solve_ref(max_iterations=0, learned_cap=2048, learned_lits=16384)
    t_ms = (time.perf_counter() - t0) * 1e3
    tst = tr["status"]
    leg["trajectory_mode"] = {
        "what": "the reference's sequential search (last-free-var decisions, decision-scheme learning, "
                "20-clause DB) on the GPU, one instance per wavefront, no iteration cap",
        "instances_per_s": N / (t_ms * 1e-3), "wall_ms_incl_workspace_alloc": t_ms,
        "decided": int((tst < 2).sum()), "verdicts_equal_cdcl": int(((tst == st) | (tst > 1) | (st > 1)).sum()),
        "capacity_stops": int((tst == 4).sum()),
    }
    # per-phase split of the device time (telemetry run on a 2,000-instance slice)
    sub = b.subset(range(min(2000, N)))
    with G.GpuBatch(sub) as gt:
        rt = gt.solve(G.SolveOpts.default(max_ms=30000, telemetry=1), want_models=False)
    ph = rt["job_stats"][:, 8:13].sum(axis=0).astype(np.float64)
    leg["phase_split"] = {k: float(v / max(ph.sum(), 1.0)) for k, v in
                          zip(("setup", "bcp", "analyze", "backtrack", "decide"), ph)}
    if cpu is not None:
        rate, ns, dt, cres, cores = cpu
        agree_cpu = sum(int(cres[i] == st[i]) for i in range(ns) if cres[i] != 2)
        leg["cpu_baseline"] = {
            "value": rate, "unit": "instances/s", "cores": cores, "kind": "port",
            "sample": f"first {ns} of the same instances, {dt:.1f} s wall on {cores} host cores "
                      f"(reference scheme restated in oracle/, one process pinned to each CPU); "
                      f"{agree_cpu}/{sum(1 for r in cres if r != 2)} decided verdicts agree with the GPU",
        }
        leg["speedup_vs_cpu"] = leg["verified_instances_per_s"] / rate
        leg["speedup_vs_cpu_wall"] = leg["verified_instances_per_s_wall"] / rate
    gb.close()
    return leg


def log(msg: str):
    """Progress on stderr; the JSON line alone goes to stdout."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cube_depth_for(n: int) -> int:
    """Cube depth 8..12 growing with the instance size (n in [100, 400])."""
    return 8 + min(4, max(0, (n - 100) * 5 // 301))


def mixed_leg(G, W, N: int, deadline_ms: float, pool: int):
    """4,096 mixed random k-SAT (k in {3,4,5} at ratios {4.26, 9.9, 21.1}, n in [100, 400])
    split by the reference's MaxClauseJobChooser into 2^8..2^12 cubes each, all in one
    persistent queue; launch shape from the autotuner; restart epochs as a replayed HIP
    graph; a conflict pool per instance and a whole-run deadline."""
    t0 = time.perf_counter()
    b, ks = W.mixed_ksat_batch(N, seed=3)
    ns = b.n_vars.astype(np.int64)
    vs = []
    for i in range(N):
        d = cube_depth_for(int(ns[i]))
        # the reference's depth rule: floor(log2(T*B*10)) + 1 = d for T*B below
        vs.append(G.cube_vars(b.instance(i), threads=[13, 26, 52, 103, 205][d - 8], blocks=1))
    jobs = G.cube_jobs(b, vs)
    log(f"mixed: {N} instances, {len(jobs[0])} cube jobs ({time.perf_counter() - t0:.1f} s host)")
    base = G.SolveOpts.default(inst_conflicts=pool, job_order=1, deadline_ms=deadline_ms, epoch_conflicts=500,
                               epoch_growth=2.0, max_epochs=8)
    with G.GpuBatch(b) as gb:
        t1 = time.perf_counter()
        best, rep = gb.autotune_solve(base, jobs=jobs, sample_jobs=4096, trial_ms=150.0)
        tune_s = time.perf_counter() - t1
        log(f"mixed: autotune {tune_s:.1f} s -> waves {best.waves} mem {best.clause_mem} chunk {best.chunk}")
        res = gb.solve(best, jobs=jobs, want_models=False)
    st, js, inf = res["status"], res["job_stats"], res["info"]
    decided = int((st != 2).sum())
    ms = inf["kernel_ms"]
    per_k = {str(k): {"instances": int((ks == k).sum()), "sat": int(((st == 0) & (ks == k)).sum()),
                      "unsat": int(((st == 1) & (ks == k)).sum())} for k in (3, 4, 5)}
    log(f"mixed: decided {decided}/{N} in {ms:.0f} ms, {inf['epochs']} epochs")
    return {
        "workload": f"{N} mixed random k-SAT (k 3/4/5 at 4.26/9.9/21.1, n 100-400), MaxClauseJobChooser cubes "
                    f"of depth 8-12 ({len(jobs[0])} jobs), one persistent queue, restart epochs as a HIP graph",
        "budget": f"{pool} conflicts per instance over its cubes, {deadline_ms / 1e3:.0f} s for the run",
        "decided": decided, "decided_frac": decided / N, "decided_per_s": decided / (ms * 1e-3),
        "sat": int((st == 0).sum()), "unsat": int((st == 1).sum()), "per_k": per_k,
        "sat_models_verified_on_gpu": int(res["verified"].sum()),
        "kernel_ms": ms, "epochs": inf["epochs"], "jobs": int(len(jobs[0])),
        "jobs_finished": int((js[:, 0] != 2).sum()),
        "props_per_s": float(js[:, 3].sum()) / (ms * 1e-3), "conflicts_per_s": float(js[:, 1].sum()) / (ms * 1e-3),
        "launch_shape": {"waves": inf["waves"], "clause_mem": {1: "LDS", 2: "HBM", 3: "split", 4: "HBM-all"}[inf["clause_mem"]],
                         "chunk": inf["chunk"], "autotune_s": tune_s},
        "autotune_report": rep.strip().splitlines(),
    }


def structured_leg(G, W, n_color: int, n_php: int):
    """2,000 4-colourings of ER(150, d in [4.2, 4.8]) and 200 PHP(9,8): exact verdicts (every
    colouring model re-verified on the GPU; PHP(9,8) is UNSAT by the pigeonhole principle)
    and learned-clause counts."""
    from gpupsat_amd.formula import Batch
    t0 = time.perf_counter()
    col = W.coloring_batch(n_color, seed=4)
    fs = [col.instance(i) for i in range(n_color)] + [W.pigeonhole_formula(9, 8)] * n_php
    b = Batch.from_formulas(fs)
    log(f"structured: {b.n_inst} instances ({time.perf_counter() - t0:.1f} s host)")
    with G.GpuBatch(b) as gb:
        res = gb.solve(G.SolveOpts.default(deadline_ms=60000), want_models=False)
    st, js, inf = res["status"], res["job_stats"], res["info"]
    c, p = slice(0, n_color), slice(n_color, None)
    out = {
        "workload": f"{n_color} 4-colourings of ER(150, d 4.2-4.8) + {n_php} PHP(9,8), one wavefront per instance",
        "coloring": {"sat": int((st[c] == 0).sum()), "unsat": int((st[c] == 1).sum()), "undef": int((st[c] == 2).sum()),
                     "models_verified_on_gpu": int(res["verified"][c].sum()),
                     "learned_clauses_total": int(js[c, 4].sum()), "learned_clauses_mean": float(js[c, 4].mean()),
                     "conflicts_mean": float(js[c, 1].mean())},
        "php_9_8": {"unsat": int((st[p] == 1).sum()), "sat": int((st[p] == 0).sum()), "undef": int((st[p] == 2).sum()),
                    "learned_clauses_total": int(js[p, 4].sum()), "learned_clauses_mean": float(js[p, 4].mean()),
                    "conflicts_mean": float(js[p, 1].mean()),
                    "solve_ms_mean": float(js[p, 7].mean() / 1e6)},
        "verdicts_exact": bool((st[p] == 1).all() and (st[c] != 1).all()
                               and int(res["verified"][c].sum()) == int((st[c] == 0).sum())),
        "kernel_ms": inf["kernel_ms"], "instances_per_s": b.n_inst / (inf["kernel_ms"] * 1e-3),
        "waves": inf["waves"],
    }
    log(f"structured: {inf['kernel_ms']:.0f} ms, PHP UNSAT {out['php_9_8']['unsat']}/{n_php}")
    return out


def learned_leg(n: int, cube_ms: float):
    """Learned split vs plain JobChooser/VSIDS on held-out random 3-SAT (n=300, m=1278):
    depth-4 cubes from the MaxClause order vs the network's top-scored variables (plus its
    initial scores and phases), same GPU and budgets; medians over instances both decide."""
    from gpupsat_amd import train_branching as T
    from gpupsat_amd import workloads as W
    if not os.path.exists(T.DEFAULT_WEIGHTS):
        return {"skipped": "no trained weights"}
    theta = np.load(T.DEFAULT_WEIGHTS)
    b = W.random_3sat_batch(n, 300, 1278, seed=777)
    import gpupsat_amd as G
    opts = G.SolveOpts.default(max_conflicts=200_000, max_ms=cube_ms)
    log(f"learned: {n} held-out instances, depth-4 cubes, {cube_ms:.0f} ms per cube")
    out = T.compare(b, theta, opts, cube_depth=4)
    out["workload"] = f"{n} held-out random 3-SAT (n=300, m=1278), 16 cubes each"
    out["weights"] = "gpupsat_amd/branching_weights.npy (trained on the GPU by gpsat_gnn_train_step)"
    from gpupsat_amd import learned as Lr
    ms, flops = Lr.time_train_steps(n_inst=512, steps=10)
    out["train_step"] = {"what": "one training step (forward, loss, backward, Adam) on 512 random 3-SAT instances "
                                 "(n=300, m=1278), wall clock around the synchronous call",
                         "ms": ms, "algorithmic_tflops": flops / 1e12, "frac_of_bf16_peak": flops / 2.5e15,
                         "flops_rule": "3 x 512 x 3 rounds x (2n+m) nodes x 2*64*64*2"}
    return out


def general_leg(G, W, N: int, steps: int, densities=(0.02, 0.05, 0.3), with_csr: bool = True):
    """The general eval/BCP kernel (any clause length, n and m) on the mixed k-SAT batch of
    the cube workload: 4,096 instances, k in {3,4,5} at ratios {4.26, 9.9, 21.1}, n in
    [100, 400], 64 rows each.  GB/s by the survey's byte accounting per instance."""
    b, ks = W.mixed_ksat_batch(N, seed=3)
    A = 64
    Ls = np.diff(b.clause_off[b.inst_clause_off])
    ms_, ns_ = np.diff(b.inst_clause_off), b.n_vars
    byts = int(sum(survey_bytes_per_instance(int(n), int(m), int(L), A, True, True)
                   for n, m, L in zip(ns_, ms_, Ls)))
    # compulsory bytes of the packed kernel: one 8-byte word pair per clause, rows in and out,
    # counts, conflict flags and props, the instance's metadata and longest-clause byte
    Ws = (2 * ns_.astype(np.int64) + 31) // 32
    packed = int(np.sum(8 * ms_ + 8 * A * Ws + 20 * A + 33))
    out = {"workload": f"{N} mixed random k-SAT (k 3/4/5 at 4.26/9.9/21.1, n 100-400, m up to {int(ms_.max())}), "
                       f"{A} rows each, fused eval + BCP through the packed general kernel (clauses of up to 6 "
                       f"literals, n <= 496); the CSR general kernel timed beside it",
           "survey_bytes_per_launch": byts, "packed_compulsory_bytes_per_launch": packed, "densities": {}}

    def timed(gb, d_r, d_c, d_o, d_f, d_p):
        run = lambda: gb.eval_bcp_dev(d_r, A, d_c, d_o, d_f, d_p)
        run()
        gb.sync()
        e0, e1 = G.Event(), G.Event()
        e0.record(gb)
        for _ in range(steps):
            run()
        e1.record(gb)
        gb.sync()
        return e0.elapsed_ms(e1) / steps

    with G.GpuBatch(b) as gb:
        out["kernel"] = gb.eval_kernel()
        for pd in densities:
            rows = W.random_assignments(b, A, pd, seed=9)
            d_r = G.DevBuf.from_host(rows)
            d_c, d_o = G.DevBuf(N * A * 12), G.DevBuf(rows.nbytes)
            d_f, d_p = G.DevBuf(N * A * 4), G.DevBuf(N * A * 4)
            ms = timed(gb, d_r, d_c, d_o, d_f, d_p)
            got = (d_c.to_host(np.int32, N * A * 3), d_o.to_host(np.uint32, rows.size),
                   d_f.to_host(np.int32, N * A), d_p.to_host(np.int32, N * A))
            ms_csr = same = None
            if with_csr:
                with G.tuned(eval_pkk=0):
                    ms_csr = timed(gb, d_r, d_c, d_o, d_f, d_p)
                    same = all(np.array_equal(x, y) for x, y in zip(got, (
                        d_c.to_host(np.int32, N * A * 3), d_o.to_host(np.uint32, rows.size),
                        d_f.to_host(np.int32, N * A), d_p.to_host(np.int32, N * A))))
            cf, pr = got[2], got[3]
            gbs = byts / (ms * 1e-3) / 1e9
            out["densities"][f"p{pd}"] = {"fused_ms": ms, "instances_per_s": N / (ms * 1e-3), "survey_gbs": gbs,
                                          "frac_of_8tbs": gbs / HBM_PEAK_GBS,
                                          "packed_compulsory_gbs": packed / (ms * 1e-3) / 1e9,
                                          "conflict_frac": float(cf.mean()),
                                          "implied_literals_per_s": float(pr.sum()) / (ms * 1e-3),
                                          "csr_kernel_ms": ms_csr,
                                          "speedup_vs_csr": ms_csr / ms if ms_csr else None,
                                          "outputs_equal_csr": same}
            for d in (d_r, d_c, d_o, d_f, d_p):
                d.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--instances", type=int, default=10_000)
    ap.add_argument("--p", type=float, default=0.3,
                    help="trail density of the headline step (the workload's is 0.3; profiling runs at other "
                         "densities set it with --no-sweeps so every eval/BCP launch is at that density)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solve-instances", type=int, default=20_000)
    ap.add_argument("--no-solve", action="store_true", help="skip every solve leg")
    ap.add_argument("--no-legs", action="store_true", help="skip the mixed, structured and learned legs")
    ap.add_argument("--no-sweeps", action="store_true",
                    help="skip the low-density BCP leg, the fused density sweep and the general-kernel leg (every "
                         "3-literal eval/BCP launch is then at --p, so a profiler's per-kernel averages are the "
                         "headline step's)")
    ap.add_argument("--mixed-instances", type=int, default=4096)
    ap.add_argument("--mixed-deadline-ms", type=float, default=20000.0)
    ap.add_argument("--learned-instances", type=int, default=256)
    ap.add_argument("--compare-packed", action="store_true",
                    help="time the packed general kernel beside the 3-literal kernel on the headline batch over the "
                         "sweep's trail densities (sets GPSAT_PKK_FOR_FAST3) and print that comparison only")
    ap.add_argument("--general-only", type=str, default=None, metavar="P[,P...]",
                    help="run only the general-kernel leg (packed kernel, no CSR comparison) at these trail "
                         "densities and print its JSON (for profiling that kernel)")
    args = ap.parse_args()
    if args.compare_packed:
        os.environ["GPSAT_PKK_FOR_FAST3"] = "1"
        import gpupsat_amd as G
        from gpupsat_amd import workloads as W
        b = W.random_3sat_batch(args.instances, 200, 852, seed=0)
        A, res = 64, {}
        with G.GpuBatch(b) as gb:
            d_c, d_f, d_p = G.DevBuf(b.n_inst * A * 12), G.DevBuf(b.n_inst * A * 4), G.DevBuf(b.n_inst * A * 4)
            for pd in (0.02, 0.05, 0.1, 0.3):
                r = W.random_assignments(b, A, pd, seed=2)
                d_r, d_o = G.DevBuf.from_host(r), G.DevBuf(r.nbytes)
                row = {}
                for name, kw in (("fast3", {}), ("packed_nw4", dict(eval_fast3=0, eval_pkk_waves=4)),
                                 ("packed_nw2", dict(eval_fast3=0, eval_pkk_waves=2)),
                                 ("packed_nw1", dict(eval_fast3=0, eval_pkk_waves=1))):
                    with G.tuned(**kw):
                        run = lambda: gb.eval_bcp_dev(d_r, A, d_c, d_o, d_f, d_p)
                        run()
                        gb.sync()
                        e0, e1 = G.Event(), G.Event()
                        e0.record(gb)
                        for _ in range(args.steps):
                            run()
                        e1.record(gb)
                        gb.sync()
                        row[name] = {"kernel": gb.eval_kernel(), "ms": e0.elapsed_ms(e1) / args.steps,
                                     "props": int(d_p.to_host(np.int32, b.n_inst * A).sum())}
                res[f"p{pd}"] = row
                d_r.close(); d_o.close()
        print(json.dumps(res))
        return
    if args.general_only:
        import gpupsat_amd as G
        from gpupsat_amd import workloads as W
        dens = tuple(float(x) for x in args.general_only.split(","))
        print(json.dumps(general_leg(G, W, args.mixed_instances, args.steps, dens, with_csr=False)))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the instance batch would shard across GPUs with no "
              "exchange between them, and multi-GPU runs are not measured here.", file=sys.stderr)
        sys.exit(2)

    import gpupsat_amd as G
    from gpupsat_amd import workloads as W

    N, n, m, A, p = args.instances, 200, 852, 64, args.p
    b = W.random_3sat_batch(N, n, m, seed=0)
    rows = W.random_assignments(b, A, p, seed=0)
    L = int(b.clause_off[-1]) // N

    # The multi-process CPU baselines run first, before this process touches the GPU, so
    # the forked workers inherit no device state.
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    solve_batch = solve_cpu = cpu = None
    if not args.no_cpu:
        log(f"CPU baseline of the headline step on {cores} cores")
        rate, done, dt, per = cpu_baseline_eval_bcp(b, rows, A, args.cpu_seconds, cores)
        cpu = {"value": rate, "unit": "instances/s", "cores": cores, "kind": "port",
               "sample": f"the oracle's clause evaluation + UP closure (oracle/oracle.c) of {done} instances x {A} rows "
                         f"(the same {b.n_inst} instances and rows, cycled in 32-instance slices), {dt:.1f} s wall on "
                         f"{cores} host processes pinned to one CPU each ({1e3 * per:.2f} ms per instance on one core)"}
    if not args.no_solve:
        solve_batch = W.random_3sat_batch(args.solve_instances, 50, 213, seed=100)
        if not args.no_cpu:
            solve_cpu = cpu_baseline_solve(solve_batch, args.cpu_seconds, cores) + (cores,)

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
    step()
    gb.sync()
    conflict = d_conf.to_host(np.int32, N * A).reshape(N, A)
    props = d_props.to_host(np.int32, N * A).reshape(N, A)

    ms_bcp_lo, props_lo, conf_lo, sweep, general = None, None, None, None, None
    if not args.no_sweeps:
        # low-density trails, where BCP actually implies literals (props/s)
        rows_lo = W.random_assignments(b, A, 0.05, seed=1)
        d_rows_lo = G.DevBuf.from_host(rows_lo)
        gb.bcp_fixpoint_dev(d_rows_lo, A, d_out, d_conf, d_props)
        ms_bcp_lo = timed(lambda: gb.bcp_fixpoint_dev(d_rows_lo, A, d_out, d_conf, d_props), args.steps)
        props_lo = d_props.to_host(np.int32, N * A)
        conf_lo = d_conf.to_host(np.int32, N * A)
        d_rows_lo.close()

        # the fused launch over the trail densities the low-density sweep asks for
        sweep = {}
        for pd in (0.02, 0.05, 0.1, 0.3):
            r_p = W.random_assignments(b, A, pd, seed=2)
            d_p = G.DevBuf.from_host(r_p)
            gb.eval_bcp_dev(d_p, A, d_counts, d_out, d_conf, d_props)
            ms_p = timed(lambda: gb.eval_bcp_dev(d_p, A, d_counts, d_out, d_conf, d_props), args.steps)
            cf = d_conf.to_host(np.int32, N * A)
            pr = d_props.to_host(np.int32, N * A)
            gbs = N * survey_bytes_per_instance(n, m, L, A, True, True) / (ms_p * 1e-3) / 1e9
            sweep[f"p{pd}"] = {"fused_ms": ms_p, "instances_per_s": N / (ms_p * 1e-3), "survey_gbs": gbs,
                               "frac_of_8tbs": gbs / HBM_PEAK_GBS, "conflict_frac": float(cf.mean()),
                               "implied_literals_per_s": float(pr.sum()) / (ms_p * 1e-3),
                               "traffic_bytes": traffic_for(pd)}
            d_p.close()
        log("general eval/BCP kernel on the mixed k-SAT batch")
        general = general_leg(G, W, 4096, args.steps)

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

    bytes_fused = N * survey_bytes_per_instance(n, m, L, A, True, True)
    bytes_eval = N * survey_bytes_per_instance(n, m, L, A, True, False)
    bytes_bcp = N * survey_bytes_per_instance(n, m, L, A, False, True)
    bytes_packed = N * packed_bytes_per_instance(n, m, A, True, True)
    achieved = bytes_fused / (ms_fused * 1e-3) / 1e9
    traffic = traffic_for(p)

    legs = {}
    if solve_batch is not None:
        log("solve leg")
        legs["solve"] = solve_leg(G, W, solve_batch, solve_cpu)
    if not args.no_solve and not args.no_legs:
        legs["mixed_ksat_cubes"] = mixed_leg(G, W, args.mixed_instances, args.mixed_deadline_ms, 100_000)
        legs["structured"] = structured_leg(G, W, 2000, 200)
        legs["learned_split"] = learned_leg(args.learned_instances, 3000.0)

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
        "config": {"workload": f"batched clause evaluation + BCP fixpoints: {N} random 3-SAT (n={n}, m={m}), "
                               f"{A} partial assignments each (p={p}); value = instances whose {A} rows all got "
                               f"exact #SAT/#UNSAT/#UNIT counts and exact BCP fixpoints, per second (no instance "
                               f"is solved or verified in this step: the verified-instance rate of the solve path "
                               f"is legs.solve)",
                   "instances": N, "n": n, "m": m, "assignments": A, "p": p},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": bytes_fused,
                     "bytes_rule": "survey accounting, 21,060 B per instance: arena 4(L+m+1), rows A*ceil(n/4) in, "
                                   "counts 12A, closure rows and flags A*(ceil(n/4)+4) out",
                     "kernel_compulsory_bytes": bytes_packed,
                     "kernel_compulsory_gbs": bytes_packed / (ms_fused * 1e-3) / 1e9,
                     "traffic_gbs": None if traffic is None else traffic / (ms_fused * 1e-3) / 1e9,
                     "traffic_frac": None if traffic is None else traffic / (ms_fused * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "kernels": {
            "eval_bcp_fused_ms": ms_fused, "eval_ms": ms_eval, "bcp_ms": ms_bcp,
            "eval_survey_gbs": bytes_eval / (ms_eval * 1e-3) / 1e9, "bcp_survey_gbs": bytes_bcp / (ms_bcp * 1e-3) / 1e9,
            "survey_bytes_per_launch": {"fused": bytes_fused, "eval": bytes_eval, "bcp": bytes_bcp},
            "copy_peak_gbs": copy_gbs,
        },
        "bcp": {
            f"p{p}_conflict_frac": float(conflict.mean()),
            f"p{p}_implied_literals_per_s": float(props.sum()) / (ms_fused * 1e-3),
            "p0.05_ms": ms_bcp_lo,
            "p0.05_conflict_frac": None if conf_lo is None else float(conf_lo.mean()),
            "p0.05_props_per_s": None if props_lo is None else float(props_lo.sum()) / (ms_bcp_lo * 1e-3),
            "fused_density_sweep": sweep,
        },
        "general_kernel": general,
        "parity_sample_ok": parity,
        "legs": legs,
    }
    print(json.dumps(out))


def traffic_for(p: float):
    """HBM bytes per fused launch at trail density p, from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "eval_bcp_3sat_n200_m852_A64_traffic.json")
    try:
        d = json.load(open(path))
    except Exception:
        return None
    by_p = d.get("bytes_per_launch_by_p") or {}
    return by_p.get(f"p{p}")


if __name__ == "__main__":
    main()
