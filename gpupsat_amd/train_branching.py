"""Train the learned branching network on labels from the plain solver, then compare
the solver with and without its proposals on held-out instances.

    python -m gpupsat_amd.train_branching --train 50000 --epochs 25 --out gpupsat_amd/branching_weights.npy
    python -m gpupsat_amd.train_branching --eval-only --weights gpupsat_amd/branching_weights.npy

Training instances: random 3-SAT, n uniform in [100, 300], ratio uniform in [4.2, 4.3].
Evaluation: random 3-SAT (n=300, m=1278) and 4-colouring of ER(150, d in [4.2, 4.8]).
Every comparison is made on instances that both runs decide, whose verdicts must agree.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import GpuBatch, SolveOpts, cube_jobs, cube_vars
from . import learned as Lr
from . import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "branching_weights.npy")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_training_sets(n_train: int, chunk: int, seed: int, budget_conflicts: int, budget_ms: float):
    """Label n_train random 3-SAT instances with the plain solver, chunk instances per
    batch.  Returns ([(GpuBatch, labels)], stats)."""
    sets = []
    t0 = time.perf_counter()
    stats = {"instances": n_train, "sat": 0, "unsat": 0, "undef": 0}
    for s in range(0, n_train, chunk):
        b = W.training_3sat_batch(min(chunk, n_train - s), seed=seed * 100_000 + s)
        gb, res, lab = Lr.generate_labels(b, SolveOpts.default(max_conflicts=budget_conflicts, max_ms=budget_ms))
        st = res["status"]
        stats["sat"] += int((st == 0).sum())
        stats["unsat"] += int((st == 1).sum())
        stats["undef"] += int((st == 2).sum())
        sets.append((gb, lab))
        log(f"labels: {s + b.n_inst}/{n_train} instances, {time.perf_counter() - t0:.0f} s")
    stats["label_seconds"] = time.perf_counter() - t0
    return sets, stats


def _per_instance(res, job_inst, n_inst):
    js = res["job_stats"]
    conf = np.zeros(n_inst)
    tns = np.zeros(n_inst)
    np.add.at(conf, job_inst, js[:, 1])
    np.add.at(tns, job_inst, js[:, 7])
    return conf, tns / 1e6


def compare(batch, theta, opts: SolveOpts, cube_depth: int = 0):
    """Plain solver vs solver with learned proposals on one batch.  Returns a dict of
    medians over instances both runs decide (verdicts must agree)."""
    gb = GpuBatch(batch)
    N = batch.n_inst
    model = Lr.GnnModel(theta)
    t0 = time.perf_counter()
    act, phase, per = Lr.propose(model, gb)
    infer_s = time.perf_counter() - t0
    model.close()
    if cube_depth:
        plain_vars = [cube_vars(batch.instance(i), strategy=1)[:cube_depth] for i in range(N)]
        learned_vars = Lr.learned_cubes(per, cube_depth)
        jp = cube_jobs(batch, plain_vars)
        jl = cube_jobs(batch, learned_vars)
        rp = gb.solve(opts, jobs=jp, want_models=False)
        log(f"  plain cubes: {rp['info']['kernel_ms']:.0f} ms")
        rl = gb.solve(opts, jobs=jl, act_init=act, phase_init=phase, want_models=False)
        log(f"  learned cubes: {rl['info']['kernel_ms']:.0f} ms")
        cp, tp = _per_instance(rp, jp[0], N)
        cl, tl = _per_instance(rl, jl[0], N)
    else:
        rp = gb.solve(opts, want_models=False)
        log(f"  plain: {rp['info']['kernel_ms']:.0f} ms")
        rl = gb.solve(opts, act_init=act, phase_init=phase, want_models=False)
        log(f"  learned: {rl['info']['kernel_ms']:.0f} ms")
        ar = np.arange(N)
        cp, tp = _per_instance(rp, ar, N)
        cl, tl = _per_instance(rl, ar, N)
    sp, sl = rp["status"], rl["status"]
    both = (sp != 2) & (sl != 2)
    agree = bool((sp[both] == sl[both]).all())
    gb.close()
    out = {
        "instances": N, "decided_plain": int((sp != 2).sum()), "decided_learned": int((sl != 2).sum()),
        "decided_both": int(both.sum()), "verdicts_identical": agree,
        "median_ms_plain": float(np.median(tp[both])) if both.any() else None,
        "median_ms_learned": float(np.median(tl[both])) if both.any() else None,
        "median_conflicts_plain": float(np.median(cp[both])) if both.any() else None,
        "median_conflicts_learned": float(np.median(cl[both])) if both.any() else None,
        "kernel_ms_plain": rp["info"]["kernel_ms"], "kernel_ms_learned": rl["info"]["kernel_ms"],
        "inference_s": infer_s,
    }
    if both.any():
        out["median_time_ratio"] = out["median_ms_learned"] / max(out["median_ms_plain"], 1e-9)
        out["median_conflict_ratio"] = out["median_conflicts_learned"] / max(out["median_conflicts_plain"], 1e-9)
    return out


def evaluate(theta, n_random: int, n_color: int, budget_conflicts: int, budget_ms: float, cube_depth: int = 4):
    opts = SolveOpts.default(max_conflicts=budget_conflicts, max_ms=budget_ms)
    rnd = W.random_3sat_batch(n_random, 300, 1278, seed=777)
    col = W.coloring_batch(n_color, seed=778)
    res = {}
    log(f"eval: {n_random} random 3-SAT n=300 m=1278")
    res["random_3sat_n300"] = compare(rnd, theta, opts)
    log(json.dumps(res["random_3sat_n300"]))
    log(f"eval: {n_color} colourings")
    res["coloring_er150"] = compare(col, theta, opts)
    log(json.dumps(res["coloring_er150"]))
    if cube_depth:
        # every cube job gets a quarter of the single-job time budget
        copts = SolveOpts.default(max_conflicts=budget_conflicts, max_ms=budget_ms / 4)
        log(f"eval: cubes of depth {cube_depth}")
        res["random_3sat_n300_cubes"] = compare(rnd, theta, copts, cube_depth=cube_depth)
        log(json.dumps(res["random_3sat_n300_cubes"]))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=8192)
    ap.add_argument("--chunk", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--label-conflicts", type=int, default=20000)
    ap.add_argument("--label-ms", type=float, default=4000)
    ap.add_argument("--eval-random", type=int, default=1000)
    ap.add_argument("--eval-color", type=int, default=500)
    ap.add_argument("--eval-conflicts", type=int, default=200000)
    ap.add_argument("--eval-ms", type=float, default=20000)
    ap.add_argument("--cube-depth", type=int, default=4)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--report", default=None)
    ap.add_argument("--time-steps", type=int, default=0,
                    help="only time this many training steps on 512 instances of n=300, m=1278 and exit")
    a = ap.parse_args()
    if a.time_steps:
        ms, flops = Lr.time_train_steps(steps=a.time_steps)
        print(json.dumps({"train_step_ms": ms, "algorithmic_tflops": flops / 1e12,
                          "frac_of_bf16_peak": flops / 2.5e15}))
        return
    report = {}
    if a.eval_only:
        theta = np.load(a.weights)
    else:
        sets, st = build_training_sets(a.train, a.chunk, a.seed, a.label_conflicts, a.label_ms)
        log("labels:", st)
        t0 = time.perf_counter()
        theta = Lr.train_sets(sets, Lr.init_params(a.seed), epochs=a.epochs, lr=a.lr, log=log)
        st["train_seconds"] = time.perf_counter() - t0
        st["loss_history"] = Lr.train_sets.history
        np.save(a.out, theta)
        report["training"] = st
        for gb, _ in sets:
            gb.close()
    report["evaluation"] = evaluate(theta, a.eval_random, a.eval_color, a.eval_conflicts, a.eval_ms, a.cube_depth)
    print(json.dumps(report))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
