"""Verdict parity of the device CDCL solver (through gpsat_solve) with the oracle's
solvers, GPU re-verification of every SAT model, and the cube/job-queue protocol."""
import glob
import os

import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd import workloads as W
from gpupsat_amd.formula import Batch, Formula, read_dimacs

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

BUDGET_MS = 4000.0


def opts(**kw):
    kw.setdefault("max_ms", BUDGET_MS)
    return G.SolveOpts.default(**kw)


def truth(orc, f):
    st, _, _ = orc.solve_cdcl(f)
    return st


def check_models(orc, gb, res, b):
    models = gb.models(res["model_rows"])
    for i in range(b.n_inst):
        if res["status"][i] == G.SAT:
            assert res["verified"][i] == 1
            assert orc.verify_model(b.instance(i), models[i]), i


def test_reference_fixtures(orc):
    print("[solve] reference fixtures", flush=True)
    paths = sorted(glob.glob(os.path.join(GOLDEN, "cnf", "*.cnf")))
    fs = [read_dimacs(p) for p in paths]
    exp = [{"SAT": 0, "UNSAT": 1}[open(p[:-4] + ".expected").read().strip()] for p in paths]
    b = Batch.from_formulas(fs)
    for mem in (1, 2):
        with G.GpuBatch(b) as gb:
            res = gb.solve(opts(clause_mem=mem))
            assert res["status"].tolist() == exp
            check_models(orc, gb, res, b)


@pytest.mark.parametrize("mem", [1, 2, 3])
def test_3sat_n50_verdicts(orc, mem):
    print(f"[solve] 100 x n=50 m=213 clause_mem={mem}", flush=True)
    b = W.random_3sat_batch(100, 50, 213, seed=0)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(clause_mem=mem))
        check_models(orc, gb, res, b)
    exp = [truth(orc, b.instance(i)) for i in range(b.n_inst)]
    assert res["status"].tolist() == exp
    assert res["info"]["clause_mem"] == mem


def test_reference_decision_rule(orc):
    print("[solve] reference decision rule", flush=True)
    b = W.random_3sat_batch(40, 40, 170, seed=3)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(decision=1, learnt_cap=20))
        check_models(orc, gb, res, b)
    assert res["status"].tolist() == [truth(orc, b.instance(i)) for i in range(b.n_inst)]


@pytest.mark.parametrize("mem", [0, 3])
def test_tiny_learned_db_and_watch_slack(orc, mem):
    """learnt_cap 4 and one learned watch slot per literal force the DB-full restart
    and the watch-overflow detach paths (automatic and split clause memory)."""
    print(f"[solve] tiny learned DB clause_mem={mem}", flush=True)
    b = W.random_3sat_batch(48, 60, 256, seed=5)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(learnt_cap=4, larena_cap=200, watch_slack=1, clause_mem=mem))
        check_models(orc, gb, res, b)
    assert res["status"].tolist() == [truth(orc, b.instance(i)) for i in range(b.n_inst)]


def test_pigeonhole_and_coloring(orc):
    print("[solve] pigeonhole + colouring", flush=True)
    fs = [W.pigeonhole_formula(p + 1, p) for p in (3, 4, 5, 6)]
    fs += [W.pigeonhole_formula(5, 5)]
    fs += [W.coloring_batch(4, n_nodes=60, seed=2).instance(i) for i in range(4)]
    b = Batch.from_formulas(fs)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts())
        check_models(orc, gb, res, b)
    st = res["status"].tolist()
    assert st[:4] == [G.UNSAT] * 4          # PHP(p+1, p) is UNSAT (theorem)
    assert st[4] == G.SAT
    assert st[5:] == [truth(orc, f) for f in fs[5:]]


@pytest.mark.parametrize("chunk", [0, 5, 16])
def test_cubes_cover_the_search_space(orc, chunk):
    """Every instance split into 2^d cubes: SAT iff some cube is SAT, UNSAT only when
    every cube is refuted (the reference's parallel path can stop early,
    src/SATSolver/Parallelizer.cu:60-74).  chunk > 1 makes a wavefront run several
    cubes of one instance in a row, reusing its watch lists."""
    print(f"[solve] cube jobs chunk={chunk}", flush=True)
    b = W.random_3sat_batch(30, 50, 213, seed=11)
    ji, co, cl = [], [0], []
    for i in range(b.n_inst):
        f = b.instance(i)
        vs = orc.choose_cube_vars(f, 0, 1, 1, 0)          # d = 4 -> 16 cubes
        for cube in orc.enumerate_cubes(vs):
            ji.append(i)
            cl.extend(cube.tolist())
            co.append(len(cl))
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(chunk=chunk), jobs=(ji, co, cl))
        check_models(orc, gb, res, b)
    assert res["status"].tolist() == [truth(orc, b.instance(i)) for i in range(b.n_inst)]
    js = res["job_stats"]
    for i in range(b.n_inst):
        mine = js[np.array(ji) == i][:, 0]
        if res["status"][i] == G.UNSAT:
            assert (mine == G.UNSAT).all()
        else:
            assert (mine == G.SAT).any()


def test_cube_reuse_with_unit_clauses(orc):
    """Formulas with unit clauses, split into cubes and run 8 cubes per wavefront: the
    units stay at the base of the trail across cubes."""
    print("[solve] cube reuse with units", flush=True)
    from gpupsat_amd.formula import Formula
    rng = np.random.default_rng(3)
    fs = []
    for i in range(12):
        f = W.random_3sat_batch(1, 40, 150 + 5 * i, seed=50 + i).instance(0)
        units = [[int(2 * v + rng.integers(0, 2))] for v in rng.choice(40, 3, replace=False)]
        fs.append(Formula.from_clauses(40, f.clauses() + units))
    b = Batch.from_formulas(fs)
    vars_ = [np.array([5, 9, 13], dtype=np.int32)] * b.n_inst
    jobs = G.cube_jobs(b, vars_)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(chunk=8), jobs=jobs)
        check_models(orc, gb, res, b)
    assert res["status"].tolist() == [truth(orc, f) for f in fs]


def test_budget_gives_undef(orc):
    print("[solve] budget", flush=True)
    b = Batch.from_formulas([W.pigeonhole_formula(9, 8)])
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(max_conflicts=3))
    assert res["status"].tolist() == [G.UNDEF]
    assert res["job_stats"][0, 0] == G.UNDEF


def test_mixed_sizes_with_budget(orc):
    print("[solve] mixed k-SAT n in [100,400]", flush=True)
    b, ks = W.mixed_ksat_batch(24, seed=1)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(max_conflicts=20000))
        check_models(orc, gb, res, b)
    for i in range(b.n_inst):
        if res["status"][i] == G.UNSAT:
            st, _, _ = orc.solve_cdcl(b.instance(i), max_conflicts=2_000_000)
            assert st == G.UNSAT


@pytest.mark.parametrize("cubes", [False, True])
def test_restart_epochs_graph(orc, cubes):
    """Restart epochs: the solve is a HIP graph replayed per epoch with a doubling
    conflict budget; verdicts equal the oracle's and the single-launch solve's."""
    print(f"[solve] restart epochs cubes={cubes}", flush=True)
    fs = [W.random_3sat_batch(1, 90, 383, seed=200 + i).instance(0) for i in range(40)]
    fs.append(W.pigeonhole_formula(7, 6))
    b = Batch.from_formulas(fs)
    jobs = None
    if cubes:
        jobs = G.cube_jobs(b, [np.array([0, 1], dtype=np.int32)] * b.n_inst)
    with G.GpuBatch(b) as gb:
        eager = gb.solve(opts(), jobs=jobs)
        res = gb.solve(opts(epoch_conflicts=40, epoch_growth=2.0, max_epochs=20), jobs=jobs)
        check_models(orc, gb, res, b)
    assert res["info"]["epochs"] > 1
    exp = [truth(orc, f) for f in fs]
    assert res["status"].tolist() == exp
    assert eager["status"].tolist() == exp


@pytest.mark.parametrize("epochs", [False, True])
def test_deadline_stops_the_launch(orc, epochs):
    """A whole-solve deadline: the hard PHP(9,8) instances end UNDEF when it passes, the
    easy ones are decided exactly, and the launch does not outlive the deadline by much."""
    print(f"[solve] deadline epochs={epochs}", flush=True)
    fs = [W.pigeonhole_formula(9, 8)] * 6
    fs += [W.random_3sat_batch(1, 50, 213, seed=300 + i).instance(0) for i in range(40)]
    b = Batch.from_formulas(fs)
    kw = dict(deadline_ms=400.0, max_ms=0.0)
    if epochs:
        kw.update(epoch_conflicts=500, max_epochs=30)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(**kw))
        check_models(orc, gb, res, b)
    st = res["status"].tolist()
    assert st[:6] == [G.UNDEF] * 6
    exp = [truth(orc, f) for f in fs[6:]]
    assert st[6:] == exp
    assert res["info"]["kernel_ms"] < 3000.0


@pytest.mark.parametrize("order", [0, 1, 2])
def test_instance_conflict_pool_and_order(orc, order):
    """A conflict pool per instance shared by its cubes: PHP(8,7) spends it and ends UNDEF,
    the easy instances are decided exactly, and no instance's jobs overrun the pool by
    more than one check interval (64 conflicts) per job; every queue order gives the same
    verdicts."""
    print(f"[solve] instance conflict pool, job order {order}", flush=True)
    fs = [W.pigeonhole_formula(8, 7)] * 3
    fs += [W.random_3sat_batch(1, 60, 256, seed=400 + i).instance(0) for i in range(30)]
    b = Batch.from_formulas(fs)
    jobs = G.cube_jobs(b, [np.array([0, 5, 9], dtype=np.int32)] * b.n_inst)
    pool = 3000
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(inst_conflicts=pool, job_order=order, chunk=4), jobs=jobs)
        check_models(orc, gb, res, b)
    st = res["status"].tolist()
    assert st[:3] == [G.UNDEF] * 3
    exp = [truth(orc, f) for f in fs[3:]]
    for i, e in enumerate(exp):
        assert st[3 + i] in (e, G.UNDEF)
    assert sum(x != G.UNDEF for x in st[3:]) >= 25
    js = res["job_stats"]
    ji = np.asarray(jobs[0])
    for i in range(3):
        assert js[ji == i, 1].sum() <= pool + 64 * int((ji == i).sum())


def test_telemetry_phases(orc):
    """Per-phase times (set-up, BCP, analysis, backtrack, decide): off by default, and when
    on they add up to the job's own time and do not change a verdict."""
    print("[solve] telemetry", flush=True)
    b = W.random_3sat_batch(64, 60, 256, seed=77)
    with G.GpuBatch(b) as gb:
        r0 = gb.solve(opts())
        r1 = gb.solve(opts(telemetry=1))
    assert r0["status"].tolist() == r1["status"].tolist()
    js0, js1 = r0["job_stats"], r1["job_stats"]
    assert (js0[:, 8:] == 0).all()
    ph = js1[:, 8:13]
    assert (ph >= 0).all()
    tot = ph.sum(axis=1)
    t = js1[:, 7]
    # the phases cover the job's time up to the clock reads around it
    assert (tot <= t + 2000).all() and (tot >= 0.9 * t - 2000).all()
    busy = js1[:, 3] > 0
    assert (ph[busy, 1] > 0).all()                       # BCP time where there were propagations
    assert (ph[js1[:, 1] > 0, 2] > 0).all()              # analysis time where there were conflicts


@pytest.mark.parametrize("mem", [0, 2, 3])
def test_learned_clauses_are_implied(orc, mem):
    """Soundness of learning: every learned clause a job still holds at its end, and every
    level-0 literal, is implied by the formula -- the formula plus the clause's negation as
    unit clauses is UNSAT for the oracle's CDCL.  A small learned DB (16 clauses) forces
    reductions; random two-literal clauses over the SAT instances show the check can fail."""
    print(f"[solve] learned clauses implied: 40 x random 3-SAT n=50, learnt_cap 16, clause_mem={mem}", flush=True)
    b = W.random_3sat_batch(40, 50, 213, seed=77)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(learnt_cap=16, clause_mem=mem), learned_cap=1 << 20)
    info = res["info"]
    assert info["learned_len"] == info["learned_need"] > 0
    blocks = res["learned"]
    assert sorted(i for i, _ in blocks) == list(range(b.n_inst))
    checked = units = not_implied = 0
    rng = np.random.default_rng(3)
    for inst, clauses in blocks:
        f = b.instance(inst)
        base = f.clauses()
        assert all(len(c) >= 1 for c in clauses)
        # up to 12 learned clauses per instance, and every level-0 unit
        long_ = [c for c in clauses if len(c) > 1]
        picks = [long_[k] for k in rng.permutation(len(long_))[:12]] + [c for c in clauses if len(c) == 1]
        for c in picks:
            g = Formula.from_clauses(f.n_vars, base + [[l ^ 1] for l in c])
            assert truth(orc, g) == G.UNSAT, (inst, c)
            checked += 1
            units += len(c) == 1
        if res["status"][inst] == G.SAT:
            for _ in range(4):
                v = rng.choice(f.n_vars, size=2, replace=False)
                c = [2 * int(x) + int(rng.integers(2)) for x in v]
                g = Formula.from_clauses(f.n_vars, base + [[l ^ 1] for l in c])
                not_implied += truth(orc, g) == G.SAT
    assert checked > 100 and not_implied > 0
    print(f"[solve] {checked} learned clauses checked ({units} level-0 units); "
          f"{not_implied} random clauses correctly found not implied", flush=True)


def test_epoch_budget_respects_max_conflicts(orc):
    """With restart epochs, a job's conflicts over all its epochs stay within the per-job
    max_conflicts; jobs that reach it end UNDEF and are not re-queued."""
    print("[solve] epoch budget bounded by max_conflicts", flush=True)
    fs = [W.pigeonhole_formula(8, 7)] * 4
    fs += [W.random_3sat_batch(1, 60, 256, seed=900 + i).instance(0) for i in range(20)]
    b = Batch.from_formulas(fs)
    cap = 700
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(max_ms=0.0, max_conflicts=cap, epoch_conflicts=100, epoch_growth=2.0, max_epochs=12))
        check_models(orc, gb, res, b)
    js = res["job_stats"]
    assert (js[:, 1] <= cap).all(), js[:, 1].max()
    assert (res["status"][:4] == G.UNDEF).all()          # PHP(8,7) needs far more than 700 conflicts
    assert (js[:4, 1] == cap).all()
    for i in range(4, b.n_inst):
        st = res["status"][i]
        if st != G.UNDEF:
            assert st == truth(orc, b.instance(i))
        else:
            assert js[i, 1] == cap


def test_all_hbm_state_mode(orc):
    """clause_mem=4 keeps the per-var state in HBM too: verdicts equal the oracle's, the
    same as the all-LDS layout gives."""
    print("[solve] all-HBM job state, n=50", flush=True)
    b = W.random_3sat_batch(60, 50, 213, seed=41)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(clause_mem=4))
        check_models(orc, gb, res, b)
        ref = gb.solve(opts(clause_mem=1))
    assert res["info"]["clause_mem"] == 4 and res["info"]["lds_bytes"] == 0
    assert res["status"].tolist() == [truth(orc, b.instance(i)) for i in range(b.n_inst)]
    assert res["status"].tolist() == ref["status"].tolist()


def test_large_formula_beyond_lds(orc):
    """Formulas whose per-var state exceeds a workgroup's LDS (n = 4,000 and 12,000) are
    solved with the whole job state in HBM instead of being refused."""
    print("[solve] n=4000/12000 with the job state in HBM", flush=True)
    fs = [W.random_3sat_batch(1, 4000, 12000, seed=51).instance(0),
          W.random_3sat_batch(1, 12000, 36000, seed=52).instance(0),
          W.pigeonhole_formula(6, 5)]
    b = Batch.from_formulas(fs)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(max_ms=20000.0))
        check_models(orc, gb, res, b)
    assert res["info"]["clause_mem"] == 4
    assert res["status"].tolist() == [G.SAT, G.SAT, G.UNSAT]


def restart_points(first, factor, upto):
    """Conflict counts T_1, T_2, ... at which the geometric schedule (first, first*factor,
    ... conflicts apart, each step truncated to an integer in FP32 as the solver computes
    it) allows restart 1, 2, ...; the schedule of Configs.cuh:104-105 is first=100, 1.3."""
    pts, t, thr = [], 0, first
    while True:
        t += thr
        if t > upto:
            return pts
        pts.append(t)
        thr = int(np.float32(thr) * np.float32(factor))


def test_repeated_solves_on_reused_blocks(orc):
    """One batch solved twice in a row (the second call on the device blocks the first left
    for reuse), then a batch of another shape, then the first again: every call gives the
    oracle's verdicts and GPU-verified models, so no stale block contents reach a result."""
    print("[solve] repeated solves on reused device blocks", flush=True)
    b1 = W.random_3sat_batch(300, 50, 213, seed=21)
    b2 = W.random_3sat_batch(120, 60, 256, seed=22)
    exp1 = [truth(orc, b1.instance(i)) for i in range(b1.n_inst)]
    exp2 = [truth(orc, b2.instance(i)) for i in range(b2.n_inst)]
    with G.GpuBatch(b1) as g1, G.GpuBatch(b2) as g2:
        for gb, b, exp in ((g1, b1, exp1), (g1, b1, exp1), (g2, b2, exp2), (g1, b1, exp1)):
            res = gb.solve(opts())
            check_models(orc, gb, res, b)
            assert res["status"].tolist() == exp


@pytest.mark.parametrize("first,factor", [(100, 1.3), (40, 2.0)])
def test_restart_schedule(orc, first, factor):
    """k_solve's restart count follows the geometric schedule: restart k needs the sum of the
    first k intervals in conflicts, so R <= #{T_k <= C}; a SAT job ends right after a
    conflict-free propagation, where a due restart would already have run, so R falls short
    of that count only when conflicts ran on past a due restart point."""
    print(f"[solve] restart schedule first={first} factor={factor}", flush=True)
    b = W.random_3sat_batch(150, 110, 469, seed=77)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(restart_first=first, restart_factor=factor, learnt_cap=4096, max_ms=0.0))
        check_models(orc, gb, res, b)
    js = res["job_stats"]
    st = res["status"]
    assert (st != G.UNDEF).all()
    exact = sat = 0
    for i in range(b.n_inst):
        c, r = int(js[i, 1]), int(js[i, 5])
        bound = len(restart_points(first, factor, c))
        assert r <= bound, (i, c, r, bound)
        if st[i] == G.SAT:
            sat += 1
            assert r >= bound - 1, (i, c, r, bound)
            exact += r == bound
    assert js[:, 5].sum() >= 100
    assert sat >= 20 and exact >= 0.8 * sat, (exact, sat)


def test_structured_workload_full_size(orc):
    """The structured workload at its stated size: 2,000 4-colourings of ER(150, average degree
    4.2-4.8) and 200 PHP(9,8), one launch.  Exact verdicts: every colouring model verified on the
    GPU and by the oracle, any UNSAT colouring confirmed by the oracle, every PHP(9,8) UNSAT (the
    pigeonhole principle); learned-clause counts reported, never above the conflicts."""
    print("[solve] 2,000 colourings of ER(150) + 200 PHP(9,8)", flush=True)
    n_col, n_php = 2000, 200
    col = W.coloring_batch(n_col, seed=4)
    fs = [col.instance(i) for i in range(n_col)] + [W.pigeonhole_formula(9, 8)] * n_php
    b = Batch.from_formulas(fs)
    with G.GpuBatch(b) as gb:
        res = gb.solve(opts(max_ms=0.0))
        models = gb.models(res["model_rows"])
    st, js = res["status"], res["job_stats"]
    assert (st[n_col:] == G.UNSAT).all()
    assert (st[:n_col] != G.UNDEF).all()
    for i in range(n_col):
        if st[i] == G.SAT:
            assert res["verified"][i] == 1 and orc.verify_model(b.instance(i), models[i]), i
        else:
            assert truth(orc, b.instance(i)) == G.UNSAT, i
    learned, conflicts = js[:, 4], js[:, 1]
    assert (learned <= conflicts).all() and (learned >= 0).all()
    assert (learned[n_col:] > 1000).all()
    print(f"[solve] colourings SAT {(st[:n_col] == G.SAT).sum()}/{n_col}, learned mean "
          f"{learned[:n_col].mean():.2f}; PHP(9,8) UNSAT {n_php}/{n_php}, learned mean {learned[n_col:].mean():.0f}, "
          f"kernel {res['info']['kernel_ms']:.0f} ms", flush=True)
