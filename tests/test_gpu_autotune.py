"""The launch-shape autotuner: it picks a shape from timed trials, and the verdicts and
fixpoints under the chosen shape are the oracle's."""
import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_autotune_eval_keeps_parity(orc):
    print("[autotune] eval/BCP shapes on 512 x n=200 3-SAT", flush=True)
    b = W.random_3sat_batch(512, 200, 852, seed=21)
    A = 64
    rows = W.random_assignments(b, A, 0.1, seed=22)
    keys = ("eval_fast3", "eval_pkk", "eval_stage_arena", "eval_waves_per_block", "eval_fast3_small64",
            "eval_pkk_waves", "eval_pkk_compact")
    saved = {k: G.tune_get(k) for k in keys}
    try:
        with G.GpuBatch(b) as gb:
            rb = G.DevBuf.from_host(rows)
            rep = gb.autotune_eval(rb, A, reps=3)
            # 9 shapes of the 3-literal kernel and 6 of the CSR kernel, each timed
            assert "chosen:" in rep and rep.count("kernel=3-literal") >= 9 and rep.count("kernel=CSR") >= 6
            counts = gb.clause_eval(rows, A)
            out, conflict, props = gb.bcp_fixpoint(rows, A)
        np.testing.assert_array_equal(counts, orc.eval_batch(b, rows, A))
        e_out, e_conf, e_props = orc.bcp_batch(b, rows, A)
        np.testing.assert_array_equal(conflict, e_conf)
        np.testing.assert_array_equal(props, e_props)
        np.testing.assert_array_equal(out, e_out)
    finally:
        for k, v in saved.items():
            G.tune_set(k, v)


def test_autotune_eval_packed_shapes(orc):
    """On a mixed k-SAT batch the eval autotuner times the packed kernel's shapes (waves per
    instance 1-4, handoff to 32-bit masks off and on) beside the CSR kernel's, and the
    fixpoints under the chosen shape are the oracle's."""
    print("[autotune] eval/BCP shapes on 48 mixed k-SAT", flush=True)
    mixed, ks = W.mixed_ksat_batch(4096, seed=3)
    idx = [int(i) for k in (3, 4, 5) for i in np.nonzero(ks == k)[0][:16]]
    from gpupsat_amd.formula import Batch
    b = Batch.from_formulas([mixed.instance(i) for i in idx])
    A = 64
    rows = W.random_assignments(b, A, 0.3, seed=23)
    keys = ("eval_fast3", "eval_pkk", "eval_stage_arena", "eval_waves_per_block", "eval_fast3_small64",
            "eval_pkk_waves", "eval_pkk_compact")
    saved = {k: G.tune_get(k) for k in keys}
    try:
        with G.GpuBatch(b) as gb:
            rb = G.DevBuf.from_host(rows)
            rep = gb.autotune_eval(rb, A, reps=2)
            assert "chosen:" in rep and rep.count("kernel=packed") >= 8 and rep.count("kernel=CSR") >= 6
            counts = gb.clause_eval(rows, A)
            out, conflict, props = gb.bcp_fixpoint(rows, A)
        np.testing.assert_array_equal(counts, orc.eval_batch(b, rows, A))
        e_out, e_conf, e_props = orc.bcp_batch(b, rows, A)
        np.testing.assert_array_equal(conflict, e_conf)
        np.testing.assert_array_equal(props, e_props)
        np.testing.assert_array_equal(out, e_out)
    finally:
        for k, v in saved.items():
            G.tune_set(k, v)


def test_autotune_solve_same_verdicts(orc):
    print("[autotune] solve shapes on mixed k-SAT with cubes", flush=True)
    b, _ = W.mixed_ksat_batch(16, seed=5)
    vs = [G.cube_vars(b.instance(i), threads=1, blocks=1) for i in range(b.n_inst)]   # depth 4
    jobs = G.cube_jobs(b, vs)
    base = G.SolveOpts.default(max_conflicts=3000, job_order=1)
    with G.GpuBatch(b) as gb:
        best, rep = gb.autotune_solve(base, jobs=jobs, trial_ms=150.0)
        assert "chosen:" in rep
        # the returned shape is explicit and is one a trial ran
        assert best.waves > 0 and best.chunk > 0 and best.clause_mem in (1, 2, 3, 4)
        ran = [ln for ln in rep.splitlines() if ln.startswith("  mem=") and "finished" in ln]
        assert any(f"waves={best.waves} chunk={best.chunk}:" in ln for ln in ran), rep
        assert best.max_conflicts == 3000 and best.job_order == 1
        r1 = gb.solve(base, jobs=jobs)
        r2 = gb.solve(best, jobs=jobs)
        models = gb.models(r2["model_rows"])
    # the full run launched the shape the autotuner measured
    assert r2["info"]["waves"] == best.waves and r2["info"]["chunk"] == best.chunk
    assert r2["info"]["clause_mem"] == best.clause_mem
    # budgets are per job and deterministic, so every decided verdict agrees
    for x, y in zip(r1["status"].tolist(), r2["status"].tolist()):
        assert x == y or G.UNDEF in (x, y)
    for i, x in enumerate(r2["status"].tolist()):
        if x == G.SAT:
            assert orc.verify_model(b.instance(i), models[i])
        elif x == G.UNSAT:
            # strict: the oracle's own CDCL must prove it (an oracle UNDEF fails the test)
            st, _, _ = orc.solve_cdcl(b.instance(i), max_conflicts=2_000_000)
            assert st == G.UNSAT, i


def test_autotune_solve_no_cubes_runs_full_shape(orc):
    """One job per instance: the wave count the tuner returns is the one the full run
    uses, even when the sampled instances are fewer than the candidate wavefronts."""
    print("[autotune] solve shapes, one job per instance", flush=True)
    b = W.random_3sat_batch(3000, 50, 213, seed=8)
    with G.GpuBatch(b) as gb:
        best, rep = gb.autotune_solve(G.SolveOpts.default(), sample_jobs=512, trial_ms=60.0)
        res = gb.solve(best, want_models=True)
        check = gb.models(res["model_rows"])
    assert res["info"]["waves"] == best.waves and res["info"]["chunk"] == best.chunk, rep
    assert "not ranked" not in rep, rep
    idx = range(0, b.n_inst, 150)
    for i in idx:
        st = res["status"][i]
        exp, _, _ = orc.solve_cdcl(b.instance(i))
        assert st == exp, i
        if st == G.SAT:
            assert orc.verify_model(b.instance(i), check[i])
