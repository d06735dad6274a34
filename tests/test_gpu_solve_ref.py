"""Trajectory mode (gpsat_solve_ref): the reference's sequential search on the device gives
the oracle restatement's (oracle/refsolve.c) status, model and every counter -- decisions,
conflicts, learned clauses, iterations, restarts -- instance by instance, including which
instances the MAX_ITERATIONS cap turns into UNDEF."""
import glob
import os

import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd import workloads as W
from gpupsat_amd.formula import Batch, Formula, read_dimacs

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def check(orc, fs, max_iterations=0, dead=None):
    b = Batch.from_formulas(fs)
    with G.GpuBatch(b) as gb:
        res = gb.solve_ref(dead=None if dead is None else np.concatenate(dead), max_iterations=max_iterations)
        models = gb.models(res["model_rows"])
    for i, f in enumerate(fs):
        st, model, stats = orc.solve_ref(f, dead=None if dead is None else dead[i], max_iterations=max_iterations)
        assert res["status"][i] == st, (i, res["status"][i], st)
        np.testing.assert_array_equal(res["stats"][i], stats, err_msg=str(i))
        if st == G.SAT:
            np.testing.assert_array_equal(models[i], model, err_msg=str(i))
    return res


def test_3sat_n50_uncapped(orc):
    print("[solve_ref] 300 x n=50, no cap", flush=True)
    b = W.random_3sat_batch(300, 50, 213, seed=100)
    res = check(orc, [b.instance(i) for i in range(b.n_inst)])
    assert (res["status"] != 2).all()


def test_3sat_n100_capped(orc):
    """n=100 with MAX_ITERATIONS = 1000: most instances end UNDEF, the same ones as the
    oracle's, after the same number of conflicts."""
    print("[solve_ref] 60 x n=100, MAX_ITERATIONS 1000", flush=True)
    b = W.random_3sat_batch(60, 100, 426, seed=5)
    res = check(orc, [b.instance(i) for i in range(b.n_inst)], max_iterations=1000)
    assert (res["status"] == 2).any()


def test_fixtures_units_and_dead_vars(orc):
    print("[solve_ref] fixtures, unit clauses, preprocessed dead vars", flush=True)
    fs = [read_dimacs(p) for p in sorted(glob.glob(os.path.join(GOLDEN, "cnf", "*.cnf")))]
    fs += [W.pigeonhole_formula(4, 3), W.pigeonhole_formula(5, 4),
           Formula.from_clauses(3, [[1], [0, 3], [2, 5]]),
           Formula.from_clauses(2, [[], [1, 3]]),
           Formula.from_clauses(4, [])]
    check(orc, fs)
    # preprocessed formulas keep their solved vars out of the free list (dead)
    rng = np.random.default_rng(7)
    pf, dead = [], []
    for i in range(40):
        f = W.random_3sat_batch(1, 40, 160, seed=500 + i).instance(0)
        units = [[int(2 * v + rng.integers(0, 2))] for v in rng.choice(40, 2, replace=False)]
        g0 = Formula.from_clauses(40, f.clauses() + units)
        st, g, solved = orc.preprocess(g0)
        if st != G.UNDEF:
            continue
        d = np.zeros(40, dtype=np.uint8)
        d[np.asarray(solved) >> 1] = 1
        pf.append(g)
        dead.append(d)
    assert len(pf) >= 10
    check(orc, pf, dead=dead)


def test_structured_and_mixed(orc):
    """Colourings, PHP(7,6) (thousands of conflicts: the 20-slot learned repository wraps
    around many times), and mixed k-SAT under a cap."""
    print("[solve_ref] colouring, PHP(7,6), mixed k-SAT", flush=True)
    fs = [W.coloring_batch(6, n_nodes=30, seed=1).instance(i) for i in range(6)]
    fs.append(W.pigeonhole_formula(7, 6))
    b, _ = W.mixed_ksat_batch(8, seed=2)
    fs += [b.instance(i) for i in range(8) if b.instance(i).n_vars <= 200]
    check(orc, fs, max_iterations=400)
