"""Device preprocessing of a whole batch (gpsat_preprocess_batch) against the oracle's
restatement of the reference's two removers (FormulaData.cu:82-105): identical status,
identical remaining clauses in the same order, identical solved-literal set."""
import glob
import os

import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd.formula import Batch, Formula, read_dimacs, unpack_rows

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def random_messy_formula(rng, n, m):
    """Clauses of length 0..5 with repeated literals, tautologies, unit and empty clauses."""
    cl = []
    for _ in range(m):
        r = rng.random()
        if r < 0.12:
            k = 1
        elif r < 0.14:
            k = 0
        else:
            k = int(rng.integers(2, 6))
        c = [int(2 * rng.integers(0, n) + rng.integers(0, 2)) for _ in range(k)]
        if k >= 2 and rng.random() < 0.15:
            c.append(c[0])              # repeated literal
        if k >= 2 and rng.random() < 0.05:
            c.append(c[1] ^ 1)          # tautology
        cl.append(c)
    return Formula.from_clauses(n, cl)


def solved_set(row, n):
    t, f = unpack_rows(row[None, :], n)
    return set((2 * np.nonzero(t[0])[0] + 1).tolist()) | set((2 * np.nonzero(f[0])[0]).tolist())


def check(orc, fs):
    b = Batch.from_formulas(fs)
    with G.GpuBatch(b) as gb:
        st, solved, rest = gb.preprocess()
    base = b.row_base(1)
    for i, f in enumerate(fs):
        est, eg, esolved = orc.preprocess(f)
        assert st[i] == est, (i, st[i], est)
        if est == G.UNSAT:
            continue
        got = solved_set(solved[base[i]:base[i + 1]], f.n_vars)
        assert got == set(esolved.tolist()), i
        g = rest.instance(i)
        if est == G.UNDEF:
            np.testing.assert_array_equal(g.lits, eg.lits)
            np.testing.assert_array_equal(g.clause_off, eg.clause_off)
        else:
            assert g.n_clauses == 0
    return st


def test_random_messy_formulas(orc):
    print("[preprocess] 400 random formulas with units, repeats, tautologies", flush=True)
    rng = np.random.default_rng(31)
    fs = [random_messy_formula(rng, int(rng.integers(2, 40)), int(rng.integers(0, 120))) for _ in range(400)]
    st = check(orc, fs)
    # the mix exercises every outcome
    assert (st == 0).any() and (st == 1).any() and (st == 2).any()


def test_edge_cases(orc):
    print("[preprocess] edge cases", flush=True)
    fs = [
        Formula.from_clauses(3, []),                         # no clauses: SAT
        Formula.from_clauses(3, [[]]),                       # empty clause, no unit: kept, UNDEF
        Formula.from_clauses(3, [[], [1]]),                  # empty clause and a unit: UNSAT
        Formula.from_clauses(2, [[1], [0]]),                 # complementary units: UNSAT
        Formula.from_clauses(2, [[1], [1], [1, 3]]),         # repeated unit clause: SAT
        Formula.from_clauses(3, [[0, 1], [2, 3, 3]]),        # tautology dropped, repeat collapsed
        Formula.from_clauses(4, [[1], [0, 3], [2, 5], [4, 7, 6]]),   # unit chain to a tautology-free end
        Formula.from_clauses(200, [[1]] + [[2 * i, 2 * i + 3] for i in range(199)] + [[398, 0, 5]]),
    ]
    st = check(orc, fs)
    assert st.tolist()[:5] == [G.SAT, G.UNDEF, G.UNSAT, G.UNSAT, G.SAT]


def test_reference_fixtures(orc):
    print("[preprocess] reference fixtures", flush=True)
    paths = sorted(glob.glob(os.path.join(GOLDEN, "cnf", "*.cnf")))
    fs = [read_dimacs(p) for p in paths]
    st = check(orc, fs)
    names = [os.path.basename(p) for p in paths]
    if "unsat_5v_chain.cnf" in names:
        assert st[names.index("unsat_5v_chain.cnf")] == G.UNSAT
