"""Pins the CPU oracle before it is trusted as the checker.

Anchors, in order of strength:
  * the reference's own known-answer fixtures tests/cnf/*.expected (copied as data
    into tests/golden/cnf/);
  * theorems (pigeonhole PHP(p+1, p) is UNSAT) and brute force over all 2^n
    assignments on small n (semantic checks: an implied literal holds in every model,
    a UP conflict means no model);
  * agreement of three solvers that share no code (reference-scheme CDCL, textbook
    CDCL, DPLL) plus brute force, with every SAT model checked clause by clause;
  * closed forms (restart schedule, cube enumeration as a positive-first DFS).
"""
import glob
import itertools
import os

import numpy as np
import pytest

from gpupsat_amd import workloads as W
from gpupsat_amd.formula import Formula, pack_rows, read_dimacs, unpack_rows

from conftest import GOLDEN

CNFS = sorted(glob.glob(os.path.join(GOLDEN, "cnf", "*.cnf")))


def expected(path):
    return {"SAT": 0, "UNSAT": 1}[open(path[:-4] + ".expected").read().strip()]


def models_of(f):
    """All models of f by brute force (small n), as bool arrays."""
    n = f.n_vars
    out = []
    cl = f.clauses()
    for bits in itertools.product([False, True], repeat=n):
        if all(any(bits[l >> 1] == bool(l & 1) for l in c) for c in cl):
            out.append(np.array(bits))
    return out


@pytest.mark.parametrize("path", CNFS, ids=[os.path.basename(p) for p in CNFS])
def test_reference_fixtures_all_solvers(orc, path):
    f = read_dimacs(path)
    exp = expected(path)
    st, g, solved = orc.preprocess(f)
    ref = st if st != orc.UNDEF else orc.solve_ref(g)[0]
    assert ref == exp
    assert orc.solve_dpll(f)[0] == exp
    assert orc.solve_cdcl(f)[0] == exp
    assert orc.solve_brute(f)[0] == exp


def test_reference_fixture_unsat_in_preprocessing(orc):
    # unsat_5v_chain: the unit 1 and the chain force x5 while the unit -5 forbids it;
    # the reference answers it in preprocessing (main.cu:154-163)
    f = read_dimacs(os.path.join(GOLDEN, "cnf", "unsat_5v_chain.cnf"))
    st, g, _ = orc.preprocess(f)
    assert st == orc.UNSAT and g is None


def test_clause_status_known_answers(orc):
    n = 3
    row = lambda t, fl: pack_rows(np.array(t), np.array(fl))
    a, na, b, nb, c = 1, 0, 3, 2, 5          # x0, ~x0, x1, ~x1, x2
    empty = row([False] * n, [False] * n)
    x0f_x1f = row([False] * n, [True, True, False])
    x0t = row([True, False, False], [False, True, True])
    assert orc.clause_status([a, b, c], empty) == (orc.UNDEF, -1)
    assert orc.clause_status([a, b, c], x0f_x1f) == (orc.UNDEF, c)       # unit on x2
    assert orc.clause_status([a, b], x0f_x1f) == (orc.UNSAT, -1)
    assert orc.clause_status([na, b], x0f_x1f) == (orc.SAT, -1)
    assert orc.clause_status([a, b, c], x0t) == (orc.SAT, -1)
    assert orc.clause_status([c], empty) == (orc.UNDEF, c)                # single literal is unit
    assert orc.clause_status([], empty) == (orc.UNSAT, -1)                # empty clause
    # occurrences are counted: [x0, x0, x2] with x0 false has 2 false of 3 -> unit x2
    assert orc.clause_status([a, a, c], x0f_x1f) == (orc.UNDEF, c)
    # [x2, x2, x0] with x0 false: one false occurrence of three -> no unit
    assert orc.clause_status([c, c, a], x0f_x1f) == (orc.UNDEF, -1)


def test_eval_batch_semantics(orc):
    rng = np.random.default_rng(5)
    b = W.random_3sat_batch(20, 12, 50, seed=7)
    A = 9
    for trial in range(3):
        rows = W.random_assignments(b, A, p=[0.0, 0.5, 1.0][trial], seed=trial)
        counts = orc.eval_batch(b, rows, A)
        base = b.row_base(A)
        for i in range(b.n_inst):
            f = b.instance(i)
            n = f.n_vars
            Wd = (2 * n + 31) // 32
            for a in range(A):
                r = rows[base[i] + a * Wd: base[i] + (a + 1) * Wd]
                t, fl = unpack_rows(r[None, :], n)
                t, fl = t[0], fl[0]
                sat = unsat = unit = 0
                for c in f.clauses():
                    tr = [t[l >> 1] if l & 1 else fl[l >> 1] for l in c]
                    fa = [fl[l >> 1] if l & 1 else t[l >> 1] for l in c]
                    if any(tr):
                        sat += 1
                    elif all(fa):
                        unsat += 1
                    elif sum(fa) == len(c) - 1:
                        unit += 1
                assert tuple(counts[i, a]) == (sat, unsat, unit)
                if trial == 0:
                    assert tuple(counts[i, a]) == (0, 0, 0)
                if trial == 2:
                    assert counts[i, a, 2] == 0 and counts[i, a, 0] + counts[i, a, 1] == f.n_clauses


def test_up_closure_against_brute_force_models(orc):
    """Every literal UP implies holds in every model of F and the trail; a UP conflict
    means F and the trail have no model (soundness); and on these 2-CNF-free random
    3-SAT samples every var fixed across all models that UP does not reach is allowed,
    so only soundness plus the properties below are asserted."""
    rng = np.random.default_rng(11)
    for inst in range(30):
        n, m = 10, 38
        f = W.random_3sat_batch(1, n, m, seed=100 + inst).instance(0)
        models = models_of(f)
        for _ in range(8):
            assigned = rng.random(n) < 0.3
            pol = rng.random(n) < 0.5
            row = pack_rows(assigned & pol, assigned & ~pol)
            conflict, out, props = orc.up_closure(f, row)
            ms = [mm for mm in models if all(mm[v] == pol[v] for v in range(n) if assigned[v])]
            if conflict:
                assert not ms
                assert props == 0 and (out == row).all()
                continue
            t, fl = unpack_rows(out[None], n)
            t, fl = t[0], fl[0]
            assert not (t & fl).any()
            assert ((t | fl) & assigned).sum() == assigned.sum()
            assert props == (t | fl).sum() - assigned.sum()
            for mm in ms:
                assert all(mm[v] for v in range(n) if t[v]) and not any(mm[v] for v in range(n) if fl[v])
            # fixpoint: no clause left unit or all-false under the closure
            for c in f.clauses():
                st, u = orc.clause_status(c, out)
                assert st != orc.UNSAT and u == -1
            # idempotent
            c2, out2, p2 = orc.up_closure(f, out)
            assert c2 == 0 and p2 == 0 and (out2 == out).all()


def test_up_closure_watched_literal_property(orc):
    """WatchedClausesListTester.cu:22-109: falsifying every literal of a clause but the
    last implies the last literal, unless that already conflicts."""
    b = W.random_3sat_batch(10, 30, 120, seed=3)
    for i in range(b.n_inst):
        f = b.instance(i)
        for c in f.clauses()[:30]:
            n = f.n_vars
            t = np.zeros(n, bool)
            fl = np.zeros(n, bool)
            for l in c[:-1]:
                (fl if l & 1 else t)[l >> 1] = True
            row = pack_rows(t, fl)
            conflict, out, _ = orc.up_closure(f, row)
            if not conflict:
                last = c[-1]
                assert (out[last >> 5] >> (last & 31)) & 1


def test_up_closure_propagation_reason_property(orc):
    """PropagationTester.cu:24-112: every implied literal has a clause whose other
    literals are all false under the closure."""
    rng = np.random.default_rng(2)
    b = W.random_3sat_batch(5, 60, 255, seed=9)
    for i in range(b.n_inst):
        f = b.instance(i)
        n = f.n_vars
        for _ in range(10):
            assigned = rng.random(n) < 0.1
            pol = rng.random(n) < 0.5
            row = pack_rows(assigned & pol, assigned & ~pol)
            conflict, out, props = orc.up_closure(f, row)
            if conflict:
                continue
            def true_(l):
                return (out[l >> 5] >> (l & 31)) & 1
            implied = [l for l in range(2 * n) if true_(l) and not (row[l >> 5] >> (l & 31)) & 1]
            assert len(implied) == props
            for l in implied:
                assert any(l in c and all(true_(q ^ 1) for q in c if q != l) for c in f.clauses())


def test_contradictory_row_is_conflict(orc):
    f = Formula.from_clauses(2, [[1, 3]])
    row = pack_rows(np.array([True, False]), np.array([True, False]))
    conflict, out, props = orc.up_closure(f, row)
    assert conflict == 1 and props == 0


def test_preprocess_equisatisfiable_and_clean(orc):
    rng = np.random.default_rng(4)
    for t in range(60):
        n = 8
        cl = []
        for _ in range(rng.integers(4, 24)):
            k = int(rng.integers(1, 4))
            cl.append([int(2 * rng.integers(0, n) + rng.integers(0, 2)) for _ in range(k)])
        f = Formula.from_clauses(n, cl)
        st, g, solved = orc.preprocess(f)
        truth = orc.solve_brute(f)[0]
        if st != orc.UNDEF:
            assert st == truth
            continue
        solved_vars = {l >> 1 for l in solved}
        assert len(solved_vars) == len(solved)
        for c in g.clauses():
            assert len(c) >= 2 and len(set(c)) == len(c)
            assert not any((l >> 1) in solved_vars for l in c)
            assert not any((l ^ 1) in c for l in c)
        # g plus the solved literals is equisatisfiable with f
        h = Formula.from_clauses(n, g.clauses() + [[l] for l in solved])
        assert orc.solve_brute(h)[0] == truth


def test_preprocess_duplicate_unit_quirk(orc):
    # UnaryClausesRemover::add refuses a literal already solved, so a duplicate unit
    # clause survives process_unary_clauses and is erased in the first propagation pass
    f = Formula.from_clauses(3, [[1], [1], [0, 3, 5], [2, 4]])
    st, g, solved = orc.preprocess(f)
    assert st == orc.UNDEF and list(solved) == [1]
    assert g.clauses() == [[3, 5], [2, 4]]


def test_restart_schedule_closed_form(orc):
    # t_{k+1} = (int)(t_k * 1.3f): 100*1.3f = 130.0000048 -> 130, 169*1.3f = 219.7 -> 219
    assert orc.restart_schedule(100, 1.3, 8).tolist() == [100, 130, 169, 219, 284, 369, 479, 622]


def test_cube_enumeration_is_positive_first_dfs(orc):
    vars_ = [7, 2, 9]
    got = orc.enumerate_cubes(vars_).tolist()
    dfs = []

    def rec(prefix, j):
        if j == len(vars_):
            dfs.append(prefix)
            return
        rec(prefix + [2 * vars_[j] + 1], j + 1)
        rec(prefix + [2 * vars_[j]], j + 1)

    rec([], 0)
    assert got == dfs


def test_cube_var_choice(orc):
    # score(v) = v * occurrences(v) (VariableChooser.cu:30); depth for T=B=32:
    # floor(log2(32*32*10)) + 1 = 14 (JobChooser.cu:115-125)
    f = W.random_3sat_batch(1, 200, 852, seed=1).instance(0)
    vs = orc.choose_cube_vars(f, 0, 32, 32, 0)
    assert len(vs) == 14
    occ = np.bincount(f.lits >> 1, minlength=f.n_vars)
    score = occ * np.arange(f.n_vars)
    order = sorted(range(f.n_vars), key=lambda v: (-score[v], v))
    assert vs.tolist() == order[:14]
    assert len(orc.choose_cube_vars(f, 0, 1, 1, 0)) == 4          # floor(log2 10) + 1
    assert len(orc.choose_cube_vars(f, 0, 1, 1, 1)) == 7          # uniform: 7 vars
    assert len(orc.choose_cube_vars(f, 0, 10_000, 10_000, 0)) == 15   # MAX_VARS cap
    tiny = Formula.from_clauses(4, [[1, 3], [5, 7]])
    assert len(orc.choose_cube_vars(tiny, 0, 32, 32, 0)) == 3     # 2^(4-2) = 4 -> log2 4 + 1


def test_3sat_n50_solvers_agree(orc):
    b = W.random_3sat_batch(100, 50, 213, seed=0)
    n_sat = 0
    for i in range(b.n_inst):
        f = b.instance(i)
        st, g, solved = orc.preprocess(f)
        r = orc.solve_ref(g)[0] if st == orc.UNDEF else st
        c, model, _ = orc.solve_cdcl(f)
        d, dm = orc.solve_dpll(f)
        assert r == c == d
        if c == orc.SAT:
            n_sat += 1
            assert orc.verify_model(f, model) and orc.verify_model(f, dm)
    assert 20 < n_sat < 95


def test_ref_solver_models_verify(orc):
    b = W.random_3sat_batch(40, 40, 160, seed=8)
    for i in range(b.n_inst):
        f = b.instance(i)
        st, model, _ = orc.solve_ref(f)
        assert st == orc.solve_brute(f)[0] if f.n_vars <= 26 else True
        if st == orc.SAT:
            assert orc.verify_model(f, model)


def test_pigeonhole_unsat_theorem(orc):
    for p in (3, 4, 5, 6):
        f = W.pigeonhole_formula(p + 1, p)
        assert orc.solve_ref(f)[0] == orc.UNSAT
        assert orc.solve_cdcl(f)[0] == orc.UNSAT
    f = W.pigeonhole_formula(5, 5)
    st, model, _ = orc.solve_ref(f)
    assert st == orc.SAT and orc.verify_model(f, model)


def test_coloring_sat_models(orc):
    b = W.coloring_batch(3, n_nodes=30, seed=1)
    for i in range(b.n_inst):
        f = b.instance(i)
        st, model, _ = orc.solve_ref(f)
        c = orc.solve_cdcl(f)[0]
        assert st == c
        if st == orc.SAT:
            assert orc.verify_model(f, model)


def test_iteration_cap_gives_undef(orc):
    f = W.pigeonhole_formula(7, 6)
    assert orc.solve_ref(f, max_iterations=5)[0] == orc.UNDEF
