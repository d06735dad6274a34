"""Parity of the HIP clause-evaluation and BCP-fixpoint kernels (through the C ABI)
with the CPU oracle: bit-exact counts, closure rows, conflict flags and props."""
import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd import workloads as W
from gpupsat_amd.formula import Batch, Formula, pack_rows

pytestmark = pytest.mark.gpu


def check_batch(orc, b, rows, A):
    with G.GpuBatch(b) as gb:
        counts = gb.clause_eval(rows, A)
        out, conflict, props = gb.bcp_fixpoint(rows, A)
    exp_counts = orc.eval_batch(b, rows, A)
    np.testing.assert_array_equal(counts, exp_counts)
    e_out, e_conf, e_props = orc.bcp_batch(b, rows, A)
    np.testing.assert_array_equal(conflict, e_conf)
    np.testing.assert_array_equal(props, e_props)
    np.testing.assert_array_equal(out, e_out)
    return counts, conflict, props


@pytest.mark.parametrize("p", [0.02, 0.05, 0.1, 0.3])
def test_config2_shape_subset(orc, p):
    print(f"[eval/bcp] n=200 m=852 N=96 A=64 p={p}", flush=True)
    b = W.random_3sat_batch(96, 200, 852, seed=2)
    rows = W.random_assignments(b, 64, p, seed=int(p * 1000))
    counts, conflict, props = check_batch(orc, b, rows, 64)
    if p <= 0.05:
        assert conflict.mean() < 0.5 and props.sum() > 0


@pytest.mark.parametrize("A", [1, 7, 33, 64])
def test_mixed_sizes(orc, A):
    print(f"[eval/bcp] mixed k-SAT N=24 A={A}", flush=True)
    b, _ = W.mixed_ksat_batch(24, seed=A)
    rows = W.random_assignments(b, A, 0.04, seed=A)
    check_batch(orc, b, rows, A)


def test_structured_and_edge_cases(orc):
    print("[eval/bcp] structured + edge cases", flush=True)
    fs = [
        W.pigeonhole_formula(4, 3),
        W.coloring_batch(1, n_nodes=40, seed=3).instance(0),
        Formula.from_clauses(3, []),                              # no clauses
        Formula.from_clauses(3, [[], [1, 3]]),                    # empty clause
        Formula.from_clauses(4, [[1], [0, 3], [2, 5], [4, 7]]),   # unit chain
        Formula.from_clauses(3, [[1, 1, 3], [0, 1, 5], [2, 2]]),  # duplicates, tautology
        Formula.from_clauses(1, [[1], [0]]),
        Formula.from_clauses(40, [list(range(0, 80, 2))]),        # one long clause
    ]
    b = Batch.from_formulas(fs)
    A = 64
    rng = np.random.default_rng(0)
    rows = []
    for f in fs:
        n = f.n_vars
        t = rng.random((A, n)) < 0.15
        fl = (rng.random((A, n)) < 0.15) & ~t
        t[5] |= fl[5]                                             # row 5: contradictory codes
        rows.append(pack_rows(t, fl).reshape(-1))
    rows = np.concatenate(rows)
    counts, conflict, _ = check_batch(orc, b, rows, A)
    assert conflict[6].all()                                      # x0 and ~x0 both units


def test_reference_fixtures_full_models(orc):
    import glob, os
    from conftest import GOLDEN
    from gpupsat_amd.formula import read_dimacs
    fs = [read_dimacs(p) for p in sorted(glob.glob(os.path.join(GOLDEN, "cnf", "*.cnf")))]
    b = Batch.from_formulas(fs)
    A = 64
    rows = []
    for f in fs:
        # the first min(64, 2^n) full assignments
        bits = (np.arange(A)[:, None] >> np.arange(f.n_vars)[None, :]) & 1
        rows.append(pack_rows(bits == 1, bits == 0).reshape(-1))
    rows = np.concatenate(rows)
    counts, _, _ = check_batch(orc, b, rows, A)


def test_config2_full_size_properties(orc):
    """10,000 x (n=200, m=852), A=64, p=0.3: oracle parity on a 256-instance sample,
    size-independent properties everywhere."""
    print("[eval/bcp] config-2 full size", flush=True)
    N, n, m, A = 10_000, 200, 852, 64
    b = W.random_3sat_batch(N, n, m, seed=0)
    rows = W.random_assignments(b, A, 0.3, seed=0)
    with G.GpuBatch(b) as gb:
        counts = gb.clause_eval(rows, A)
        out, conflict, props = gb.bcp_fixpoint(rows, A)
        # idempotence: the fixpoint of a closure is itself
        out2, conflict2, props2 = gb.bcp_fixpoint(out, A)
        counts2 = gb.clause_eval(out, A)
    assert (counts.sum(-1) <= m).all()
    assert (conflict2 >= conflict).all()
    ok = conflict == 0
    assert (props2[ok] == 0).all()
    assert (counts2[ok][:, 1] == 0).all() and (counts2[ok][:, 2] == 0).all()
    # closure rows contain the trail
    assert ((out & rows) == rows).all()
    sample = np.random.default_rng(1).choice(N, 256, replace=False)
    sb = b.subset(sample)
    base = b.row_base(A)
    srows = np.concatenate([rows[base[i]:base[i + 1]] for i in sample])
    np.testing.assert_array_equal(counts[sample], orc.eval_batch(sb, srows, A))
    e_out, e_conf, e_props = orc.bcp_batch(sb, srows, A)
    np.testing.assert_array_equal(conflict[sample], e_conf)
    np.testing.assert_array_equal(props[sample], e_props)
    np.testing.assert_array_equal(np.concatenate([out[base[i]:base[i + 1]] for i in sample]), e_out)
