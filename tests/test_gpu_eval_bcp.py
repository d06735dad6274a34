"""Parity of the HIP clause-evaluation and BCP-fixpoint kernels (through the C ABI)
with the CPU oracle: bit-exact counts, closure rows, conflict flags and props."""
import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd import workloads as W
from gpupsat_amd.formula import Batch, Formula, pack_rows

pytestmark = pytest.mark.gpu


def fused(gb, rows, A):
    """One eval+BCP launch on device buffers."""
    rb = G.DevBuf.from_host(np.ascontiguousarray(rows, dtype=np.uint32))
    nI = gb.n_inst
    cb, ob = G.DevBuf(nI * A * 12), G.DevBuf(rows.size * 4)
    fb, pb = G.DevBuf(nI * A * 4), G.DevBuf(nI * A * 4)
    gb.eval_bcp_dev(rb, A, cb, ob, fb, pb)
    gb.sync()
    return (cb.to_host(np.int32, nI * A * 3).reshape(nI, A, 3), ob.to_host(np.uint32, rows.size),
            fb.to_host(np.int32, nI * A).reshape(nI, A), pb.to_host(np.int32, nI * A).reshape(nI, A))


def check_batch(orc, b, rows, A):
    with G.GpuBatch(b) as gb:
        counts = gb.clause_eval(rows, A)
        out, conflict, props = gb.bcp_fixpoint(rows, A)
        f_counts, f_out, f_conflict, f_props = fused(gb, rows, A)
    np.testing.assert_array_equal(f_counts, counts)
    np.testing.assert_array_equal(f_out, out)
    np.testing.assert_array_equal(f_conflict, conflict)
    np.testing.assert_array_equal(f_props, props)
    exp_counts = orc.eval_batch(b, rows, A)
    np.testing.assert_array_equal(counts, exp_counts)
    e_out, e_conf, e_props = orc.bcp_batch(b, rows, A)
    np.testing.assert_array_equal(conflict, e_conf)
    np.testing.assert_array_equal(props, e_props)
    np.testing.assert_array_equal(out, e_out)
    return counts, conflict, props


@pytest.mark.parametrize("fast3", [1, 0])
@pytest.mark.parametrize("p", [0.02, 0.05, 0.1, 0.2, 0.25, 0.3])
def test_3sat_n200_subset(orc, p, fast3):
    """n=200 m=852 3-SAT, through the register-resident kernel (fast3=1) and the
    general one (fast3=0).  In the fused launch the rows still open after the evaluation
    number above 32 for p <= 0.1, 21-41 for p = 0.2 (72 of the 96 instances between 9 and
    32), and mostly at most 8 for p = 0.3: all three ways the register-resident kernel
    runs its BCP passes (64-bit, compacted by transposes, compacted bit by bit)."""
    print(f"[eval/bcp] n=200 m=852 N=96 A=64 p={p} fast3={fast3}", flush=True)
    b = W.random_3sat_batch(96, 200, 852, seed=2)
    rows = W.random_assignments(b, 64, p, seed=int(p * 1000))
    with G.tuned(eval_fast3=fast3):
        counts, conflict, props = check_batch(orc, b, rows, 64)
    if p <= 0.05:
        assert conflict.mean() < 0.5 and props.sum() > 0


@pytest.mark.parametrize("A", [1, 33, 64])
def test_fast3_limits(orc, A):
    """Batches at the register-resident kernel's limits: n=256 (16 words per row),
    m=1024 (16 full clause groups), a clause-free instance, repeated and complementary
    literals inside a clause, unit chains that cross groups."""
    print(f"[eval/bcp] 3-literal kernel limits A={A}", flush=True)
    from gpupsat_amd.formula import Formula
    fs = [W.random_3sat_batch(1, 256, 1024, seed=9).instance(0),
          W.random_3sat_batch(1, 100, 430, seed=10).instance(0),
          Formula.from_clauses(5, []),
          Formula.from_clauses(4, [[1, 1, 3], [0, 1, 5], [2, 2, 6], [7, 7, 7]]),
          Formula.from_clauses(1, [[1, 1, 1]]),
          W.random_3sat_batch(1, 256, 700, seed=11).instance(0)]
    # an implication chain x0 -> x1 -> ... -> x199 through clauses (~x_i | x_{i+1} | z),
    # z = x200 false in every row; the clauses are listed last-link first so each
    # Gauss-Seidel pass over the groups advances the chain by one group at most
    chain = [[2 * i, 2 * i + 3, 401] for i in range(198, -1, -1)]
    fs.append(Formula.from_clauses(201, chain))
    b = Batch.from_formulas(fs)
    rng = np.random.default_rng(A)
    rows = []
    for k, f in enumerate(fs):
        n = f.n_vars
        t = rng.random((A, n)) < 0.02
        fl = (rng.random((A, n)) < 0.02) & ~t
        if k == len(fs) - 1:
            t[:] = False
            fl[:] = False
            t[:, 0] = True
            fl[:, 200] = True
        rows.append(pack_rows(t, fl).reshape(-1))
    rows = np.concatenate(rows)
    for fast3 in (1, 0):
        with G.tuned(eval_fast3=fast3):
            _, conflict, props = check_batch(orc, b, rows, A)
    assert (props[-1] == 199).all() and (conflict[-1] == 0).all()


@pytest.mark.parametrize("A,k", [(64, 20), (40, 12), (64, 33), (48, 5)])
def test_bcp_handoff(orc, A, k):
    """A 200-link implication chain started in k of the A rows (x0 true there, the other rows
    leave x0 open), beside a random instance at p = 0.1.  More than 32 rows are open, so the
    passes start on the 64-bit table; the rows without the chain stop moving after the first
    pass, and with 8 to 32 chain rows the passes continue on a compacted table of those rows
    (k = 33: all 64-bit; k = 5: too few moving rows to hand off)."""
    print(f"[eval/bcp] BCP handoff A={A} k={k}", flush=True)
    from gpupsat_amd.formula import Formula
    chain = [[2 * i, 2 * i + 3, 401] for i in range(198, -1, -1)]
    fs = [Formula.from_clauses(201, chain), W.random_3sat_batch(1, 200, 852, seed=21).instance(0)]
    b = Batch.from_formulas(fs)
    t = np.zeros((A, 201), dtype=bool)
    fl = np.zeros((A, 201), dtype=bool)
    fl[:, 200] = True
    t[:k, 0] = True
    r0 = pack_rows(t, fl).reshape(-1)
    rng = np.random.default_rng(k)
    t1 = rng.random((A, 200)) < 0.05
    f1 = (rng.random((A, 200)) < 0.05) & ~t1
    rows = np.concatenate([r0, pack_rows(t1, f1).reshape(-1)])
    with G.tuned(eval_fast3=1):
        _, conflict, props = check_batch(orc, b, rows, A)
    assert (props[0, :k] == 199).all() and (props[0, k:] == 0).all() and (conflict[0] == 0).all()


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


def test_3sat_n200_full_batch_properties(orc):
    """10,000 x (n=200, m=852), A=64, p=0.3: oracle parity on a 256-instance sample,
    size-independent properties everywhere."""
    print("[eval/bcp] 10,000 x n=200 full batch", flush=True)
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


@pytest.mark.parametrize("p", [0.3, 0.05])
def test_timed_fused_kernel_full_size(orc, p):
    """The benchmark's own launch, the fused register-resident kernel on 10,000 x (n=200,
    m=852) with 64 rows through the device entry point, equals the separate evaluation and
    BCP launches output for output over the whole batch, and the oracle on a 256-instance
    sample (VariablesStateHandler.cu:180-206, WatchedClausesList.cu:46-309)."""
    print(f"[eval/bcp] fused launch at full size, p={p}", flush=True)
    N, n, m, A = 10_000, 200, 852, 64
    b = W.random_3sat_batch(N, n, m, seed=0)
    rows = W.random_assignments(b, A, p, seed=0)
    with G.GpuBatch(b) as gb:
        assert G.tune_get("eval_fast3") == 1
        f_counts, f_out, f_conflict, f_props = fused(gb, rows, A)
        rb = G.DevBuf.from_host(rows)
        cb, ob = G.DevBuf(N * A * 12), G.DevBuf(rows.size * 4)
        fb, pb = G.DevBuf(N * A * 4), G.DevBuf(N * A * 4)
        gb.clause_eval_dev(rb, A, cb)
        gb.bcp_fixpoint_dev(rb, A, ob, fb, pb)
        gb.sync()
        counts = cb.to_host(np.int32, N * A * 3).reshape(N, A, 3)
        out = ob.to_host(np.uint32, rows.size)
        conflict = fb.to_host(np.int32, N * A).reshape(N, A)
        props = pb.to_host(np.int32, N * A).reshape(N, A)
    np.testing.assert_array_equal(f_counts, counts)
    np.testing.assert_array_equal(f_out, out)
    np.testing.assert_array_equal(f_conflict, conflict)
    np.testing.assert_array_equal(f_props, props)
    sample = np.random.default_rng(2).choice(N, 256, replace=False)
    sb = b.subset(sample)
    base = b.row_base(A)
    srows = np.concatenate([rows[base[i]:base[i + 1]] for i in sample])
    np.testing.assert_array_equal(f_counts[sample], orc.eval_batch(sb, srows, A))
    e_out, e_conf, e_props = orc.bcp_batch(sb, srows, A)
    np.testing.assert_array_equal(f_conflict[sample], e_conf)
    np.testing.assert_array_equal(f_props[sample], e_props)
    np.testing.assert_array_equal(np.concatenate([f_out[base[i]:base[i + 1]] for i in sample]), e_out)


@pytest.mark.parametrize("A", [1, 8, 64])
def test_large_instances_hbm_mask_table(orc, A):
    """Instances whose 2n literal masks exceed a workgroup's LDS (n = 2,500 and 5,000,
    beside a small one): the general kernel keeps its mask table in an HBM slice per
    instance, with the same bit-exact counts and fixpoints."""
    print(f"[eval/bcp] n=2500/5000 mask tables in HBM, A={A}", flush=True)
    fs = [W.random_3sat_batch(1, 2500, 7500, seed=31).instance(0),
          W.random_3sat_batch(1, 60, 255, seed=32).instance(0),
          W.mixed_ksat_batch(1, seed=33)[0].instance(0),
          W.random_3sat_batch(1, 5000, 21000, seed=34).instance(0)]
    b = Batch.from_formulas(fs)
    rows = W.random_assignments(b, A, 0.02, seed=35)
    counts, conflict, props = check_batch(orc, b, rows, A)
    assert props.sum() > 0


def mixed_length_formula(n, m, rng):
    """m clauses over n vars with lengths 1-6 (1% units, 10% binary, the rest 3-6)."""
    cls = []
    for _ in range(m):
        k = int(rng.choice([1, 2, 3, 4, 5, 6], p=[0.01, 0.1, 0.22, 0.23, 0.22, 0.22]))
        vs = rng.choice(n, size=k, replace=False)
        cls.append([2 * int(v) + int(rng.integers(2)) for v in vs])
    return Formula.from_clauses(n, cls)


def packed_batch():
    rng = np.random.default_rng(61)
    mixed, _ = W.mixed_ksat_batch(8, seed=62)
    fs = [mixed.instance(i) for i in range(mixed.n_inst)]
    fs += [
        W.random_3sat_batch(1, 300, 1278, seed=63).instance(0),   # 3-SAT beyond the 3-literal kernel
        Batch.uniform(W.random_ksat_clauses(np.random.default_rng(64), 1, 70, 2900, 6), 70).instance(0),  # 6-SAT
        mixed_length_formula(496, 1500, rng),                        # W = 31, the largest packed n
        mixed_length_formula(37, 90, rng),
        Formula.from_clauses(3, []),                                 # no clauses
        Formula.from_clauses(3, [[], [1, 3]]),                       # empty clause
        Formula.from_clauses(4, [[1], [0, 3], [2, 5], [4, 7]]),      # unit chain
        Formula.from_clauses(3, [[1, 1, 3], [0, 1, 5], [2, 2]]),     # duplicates, tautology
        Formula.from_clauses(1, [[1], [0]]),
        Formula.from_clauses(6, [[0, 2, 4, 6, 8, 10], [1, 3, 5, 7, 9, 11], [11, 9]]),
    ]
    return Batch.from_formulas(fs)


def packed_rows(b, A, p, seed):
    rows = W.random_assignments(b, A, p, seed=seed)
    # row 3 of every instance (A > 3): a contradictory input
    if A > 3:
        base = 0
        for i in range(b.n_inst):
            w = (2 * int(b.n_vars[i]) + 31) // 32
            if w:
                rows[base + 3 * w] |= 0x3
            base += A * w
    return rows


@pytest.mark.parametrize("A,p,nw", [(64, 0.02, 4), (64, 0.05, 4), (64, 0.15, 4), (64, 0.3, 4), (1, 0.05, 4),
                                    (37, 0.1, 4), (64, 0.05, 1), (64, 0.1, 3), (64, 0.3, 2)])
def test_packed_general_kernel(orc, A, p, nw):
    """The packed general kernel (clauses of 1-6 literals in any mix, n up to 496, m up to
    2,900, empty and unit clauses, duplicate literals, contradictory inputs; nw wavefronts
    sharing each instance) gives the oracle's counts, closure rows, conflicts and props bit
    for bit, and the same outputs as the CSR general kernel."""
    print(f"[eval/bcp] packed general kernel A={A} p={p} waves={nw}", flush=True)
    b = packed_batch()
    rows = packed_rows(b, A, p, seed=int(1000 * p) + A)
    with G.GpuBatch(b) as gb:
        assert gb.eval_kernel() == "packed"
        with G.tuned(eval_pkk=0):
            assert gb.eval_kernel() == "csr"
            c_counts, c_out, c_conflict, c_props = fused(gb, rows, A)
    with G.tuned(eval_pkk_waves=nw):
        counts, conflict, props = check_batch(orc, b, rows, A)
    np.testing.assert_array_equal(counts, c_counts)
    np.testing.assert_array_equal(conflict, c_conflict)
    np.testing.assert_array_equal(props, c_props)
    if p <= 0.05 and A == 64:
        assert props.sum() > 0 and conflict.mean() < 0.9


@pytest.mark.parametrize("cmax,nw", [(0, 4), (8, 4), (32, 4), (32, 1), (32, 3)])
@pytest.mark.parametrize("p", [0.1, 0.3])
def test_packed_compacted_passes(orc, p, cmax, nw):
    """The packed kernel's handoff to 32-bit masks once at most cmax rows still move (0: never):
    on the mixed test batch and on 48 instances of the mixed k-SAT workload (k 3/4/5, n up to
    400, where 5-SAT rows at p = 0.3 keep moving for many passes), the oracle's counts, closure
    rows, conflicts and props bit for bit."""
    print(f"[eval/bcp] packed compacted passes p={p} cmax={cmax} waves={nw}", flush=True)
    mixed, ks = W.mixed_ksat_batch(4096, seed=3)
    idx = [int(i) for k in (3, 4, 5) for i in np.nonzero(ks == k)[0][:16]]
    pb = packed_batch()
    b = Batch.from_formulas([pb.instance(i) for i in range(pb.n_inst)] + [mixed.instance(i) for i in idx])
    rows = packed_rows(b, 64, p, seed=int(100 * p) + cmax + nw)
    with G.tuned(eval_pkk_compact=cmax, eval_pkk_waves=nw):
        with G.GpuBatch(b) as gb:
            assert gb.eval_kernel() == "packed"
        check_batch(orc, b, rows, 64)


@pytest.mark.parametrize("small64", [0, 8, 32])
@pytest.mark.parametrize("p", [0.2, 0.3])
def test_3literal_bcp_modes(orc, p, small64):
    """The 3-literal kernel's BCP modes for few open rows: compacted bit by bit (small64 = 0),
    the default (at most 8 open rows stay on the 64-bit table), and the 64-bit table for up to
    32 open rows; the same oracle fixpoints in every mode."""
    print(f"[eval/bcp] 3-literal BCP modes p={p} small64={small64}", flush=True)
    b = W.random_3sat_batch(96, 200, 852, seed=5)
    rows = W.random_assignments(b, 64, p, seed=int(p * 100) + small64)
    with G.tuned(eval_fast3=1, eval_fast3_small64=small64):
        check_batch(orc, b, rows, 64)
