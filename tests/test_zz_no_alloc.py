"""gpsat_solve_noalloc: the solve with every device buffer carved from the library's static
scratch region.  Called through the C ABI, its outputs equal gpsat_solve's on two calls in a
row (the second with a smaller batch, so stale scratch contents would show), and sizes
beyond the region are refused with GPSAT_ERR_CAPACITY before anything runs."""
import ctypes

import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd import workloads as W

pytestmark = pytest.mark.gpu


def _call(sym, gb, opts, jobs=None):
    """One gpsat_solve / gpsat_solve_noalloc call with host buffers; returns
    (rc, status, model rows, job stats)."""
    N = gb.n_inst
    status = np.full(N, 7, dtype=np.int32)
    rows = np.zeros(max(gb.row_words(1), 1), dtype=np.uint32)
    verified = np.zeros(N, dtype=np.int32)
    if jobs is None:
        ji = co = cl = None
        J = N
    else:
        ji, co, cl = (np.ascontiguousarray(x, dtype=np.int32) for x in jobs)
        J = len(ji)
    js = np.zeros((max(J, 1), len(G.JOB_STAT_FIELDS)), dtype=np.int64)
    info = G.SolveInfo()
    P = lambda a: G._p(a) if a is not None else None
    fn = getattr(G.lib(), sym)
    rc = fn(gb.h, ctypes.byref(opts), P(ji), P(co), P(cl), ctypes.c_int32(J), None, None, G._p(status), G._p(rows),
            G._p(verified), G._p(js), None, ctypes.byref(info))
    return rc, status, rows, js[:J], verified, info


def _same(orc, b, opts):
    with G.GpuBatch(b) as gb:
        rc0, st0, rows0, js0, v0, _ = _call("gpsat_solve", gb, opts)
        assert rc0 == 0, G.lib().gpsat_last_error()
        rc1, st1, rows1, js1, v1, info = _call("gpsat_solve_noalloc", gb, opts)
        assert rc1 == 0, G.lib().gpsat_last_error()
    np.testing.assert_array_equal(st1, st0)
    np.testing.assert_array_equal(rows1, rows0)
    np.testing.assert_array_equal(v1, v0)
    # status, conflicts, decisions, propagations, learned, restarts, max level
    np.testing.assert_array_equal(js1[:, :7], js0[:, :7])
    assert int(v1.sum()) == int((st1 == G.SAT).sum())
    for i in range(0, b.n_inst, 10):
        exp, _, _ = orc.solve_cdcl(b.instance(i))
        assert st1[i] == exp, i
    return info


def test_noalloc_equals_solve_twice(orc):
    print("[noalloc] 240 x n=50, then 90 x n=40", flush=True)
    o = G.SolveOpts.default()
    info = _same(orc, W.random_3sat_batch(240, 50, 213, seed=71), o)
    assert info.n_sat + info.n_unsat == 240
    info = _same(orc, W.random_3sat_batch(90, 40, 170, seed=72), o)
    assert info.n_sat + info.n_unsat == 90


def test_noalloc_hbm_workspace(orc):
    """Clause part in HBM: the per-wavefront workspace slices come from the scratch too."""
    print("[noalloc] clause part in HBM, 120 x n=60", flush=True)
    _same(orc, W.random_3sat_batch(120, 60, 256, seed=73), G.SolveOpts.default(clause_mem=2))


def test_noalloc_refuses_oversize_and_epochs():
    print("[noalloc] oversize job table and epochs refused", flush=True)
    b = W.random_3sat_batch(1, 50, 213, seed=74)
    J = 700_000          # job table + job stats (120 B per job) exceed the 64 MiB region
    ji = np.zeros(J, dtype=np.int32)
    co = np.arange(J + 1, dtype=np.int32)
    cl = (np.arange(J, dtype=np.int32) % 2).astype(np.int32)
    with G.GpuBatch(b) as gb:
        rc, *_ = _call("gpsat_solve_noalloc", gb, G.SolveOpts.default(), jobs=(ji, co, cl))
        assert rc == -4, rc
        assert "scratch" in G.lib().gpsat_last_error().decode()
        rc, *_ = _call("gpsat_solve_noalloc", gb, G.SolveOpts.default(epoch_conflicts=100))
        assert rc == -1, rc
