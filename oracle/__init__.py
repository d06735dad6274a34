"""ctypes wrapper over the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package as the checker.  The product package
(gpupsat_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GPSAT_ORACLE_LIB selects another build of the oracle (the sanitizer build of
# tests/run_sanitized.sh)
_LIB_PATH = os.environ.get("GPSAT_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

SAT, UNSAT, UNDEF = 0, 1, 2


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                              stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def eval_batch(batch, rows: np.ndarray, A: int) -> np.ndarray:
    counts = np.zeros((batch.n_inst, A, 3), dtype=np.int32)
    lib().orc_eval_batch(ctypes.c_int32(batch.n_inst), _p(_i32(batch.lits)), _p(_i64(batch.clause_off)),
                         _p(_i64(batch.inst_clause_off)), _p(_i32(batch.n_vars)),
                         _p(np.ascontiguousarray(rows, dtype=np.uint32)), ctypes.c_int32(A), _p(counts))
    return counts


def bcp_batch(batch, rows: np.ndarray, A: int):
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    out = np.zeros_like(rows)
    conflict = np.zeros((batch.n_inst, A), dtype=np.int32)
    props = np.zeros((batch.n_inst, A), dtype=np.int32)
    lib().orc_bcp_batch(ctypes.c_int32(batch.n_inst), _p(_i32(batch.lits)), _p(_i64(batch.clause_off)),
                        _p(_i64(batch.inst_clause_off)), _p(_i32(batch.n_vars)), _p(rows),
                        ctypes.c_int32(A), _p(out), _p(conflict), _p(props))
    return out, conflict, props


def clause_status(lits, row):
    lits = _i32(lits)
    u = ctypes.c_int32(0)
    st = lib().orc_clause_status(_p(lits), ctypes.c_int32(len(lits)), _p(np.ascontiguousarray(row, dtype=np.uint32)),
                                 ctypes.byref(u))
    return st, u.value


def up_closure(f, row):
    row = np.ascontiguousarray(row, dtype=np.uint32)
    out = np.zeros_like(row)
    props = ctypes.c_int32(0)
    c = lib().orc_up_closure(_p(_i32(f.lits)), _p(_i64(f.clause_off)), ctypes.c_int32(f.n_clauses),
                             ctypes.c_int32(f.n_vars), _p(row), _p(out), ctypes.byref(props))
    return c, out, props.value


def preprocess(f):
    """Returns (status, Formula after preprocessing or None, solved literal codes)."""
    from gpupsat_amd.formula import Formula  # data class only, no native code
    L = max(len(f.lits), 1)
    out_lits = np.zeros(L, dtype=np.int32)
    out_off = np.zeros(f.n_clauses + 1, dtype=np.int64)
    out_m = ctypes.c_int32(0)
    solved = np.zeros(max(f.n_clauses, 1) + 1, dtype=np.int32)
    ns = ctypes.c_int32(0)
    st = lib().orc_preprocess(_p(_i32(f.lits)), _p(_i64(f.clause_off)), ctypes.c_int32(f.n_clauses),
                              ctypes.c_int32(f.n_vars), _p(out_lits), _p(out_off), ctypes.byref(out_m),
                              _p(solved), ctypes.byref(ns))
    g = None
    if st == UNDEF:
        m = out_m.value
        g = Formula(f.n_vars, out_lits[: out_off[m]].copy(), out_off[: m + 1].copy())
    return st, g, solved[: ns.value].copy()


def solve_ref(f, dead=None, max_iterations: int = 0):
    model = np.zeros(max(f.n_vars, 1), dtype=np.int8)
    stats = np.zeros(5, dtype=np.int64)
    d = None if dead is None else np.ascontiguousarray(dead, dtype=np.uint8)
    st = lib().orc_solve_ref(_p(_i32(f.lits)), _p(_i64(f.clause_off)), ctypes.c_int32(f.n_clauses),
                             ctypes.c_int32(f.n_vars), _p(d) if d is not None else None,
                             ctypes.c_int32(max_iterations), _p(model), _p(stats))
    return st, model[: f.n_vars], stats


def solve_ref_batch(batch, max_iterations: int = 0) -> np.ndarray:
    status = np.zeros(batch.n_inst, dtype=np.int32)
    lib().orc_solve_ref_batch(ctypes.c_int32(batch.n_inst), _p(_i32(batch.lits)), _p(_i64(batch.clause_off)),
                              _p(_i64(batch.inst_clause_off)), _p(_i32(batch.n_vars)),
                              ctypes.c_int32(max_iterations), _p(status))
    return status


def solve_dpll(f, max_nodes: int = 0):
    model = np.zeros(max(f.n_vars, 1), dtype=np.int8)
    st = lib().orc_solve_dpll(_p(_i32(f.lits)), _p(_i64(f.clause_off)), ctypes.c_int32(f.n_clauses),
                              ctypes.c_int32(f.n_vars), _p(model), ctypes.c_int64(max_nodes))
    return st, model[: f.n_vars]


def solve_brute(f):
    model = np.zeros(max(f.n_vars, 1), dtype=np.int8)
    st = lib().orc_solve_brute(_p(_i32(f.lits)), _p(_i64(f.clause_off)), ctypes.c_int32(f.n_clauses),
                               ctypes.c_int32(f.n_vars), _p(model))
    return st, model[: f.n_vars]


def solve_cdcl(f, max_conflicts: int = -1):
    model = np.zeros(max(f.n_vars, 1), dtype=np.int8)
    stats = np.zeros(5, dtype=np.int64)
    st = lib().orc_solve_cdcl(_p(_i32(f.lits)), _p(_i64(f.clause_off)), ctypes.c_int32(f.n_clauses),
                              ctypes.c_int32(f.n_vars), _p(model), ctypes.c_int64(max_conflicts), _p(stats))
    return st, model[: f.n_vars], stats


def verify_model(f, model) -> bool:
    m = np.ascontiguousarray(model, dtype=np.int8)
    return bool(lib().orc_verify_model(_p(_i32(f.lits)), _p(_i64(f.clause_off)), ctypes.c_int32(f.n_clauses), _p(m)))


def choose_cube_vars(f, n_dead: int, threads: int, blocks: int, strategy: int = 0):
    out = np.zeros(16, dtype=np.int32)
    d = lib().orc_choose_cube_vars(_p(_i32(f.lits)), _p(_i64(f.clause_off)), ctypes.c_int32(f.n_clauses),
                                   ctypes.c_int32(f.n_vars), ctypes.c_int32(n_dead), ctypes.c_int64(threads),
                                   ctypes.c_int64(blocks), ctypes.c_int32(strategy), _p(out))
    return out[:d].copy()


def enumerate_cubes(vars_) -> np.ndarray:
    v = _i32(vars_)
    d = len(v)
    out = np.zeros((1 << d, d), dtype=np.int32)
    lib().orc_enumerate_cubes(_p(v), ctypes.c_int32(d), _p(out))
    return out


def restart_schedule(first: int, factor: float, k: int) -> np.ndarray:
    out = np.zeros(k, dtype=np.int32)
    lib().orc_restart_schedule(ctypes.c_int32(first), ctypes.c_float(factor), ctypes.c_int32(k), _p(out))
    return out
