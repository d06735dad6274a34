"""gpupsat_amd -- MI355X SAT engine: Python binding over the C ABI of libgpsat.so.

The compute path is the HIP library built from gpupsat_amd/csrc for gfx950; this
module only marshals numpy arrays through ctypes.  There is no CPU fallback: if the
library is missing, or no GPU is visible, calls raise GpsatError.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from .formula import Batch, Formula, pack_rows, unpack_rows, words_per_row  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
# GPSAT_LIB selects another build of the library (the sanitizer build of
# tests/run_sanitized.sh)
LIB_PATH = os.environ.get("GPSAT_LIB") or os.path.join(_HERE, "libgpsat.so")
CLI_PATH = os.path.join(_HERE, "bin", "gpupsat")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "gpsat.h")

SAT, UNSAT, UNDEF = 0, 1, 2


class GpsatError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libgpsat.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GpsatError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.gpsat_last_error.restype = ctypes.c_char_p
        _lib.gpsat_version.restype = ctypes.c_char_p
    return _lib


def _check(rc: int):
    if rc != 0:
        raise GpsatError(f"gpsat error {rc}: {lib().gpsat_last_error().decode()}")


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def device_count() -> int:
    c = ctypes.c_int32(0)
    _check(lib().gpsat_device_count(ctypes.byref(c)))
    return c.value


def synchronize():
    _check(lib().gpsat_device_synchronize())


def tune_get(key: str) -> int:
    v = ctypes.c_int32()
    _check(lib().gpsat_tune_get(key.encode(), ctypes.byref(v)))
    return v.value


def tune_set(key: str, value: int):
    _check(lib().gpsat_tune_set(key.encode(), ctypes.c_int32(int(value))))


class tuned:
    """Context manager: set launch-shape keys (gpsat_tune_set) and restore them on exit."""

    def __init__(self, **kw):
        self.kw = kw
        self.old = {}

    def __enter__(self):
        for k, v in self.kw.items():
            self.old[k] = tune_get(k)
            tune_set(k, v)
        return self

    def __exit__(self, *a):
        for k, v in self.old.items():
            tune_set(k, v)


class DevBuf:
    """A device allocation owned by Python (freed on close / garbage collection)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().gpsat_dev_alloc(ctypes.c_int64(self.nbytes), ctypes.byref(p)))
        self.ptr = p

    @staticmethod
    def from_host(a: np.ndarray) -> "DevBuf":
        a = np.ascontiguousarray(a)
        b = DevBuf(max(a.nbytes, 4))
        _check(lib().gpsat_memcpy_h2d(b.ptr, _p(a), ctypes.c_int64(a.nbytes)))
        return b

    def to_host(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        _check(lib().gpsat_memcpy_d2h(_p(out), self.ptr, ctypes.c_int64(out.nbytes)))
        return out

    def zero(self):
        _check(lib().gpsat_memset_dev(self.ptr, 0, ctypes.c_int64(self.nbytes)))

    def close(self):
        if self.ptr:
            lib().gpsat_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib().gpsat_event_create(ctypes.byref(p)))
        self.ptr = p

    def record(self, batch: Optional["GpuBatch"] = None):
        _check(lib().gpsat_event_record(self.ptr, batch.h if batch is not None else None))

    def elapsed_ms(self, stop: "Event") -> float:
        ms = ctypes.c_float(0)
        _check(lib().gpsat_event_elapsed_ms(self.ptr, stop.ptr, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().gpsat_event_destroy(self.ptr)
        except Exception:
            pass


# ---- host-side boundary functions (no GPU needed)
def read_dimacs(path: str):
    """DIMACS through the library's reader: (Formula, header_vars, header_clauses)."""
    f = ctypes.c_void_p()
    _check(lib().gpsat_dimacs_read(path.encode(), ctypes.byref(f)))
    return _take_formula(f)


def parse_dimacs(text: str):
    f = ctypes.c_void_p()
    b = text.encode()
    _check(lib().gpsat_dimacs_parse(b, ctypes.c_int64(len(b)), ctypes.byref(f)))
    return _take_formula(f)


def _take_formula(f):
    n, m, hv, hc = (ctypes.c_int32() for _ in range(4))
    L = ctypes.c_int64()
    lib().gpsat_formula_info(f, ctypes.byref(n), ctypes.byref(m), ctypes.byref(L), ctypes.byref(hv), ctypes.byref(hc))
    lits = np.zeros(max(L.value, 1), dtype=np.int32)
    off = np.zeros(m.value + 1, dtype=np.int64)
    lib().gpsat_formula_arrays(f, _p(lits), _p(off))
    lib().gpsat_formula_free(f)
    return Formula(n.value, lits[: L.value], off), hv.value, hc.value


def preprocess(f: Formula):
    """The reference's preprocessing: (status, Formula or None, solved literals)."""
    L = max(len(f.lits), 1)
    out_lits = np.zeros(L, dtype=np.int32)
    out_off = np.zeros(f.n_clauses + 1, dtype=np.int64)
    solved = np.zeros(f.n_clauses + 1, dtype=np.int32)
    om, ns, st = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    lits = np.ascontiguousarray(f.lits, dtype=np.int32)
    off = np.ascontiguousarray(f.clause_off, dtype=np.int64)
    _check(lib().gpsat_preprocess(_p(lits), _p(off), ctypes.c_int32(f.n_clauses), ctypes.c_int32(f.n_vars),
                                  _p(out_lits), _p(out_off), ctypes.byref(om), _p(solved), ctypes.byref(ns),
                                  ctypes.byref(st)))
    g = None
    if st.value == UNDEF:
        g = Formula(f.n_vars, out_lits[: out_off[om.value]].copy(), out_off[: om.value + 1].copy())
    return st.value, g, solved[: ns.value].copy()


def cube_vars(f: Formula, threads: int = 32, blocks: int = 32, strategy: int = 0, dead=()):
    out = np.zeros(16, dtype=np.int32)
    d = ctypes.c_int32()
    dv = np.ascontiguousarray(list(dead) or [0], dtype=np.int32)
    lits = np.ascontiguousarray(f.lits, dtype=np.int32)
    off = np.ascontiguousarray(f.clause_off, dtype=np.int64)
    _check(lib().gpsat_cube_vars(_p(lits), _p(off), ctypes.c_int32(f.n_clauses), ctypes.c_int32(f.n_vars), _p(dv),
                                 ctypes.c_int32(len(dead)), ctypes.c_int64(threads), ctypes.c_int64(blocks),
                                 ctypes.c_int32(strategy), _p(out), ctypes.byref(d)))
    return out[: d.value].copy()


def cube_vars_scored(scores, depth: int, dead=()):
    s = np.ascontiguousarray(scores, dtype=np.float32)
    out = np.zeros(max(depth, 1), dtype=np.int32)
    d = ctypes.c_int32()
    dv = np.ascontiguousarray(list(dead) or [0], dtype=np.int32)
    _check(lib().gpsat_cube_vars_scored(_p(s), ctypes.c_int32(len(s)), _p(dv), ctypes.c_int32(len(dead)),
                                        ctypes.c_int32(depth), _p(out), ctypes.byref(d)))
    return out[: d.value].copy()


def cubes(vars_) -> np.ndarray:
    v = np.ascontiguousarray(vars_, dtype=np.int32)
    d = len(v)
    out = np.zeros(((1 << d), max(d, 1)), dtype=np.int32)
    _check(lib().gpsat_cubes(_p(v) if d else None, ctypes.c_int32(d), _p(out)))
    return out[:, :d]


def cube_jobs(batch: Batch, vars_per_inst) -> tuple:
    """Job arrays (job_inst, cube_off, cube_lits) splitting instance i on vars_per_inst[i]
    into its 2^d cubes, in the reference's JobChooser order (gpsat_cubes)."""
    ji, cl, lens = [], [], []
    for i, vs in enumerate(vars_per_inst):
        c = cubes(vs)
        ji.append(np.full(c.shape[0], i, dtype=np.int32))
        cl.append(c.reshape(-1))
        lens.append(np.full(c.shape[0], c.shape[1], dtype=np.int64))
    ji = np.concatenate(ji) if ji else np.zeros(0, dtype=np.int32)
    lens = np.concatenate(lens) if lens else np.zeros(0, dtype=np.int64)
    co = np.zeros(len(ji) + 1, dtype=np.int64)
    np.cumsum(lens, out=co[1:])
    if co[-1] > np.iinfo(np.int32).max:
        raise GpsatError("cube literal arena exceeds 2^31 entries")
    cl = np.concatenate(cl).astype(np.int32) if cl and co[-1] else np.zeros(1, dtype=np.int32)
    return ji, co.astype(np.int32), cl


class SolveOpts(ctypes.Structure):
    """Mirror of gpsat_solve_opts (include/gpsat.h)."""
    _fields_ = [
        ("decision", ctypes.c_int32), ("restart_first", ctypes.c_int32), ("restart_factor", ctypes.c_float),
        ("learnt_cap", ctypes.c_int32), ("larena_cap", ctypes.c_int32), ("watch_slack", ctypes.c_int32),
        ("max_conflicts", ctypes.c_int64), ("max_decisions", ctypes.c_int64), ("var_decay", ctypes.c_float),
        ("phase_init", ctypes.c_int32), ("max_ms", ctypes.c_double), ("waves", ctypes.c_int32),
        ("clause_mem", ctypes.c_int32), ("verify", ctypes.c_int32), ("chunk", ctypes.c_int32),
        ("epoch_conflicts", ctypes.c_int64), ("epoch_growth", ctypes.c_float), ("max_epochs", ctypes.c_int32),
        ("deadline_ms", ctypes.c_double), ("inst_conflicts", ctypes.c_int64), ("job_order", ctypes.c_int32),
        ("telemetry", ctypes.c_int32), ("learned_out", ctypes.POINTER(ctypes.c_int32)),
        ("learned_cap", ctypes.c_int64),
    ]

    @staticmethod
    def default(**kw) -> "SolveOpts":
        o = SolveOpts()
        _check(lib().gpsat_solve_opts_default(ctypes.byref(o)))
        for k, v in kw.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        return o


def parse_learned(d: np.ndarray) -> list:
    """Split a learned-clause dump into (instance, [clause literal lists]) blocks."""
    out, p = [], 0
    while p < len(d):
        inst, n = int(d[p]), int(d[p + 1])
        body = d[p + 2:p + 2 + n]
        ends = np.flatnonzero(body == -1)
        clauses, s = [], 0
        for e in ends:
            clauses.append(body[s:e].tolist())
            s = e + 1
        out.append((inst, clauses))
        p += 2 + n
    return out


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double), ("verify_ms", ctypes.c_double), ("waves", ctypes.c_int32),
        ("clause_mem", ctypes.c_int32), ("lds_bytes", ctypes.c_int64), ("workspace_bytes", ctypes.c_int64),
        ("n_sat", ctypes.c_int32), ("n_unsat", ctypes.c_int32), ("n_undef", ctypes.c_int32),
        ("n_verified", ctypes.c_int32), ("n_errors", ctypes.c_int32), ("epochs", ctypes.c_int32),
        ("chunk", ctypes.c_int32), ("learned_len", ctypes.c_int64), ("learned_need", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


JOB_STAT_FIELDS = ("status", "conflicts", "decisions", "props", "learned", "restarts", "max_level", "time_ns",
                   "setup_ns", "bcp_ns", "analyze_ns", "backtrack_ns", "decide_ns")


class GpuBatch:
    """A formula batch resident in HBM (gpsat_batch handle)."""

    def __init__(self, batch: Batch):
        self.batch = batch
        self.h = ctypes.c_void_p()
        self._keep = [np.ascontiguousarray(batch.lits, dtype=np.int32),
                      np.ascontiguousarray(batch.clause_off, dtype=np.int64),
                      np.ascontiguousarray(batch.inst_clause_off, dtype=np.int64),
                      np.ascontiguousarray(batch.n_vars, dtype=np.int32)]
        _check(lib().gpsat_batch_create(*(_p(a) for a in self._keep), ctypes.c_int32(batch.n_inst),
                                        ctypes.byref(self.h)))

    @property
    def n_inst(self) -> int:
        return self.batch.n_inst

    def eval_kernel(self) -> str:
        """The eval/BCP kernel launches on this batch run under the current tuning:
        "fast3", "packed", "csr" or "csr_hbm" (gpsat_batch_eval_kernel)."""
        k = ctypes.c_int32()
        _check(lib().gpsat_batch_eval_kernel(self.h, ctypes.byref(k)))
        return ("fast3", "packed", "csr", "csr_hbm")[k.value]

    def row_words(self, A: int) -> int:
        w = ctypes.c_int64(0)
        _check(lib().gpsat_batch_row_words(self.h, ctypes.c_int32(A), ctypes.byref(w)))
        return w.value

    def sync(self):
        _check(lib().gpsat_batch_sync(self.h))

    # ---- host-buffer forms
    def clause_eval(self, rows: np.ndarray, A: int) -> np.ndarray:
        rows = np.ascontiguousarray(rows, dtype=np.uint32)
        assert rows.size == self.row_words(A)
        counts = np.zeros((self.n_inst, A, 3), dtype=np.int32)
        _check(lib().gpsat_clause_eval(self.h, _p(rows), ctypes.c_int32(A), _p(counts)))
        return counts

    def bcp_fixpoint(self, rows: np.ndarray, A: int):
        rows = np.ascontiguousarray(rows, dtype=np.uint32)
        assert rows.size == self.row_words(A)
        out = np.zeros_like(rows)
        conflict = np.zeros((self.n_inst, A), dtype=np.int32)
        props = np.zeros((self.n_inst, A), dtype=np.int32)
        _check(lib().gpsat_bcp_fixpoint(self.h, _p(rows), ctypes.c_int32(A), _p(out), _p(conflict), _p(props)))
        return out, conflict, props

    # ---- device-buffer forms (asynchronous on the batch stream)
    def clause_eval_dev(self, rows: DevBuf, A: int, counts: DevBuf):
        _check(lib().gpsat_clause_eval_dev(self.h, rows.ptr, ctypes.c_int32(A), counts.ptr))

    def bcp_fixpoint_dev(self, rows: DevBuf, A: int, out: DevBuf, conflict: DevBuf, props: DevBuf):
        _check(lib().gpsat_bcp_fixpoint_dev(self.h, rows.ptr, ctypes.c_int32(A), out.ptr, conflict.ptr, props.ptr))

    def eval_bcp_dev(self, rows: DevBuf, A: int, counts: DevBuf, out: DevBuf, conflict: DevBuf, props: DevBuf):
        _check(lib().gpsat_eval_bcp_dev(self.h, rows.ptr, ctypes.c_int32(A), counts.ptr, out.ptr, conflict.ptr,
                                        props.ptr))

    # ---- trajectory mode
    def solve_ref(self, dead=None, max_iterations: int = 0, learned_cap: int = 0, learned_lits: int = 0) -> dict:
        """The reference's own sequential search, one instance per wavefront (gpsat_solve_ref):
        status (0 SAT, 1 UNSAT, 2 UNDEF, 4 a graph vertex past the reference's edge capacity),
        model rows and per-instance stats (decisions, conflicts, learned, iterations,
        restarts).  learned_cap / learned_lits are accepted and not used."""
        N = self.n_inst
        status = np.zeros(N, dtype=np.int32)
        rows = np.zeros(max(self.row_words(1), 1), dtype=np.uint32)
        stats = np.zeros((max(N, 1), 5), dtype=np.int64)
        d = None
        if dead is not None:
            d = np.ascontiguousarray(dead, dtype=np.uint8)
            assert d.size == int(self.batch.n_vars.astype(np.int64).sum())
        _check(lib().gpsat_solve_ref(self.h, _p(d) if d is not None else None, ctypes.c_int32(max_iterations),
                                     ctypes.c_int32(learned_cap), ctypes.c_int32(learned_lits), _p(status), _p(rows),
                                     _p(stats)))
        return {"status": status, "model_rows": rows, "stats": stats[:N]}

    # ---- device preprocessing
    def preprocess(self):
        """The reference's preprocessing (duplicate/tautology removal, unit propagation)
        for every instance on the device (gpsat_preprocess_batch).  Returns
        (status int32[N], solved literal-bitmap rows, Batch of the remaining formulas)."""
        b = self.batch
        N = b.n_inst
        status = np.zeros(N, dtype=np.int32)
        solved = np.zeros(max(self.row_words(1), 1), dtype=np.uint32)
        L = max(len(b.lits), 1)
        out_lits = np.zeros(L, dtype=np.int32)
        out_off = np.zeros(int(b.inst_clause_off[-1]) + 1, dtype=np.int64)
        out_ioff = np.zeros(N + 1, dtype=np.int64)
        _check(lib().gpsat_preprocess_batch(self.h, _p(status), _p(solved), _p(out_lits), _p(out_off), _p(out_ioff)))
        from .formula import Batch as _Batch
        nc = int(out_ioff[-1])
        rest = _Batch(out_lits[: out_off[nc]].copy(), out_off[: nc + 1].copy(), out_ioff.copy(),
                      b.n_vars.copy())
        return status, solved, rest

    # ---- autotuner
    def autotune_solve(self, opts: Optional["SolveOpts"] = None, jobs=None, sample_jobs: int = 0,
                       trial_ms: float = 0.0):
        """Launch shape (waves, clause_mem, chunk) for this batch and job list from short
        trial solves (gpsat_autotune_solve).  Returns (SolveOpts, report text)."""
        o = opts if opts is not None else SolveOpts.default()
        best = SolveOpts()
        rep = ctypes.create_string_buffer(8192)
        if jobs is None:
            ji = co = cl = None
            nj = 0
        else:
            ji = np.ascontiguousarray(jobs[0], dtype=np.int32)
            co = np.ascontiguousarray(jobs[1], dtype=np.int32)
            cl = np.ascontiguousarray(jobs[2], dtype=np.int32)
            nj = len(ji)
        _check(lib().gpsat_autotune_solve(self.h, ctypes.byref(o), _p(ji) if ji is not None else None,
                                          _p(co) if co is not None else None, _p(cl) if cl is not None else None,
                                          ctypes.c_int32(nj), ctypes.c_int32(sample_jobs), ctypes.c_double(trial_ms),
                                          ctypes.byref(best), rep, ctypes.c_int32(len(rep))))
        return best, rep.value.decode()

    def autotune_eval(self, rows: "DevBuf", A: int, do_eval: bool = True, do_bcp: bool = True, reps: int = 0) -> str:
        """Time the eval/BCP launch shapes on these rows and keep the fastest
        (gpsat_autotune_eval).  Returns the report text."""
        rep = ctypes.create_string_buffer(8192)
        _check(lib().gpsat_autotune_eval(self.h, rows.ptr, ctypes.c_int32(A), ctypes.c_int32(int(do_eval)),
                                         ctypes.c_int32(int(do_bcp)), ctypes.c_int32(reps), rep,
                                         ctypes.c_int32(len(rep))))
        return rep.value.decode()

    # ---- solve path
    def solve(self, opts: Optional[SolveOpts] = None, jobs=None, act_init=None, phase_init=None,
              want_models: bool = True, want_bumps: bool = False, learned_cap: int = 0) -> dict:
        """jobs: None (one job per instance) or (job_inst int32[J], cube_off int32[J+1], cube_lits int32[]).
        learned_cap > 0: also return "learned", a list of (instance, clauses) blocks, one per job
        that ran (its held learned clauses and level-0 literals, each implied by the formula and
        the job's cube), as many as fit in learned_cap int32 slots."""
        o = SolveOpts.from_buffer_copy(opts) if opts is not None else SolveOpts.default()
        dump = None
        if learned_cap > 0:
            dump = np.zeros(int(learned_cap), dtype=np.int32)
            o.learned_out = dump.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            o.learned_cap = int(learned_cap)
        N = self.n_inst
        status = np.full(N, 2, dtype=np.int32)
        verified = np.zeros(N, dtype=np.int32)
        rows = np.zeros(max(self.row_words(1), 1), dtype=np.uint32)
        if jobs is None:
            ji = co = cl = None
            J = N
        else:
            ji = np.ascontiguousarray(jobs[0], dtype=np.int32)
            co = np.ascontiguousarray(jobs[1], dtype=np.int32)
            cl = np.ascontiguousarray(jobs[2] if len(jobs[2]) else np.zeros(1, np.int32), dtype=np.int32)
            J = len(ji)
        js = np.zeros((max(J, 1), len(JOB_STAT_FIELDS)), dtype=np.int64)
        nv = int(self.batch.n_vars.astype(np.int64).sum())
        bumps = np.zeros(max(nv, 1), dtype=np.int32) if want_bumps else None
        act = None if act_init is None else np.ascontiguousarray(act_init, dtype=np.float32)
        ph = None if phase_init is None else np.ascontiguousarray(phase_init, dtype=np.int8)
        info = SolveInfo()
        P = lambda a: _p(a) if a is not None else None
        rc = lib().gpsat_solve(self.h, ctypes.byref(o), P(ji), P(co), P(cl), ctypes.c_int32(J), P(act), P(ph),
                               _p(status), _p(rows), _p(verified), _p(js), P(bumps), ctypes.byref(info))
        _check(rc)
        out = {"status": status, "verified": verified, "job_stats": js[:J], "info": info.as_dict()}
        if want_models:
            out["model_rows"] = rows
        if want_bumps:
            out["var_bumps"] = bumps[:nv]
        if dump is not None:
            out["learned"] = parse_learned(dump[:info.learned_len])
        return out

    def models(self, model_rows: np.ndarray) -> list:
        """Split solve() model rows into int8[n] models (1 true, 0 false, -1 unassigned)."""
        base = self.batch.row_base(1)
        out = []
        for i in range(self.n_inst):
            n = int(self.batch.n_vars[i])
            t, f = unpack_rows(model_rows[base[i]:base[i + 1]][None, :], n)
            m = np.full(n, -1, dtype=np.int8)
            m[t[0]] = 1
            m[f[0]] = 0
            out.append(m)
        return out

    def close(self):
        if self.h:
            lib().gpsat_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
