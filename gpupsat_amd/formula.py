"""CNF formulas and instance batches in the int32 CSR literal layout.

Literal code ``l = 2*v + s`` with ``s = 1`` for the positive literal, as in the
reference's ``mkLit`` (src/SATSolver/SolverTypes.cu:6-11); DIMACS ``x > 0`` maps to
``2*(x-1)+1`` and ``x < 0`` to ``2*(|x|-1)`` (src/FileManager/CnfReader.cpp:106-112).

A batch holds many instances back to back:

* ``lits``            int32[L]     literal codes of every clause of every instance
* ``clause_off``      int64[M+1]   global literal offset of each clause
* ``inst_clause_off`` int64[N+1]   global clause offset of each instance
* ``n_vars``          int32[N]     variables of each instance

Assignment rows: one row per (instance, assignment) is a bitmap indexed by
literal code with ``W = ceil(2n/32)`` uint32 words; bit ``l`` set means literal
``l`` is true, so var ``v`` has the 2-bit code (bit 2v = false, bit 2v+1 = true):
00 unassigned, 01 false, 10 true, 11 contradictory.  The rows of instance ``i``
start at word ``A * sum_{j<i} W_j``.
"""
from __future__ import annotations

import dataclasses
import hashlib
from typing import Iterable, List, Sequence

import numpy as np


def words_per_row(n: int) -> int:
    return (2 * int(n) + 31) // 32


@dataclasses.dataclass
class Formula:
    """One CNF formula in CSR form."""
    n_vars: int
    lits: np.ndarray          # int32[L]
    clause_off: np.ndarray    # int64[m+1]

    @property
    def n_clauses(self) -> int:
        return len(self.clause_off) - 1

    def clauses(self) -> List[List[int]]:
        return [self.lits[self.clause_off[c]:self.clause_off[c + 1]].tolist() for c in range(self.n_clauses)]

    @staticmethod
    def from_clauses(n_vars: int, clauses: Iterable[Sequence[int]]) -> "Formula":
        cl = [list(c) for c in clauses]
        off = np.zeros(len(cl) + 1, dtype=np.int64)
        off[1:] = np.cumsum([len(c) for c in cl]) if cl else []
        lits = np.fromiter((x for c in cl for x in c), dtype=np.int32, count=int(off[-1]))
        return Formula(int(n_vars), lits, off)

    @staticmethod
    def from_dimacs_clauses(clauses: Iterable[Sequence[int]], n_vars: int | None = None) -> "Formula":
        cl = [[2 * (abs(x) - 1) + (1 if x > 0 else 0) for x in c] for c in clauses]
        mx = max((abs(x) for c in clauses for x in c), default=0)
        return Formula.from_clauses(n_vars if n_vars is not None else mx, cl)

    def to_dimacs(self) -> str:
        out = [f"p cnf {self.n_vars} {self.n_clauses}"]
        for c in self.clauses():
            out.append(" ".join(str((l >> 1) + 1 if l & 1 else -((l >> 1) + 1)) for l in c) + " 0")
        return "\n".join(out) + "\n"


def parse_dimacs(text: str):
    """Parse DIMACS text.  Returns (header_vars, header_clauses, clauses as signed ints).

    Follows the reference grammar (src/FileManager/CnfReader.cpp:30-50): comment
    lines start with 'c', one 'p cnf V C' header, then clauses terminated by 0."""
    hv = hc = None
    clauses: List[List[int]] = []
    cur: List[int] = []
    for line in text.splitlines():
        s = line.strip()
        if not s or s.startswith("c") or s.startswith("%"):
            continue
        if s.startswith("p"):
            parts = s.split()
            if len(parts) < 4 or parts[1] != "cnf":
                raise ValueError("bad DIMACS header: " + s)
            hv, hc = int(parts[2]), int(parts[3])
            continue
        for tok in s.split():
            x = int(tok)
            if x == 0:
                clauses.append(cur)
                cur = []
            else:
                cur.append(x)
    if cur:
        clauses.append(cur)
    if hv is None:
        raise ValueError("missing DIMACS header")
    return hv, hc, clauses


def read_dimacs(path: str) -> Formula:
    """n_vars = highest var seen, as the reference does (CnfReader.cpp:117-123)."""
    with open(path) as f:
        _, _, clauses = parse_dimacs(f.read())
    return Formula.from_dimacs_clauses(clauses)


@dataclasses.dataclass
class Batch:
    """Many instances in one CSR literal arena (host arrays)."""
    lits: np.ndarray
    clause_off: np.ndarray
    inst_clause_off: np.ndarray
    n_vars: np.ndarray

    @property
    def n_inst(self) -> int:
        return len(self.n_vars)

    def words(self) -> np.ndarray:
        return (2 * self.n_vars.astype(np.int64) + 31) // 32

    def row_base(self, A: int) -> np.ndarray:
        w = self.words()
        base = np.zeros(self.n_inst + 1, dtype=np.int64)
        base[1:] = np.cumsum(w)
        return base * A

    def instance(self, i: int) -> Formula:
        c0, c1 = int(self.inst_clause_off[i]), int(self.inst_clause_off[i + 1])
        l0, l1 = int(self.clause_off[c0]), int(self.clause_off[c1])
        return Formula(int(self.n_vars[i]), self.lits[l0:l1].copy(), self.clause_off[c0:c1 + 1] - l0)

    def subset(self, idx: Sequence[int]) -> "Batch":
        return Batch.from_formulas([self.instance(int(i)) for i in idx])

    def manifest_hash(self) -> str:
        h = hashlib.sha256()
        for a in (self.lits, self.clause_off, self.inst_clause_off, self.n_vars):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()[:16]

    @staticmethod
    def from_formulas(fs: Sequence[Formula]) -> "Batch":
        lits = np.concatenate([f.lits for f in fs]).astype(np.int32) if fs else np.zeros(0, np.int32)
        offs = [np.zeros(1, np.int64)]
        icoff = np.zeros(len(fs) + 1, dtype=np.int64)
        base = 0
        for i, f in enumerate(fs):
            offs.append(f.clause_off[1:].astype(np.int64) + base)
            base += int(f.clause_off[-1])
            icoff[i + 1] = icoff[i] + f.n_clauses
        return Batch(lits, np.concatenate(offs), icoff, np.array([f.n_vars for f in fs], dtype=np.int32))

    @staticmethod
    def uniform(clauses: np.ndarray, n_vars: int) -> "Batch":
        """clauses: int32[N, m, k] literal codes, every instance with n_vars vars."""
        N, m, k = clauses.shape
        lits = np.ascontiguousarray(clauses.reshape(-1), dtype=np.int32)
        clause_off = np.arange(0, N * m * k + 1, k, dtype=np.int64)
        icoff = np.arange(0, N * m + 1, m, dtype=np.int64)
        return Batch(lits, clause_off, icoff, np.full(N, n_vars, dtype=np.int32))


def pack_rows(true_mask: np.ndarray, false_mask: np.ndarray) -> np.ndarray:
    """Pack boolean [.., n] true/false masks into literal-bitmap rows uint32[.., W]."""
    shp = true_mask.shape
    n = shp[-1]
    bits = np.zeros(shp[:-1] + (2 * n,), dtype=bool)
    bits[..., 1::2] = true_mask
    bits[..., 0::2] = false_mask
    W = words_per_row(n)
    pad = W * 32 - 2 * n
    if pad:
        bits = np.concatenate([bits, np.zeros(shp[:-1] + (pad,), dtype=bool)], axis=-1)
    by = np.packbits(bits, axis=-1, bitorder="little")
    return np.ascontiguousarray(by).view(np.uint32).reshape(shp[:-1] + (W,))


def unpack_rows(rows: np.ndarray, n: int):
    """Inverse of pack_rows: returns (true_mask, false_mask) boolean [.., n]."""
    by = np.ascontiguousarray(rows).view(np.uint8)
    bits = np.unpackbits(by, axis=-1, bitorder="little")[..., : 2 * n].astype(bool)
    return bits[..., 1::2], bits[..., 0::2]


def model_to_row(model: np.ndarray) -> np.ndarray:
    """model int8[n] (1 true, 0 false, -1 unassigned) -> one row."""
    m = np.asarray(model)
    return pack_rows(m == 1, m == 0)
