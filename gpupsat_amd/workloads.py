"""Seeded synthetic workloads: uniform random k-SAT, graph colouring and pigeonhole.

Every generator takes an explicit seed; instance i of a family is generated from
``np.random.SeedSequence([family_id, seed])`` so a workload is reproducible from
(family, seed, sizes) alone.  The reference ships no generator.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from .formula import Batch, Formula, pack_rows, words_per_row

FAMILY_RANDOM = 1
FAMILY_ASSIGN = 2
FAMILY_MIXED = 3
FAMILY_COLOR = 4
FAMILY_PHP = 5
FAMILY_TRAIN = 6


def _rng(family: int, seed: int) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence([family, seed]))


def random_ksat_clauses(rng: np.random.Generator, N: int, n: int, m: int, k: int) -> np.ndarray:
    """int32[N, m, k]: each clause has k distinct vars drawn uniformly, each sign
    Bernoulli(0.5)."""
    v = rng.integers(0, n, size=(N, m, k), dtype=np.int64)
    while True:
        s = np.sort(v, axis=-1)
        bad = (s[..., 1:] == s[..., :-1]).any(axis=-1)
        nb = int(bad.sum())
        if nb == 0:
            break
        v[bad] = rng.integers(0, n, size=(nb, k), dtype=np.int64)
    sign = rng.integers(0, 2, size=(N, m, k), dtype=np.int64)
    return (2 * v + sign).astype(np.int32)


def random_3sat_batch(N: int, n: int, m: int, seed: int = 0) -> Batch:
    """Uniform random 3-SAT (n vars, m clauses), N instances."""
    return Batch.uniform(random_ksat_clauses(_rng(FAMILY_RANDOM, seed), N, n, m, 3), n)


def random_assignments(batch: Batch, A: int, p: float, seed: int = 0) -> np.ndarray:
    """A partial assignments per instance: each var assigned with probability p,
    uniform polarity.  Returns the packed rows, uint32[sum_i A*W_i]."""
    rng = _rng(FAMILY_ASSIGN, seed)
    nv = batch.n_vars
    if (nv == nv[0]).all():
        n = int(nv[0])
        assigned = rng.random((batch.n_inst, A, n)) < p
        pol = rng.random((batch.n_inst, A, n)) < 0.5
        return pack_rows(assigned & pol, assigned & ~pol).reshape(-1)
    out = []
    for i in range(batch.n_inst):
        n = int(nv[i])
        assigned = rng.random((A, n)) < p
        pol = rng.random((A, n)) < 0.5
        out.append(pack_rows(assigned & pol, assigned & ~pol).reshape(-1))
    return np.concatenate(out)


def mixed_ksat_batch(N: int, seed: int = 0) -> Tuple[Batch, np.ndarray]:
    """Mixed random k-SAT: k uniform in {3,4,5} at clause ratio {4.26, 9.9, 21.1},
    n uniform in [100, 400], m = round(ratio * n).  Returns (batch, k per instance)."""
    rng = _rng(FAMILY_MIXED, seed)
    ratios = {3: 4.26, 4: 9.9, 5: 21.1}
    ks = rng.integers(3, 6, size=N)
    ns = rng.integers(100, 401, size=N)
    fs = []
    for i in range(N):
        k, n = int(ks[i]), int(ns[i])
        m = int(round(ratios[k] * n))
        cl = random_ksat_clauses(rng, 1, n, m, k)[0]
        fs.append(Formula(n, cl.reshape(-1), np.arange(0, m * k + 1, k, dtype=np.int64)))
    return Batch.from_formulas(fs), ks


def coloring_formula(n_nodes: int, edges: np.ndarray, colors: int = 4) -> Formula:
    """Graph colouring: var(node, c) = node*colors + c; at-least-one colour per node,
    pairwise at-most-one, and for every edge and colour the two ends differ."""
    cl: List[List[int]] = []
    pos = lambda v: 2 * v + 1
    neg = lambda v: 2 * v
    for u in range(n_nodes):
        cl.append([pos(u * colors + c) for c in range(colors)])
    for u in range(n_nodes):
        for a in range(colors):
            for b in range(a + 1, colors):
                cl.append([neg(u * colors + a), neg(u * colors + b)])
    for (u, w) in edges:
        for c in range(colors):
            cl.append([neg(int(u) * colors + c), neg(int(w) * colors + c)])
    return Formula.from_clauses(n_nodes * colors, cl)


def er_graph(rng: np.random.Generator, n: int, avg_deg: float) -> np.ndarray:
    p = avg_deg / (n - 1)
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(len(iu)) < p
    return np.stack([iu[keep], ju[keep]], axis=1)


def coloring_batch(N: int, n_nodes: int = 150, deg_lo: float = 4.2, deg_hi: float = 4.8,
                   seed: int = 0) -> Batch:
    rng = _rng(FAMILY_COLOR, seed)
    fs = []
    for _ in range(N):
        d = rng.uniform(deg_lo, deg_hi)
        fs.append(coloring_formula(n_nodes, er_graph(rng, n_nodes, d)))
    return Batch.from_formulas(fs)


def pigeonhole_formula(pigeons: int, holes: int) -> Formula:
    """PHP(p, h): var(p, h) = p*holes + h; every pigeon in some hole, no two pigeons
    share a hole.  Unsatisfiable whenever pigeons > holes."""
    cl: List[List[int]] = []
    for p in range(pigeons):
        cl.append([2 * (p * holes + h) + 1 for h in range(holes)])
    for h in range(holes):
        for p in range(pigeons):
            for q in range(p + 1, pigeons):
                cl.append([2 * (p * holes + h), 2 * (q * holes + h)])
    return Formula.from_clauses(pigeons * holes, cl)


def pigeonhole_batch(N: int, pigeons: int = 9, holes: int = 8) -> Batch:
    f = pigeonhole_formula(pigeons, holes)
    return Batch.from_formulas([f] * N)


def training_3sat_batch(N: int, n_lo: int = 100, n_hi: int = 300, r_lo: float = 4.2,
                        r_hi: float = 4.3, seed: int = 0) -> Batch:
    """Random 3-SAT with n uniform in [n_lo, n_hi] and ratio uniform in [r_lo, r_hi]."""
    rng = _rng(FAMILY_TRAIN, seed)
    fs = []
    for _ in range(N):
        n = int(rng.integers(n_lo, n_hi + 1))
        m = int(round(rng.uniform(r_lo, r_hi) * n))
        cl = random_ksat_clauses(rng, 1, n, m, 3)[0]
        fs.append(Formula(n, cl.reshape(-1), np.arange(0, 3 * m + 1, 3, dtype=np.int64)))
    return Batch.from_formulas(fs)


__all__ = [
    "random_ksat_clauses", "random_3sat_batch", "random_assignments", "mixed_ksat_batch",
    "coloring_formula", "coloring_batch", "er_graph", "pigeonhole_formula", "pigeonhole_batch",
    "training_3sat_batch", "words_per_row",
]
