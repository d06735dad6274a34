"""CPU oracle of the learned branching network (TEST INFRASTRUCTURE ONLY).

No reference counterpart: the network has no equivalent in the reference.  This module restates
the network's mathematics in numpy FP64 so the GPU kernels (gpupsat_amd/csrc/gnn.hip)
can be checked: forward values and analytic gradients, with the GPU's BF16 rounding
points reproduced when quant=True.  Its own backward is pinned by central finite
differences with quant=False (tests/test_gnn_oracle.py).

Graph: literal nodes (2n per instance) and clause nodes (m), an edge per occurrence.
    H_l^0 = occ(l)/8 * u_l + v_l                 H_c^0 = len(c)/4 * u_c + v_c
    for t = 1..R:
        M_c = 1/4 * sum_{l in c} H_l^{t-1}
        H_c^t = relu([H_c^{t-1} | M_c] W_c^T + b_c)
        M_l = 1/8 * sum_{c ni l} H_c^t
        H_l^t = relu([H_l^{t-1} | M_l | H_{l^1}^{t-1}] W_l^T + b_l)
    Z_v = [H_{2v+1}^R | H_{2v}^R];  Rv = relu(Z W_o^T + b_o);  Y = Rv W_y^T + b_y
    score_v = Y[v, 0], phase_logit_v = Y[v, 1]
    loss = mean_v [ (score_v - t_v)^2 + alpha * mask_v * BCE(phase_logit_v, p_v) ]
BF16 points (quant=True): the weights W_c, W_l, W_o, W_y and every input of the
three matrix products that run on MFMA ([H|M], [H|M|H'], Z forward; the upstream
gradients dP fed to the backward products) are rounded to BF16 (round to nearest even);
the 2-output head, biases, embeddings and aggregations stay FP32/FP64.
"""
from __future__ import annotations

import numpy as np

HID = 64
ROUNDS = 3
ALPHA = 1.0
PARAM_SHAPES = [
    ("u_l", (HID,)), ("v_l", (HID,)), ("u_c", (HID,)), ("v_c", (HID,)),
    ("W_c", (HID, 2 * HID)), ("b_c", (HID,)),
    ("W_l", (HID, 3 * HID)), ("b_l", (HID,)),
    ("W_o", (HID, 2 * HID)), ("b_o", (HID,)),
    ("W_y", (2, HID)), ("b_y", (2,)),
]


def n_params() -> int:
    return int(sum(np.prod(s) for _, s in PARAM_SHAPES))


def unflatten(theta: np.ndarray) -> dict:
    p, o = {}, 0
    for name, shp in PARAM_SHAPES:
        k = int(np.prod(shp))
        p[name] = theta[o:o + k].reshape(shp)
        o += k
    return p


def flatten(p: dict) -> np.ndarray:
    return np.concatenate([np.asarray(p[name], dtype=np.float64).reshape(-1) for name, _ in PARAM_SHAPES])


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in PARAM_SHAPES:
        if name.startswith("W"):
            p[name] = rng.normal(0, np.sqrt(2.0 / shp[1]), size=shp)
        elif name in ("u_l", "u_c"):
            p[name] = rng.normal(0, 0.5, size=shp)
        elif name in ("v_l", "v_c"):
            p[name] = np.abs(rng.normal(0, 0.5, size=shp))
        else:
            p[name] = np.zeros(shp)
    return flatten(p).astype(np.float32)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


class Graph:
    """Incidence structure of a batch (concatenated instances)."""

    def __init__(self, batch):
        lits, coff, icoff, nv = batch.lits, batch.clause_off, batch.inst_clause_off, batch.n_vars
        self.n_inst = batch.n_inst
        self.var_base = np.concatenate([[0], np.cumsum(nv.astype(np.int64))])
        self.V = int(self.var_base[-1])
        self.L = 2 * self.V
        self.C = int(icoff[-1])
        inst_of_clause = np.repeat(np.arange(batch.n_inst), np.diff(icoff))
        clen = np.diff(coff).astype(np.int64)
        edge_clause = np.repeat(np.arange(self.C), clen)
        edge_inst = inst_of_clause[edge_clause]
        edge_lit = 2 * self.var_base[edge_inst] + lits.astype(np.int64)   # global literal node
        self.edge_clause, self.edge_lit = edge_clause, edge_lit
        self.clen = clen
        self.occ = np.bincount(edge_lit, minlength=self.L)


def _relu(x):
    return np.maximum(x, 0.0)


def forward(theta, g: Graph, quant: bool, rounds: int = ROUNDS, keep: bool = False):
    p = unflatten(np.asarray(theta, dtype=np.float64))
    q = bf16 if quant else (lambda x: np.asarray(x, dtype=np.float64))
    Wc, Wl, Wo, Wy = q(p["W_c"]), q(p["W_l"]), q(p["W_o"]), q(p["W_y"])
    Hl = (g.occ / 8.0)[:, None] * p["u_l"][None] + p["v_l"][None]
    Hc = (g.clen / 4.0)[:, None] * p["u_c"][None] + p["v_c"][None]
    cache = {"Hl": [Hl], "Hc": [Hc], "Xc": [], "Pc": [], "Xl": [], "Pl": []}
    flip = np.arange(g.L) ^ 1
    for _ in range(rounds):
        Mc = np.zeros((g.C, HID))
        np.add.at(Mc, g.edge_clause, Hl[g.edge_lit])
        Mc *= 0.25
        Xc = q(np.concatenate([Hc, Mc], axis=1))
        Pc = Xc @ Wc.T + p["b_c"][None]
        Hc = _relu(Pc)
        Ml = np.zeros((g.L, HID))
        np.add.at(Ml, g.edge_lit, Hc[g.edge_clause])
        Ml *= 0.125
        Xl = q(np.concatenate([Hl, Ml, Hl[flip]], axis=1))
        Pl = Xl @ Wl.T + p["b_l"][None]
        Hl = _relu(Pl)
        for k, v in (("Xc", Xc), ("Pc", Pc), ("Xl", Xl), ("Pl", Pl)):
            cache[k].append(v)
        cache["Hl"].append(Hl)
        cache["Hc"].append(Hc)
    Z = q(np.concatenate([Hl[1::2], Hl[0::2]], axis=1))
    Po = Z @ Wo.T + p["b_o"][None]
    Rv = _relu(Po)
    Y = Rv @ Wy.T + p["b_y"][None]
    cache.update(Z=Z, Po=Po, Rv=Rv, W=(Wc, Wl, Wo, Wy), p=p)
    return (Y, cache) if keep else Y


def loss_and_dY(Y, target, phase, mask, alpha: float = ALPHA):
    V = Y.shape[0]
    s, z = Y[:, 0], Y[:, 1]
    sig = 1.0 / (1.0 + np.exp(-z))
    bce = np.maximum(z, 0) - z * phase + np.log1p(np.exp(-np.abs(z)))
    loss = float(np.sum((s - target) ** 2 + alpha * mask * bce) / V)
    dY = np.zeros_like(Y)
    dY[:, 0] = 2.0 * (s - target) / V
    dY[:, 1] = alpha * mask * (sig - phase) / V
    return loss, dY


def backward(theta, g: Graph, target, phase, mask, quant: bool, rounds: int = ROUNDS, alpha: float = ALPHA):
    """Returns (loss, Y, flat gradient)."""
    Y, c = forward(theta, g, quant, rounds, keep=True)
    loss, dY = loss_and_dY(Y, target, phase, mask, alpha)
    q = bf16 if quant else (lambda x: np.asarray(x, dtype=np.float64))
    p = c["p"]
    Wc, Wl, Wo, Wy = c["W"]
    grad = {name: np.zeros(shp) for name, shp in PARAM_SHAPES}
    grad["W_y"] = dY.T @ c["Rv"]
    grad["b_y"] = dY.sum(0)
    dRv = dY @ Wy
    dPo = dRv * (c["Po"] > 0)
    grad["W_o"] = q(dPo).T @ c["Z"]
    grad["b_o"] = dPo.sum(0)
    dZ = q(dPo) @ Wo
    dHl = np.zeros((g.L, HID))
    dHl[1::2] += dZ[:, :HID]
    dHl[0::2] += dZ[:, HID:]
    dHc = np.zeros((g.C, HID))
    flip = np.arange(g.L) ^ 1
    for t in range(rounds - 1, -1, -1):
        dPl = dHl * (c["Pl"][t] > 0)
        grad["W_l"] += q(dPl).T @ c["Xl"][t]
        grad["b_l"] += dPl.sum(0)
        dXl = q(dPl) @ Wl
        dHl_prev = dXl[:, :HID] + dXl[flip, 2 * HID:]
        dMl = dXl[:, HID:2 * HID]
        np.add.at(dHc, g.edge_clause, 0.125 * dMl[g.edge_lit])
        dPc = dHc * (c["Pc"][t] > 0)
        grad["W_c"] += q(dPc).T @ c["Xc"][t]
        grad["b_c"] += dPc.sum(0)
        dXc = q(dPc) @ Wc
        dHc = dXc[:, :HID]
        dMc = dXc[:, HID:]
        np.add.at(dHl_prev, g.edge_lit, 0.25 * dMc[g.edge_clause])
        dHl = dHl_prev
    grad["u_l"] = ((g.occ / 8.0)[:, None] * dHl).sum(0)
    grad["v_l"] = dHl.sum(0)
    grad["u_c"] = ((g.clen / 4.0)[:, None] * dHc).sum(0)
    grad["v_c"] = dHc.sum(0)
    return loss, Y, flatten(grad)


def loss_only(theta, g, target, phase, mask, quant=False, rounds=ROUNDS, alpha=ALPHA):
    Y = forward(theta, g, quant, rounds)
    return loss_and_dY(Y, target, phase, mask, alpha)[0]
