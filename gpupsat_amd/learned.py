"""Learned branching: labels from the plain solver, GPU training of the incidence-graph
network, and the proposals it feeds back to the solver (initial VSIDS scores, initial
phases, cube-split variables).  No reference counterpart.

Labels (per variable, from gpsat_solve on the training instances):
  * score target: conflict-analysis participation, log1p(bumps) / max over the instance;
  * phase target: the model's value of the var, for instances the solver proved SAT
    (masked out otherwise).
The network runs entirely in libgpsat.so (gnn.hip); this module only moves arrays.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import GpsatError, GpuBatch, SolveOpts, _check, _p, cube_vars_scored, lib
from .formula import Batch, unpack_rows

ROUNDS = 3


def param_count() -> int:
    n = ctypes.c_int32()
    _check(lib().gpsat_gnn_param_count(ctypes.byref(n)))
    return n.value


def init_params(seed: int = 0) -> np.ndarray:
    """Initial parameters for training (normal weights with variance 1/fan_in, small
    output layer)."""
    rng = np.random.default_rng(seed)
    H = 64
    shapes = [("u_l", (H,)), ("v_l", (H,)), ("u_c", (H,)), ("v_c", (H,)), ("W_c", (H, 2 * H)), ("b_c", (H,)),
              ("W_l", (H, 3 * H)), ("b_l", (H,)), ("W_o", (H, 2 * H)), ("b_o", (H,)), ("W_y", (2, H)), ("b_y", (2,))]
    out = []
    for name, shp in shapes:
        if name == "W_y":
            out.append(rng.normal(0, 0.05, size=shp))
        elif name.startswith("W"):
            # variance 1/fan_in keeps the summed messages from growing round after round
            out.append(rng.normal(0, np.sqrt(1.0 / shp[1]), size=shp))
        elif name in ("u_l", "u_c"):
            out.append(rng.normal(0, 0.3, size=shp))
        elif name in ("v_l", "v_c"):
            out.append(np.abs(rng.normal(0, 0.3, size=shp)))
        else:
            out.append(np.zeros(shp))
    theta = np.concatenate([o.reshape(-1) for o in out]).astype(np.float32)
    assert theta.size == param_count()
    return theta


class GnnModel:
    def __init__(self, theta: np.ndarray):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        if t.size != param_count():
            raise GpsatError("parameter vector has the wrong size")
        self.h = ctypes.c_void_p()
        _check(lib().gpsat_gnn_model_create(_p(t), ctypes.byref(self.h)))

    def params(self) -> np.ndarray:
        out = np.zeros(param_count(), dtype=np.float32)
        _check(lib().gpsat_gnn_model_params(self.h, _p(out)))
        return out

    def close(self):
        if self.h:
            lib().gpsat_gnn_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GnnGraph:
    def __init__(self, gb: GpuBatch, rounds: int = ROUNDS):
        self.gb = gb
        self.V = int(gb.batch.n_vars.astype(np.int64).sum())
        self.h = ctypes.c_void_p()
        _check(lib().gpsat_gnn_graph_create(gb.h, ctypes.c_int32(rounds), ctypes.byref(self.h)))

    def labels(self, target, phase, mask):
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (target, phase, mask)]
        _check(lib().gpsat_gnn_graph_labels(self.h, *(_p(x) for x in a)))

    def forward(self, model: GnnModel) -> np.ndarray:
        y = np.zeros((self.V, 2), dtype=np.float32)
        _check(lib().gpsat_gnn_forward(model.h, self.h, _p(y)))
        return y

    def loss_grad(self, model: GnnModel, alpha: float = 1.0):
        loss = ctypes.c_float()
        grad = np.zeros(param_count(), dtype=np.float32)
        y = np.zeros((self.V, 2), dtype=np.float32)
        _check(lib().gpsat_gnn_loss_grad(model.h, self.h, ctypes.c_float(alpha), ctypes.byref(loss), _p(grad), _p(y)))
        return loss.value, grad, y

    def train_step(self, model: GnnModel, lr: float, alpha: float = 1.0) -> float:
        loss = ctypes.c_float()
        _check(lib().gpsat_gnn_train_step(model.h, self.h, ctypes.c_float(alpha), ctypes.c_float(lr),
                                          ctypes.byref(loss)))
        return loss.value

    def close(self):
        if self.h:
            lib().gpsat_gnn_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def labels_from_solve(batch: Batch, res: dict, gb: GpuBatch):
    """Per-var targets from a solve run made with want_bumps=True."""
    bumps = res["var_bumps"].astype(np.float64)
    nv = batch.n_vars.astype(np.int64)
    base = np.concatenate([[0], np.cumsum(nv)])
    target = np.zeros(base[-1], dtype=np.float32)
    phase = np.zeros(base[-1], dtype=np.float32)
    mask = np.zeros(base[-1], dtype=np.float32)
    models = gb.models(res["model_rows"])
    for i in range(batch.n_inst):
        b = np.log1p(bumps[base[i]:base[i + 1]])
        mx = b.max()
        target[base[i]:base[i + 1]] = b / mx if mx > 0 else 0.0
        if res["status"][i] == 0:
            phase[base[i]:base[i + 1]] = (models[i] == 1).astype(np.float32)
            mask[base[i]:base[i + 1]] = 1.0
    return target, phase, mask


def generate_labels(batch: Batch, opts: Optional[SolveOpts] = None):
    """Solve the batch with the plain solver and turn its counters into labels."""
    gb = GpuBatch(batch)
    res = gb.solve(opts if opts is not None else SolveOpts.default(max_conflicts=20000, max_ms=20000),
                   want_bumps=True)
    lab = labels_from_solve(batch, res, gb)
    return gb, res, lab


def train(graphs: Sequence[GnnGraph], theta: Optional[np.ndarray] = None, epochs: int = 10, lr: float = 2e-3,
          alpha: float = 1.0, log=None) -> np.ndarray:
    """Adam over the given labelled graphs (each a batch of instances).  Returns the
    trained FP32 parameters."""
    model = GnnModel(theta if theta is not None else init_params(0))
    history = []
    for ep in range(epochs):
        losses = [g.train_step(model, lr, alpha) for g in graphs]
        history.append(float(np.mean(losses)))
        if log:
            log(f"epoch {ep}: loss {history[-1]:.5f}")
    out = model.params()
    model.close()
    train.history = history
    return out


def train_sets(sets, theta: Optional[np.ndarray] = None, epochs: int = 10, lr: float = 2e-3, alpha: float = 1.0,
               log=None) -> np.ndarray:
    """Adam over labelled batches [(GpuBatch, (target, phase, mask)), ...].  Each batch's
    graph (its incidence structure and labels) is built once and kept for every epoch; the
    activations of a step live in the model's arena, so device memory holds one batch's
    activations however many batches there are."""
    model = GnnModel(theta if theta is not None else init_params(0))
    graphs = []
    for gb, lab in sets:
        g = GnnGraph(gb)
        g.labels(*lab)
        graphs.append(g)
    history = []
    for ep in range(epochs):
        losses = [g.train_step(model, lr, alpha) for g in graphs]
        history.append(float(np.mean(losses)))
        if log:
            log(f"epoch {ep}: loss {history[-1]:.5f}")
    out = model.params()
    for g in graphs:
        g.close()
    model.close()
    train_sets.history = history
    return out


# Forward FLOPs of one instance as the survey counts them: 3 rounds x (2n + m) nodes x
# 2 * 64 * 64 * 2; a training step (forward + backward) is three times that.
def forward_flops(n: int, m: int, rounds: int = ROUNDS) -> float:
    return rounds * (2 * n + m) * 2 * 64 * 64 * 2.0


def time_train_steps(n_inst: int = 512, n: int = 300, m: int = 1278, steps: int = 20, seed: int = 5):
    """Wall-clock time of training steps (forward, loss, backward, Adam) on a batch of
    n_inst random 3-SAT instances of the learned-part evaluation shape, with random labels.
    Returns (ms per step, algorithmic FLOP/s by forward_flops)."""
    import time
    from . import workloads as W
    b = W.random_3sat_batch(n_inst, n, m, seed=seed)
    rng = np.random.default_rng(seed)
    V = n_inst * n
    with GpuBatch(b) as gb:
        model = GnnModel(init_params(seed))
        g = GnnGraph(gb)
        g.labels(rng.random(V), (rng.random(V) < 0.5).astype(np.float32), (rng.random(V) < 0.5).astype(np.float32))
        for _ in range(3):
            g.train_step(model, 1e-3)
        t0 = time.perf_counter()
        for _ in range(steps):
            g.train_step(model, 1e-3)
        dt = (time.perf_counter() - t0) / steps
        g.close()
        model.close()
    return dt * 1e3, 3 * n_inst * forward_flops(n, m) / dt


def propose(model: GnnModel, gb: GpuBatch, rounds: int = ROUNDS):
    """Network outputs for a batch: (act_init, phase_init, scores per instance)."""
    g = GnnGraph(gb, rounds)
    y = g.forward(model)
    g.close()
    score = y[:, 0].astype(np.float32)
    phase = (y[:, 1] > 0).astype(np.int8)
    nv = gb.batch.n_vars.astype(np.int64)
    base = np.concatenate([[0], np.cumsum(nv)])
    # initial VSIDS scores: the predicted participation, shifted to be non-negative
    act = np.zeros_like(score)
    per = []
    for i in range(gb.n_inst):
        s = score[base[i]:base[i + 1]]
        act[base[i]:base[i + 1]] = np.maximum(s - s.min(), 0.0)
        per.append(s)
    return act, phase, per


def learned_cubes(per_scores: List[np.ndarray], depth: int):
    """Cube-split vars: the depth highest-scoring vars of every instance."""
    return [cube_vars_scored(s, depth) for s in per_scores]
