"""The learned branching network on the GPU: forward values and gradients of the HIP
kernels (BF16 MFMA, FP32 accumulate) against the CPU FP64 oracle with the same BF16
rounding points, and training that lowers the loss."""
import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd import learned as Lr
from gpupsat_amd import workloads as W
from gpupsat_amd.formula import Batch
from oracle import gnn_ref as R

pytestmark = pytest.mark.gpu


def problem(seed=0):
    fs = [W.random_3sat_batch(1, 20 + 7 * i, 85 + 30 * i, seed=seed + i).instance(0) for i in range(5)]
    fs.append(W.coloring_batch(1, n_nodes=12, seed=seed).instance(0))
    fs.append(W.mixed_ksat_batch(1, seed=seed)[0].instance(0))
    b = Batch.from_formulas(fs)
    g = R.Graph(b)
    rng = np.random.default_rng(seed)
    return b, g, rng.random(g.V), (rng.random(g.V) < 0.5).astype(float), (rng.random(g.V) < 0.6).astype(float)


def blocks():
    o = 0
    for name, shp in R.PARAM_SHAPES:
        k = int(np.prod(shp))
        yield name, slice(o, o + k)
        o += k


def test_param_layout_matches_oracle():
    assert Lr.param_count() == R.n_params()


def test_forward_matches_oracle():
    print("[gnn] forward parity", flush=True)
    b, g, *_ = problem(0)
    theta = R.init_params(3)
    with G.GpuBatch(b) as gb:
        m = Lr.GnnModel(theta)
        gg = Lr.GnnGraph(gb)
        y = gg.forward(m)
    ref = R.forward(theta.astype(np.float64), g, quant=True)
    err = np.abs(y - ref).max()
    scale = np.abs(ref).max()
    rel = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print("max abs err", err, "scale", scale, "relative L2", rel)
    # same BF16 rounding points; FP32 vs FP64 accumulation may flip an occasional rounding
    assert rel <= 2e-3
    assert err <= 2e-2 * max(1.0, scale)


def test_gradient_matches_oracle():
    print("[gnn] gradient parity", flush=True)
    b, g, target, phase, mask = problem(1)
    theta = R.init_params(5)
    with G.GpuBatch(b) as gb:
        m = Lr.GnnModel(theta)
        gg = Lr.GnnGraph(gb)
        gg.labels(target, phase, mask)
        loss, grad, y = gg.loss_grad(m, alpha=1.0)
    rl, ry, rgrad = R.backward(theta.astype(np.float64), g, target, phase, mask, quant=True)
    assert abs(loss - rl) <= 1e-3 * max(1.0, abs(rl))
    for name, sl in blocks():
        a, r = grad[sl], rgrad[sl]
        rel = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-12)
        print(name, "rel err", rel)
        assert rel <= 5e-3, (name, rel)


def test_training_lowers_loss():
    print("[gnn] training", flush=True)
    b, g, target, phase, mask = problem(2)
    with G.GpuBatch(b) as gb:
        m = Lr.GnnModel(R.init_params(7))
        gg = Lr.GnnGraph(gb)
        gg.labels(target, phase, mask)
        first = gg.train_step(m, lr=3e-3)
        for _ in range(150):
            last = gg.train_step(m, lr=3e-3)
        theta = m.params()
    assert last < 0.8 * first
    # the trained parameters give the oracle the same lower loss
    assert R.loss_only(theta.astype(np.float64), g, target, phase, mask, quant=True) < 0.85 * first
