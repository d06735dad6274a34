"""Pins the CPU oracle of the learned branching network: its analytic gradient equals
central finite differences (FP64, no quantization), and the BF16 rounding helper
matches the IEEE definition on known values."""
import numpy as np

from gpupsat_amd import workloads as W
from oracle import gnn_ref as R


def small_problem(seed=0, N=3):
    b = W.random_3sat_batch(N, 12, 40, seed=seed)
    g = R.Graph(b)
    rng = np.random.default_rng(seed)
    target = rng.random(g.V)
    phase = (rng.random(g.V) < 0.5).astype(float)
    mask = (rng.random(g.V) < 0.7).astype(float)
    return g, target, phase, mask


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 3.0e38, -2.5, 1.0 + 2**-8 + 2**-9], dtype=np.float32)
    got = R.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even mantissa (1.0)
    assert got[0] == 1.0 and got[1] == 1.0
    assert got[2] == np.float32(1.0078125)
    assert got[4] == -2.5
    assert got[5] == np.float32(1.0078125)


def test_gradient_matches_finite_differences():
    g, target, phase, mask = small_problem()
    theta = R.init_params(1).astype(np.float64)
    loss, Y, grad = R.backward(theta, g, target, phase, mask, quant=False)
    rng = np.random.default_rng(3)
    idx = rng.choice(len(theta), 60, replace=False)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in R.PARAM_SHAPES])
    # every parameter block is probed at least once
    idx = np.unique(np.concatenate([idx, offs[:-1]]))
    eps = 1e-6
    for i in idx:
        tp = theta.copy(); tp[i] += eps
        tm = theta.copy(); tm[i] -= eps
        fd = (R.loss_only(tp, g, target, phase, mask) - R.loss_only(tm, g, target, phase, mask)) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_quantized_forward_close_to_exact():
    g, target, phase, mask = small_problem(seed=2)
    theta = R.init_params(4)
    Y0 = R.forward(theta, g, quant=False)
    Y1 = R.forward(theta, g, quant=True)
    assert np.abs(Y1 - Y0).max() <= 0.05 * (1 + np.abs(Y0).max())
