"""The learned branching network at the shape of the learned-part workload (n=300,
m=1278): forward and gradient parity with the FP64 oracle, and what the committed
trained weights gain over the plain method on held-out instances.

Plain method: cube-split variables from MaxClauseJobChooser's order
(src/JobsManager/JobChooser.cu:52-90 with VariableChooser.cu:23-40) and the solver's own
VSIDS from zero scores.  Learned: the network's top-scored variables as the cube split,
plus its initial VSIDS scores and phases.  Both arms run the same cubes count and the same
per-cube conflict budget on the same GPU, and their verdicts must be identical wherever both
decide.  The gain is measured on every instance, decided or not (PAR-2 over conflicts: an
undecided instance costs twice the whole budget), so it is not a statistic of survivors."""
import os

import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd import learned as Lr
from gpupsat_amd import train_branching as T
from gpupsat_amd import workloads as W
from oracle import gnn_ref as R

pytestmark = pytest.mark.gpu

DEPTH = 4                 # 16 cubes per instance
CUBE_CONFLICTS = 40_000   # per-cube conflict budget of both arms


def blocks():
    o = 0
    for name, shp in R.PARAM_SHAPES:
        k = int(np.prod(shp))
        yield name, slice(o, o + k)
        o += k


def config5_batch(n_inst, seed):
    return W.random_3sat_batch(n_inst, 300, 1278, seed=seed)


def test_gnn_forward_parity_n300():
    print("[learned] forward parity at n=300 m=1278 (4 instances)", flush=True)
    b = config5_batch(4, 778)
    g = R.Graph(b)
    theta = np.load(T.DEFAULT_WEIGHTS)
    with G.GpuBatch(b) as gb:
        m = Lr.GnnModel(theta)
        gg = Lr.GnnGraph(gb)
        y = gg.forward(m)
        m.close()
        gg.close()
    ref = R.forward(theta.astype(np.float64), g, quant=True)
    rel = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"relative L2 {rel:.2e}, max abs / max |y| {err:.2e}")
    # same BF16 rounding points on both sides; FP32 (GPU) vs FP64 (oracle) accumulation
    # moves an occasional pre-rounding value across a BF16 rounding boundary (2^-8 relative)
    assert rel <= 2e-3
    assert err <= 2e-2


def test_gnn_gradient_parity_n300():
    print("[learned] gradient parity at n=300 m=1278 (4 instances)", flush=True)
    b = config5_batch(4, 779)
    g = R.Graph(b)
    rng = np.random.default_rng(3)
    target, phase = rng.random(g.V), (rng.random(g.V) < 0.5).astype(float)
    mask = (rng.random(g.V) < 0.6).astype(float)
    theta = np.load(T.DEFAULT_WEIGHTS)
    with G.GpuBatch(b) as gb:
        m = Lr.GnnModel(theta)
        gg = Lr.GnnGraph(gb)
        gg.labels(target, phase, mask)
        loss, grad, _ = gg.loss_grad(m, alpha=1.0)
        loss2, grad2, _ = gg.loss_grad(m, alpha=1.0)
        m.close()
        gg.close()
    rl, _, rgrad = R.backward(theta.astype(np.float64), g, target, phase, mask, quant=True)
    assert abs(loss - rl) <= 1e-4 * max(1.0, abs(rl))
    # the gradient reductions run in a fixed order: two evaluations are bit-identical
    np.testing.assert_array_equal(grad, grad2)
    assert loss == loss2
    for name, sl in blocks():
        a, r = grad[sl], rgrad[sl]
        rel = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-12)
        print(f"{name}: relative L2 {rel:.2e}")
        assert rel <= 5e-3, (name, rel)


def test_learned_split_beats_plain_on_held_out():
    n_inst = 64
    print(f"[learned] A/B on {n_inst} held-out n=300 instances, {1 << DEPTH} cubes each, "
          f"{CUBE_CONFLICTS} conflicts per cube", flush=True)
    assert os.path.exists(T.DEFAULT_WEIGHTS)
    theta = np.load(T.DEFAULT_WEIGHTS)
    b = config5_batch(n_inst, 777)            # random 3-SAT family; training used another family
    opts = G.SolveOpts.default(max_conflicts=CUBE_CONFLICTS, max_ms=60_000.0)
    with G.GpuBatch(b) as gb:
        model = Lr.GnnModel(theta)
        act, phase, per = Lr.propose(model, gb)
        model.close()
        plain_vars = [G.cube_vars(b.instance(i), strategy=1)[:DEPTH] for i in range(n_inst)]
        learned_vars = Lr.learned_cubes(per, DEPTH)
        jp = G.cube_jobs(b, plain_vars)
        jl = G.cube_jobs(b, learned_vars)
        rp = gb.solve(opts, jobs=jp)
        rl = gb.solve(opts, jobs=jl, act_init=act, phase_init=phase)
        mp, ml = gb.models(rp["model_rows"]), gb.models(rl["model_rows"])
    sp, sl = rp["status"], rl["status"]
    both = (sp != G.UNDEF) & (sl != G.UNDEF)
    assert (sp[both] == sl[both]).all(), "verdicts differ"
    for i in range(n_inst):
        if sp[i] == G.SAT:
            assert rp["verified"][i] == 1
        if sl[i] == G.SAT:
            assert rl["verified"][i] == 1
    penalty = 2 * (1 << DEPTH) * CUBE_CONFLICTS

    def par2(res, jobs, st):
        conf = np.zeros(n_inst)
        np.add.at(conf, jobs[0], res["job_stats"][:, 1])
        return np.where(st != G.UNDEF, conf, penalty)

    p_plain, p_learned = par2(rp, jp, sp), par2(rl, jl, sl)
    print(f"decided plain {int((sp != 2).sum())}, learned {int((sl != 2).sum())}, both {int(both.sum())}; "
          f"PAR-2 conflicts mean plain {p_plain.mean():.0f}, learned {p_learned.mean():.0f} "
          f"(ratio {p_learned.mean() / p_plain.mean():.3f}); "
          f"median plain {np.median(p_plain):.0f}, learned {np.median(p_learned):.0f}")
    assert (sl != G.UNDEF).sum() >= (sp != G.UNDEF).sum()
    assert p_learned.mean() < p_plain.mean()
