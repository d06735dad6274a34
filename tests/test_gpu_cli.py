"""The gpupsat command line on the GPU: verdict and model lines of every reference
fixture in sequential (-b 1 -t 1) and cube (-b/-t > 1) modes."""
import glob
import os
import subprocess

import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd import workloads as W

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CNFS = sorted(glob.glob(os.path.join(GOLDEN, "cnf", "*.cnf")))


def run(*args, cwd=None):
    r = subprocess.run([G.CLI_PATH, *args], capture_output=True, text=True, cwd=cwd, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.strip().splitlines()


def verdict(lines):
    for l in lines:
        if l in ("SATISFIABLE", "UNSATISFIABLE", "UNDEFINED"):
            return l
    raise AssertionError(lines)


@pytest.mark.parametrize("mode", [["-b", "1", "-t", "1"], ["-b", "2", "-t", "2"], []], ids=["seq", "cubes4", "default"])
def test_cli_reference_fixtures(mode):
    print(f"[cli] fixtures mode={mode}", flush=True)
    for p in CNFS:
        exp = {"SAT": "SATISFIABLE", "UNSAT": "UNSATISFIABLE"}[open(p[:-4] + ".expected").read().strip()]
        lines = run("-i", p, *mode, "--timeout-ms", "5000")
        assert verdict(lines) == exp, p
        if exp == "SATISFIABLE":
            assert lines[-1] == "Solution was verified"
            assert lines[-2].endswith(" 0")


def test_cli_random_instance_model(orc, tmp_path):
    print("[cli] random 3-SAT model line", flush=True)
    b = W.random_3sat_batch(6, 60, 255, seed=4)
    for i in range(b.n_inst):
        f = b.instance(i)
        p = tmp_path / f"r{i}.cnf"
        p.write_text(f.to_dimacs())
        lines = run("-i", str(p), "-b", "1", "-t", "1", "-v", "0", "-l", cwd=tmp_path)
        st = orc.solve_cdcl(f)[0]
        assert verdict(lines) == ("SATISFIABLE" if st == 0 else "UNSATISFIABLE")
        if st == 0:
            model = np.array([1 if int(x) > 0 else 0 for x in lines[-2].split()[:-1]], dtype=np.int8)
            assert orc.verify_model(f, model)
    assert len((tmp_path / "autolog.txt").read_text().splitlines()) == b.n_inst


def test_cli_trajectory_mode(orc, tmp_path):
    """--trajectory: a sequential run follows the reference's own search; its counters and
    verdict equal the oracle restatement's on the preprocessed formula, MAX_ITERATIONS 1000
    included (n=100 instances mostly end UNDEFINED, like the reference)."""
    print("[cli] trajectory mode", flush=True)
    import re
    fs = [W.random_3sat_batch(1, 50, 213, seed=600 + i).instance(0) for i in range(4)]
    fs += [W.random_3sat_batch(1, 100, 426, seed=700 + i).instance(0) for i in range(3)]
    for i, f in enumerate(fs):
        p = tmp_path / f"t{i}.cnf"
        p.write_text(f.to_dimacs())
        lines = run("-i", str(p), "-b", "1", "-t", "1", "--trajectory")
        pst, g, solved = orc.preprocess(f)
        if pst != G.UNDEF:
            continue
        dead = np.zeros(f.n_vars, dtype=np.uint8)
        dead[np.asarray(solved, dtype=np.int64) >> 1] = 1
        st, _, stats = orc.solve_ref(g, dead=dead, max_iterations=1000)
        assert verdict(lines) == ["SATISFIABLE", "UNSATISFIABLE", "UNDEFINED"][st]
        m = [re.match(r"Decisions (\d+), conflicts (\d+), learned (\d+), iterations (\d+), restarts (\d+)", l)
             for l in lines]
        got = [list(map(int, x.groups())) for x in m if x][0]
        assert got == stats.tolist(), (i, got, stats)


def test_cli_large_formula(orc, tmp_path):
    """A 6,000-variable CNF (per-var solver state larger than the LDS) through the CLI:
    SATISFIABLE with a verified model line, not a capacity error."""
    print("[cli] n=6000 formula", flush=True)
    f = W.random_3sat_batch(1, 6000, 18000, seed=61).instance(0)
    p = tmp_path / "big.cnf"
    p.write_text(f.to_dimacs())
    lines = run("-i", str(p), "-b", "1", "-t", "1", "-v", "0")
    assert verdict(lines) == "SATISFIABLE"
    assert lines[-1] == "Solution was verified"
    model = np.array([1 if int(x) > 0 else 0 for x in lines[-2].split()[:-1]], dtype=np.int8)
    assert orc.verify_model(f, model)
