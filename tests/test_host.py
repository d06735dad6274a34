"""Host side of the boundary (DIMACS reader, preprocessing, JobChooser, CLI plumbing),
checked against the oracle without a GPU."""
import glob
import os
import subprocess

import numpy as np
import pytest

import gpupsat_amd as G
from gpupsat_amd import workloads as W
from gpupsat_amd.formula import Formula, read_dimacs

from conftest import GOLDEN, ROOT

CNFS = sorted(glob.glob(os.path.join(GOLDEN, "cnf", "*.cnf")))


def random_formula(rng, n, m_lo=3, m_hi=30, kmax=4):
    cl = []
    for _ in range(int(rng.integers(m_lo, m_hi))):
        k = int(rng.integers(0 if rng.random() < 0.02 else 1, kmax + 1))
        cl.append([int(2 * rng.integers(0, n) + rng.integers(0, 2)) for _ in range(k)])
    return Formula.from_clauses(n, cl)


@pytest.mark.parametrize("path", CNFS, ids=[os.path.basename(p) for p in CNFS])
def test_dimacs_reader_matches_reference_grammar(path):
    f, hv, hc = G.read_dimacs(path)
    g = read_dimacs(path)
    assert f.n_vars == g.n_vars and f.clauses() == g.clauses()
    assert hc == g.n_clauses


def test_dimacs_edge_cases():
    f, hv, hc = G.parse_dimacs("c x\np cnf 10 3\n1 -2 0\n\n3 0 -4\n5 0\n%\n0\n")
    assert f.n_vars == 5 and hv == 10 and hc == 3          # n_vars = highest var seen
    assert f.clauses() == [[1, 2], [5], [6, 9]]
    with pytest.raises(G.GpsatError):
        G.parse_dimacs("1 2 0\n")                          # no header
    with pytest.raises(G.GpsatError):
        G.parse_dimacs("p cnf 2 1\n1 x 0\n")


def test_preprocess_matches_oracle(orc):
    rng = np.random.default_rng(7)
    for t in range(400):
        f = random_formula(rng, int(rng.integers(2, 12)))
        a = G.preprocess(f)
        b = orc.preprocess(f)
        assert a[0] == b[0], t
        assert list(a[2]) == list(b[2]), t
        if a[0] == G.UNDEF:
            assert a[1].clauses() == b[1].clauses(), t


def test_preprocess_fixture_unsat():
    f, _, _ = G.read_dimacs(os.path.join(GOLDEN, "cnf", "unsat_5v_chain.cnf"))
    assert G.preprocess(f)[0] == G.UNSAT


def test_cube_vars_match_oracle(orc):
    for seed in range(5):
        f = W.random_3sat_batch(1, 200, 852, seed=seed).instance(0)
        for (t, b, s) in [(32, 32, 0), (1, 1, 0), (8, 4, 0), (32, 32, 1), (10000, 10000, 0)]:
            assert G.cube_vars(f, t, b, s).tolist() == orc.choose_cube_vars(f, 0, t, b, s).tolist()


def test_simple_and_scored_choosers():
    f = W.random_3sat_batch(1, 30, 120, seed=1).instance(0)
    assert G.cube_vars(f, strategy=2, dead=[0, 2, 3]).tolist() == [1, 4, 5, 6, 7, 8, 9]
    scores = np.zeros(30, np.float32)
    scores[[5, 17, 9]] = [3.0, 2.0, 1.0]
    assert G.cube_vars_scored(scores, 3).tolist() == [5, 17, 9]
    assert G.cube_vars_scored(scores, 2, dead=[5]).tolist() == [17, 9]


def test_cubes_match_oracle(orc):
    for vs in ([3], [7, 2, 9], list(range(10))):
        assert (G.cubes(vs) == orc.enumerate_cubes(vs)).all()


CLI = G.CLI_PATH


def run_cli(*args, cwd=None):
    return subprocess.run([CLI, *args], capture_output=True, text=True, cwd=cwd, timeout=120)


def test_cli_flags():
    r = run_cli("--version")
    assert r.returncode == 0 and r.stdout.strip() == "v0.0.1"
    r = run_cli("--help")
    assert r.returncode == 0 and "--number-of-threads" in r.stdout
    r = run_cli("-i", "/nonexistent.cnf")
    assert r.returncode == 255 and "was not found" in r.stdout
    r = run_cli("-i", os.path.join(GOLDEN, "cnf", "sat_4v_simple.cnf"), "-s", "bogus")
    assert "Strategy must be either distributed or uniform" in r.stderr


def test_cli_answers_in_preprocessing(tmp_path):
    r = run_cli("-i", os.path.join(GOLDEN, "cnf", "unsat_5v_chain.cnf"), "-l", cwd=tmp_path)
    assert r.returncode == 0
    assert "Solved in pre-processing." in r.stdout
    assert r.stdout.strip().splitlines()[-1] == "UNSATISFIABLE"
    assert (tmp_path / "autolog.txt").read_text().count(",") == 3
    # a formula preprocessing satisfies: the model line holds the solved literals
    p = tmp_path / "units.cnf"
    p.write_text("p cnf 3 3\n1 0\n-2 0\n1 2 3 0\n")
    r = run_cli("-i", str(p), "-v", "0")
    lines = r.stdout.strip().splitlines()
    assert lines[-3:] == ["SATISFIABLE", "1 -2 3 0", "Solution was verified"]


def test_cli_without_gpu_fails_loudly():
    if G.device_count() > 0:
        pytest.skip("a GPU is visible")
    r = run_cli("-i", os.path.join(GOLDEN, "cnf", "sat_4v_simple.cnf"))
    assert r.returncode != 0 and "no HIP device" in r.stderr
