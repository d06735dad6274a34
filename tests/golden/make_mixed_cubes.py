"""Writes tests/golden/mixed_cubes.json: instances drawn from the mixed random k-SAT
generator (k in {3,4,5} at clause ratios {4.26, 9.9, 21.1}, n uniform in [100, 400]) with the
oracle CDCL's verdicts.  The 3- and 4-SAT draws are ones the oracle decides within a conflict
cap; 5-SAT at its threshold ratio is beyond that cap even at n = 100, so the 5-SAT draws (n at
most 110) are kept with the oracle's answer as it stands at the cap (2: not decided).

The cube test runs these through depth 8-12 cubes, restart epochs and an autotuned launch
shape, and confirms every UNSAT verdict against the oracle verdict stored here (an UNSAT
that the oracle did not prove fails the test).  Instance j is
gpupsat_amd.workloads.mixed_ksat_batch(1, seed=SEED0 + s) for the stored s; its literal
array's SHA-256 is stored so a generator change is caught.

Usage: python tests/golden/make_mixed_cubes.py   (about 10-20 minutes on 7 processes)
"""
import hashlib
import json
import multiprocessing as mp
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

SEED0 = 50_000
CAP = 400_000                            # oracle conflicts per instance
WANT = {3: 6, 4: 5, 5: 4}                # instances per k
N_MAX = {3: 260, 4: 140, 5: 110}         # larger n is beyond the oracle's cap at these ratios
KEEP_UNDECIDED = {5}                     # k whose draws are kept without an oracle verdict


def job(s):
    import oracle
    from gpupsat_amd import workloads as W
    b, ks = W.mixed_ksat_batch(1, seed=SEED0 + s)
    f = b.instance(0)
    k = int(ks[0])
    if f.n_vars > N_MAX[k]:
        return None
    st, model, stats = oracle.solve_cdcl(f, max_conflicts=CAP)
    if st == 2 and k not in KEEP_UNDECIDED:
        return None
    if st == 0:
        assert oracle.verify_model(f, model)
    return {"s": s, "k": k, "n": f.n_vars, "m": f.n_clauses,
            "sha256": hashlib.sha256(np.ascontiguousarray(f.lits, dtype=np.int32).tobytes()).hexdigest(),
            "oracle": int(st), "oracle_conflicts": int(stats[1])}


def main():
    import oracle
    oracle.build()
    got = {3: [], 4: [], 5: []}
    with mp.Pool(7) as pool:
        for r in pool.imap(job, range(6000), chunksize=4):
            if r is None or len(got[r["k"]]) >= WANT[r["k"]]:
                continue
            got[r["k"]].append(r)
            print(r, flush=True)
            if all(len(got[k]) >= WANT[k] for k in WANT):
                pool.terminate()
                break
    out = {"generator": "gpupsat_amd.workloads.mixed_ksat_batch(1, seed=seed0 + s)", "seed0": SEED0,
           "oracle": "oracle.solve_cdcl", "cap": CAP,
           "instances": [r for k in (3, 4, 5) for r in got[k]]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mixed_cubes.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path, {k: len(v) for k, v in got.items()})


if __name__ == "__main__":
    main()
