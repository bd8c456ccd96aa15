"""Writes tests/golden/workload4.json: per item of the mixed-batch workload the
input hash, the oracle's skeleton hash (sha256, first 16 hex digits) and the
oracle's iteration count.  Run from the repository root:

    python tests/golden/make_workload4.py
"""
import hashlib
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def one(i):
    import _oracle as O
    from begputhinning_amd import synth
    d2, d3 = synth.batch_dims()
    a = synth.workload4_item(i, d2, d3)
    ref, it = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
    return sha(a), sha(ref), int(it), list(a.shape)


def main():
    from begputhinning_amd import synth
    d2, d3 = synth.batch_dims()
    n = len(d2) + len(d3)
    with ProcessPoolExecutor(min(8, os.cpu_count() or 2)) as ex:
        res = list(ex.map(one, range(n), chunksize=4))
    out = {"input_sha": [r[0] for r in res], "skeleton_sha": [r[1] for r in res],
           "iterations": [r[2] for r in res], "shapes": [r[3] for r in res]}
    with open(os.path.join(ROOT, "tests", "golden", "workload4.json"), "w") as f:
        json.dump(out, f)
    print(n, "items")


if __name__ == "__main__":
    main()
