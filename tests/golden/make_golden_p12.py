"""Generate p12_digests.json (run in the build container, where
oracle/_ref/libcsref3.so, the reference's LIMIT_ALL_SLOPES_P12 built at
NUMORDER 3 from its own sources by oracle/ref/Makefile, exists; the fixture
is committed and the tests only read it).

For each case: the SHA-256 of the patch's input arrays as the oracle's
p12_problem writes them (so a test can tell a changed input from a changed
routine) and of every array the reference's routine leaves, V_min_local and
V_max_local included.  tests/test_cpu_p12.py checks the oracle against these
digests bit for bit, with or without the reference build."""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402

# N, NR, NGhost, sector, seed, shock, noise
CASES = [(6, 3, 4, 1, 21, True, 0.5), (8, 2, 4, 4, 22, False, 0.0), (5, 2, 5, 2, 23, True, 1.0)]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def digests(A):
    return {k: digest(A[k]) for k in sorted(A)}


def main():
    if orc.rlib3() is None:
        raise SystemExit("oracle/_ref/libcsref3.so is not built")
    out = []
    for (N, NR, g, s, seed, shock, noise) in CASES:
        prob, A0 = orc.p12_problem(N, NR, g, sector=s, seed=seed, shock=shock, noise=noise)
        B = orc.p12_reference(prob, orc.p1_copy(A0))
        out.append({"case": [N, NR, g, s, seed, shock, noise], "inputs": digests(A0), "reference": digests(B)})
    with open(os.path.join(HERE, "p12_digests.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
