"""Writes tests/golden/trajectories.npz: SIMP trajectories of the oracle
(the FP64 CPU restatement of the reference loop in oracle/) that the GPU
tests in tests/test_gpu_loops.py compare against. Run from the repository
root after building the oracle (make -C oracle):

    python tests/golden/make_trajectories.py [case ...]

With case names, only those cases are (re)computed and the others are kept
from the existing file.

Each case stores its compliance per iteration, the change per iteration,
the CG iteration counts and the final design (reference element order).
The cases and their configurations are defined in tests/trajectory_cases.py,
which the GPU tests read too.
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import oracle as O  # noqa: E402
import trajectory_cases as TC  # noqa: E402


def main(argv):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trajectories.npz")
    out = {}
    names = TC.CASES
    if argv:
        names = argv
        for n in names:
            assert n in TC.CASES, n
        with np.load(path) as old:
            out = {k: old[k] for k in old.files}
    for name in names:
        cfg, fixed, F, pres, void, x0, niter = TC.build(name, "oracle")
        t0 = time.time()
        r = O.simp_run(cfg, fixed, F, pres, void, niter, x0)
        print("%-28s %3d iterations in %6.1f s; compliance %s" % (name, r["iterations"], time.time() - t0,
                                                                  np.array2string(r["compliance"][:4], precision=6)),
              flush=True)
        out[name + "/compliance"] = r["compliance"]
        out[name + "/change"] = r["change"]
        out[name + "/cg"] = r["cg"]
        out[name + "/x"] = r["x"]
    np.savez_compressed(path, **out)


if __name__ == "__main__":
    main(sys.argv[1:])
