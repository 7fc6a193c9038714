"""Write tests/golden/kernels_ref.npz: reference Fortran outputs for every per-block kernel case.

Runs in the build container only, where oracle/Makefile has compiled the reference's own
Fortran (src/CGMfuncs/*.f90, src/FFV/bcut.f90) into oracle/_ref.  Inputs are regenerated
from the seed by tests/kernel_cases.py, so the fixture holds outputs only:
  <case>/<output>        outputs of the all-double build (bit-exact target for the oracle)
  q16:<case>/<output>    outputs of the as-configured build where they differ (sparse: index, value)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
import kernel_cases as K  # noqa: E402

SEED, SZ, G = 3, (4, 5, 6), 2


def main():
    O.build()
    d8, q16 = O.ref("d8"), O.ref("q16")
    if d8 is None or q16 is None:
        raise SystemExit("oracle/_ref not built (needs /root/reference)")
    out = {}
    for name in K.CASES:
        a = K.run_case(name, SEED, SZ, G, d8)
        b = K.run_case(name, SEED, SZ, G, q16)
        for k, v in a.items():
            out["%s/%s" % (name, k)] = v
            diff = np.flatnonzero(v.ravel() != b[k].ravel())
            if diff.size:
                out["q16:%s/%s:idx" % (name, k)] = diff.astype(np.int32)
                out["q16:%s/%s:val" % (name, k)] = b[k].ravel()[diff]
    path = os.path.join(HERE, "kernels_ref.npz")
    np.savez_compressed(path, seed=SEED, sz=np.array(SZ), g=G, **out)
    print("wrote", path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
