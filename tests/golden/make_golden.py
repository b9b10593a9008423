"""Write the golden fixtures of tests/golden/ from the oracle (CPU restatement of the
reference, itself pinned to the reference's compiled operators by
tests/test_oracle_ops.py). Run from the repository root:

    python tests/golden/make_golden.py [config_a]

config_a: full-size acoustic forward (100^3, nt=1000), traces [1][100][1000] of the
FP32 oracle and of the FP64 oracle (same injector stencil), ~1.2 MB.
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fwi3d_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def config_a():
    s = W.config_a()
    shape = s.shape
    inj = O.Injector(0, s.src_loc, shape, s.d, (0, 0, 0), "f")
    ext = O.Injector(0, s.rcv_loc, shape, s.d, (0, 0, 0), "f")
    par = O.make_par(s.nt, 0, s.dt, bc=s.bc, taper=s.taper, taper_strength=s.taper_strength)
    t0 = time.time()
    rf, _ = O.propagate("acoustic", s.model, s.src, inj, ext, par, s.d, s="f")
    t1 = time.time()
    rd, _ = O.propagate("acoustic", s.model.astype(np.float64), s.src.astype(np.float64), inj.as_double(),
                        ext.as_double(), par, s.d, s="d")
    t2 = time.time()
    np.save(os.path.join(OUT, "config_a_traces_f32.npy"), rf.astype(np.float32))
    np.save(os.path.join(OUT, "config_a_traces_f64.npy"), rd.astype(np.float64))
    rel = np.linalg.norm(rf - rd) / np.linalg.norm(rd)
    print(f"config_a: fp32 {t1 - t0:.1f}s fp64 {t2 - t1:.1f}s  rel(f32 vs f64)={rel:.3e}")


if __name__ == "__main__":
    which = sys.argv[1:] or ["config_a"]
    for w in which:
        globals()[w]()
