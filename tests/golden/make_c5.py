"""Writes tests/golden/c5_first8.npz from the FP64 oracle: the first eight
requests of the seeded serving stream (configs.c5_requests), at their real
grid sizes, after their first 200 steps.

    python tests/golden/make_c5.py

Per request q (data only):
  N<q>, nt<q>          grid and run length of the request
  field_sha256<q>      [6] SHA-256 of Ex..Hz (little-endian float64, [i][j][k])
  field_l2<q>          [6] L2 norms
  intensity<q>         the port's per-step sum of |E|^2 over the 200 steps
  dft_sha256<q>        SHA-256 of the port's [point][freq][6] (re, im) DFT
"""
import hashlib
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402
from cpsfdtd_amd import configs  # noqa: E402

STEPS = 200
COUNT = 8


def sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest().encode(),
                         dtype=np.uint8)


def main():
    po.build()
    out = {"steps": np.array([STEPS])}
    t0 = time.time()
    for q, (_, p) in enumerate(configs.c5_requests(COUNT)):
        case = configs.c5(tempfile.mkdtemp(prefix="c5_%d_" % q), p)
        o = po.OracleSim(case)
        o.step(STEPS)
        f = o.fields()
        out["N%d" % q] = np.array(p["N"])
        out["nt%d" % q] = np.array([case.nt])
        out["field_sha256%d" % q] = np.array([sha(x) for x in f])
        out["field_l2%d" % q] = np.array([float(np.linalg.norm(x)) for x in f])
        out["intensity%d" % q] = o.port_intensity(0)
        d = o.port_dft(0)
        out["dft_sha256%d" % q] = sha(np.stack([d.real, d.imag], axis=-1))
        print("request %d %s %s: %.0f s" % (q, p["N"], p["kind"], time.time() - t0), flush=True)
        del o, f
    path = os.path.join(HERE, "c5_first8.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), flush=True)


if __name__ == "__main__":
    main()
