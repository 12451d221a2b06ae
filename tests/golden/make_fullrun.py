"""Writes the whole-run fixtures of the c1, c2 and c3 workloads from the FP64
oracle (oracle/liboracle.so), for the GPU tests that run the same workloads
to their last step (tests/test_gpu_fullrun.py).

    python tests/golden/make_fullrun.py [c1] [c2] [c3]

The oracle needs about 20 s (c1), 25 min (c2) and 10 min (c3) on 8 cores.
Each fixture fullrun_<name>.npz holds data only:
  nt, checkpoints        run length; steps after which the fields were hashed
  field_sha256           [checkpoint][6] SHA-256 of Ex..Hz as little-endian
                         float64 in [i][j][k] order (hex, as uint8)
  field_l2               [checkpoint][6] L2 norm of each component
  intensity<q>           per-step sum over port q's points of Ex^2+Ey^2+Ez^2
  samples_sha256<q>      (ports recording every step) SHA-256 of the
                         [step][point][6] samples
  dft_sha256<q>          (DFT frequencies) SHA-256 of the [point][freq][6]
                         complex accumulators as (re, im) float64 pairs
  flux<q>, freqs         spectral energy through plane port q (J/Hz,
                         2 (stride dt)^2 dA Re sum (Ex Hy* - Ey Hx*), z normal)
  gain_sha256, gain_sum, gain_sample
                         (gain) SHA-256 of the [cell][4] populations, their
                         per-level sums and every 251st cell's populations
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

# workload -> (generator keyword arguments, checkpoints); checkpoints are
# multiples of the engine's 128-step graph block so the GPU runs replay
# whole blocks between them, and the last is the run's end
WORKLOADS = {
    "c1": (configs.c1, {"nt": 1000, "record": True}, [512, 1000]),
    "c2": (configs.c2, {"nt": 4000}, [1024, 2048, 4000]),
    "c3": (configs.c3, {"nt": 20000}, [5120, 10240, 20000]),
}
GAIN_STRIDE = 251


def sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest().encode(),
                         dtype=np.uint8)


def dft_pairs(z):
    return np.stack([z.real, z.imag], axis=-1)


def plane_flux(dft, scale):
    """Spectral energy through a z-normal plane port from its DFT
    [point][freq][6]: scale * Re sum_points (Ex Hy* - Ey Hx*)."""
    return scale * np.real(np.sum(dft[:, :, 0] * np.conj(dft[:, :, 4]) - dft[:, :, 1] * np.conj(dft[:, :, 3]), axis=0))


def make(name):
    make_case, kw, checkpoints = WORKLOADS[name]
    case = make_case(tempfile.mkdtemp(prefix="fullrun_%s_" % name), **kw)
    o = po.OracleSim(case)
    out = {"nt": np.array([case.nt]), "checkpoints": np.array(checkpoints)}
    shas, l2 = [], []
    t0 = time.time()
    for cp in checkpoints:
        o.step(cp - o.steps_done)
        f = o.fields()
        shas.append([sha(x) for x in f])
        l2.append([float(np.linalg.norm(x)) for x in f])
        print("%s: step %d (%.0f s)" % (name, cp, time.time() - t0), flush=True)
    out["field_sha256"] = np.array(shas)
    out["field_l2"] = np.array(l2)
    for q in range(len(case.ports)):
        out["intensity%d" % q] = o.port_intensity(q)
        if case.record_ports:
            out["samples_sha256%d" % q] = sha(o.port_samples(q))
        if case.freqs:
            dft = o.port_dft(q)
            out["dft_sha256%d" % q] = sha(dft_pairs(dft))
            h = case.dft_stride * o.dt
            out["flux%d" % q] = plane_flux(dft, 2.0 * h * h * o.d[0] * o.d[1])
    if case.freqs:
        out["freqs"] = np.array(case.freqs)
    if case.gain:
        g = o.gain_pop()
        out["gain_sha256"] = sha(g)
        out["gain_sum"] = g.sum(axis=0)
        out["gain_sample"] = g[::GAIN_STRIDE].copy()
    path = os.path.join(HERE, "fullrun_%s.npz" % name)
    np.savez_compressed(path, **out)
    print("%s: wrote %s (%d bytes)" % (name, path, os.path.getsize(path)), flush=True)


def main():
    po.build()
    for name in sys.argv[1:] or list(WORKLOADS):
        make(name)


if __name__ == "__main__":
    main()
