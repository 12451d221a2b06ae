"""bench.py's spectral read-outs on closed forms: a sampled Lorentzian gives
back its own centre and Q (the half-maximum read-out to the sampling's
interpolation error, the least-squares fit exactly), and a spectrum whose
peak is not bracketed gives no Q."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def lorentzian(f, f0, q):
    g = f0 / (2 * q)
    return 3.0 / (1 + ((f - f0) / g) ** 2)


def test_lorentzian_recovered():
    # 64 frequencies evenly spaced in wavelength over 400-800 nm (the c2 DFT grid)
    c = 299792458.0
    f = c / np.linspace(400e-9, 800e-9, 64)
    for f0, q in ((5.0e14, 6.0), (4.4e14, 15.0), (6.3e14, 3.0)):
        p = lorentzian(f, f0, q)
        pk, qh = bench.spectrum_peak(f, p)
        assert abs(pk - f0) <= np.max(np.abs(np.diff(f)))  # the largest sample at most one step off
        assert abs(qh / q - 1) < 0.05, (q, qh)
        pf, qf = bench.lorentz_fit(f, p)
        assert abs(pf / f0 - 1) < 1e-9 and abs(qf / q - 1) < 1e-7, (pf, qf)


def test_unbracketed_peak_has_no_q():
    f = np.linspace(4e14, 7e14, 32)
    p = f - 3e14  # rises to the last sample
    assert bench.spectrum_peak(f, p)[1] is None
    assert bench.lorentz_fit(f, p) == (None, None)


def test_cpu_fifo_model_queue(monkeypatch):
    """Two requests at t = 0 and one long after: the second waits for the
    first, the third finds the CPU idle."""
    from cpsfdtd_amd import configs
    rq = [(0.0, {"N": (10, 10, 10), "steps": 100}),   # 1e5 cell-steps: 1 s at 0.1 Mcell/s
          (0.0, {"N": (20, 10, 10), "steps": 100}),   # 2 s, done at 3 s
          (50.0, {"N": (10, 10, 10), "steps": 400})]  # 4 s, done at 54 s
    monkeypatch.setattr(configs, "c5_requests", lambda n: rq[:n])
    r = bench.cpu_fifo_model(3, 0.1)
    assert r["makespan_s"] == 54.0
    assert r["p50_s"] == 3.0 and r["p99_s"] == round(float(np.percentile([1.0, 3.0, 4.0], 99)), 1)


def test_bench_refuses_more_than_one_gpu():
    """This engine runs on one GPU: --gpus 2 exits with a message (status 2)
    before touching a device."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr, (r.returncode, r.stderr[-300:])
