"""The benchmark workloads at their full sizes on the GPU: the FP64 engine
against the FP64 oracle bit for bit over the first steps (fields, port
samples, gain populations), and the FP32 engine against the FP64 engine over
a longer run (relative L2 of the port records)."""
import numpy as np
import pytest

import cpsfdtd_amd as eng
from cpsfdtd_amd import configs

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    den = np.linalg.norm(np.ravel(b))
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / (den if den > 0 else 1.0)


def bit_exact(case, steps, oracle_lib, ports=True, gain=False):
    o = oracle_lib.OracleSim(case)
    o.step(steps)
    e = eng.Engine(case, 64)
    e.step(steps)
    e.sync()
    for c in range(6):
        assert np.array_equal(e.field(c), o.field(c)), c
    if ports:
        for q in range(len(case.ports)):
            assert np.array_equal(e.port_samples(q), o.port_samples(q)), q
    if gain:
        assert np.array_equal(e.gain_pop(), o.gain_pop())
    return e, o


def test_c1_full_size_f64(tmp_path, oracle_lib):
    print("c1 128^3, 100 steps", flush=True)
    case = configs.c1(str(tmp_path), nt=100)
    e, o = bit_exact(case, 100, oracle_lib)
    assert np.abs(o.field(2)).max() > 0


def test_c3_full_size_f64(tmp_path, oracle_lib):
    print("c3 192x192x64, 30 steps", flush=True)
    case = configs.c3(str(tmp_path), nt=30, record=True)
    bit_exact(case, 30, oracle_lib, gain=True)


def test_c2_full_size_f64(tmp_path, oracle_lib):
    print("c2 256^3, 6 steps", flush=True)
    case = configs.c2(str(tmp_path), nt=6, record=True)
    bit_exact(case, 6, oracle_lib)


def test_c2_full_run_f32_vs_f64(tmp_path):
    """The whole 4000-step 256^3 workload in FP32 (FP64 accumulators) against
    the bit-exact FP64 engine: each port's |E|^2 time series and its
    64-frequency DFT within 1e-4 relative L2."""
    print("c2 256^3, 4000 steps FP32 vs FP64", flush=True)
    case = configs.c2(str(tmp_path), nt=4000)
    runs = {}
    for prec in (64, 32):
        e = eng.Engine(case, prec)
        e.step(4000)
        runs[prec] = ([e.port_intensity(q) for q in range(2)], [e.port_dft(q) for q in range(2)])
        e.close()
    for q in range(2):
        assert np.abs(runs[64][0][q]).max() > 0
        err_t = rel_l2(runs[32][0][q], runs[64][0][q])
        err_f = rel_l2(runs[32][1][q], runs[64][1][q])
        print("port %d: |E|^2 series rel L2 %.2e, DFT rel L2 %.2e" % (q, err_t, err_f), flush=True)
        assert err_t < 1e-4 and err_f < 1e-4, (q, err_t, err_f)
