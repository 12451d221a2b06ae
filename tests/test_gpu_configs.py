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


def corner_case(directory, n, nt=10):
    """PEC box of n^3 cells, dx = 10 nm, an eps_r 4 cuboid and an Ez point
    source near the low corner, no CPML (every tile on the lean path)."""
    from cpsfdtd_amd.case import CaseWriter, Shape, FIXED
    d = (10e-9, 10e-9, 10e-9)
    w = CaseWriter(directory)
    w.materials([(1, 1, 0, 0), (4.0, 1, 0, 0)])
    w.grid((n, n, n), d, [Shape(1, "cuboid", x0=4 * d[0], y0=18 * d[1], z0=6 * d[2], a=6 * d[0], b=5 * d[1], c=8 * d[2])])
    w.source(500e-9, 2e-15, 4e-15)
    w.point_source(12, 14, 16, comp=2, amp=1.0)
    return w.run(nt, flags=FIXED)


def test_grid_beyond_2_31_cells(tmp_path):
    """A 1300^3 FP32 grid (2.2e9 cells, 53 GB of fields: 64-bit cell indexing
    everywhere) against a 48^3 grid with the same corner geometry and source:
    after 10 steps influence has travelled at most 20 cells, so the corner
    region [0,24)^3 of both grids is the same, bit for bit."""
    import os
    os.makedirs(tmp_path / "big")
    os.makedirs(tmp_path / "small")
    print("1300^3 FP32 grid, 10 steps", flush=True)
    big = eng.Engine(corner_case(str(tmp_path / "big"), 1300), 32)
    assert big.N == (1300, 1300, 1300) and 1301 ** 3 > 2 ** 31
    big.step(10)
    small = eng.Engine(corner_case(str(tmp_path / "small"), 48), 32)
    small.step(10)
    for c in range(6):
        a = big.region(c, 0, 24, 0, 24, 0, 24)
        b = small.region(c, 0, 24, 0, 24, 0, 24)
        assert np.abs(b).max() > 0 or c in (3,), c
        assert np.array_equal(a, b), c
    # the far corner of the big grid is untouched
    assert np.abs(big.region(2, 1270, 1300, 1270, 1300, 1270, 1300)).max() == 0.0
    assert big.nonfinite() == 0
    big.close()
    small.close()
