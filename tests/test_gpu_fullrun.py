"""The c1, c2 and c3 workloads at full size run to their last step on the GPU
and compared with the FP64 oracle's whole runs (fixtures written by
tests/golden/make_fullrun.py; the oracle needs minutes to hours of CPU for
them, so its results travel as data).

FP64 engine: every field bit for bit at each checkpoint (SHA-256 of the
float64 arrays; checkpoints on 128-step boundaries, so replayed graph blocks
and eager steps both lie between them), recorded port samples, DFT
accumulators and gain populations bit for bit; the port |E|^2 sums and flux
spectra, whose reductions run in another order, within 1e-12 relative.

FP32 engine (FP64 DFT accumulators and populations): port |E|^2 series and
flux spectra within 1e-4 relative L2 of the FP64 oracle, the transmitted
flux spectrum's peak and Q within 1e-3 and 1e-2, gain populations within
1e-4 and their inversion N2 - N1 within 1e-3."""
import hashlib
import os

import numpy as np
import pytest

import cpsfdtd_amd as eng
from cpsfdtd_amd import configs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    path = os.path.join(GOLD, "fullrun_%s.npz" % name)
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated (tests/golden/make_fullrun.py)" % path)
    return dict(np.load(path, allow_pickle=False))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def hexs(a):
    return bytes(a).decode()


def rel_l2(a, b):
    den = np.linalg.norm(np.ravel(b))
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / (den if den > 0 else 1.0))


def make(name, directory):
    return {"c1": lambda: configs.c1(directory, nt=1000, record=True),
            "c2": lambda: configs.c2(directory, nt=4000),
            "c3": lambda: configs.c3(directory, nt=20000)}[name]()


@pytest.mark.parametrize("name", ["c1", "c2", "c3"])
def test_full_run_f64_bit_exact(tmp_path, name):
    g = load(name)
    case = make(name, str(tmp_path))
    assert case.nt == int(g["nt"][0])
    e = eng.Engine(case, 64)
    for q, cp in enumerate(g["checkpoints"]):
        print("%s FP64: to step %d" % (name, cp), flush=True)
        e.step(int(cp) - e.steps_done)
        e.sync()
        for c in range(6):
            f = e.field(c)
            assert sha(f) == hexs(g["field_sha256"][q][c]), (name, int(cp), c, np.linalg.norm(f), g["field_l2"][q][c])
    for q in range(len(case.ports)):
        assert rel_l2(e.port_intensity(q), g["intensity%d" % q]) < 1e-12, q
        if "samples_sha256%d" % q in g:
            assert sha(e.port_samples(q)) == hexs(g["samples_sha256%d" % q]), q
        if "dft_sha256%d" % q in g:
            d = e.port_dft(q)
            assert sha(np.stack([d.real, d.imag], axis=-1)) == hexs(g["dft_sha256%d" % q]), q
            fl = e.port_flux(q, 2)
            assert np.all(np.abs(fl - g["flux%d" % q]) <= 1e-12 * np.abs(g["flux%d" % q]).max()), q
    if "gain_sha256" in g:
        assert sha(e.gain_pop()) == hexs(g["gain_sha256"])
    e.close()


@pytest.mark.parametrize("name", ["c1", "c2", "c3"])
def test_full_run_f32_quality(tmp_path, name):
    import bench
    g = load(name)
    case = make(name, str(tmp_path))
    print("%s FP32: %d steps" % (name, case.nt), flush=True)
    e = eng.Engine(case, 32)
    e.step(case.nt)
    e.sync()
    for q in range(len(case.ports)):
        ref = g["intensity%d" % q]
        assert np.abs(ref).max() > 0
        err = rel_l2(e.port_intensity(q), ref)
        print("%s port %d: |E|^2 series rel L2 %.2e" % (name, q, err), flush=True)
        assert err < 1e-4, (q, err)
        if "flux%d" % q in g:
            fl = e.port_flux(q, 2)
            err = rel_l2(fl, g["flux%d" % q])
            print("%s port %d: flux spectrum rel L2 %.2e" % (name, q, err), flush=True)
            assert err < 1e-4, (q, err)
    if "flux1" in g:  # the transmitted port's spectral peak and Q
        f32, q32 = bench.spectrum_peak(g["freqs"], e.port_flux(1, 2))
        f64, q64 = bench.spectrum_peak(g["freqs"], g["flux1"])
        print("%s transmitted peak %.4g / %.4g Hz, Q %s / %s" % (name, f32, f64, q32, q64), flush=True)
        assert abs(f32 / f64 - 1) < 1e-3
        assert (q32 is None) == (q64 is None) and (q64 is None or abs(q32 / q64 - 1) < 1e-2)
    if "gain_sample" in g:
        pop = e.gain_pop()[::251]
        ref = g["gain_sample"]
        err = rel_l2(pop, ref)
        # the inversion N2 - N1 is a small difference of large densities
        inv_err = rel_l2(pop[:, 2] - pop[:, 1], ref[:, 2] - ref[:, 1])
        print("%s gain populations rel L2 %.2e, inversion rel L2 %.2e" % (name, err, inv_err), flush=True)
        assert err < 1e-4 and inv_err < 1e-3
    e.close()
