"""HIP engine (through the C ABI) against the FP64 oracle and the reference's
golden fixtures.

FP64: fields, port samples and DFT accumulators are compared bit for bit (the
engine keeps the reference's operand order and is built without FP
contraction); reductions whose order differs (port intensity sums) within a
relative 1e-12.  FP32: port time series within a relative L2 of 1e-4 of the
FP64 oracle on these short runs.
"""
import os

import numpy as np
import pytest

import cases
import cpsfdtd_amd as eng
from cpsfdtd_amd.case import FIXED, REF_ZEROHEAP, SRC_REF_PULSE

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)


def assert_fields_equal(e, o, what=""):
    for c in range(6):
        a, b = e.field(c), o.field(c)
        assert np.array_equal(a, b), "%s comp %d max|d|=%g rel=%g" % (what, c, np.abs(a - b).max(), rel_l2(a, b))


def run_both(case, steps, module=False, precision=64):
    o = cases_oracle(case)
    o.step(steps)
    e = eng.Engine(case, precision)
    if module:
        for n in range(steps):
            e.module_step(n, tfsf=case.tfsf)
    else:
        e.step(steps)
    e.sync()
    return e, o


def cases_oracle(case):
    import pyoracle
    return pyoracle.OracleSim(case)


def test_tiny_matches_reference_fixture(tmp_path):
    """40 steps of the reference's own core (fixture) reproduced on the GPU."""
    g = dict(np.load(os.path.join(GOLD, "ref_tiny.npz"), allow_pickle=False))
    case = cases.tiny(str(tmp_path))
    e = eng.Engine(case, 64)
    e.step(int(g["nt"][0]))
    for c, nm in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        assert np.array_equal(e.field(c).ravel(), g[nm]), nm
    assert np.array_equal(e.port_samples(0), g["port0"])


@pytest.mark.parametrize("module", [False, True], ids=["fused", "module"])
@pytest.mark.parametrize("flags", [REF_ZEROHEAP | SRC_REF_PULSE, FIXED], ids=["ref", "fixed"])
def test_mixed_f64_bit_exact(tmp_path, module, flags):
    case = cases.mixed(str(tmp_path), flags=flags)
    e, o = run_both(case, 60, module=module)
    assert_fields_equal(e, o)
    for q in range(2):
        assert np.array_equal(e.port_samples(q), o.port_samples(q))
        assert rel_l2(e.port_intensity(q), o.port_intensity(q)) < 1e-12


@pytest.mark.parametrize("form", ["merged", "split", "grid"])
def test_update_forms_bit_exact(tmp_path, monkeypatch, form):
    """Every fused-step form: merged (one launch per half-step, lean or CPML
    path per tile), split (lean box, shell boxes, special cells) and the single
    full-physics grid kernel per half-step."""
    monkeypatch.setenv("CFD_UPDATE_KERNEL", form)
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED)
    e, o = run_both(case, 40)
    assert_fields_equal(e, o, form)
    assert np.array_equal(e.gain_pop(), o.gain_pop())


def test_dl_gain_f64_bit_exact(tmp_path):
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED)
    e, o = run_both(case, 50)
    assert e.n_dl > 0 and e.n_gain > 0
    assert_fields_equal(e, o, "dl+gain")
    assert np.array_equal(e.gain_pop(), o.gain_pop())


def test_dl_gain_module_f64_bit_exact(tmp_path):
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED)
    e, o = run_both(case, 30, module=True)
    assert_fields_equal(e, o, "dl+gain module")
    assert np.array_equal(e.gain_pop(), o.gain_pop())


@pytest.mark.parametrize("form", ["split", "merged"])
@pytest.mark.parametrize("periodic,npml", [(1, (0, 5, 6)), (3, (0, 0, 6))], ids=["x", "xy"])
def test_periodic_f64_bit_exact(tmp_path, monkeypatch, periodic, npml, form):
    """Periodic x (and y) planes, in the split form and in the merged form
    (whose epilogue applies the TF/SF terms on the periodic E planes too)."""
    monkeypatch.setenv("CFD_UPDATE_KERNEL", form)
    case = cases.mixed(str(tmp_path), periodic=periodic, npml=npml, flags=FIXED, theta=0.0, phi=0.0, psi=0.0)
    e, o = run_both(case, 50)
    assert_fields_equal(e, o, "periodic")


def test_port_dft_f64(tmp_path):
    case = cases.mixed(str(tmp_path), freqs=(4.0e14, 5.5e14, 7.5e14), stride=3, flags=FIXED)
    e, o = run_both(case, 60)
    for q in range(2):
        assert np.array_equal(e.port_dft(q), o.port_dft(q))


def test_fp32_close_to_fp64_oracle(tmp_path):
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED)
    e, o = run_both(case, 60, precision=32)
    for q in range(2):
        err = rel_l2(e.port_samples(q), o.port_samples(q))
        assert err < 1e-4, err
    for c in range(6):
        assert rel_l2(e.field(c), o.field(c)) < 1e-4


def test_bad_input_raises(tmp_path):
    case = cases.mixed(str(tmp_path))
    with open(case.path("CFS_PML.txt"), "w") as fh:
        fh.write("100 100\n")
    with pytest.raises(eng.EngineError):
        eng.Engine(case, 64)


def test_run_length_enforced(tmp_path):
    case = cases.mixed(str(tmp_path), nt=5)
    e = eng.Engine(case, 32)
    e.step(5)
    with pytest.raises(eng.EngineError, match="run finished"):
        e.step(1)


def test_graph_replayed_blocks_bit_exact(tmp_path):
    """Step blocks replayed from a captured HIP graph (aligned 128-step blocks)
    against the oracle: eager steps, a first replay, host reads, a second
    replay, eager steps again; fields, port samples, DFT and gain populations
    compared after every stage."""
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED, nt=400, freqs=(4.0e14, 6.5e14), stride=2)
    o = cases_oracle(case)
    e = eng.Engine(case, 64)
    done = 0
    for k in (3, 253, 128, 16):  # eager to 128, replay 128-256, replay 256-384, eager to 400
        e.step(k)
        o.step(k)
        done += k
        assert e.steps_done == done
        assert_fields_equal(e, o, "after %d steps" % done)
        for q in range(2):
            assert np.array_equal(e.port_samples(q), o.port_samples(q)), (done, q)
        assert np.array_equal(e.gain_pop(), o.gain_pop()), done
    for q in range(2):
        assert np.array_equal(e.port_dft(q), o.port_dft(q))


def test_graph_replay_equals_eager_fp32(tmp_path, monkeypatch):
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED, nt=300, freqs=(5.0e14,), stride=1)
    runs = []
    for graphs in ("1", "0"):
        monkeypatch.setenv("CFD_GRAPHS", graphs)
        e = eng.Engine(case, 32)
        e.step(300)
        runs.append(([e.field(c) for c in range(6)], [e.port_dft(q) for q in range(2)]))
        e.close()
    for a, b in zip(runs[0][0] + runs[0][1], runs[1][0] + runs[1][1]):
        assert np.array_equal(a, b)
