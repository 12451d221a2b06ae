"""Batches (cfd_batch_*, cpsfdtd_amd.Batch): instances of one shape stepped
together, each pass one launch over all instances.  Every member must equal
a solo run of its own configuration bit for bit (same kernels, same operand
order), in FP64 and FP32, and the FP64 members equal the oracle."""
import os

import numpy as np
import pytest

import cpsfdtd_amd as eng
from cpsfdtd_amd import configs

pytestmark = pytest.mark.gpu


def c4_cases(root, idx, nt):
    ps = configs.c4_params()
    out = []
    for i in idx:
        d = os.path.join(str(root), "i%d" % i)
        os.makedirs(d)
        out.append(configs.c4(d, ps[i], nt=nt))
    return out


@pytest.mark.parametrize("prec", [64, 32])
def test_batch_members_equal_solo_runs(tmp_path, prec):
    idx = [0, 1, 7, 2, 9, 3]  # dielectric pillars and Drude disks (7 and 9)
    cases = c4_cases(tmp_path, idx, 80)
    kinds = [configs.c4_params()[i][0] for i in idx]
    assert "drude" in kinds and "dielectric" in kinds
    print("batch of %d c4 cells, FP%d, 80 steps" % (len(cases), prec), flush=True)
    b = eng.Batch(cases, prec)
    b.step(50)
    b.step(30)
    b.sync()
    for k, case in enumerate(cases):
        s = eng.Engine(case, prec)
        s.step(80)
        m = b.member(k)
        assert m.steps_done == 80
        for c in range(6):
            assert np.array_equal(m.field(c), s.field(c)), (k, c)
        for q in range(2):
            assert np.array_equal(m.port_dft(q), s.port_dft(q)), (k, q)
            assert np.array_equal(m.port_intensity(q), s.port_intensity(q)), (k, q)
        s.close()
    b.close()


def test_batch_member_matches_oracle(tmp_path, oracle_lib):
    cases = c4_cases(tmp_path, [7, 0], 60)
    b = eng.Batch(cases, 64)
    b.step(60)
    for k, case in enumerate(cases):
        o = oracle_lib.OracleSim(case)
        o.step(60)
        m = b.member(k)
        for c in range(6):
            assert np.array_equal(m.field(c), o.field(c)), (k, c)
        for q in range(2):
            assert np.array_equal(m.port_dft(q), o.port_dft(q)), (k, q)
    b.close()


def test_batch_rejects_mixed_shapes_and_member_stepping(tmp_path):
    a = c4_cases(tmp_path / "a", [0], 20)[0]
    os.makedirs(tmp_path / "b")
    other = configs.c4(str(tmp_path / "b"), configs.c4_params()[1], scale=0.5, nt=20)
    with pytest.raises(eng.EngineError, match="grid size"):
        eng.Batch([a, other], 32)
    b = eng.Batch([a], 32)
    m = b.member(0)
    with pytest.raises(eng.EngineError, match="batch member"):
        m.step(1)
    b.step(20)
    with pytest.raises(eng.EngineError, match="run finished"):
        b.step(1)
    b.close()
