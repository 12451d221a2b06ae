"""Packs (cfd_pack_*): members of different grid sizes and physics stepped
together, one launch per pass over all of them, give the results each gives
alone, bit for bit, in FP64 and FP32: fields, port samples, |E|^2 series,
DFT accumulators and gain populations.  Runs cross replayed 16-step
sequences and eager remainders; members end at different steps, and one joins
a running pack."""
import numpy as np
import pytest

import cases
import cpsfdtd_amd as eng
from cpsfdtd_amd import configs
from cpsfdtd_amd.case import FIXED

pytestmark = pytest.mark.gpu


def member_cases(root):
    """Four different members: a Drude-Lorentz + gain mixed case with a
    64-frequency-free DFT port set, a periodic case, a smaller mixed case with
    TF/SF only, and a c5 request (gain slab, point source, plane port DFT)."""
    out = []
    for name, make in (
            ("dlgain", lambda d: cases.mixed(d, n=(44, 38, 34), dl=True, gain=True, flags=FIXED, nt=300,
                                             freqs=(4.5e14, 6.0e14, 7.5e14), stride=2)),
            ("periodic", lambda d: cases.mixed(d, n=(32, 28, 30), npml=(0, 0, 6), periodic=3, flags=FIXED, nt=260,
                                               dl=True)),
            ("small", lambda d: cases.mixed(d, n=(24, 22, 26), npml=(4, 4, 4), box=6, flags=FIXED, nt=140, src=None)),
            ("c5gain", lambda d: configs.c5(d, {"N": (48, 40, 56), "steps": 333, "kind": "gain"})),
    ):
        d = root / name
        d.mkdir()
        out.append((name, make(str(d))))
    return out


def results(e, case):
    r = {"fields": [e.field(c) for c in range(6)]}
    for q in range(len(case.ports)):
        r["I%d" % q] = e.port_intensity(q)
        if case.record_ports:
            r["S%d" % q] = e.port_samples(q)
        if case.freqs:
            r["D%d" % q] = e.port_dft(q)
    if case.gain:
        r["gain"] = e.gain_pop()
    return r


def assert_same(a, b, what):
    assert sorted(a) == sorted(b)
    for k in a:
        if k == "fields":
            for c in range(6):
                assert np.array_equal(a[k][c], b[k][c]), (what, k, c)
        else:
            assert np.array_equal(a[k], b[k]), (what, k)


@pytest.mark.parametrize("precision", [64, 32])
def test_pack_members_equal_solo_runs(tmp_path, precision):
    mc = member_cases(tmp_path)
    solo = {}
    for name, case in mc:
        e = eng.Engine(case, precision)
        e.step(case.nt)
        solo[name] = results(e, case)
        e.close()
    print("pack FP%d" % precision, flush=True)
    pack = eng.Pack(precision)
    members = {name: (eng.Engine(case, precision), case) for name, case in mc}
    for name in ("dlgain", "periodic", "small"):
        pack.add(members[name][0])
    assert len(pack) == 3
    pack.step(160)          # ten replayed 16-step sequences; "small" ends at 140
    pack.sync()
    assert members["small"][0].steps_done == 140
    pack.remove(members["small"][0])
    pack.add(members["c5gain"][0])  # joins at its step 0 while the others run on
    pack.step(180)          # 11 replays and 4 eager steps; "periodic" ends at 260 inside this
    pack.sync()
    pack.step(200)
    pack.sync()
    assert pack.check_finite() == [0] * len(pack)  # the serving watchdog's count
    for name, (e, case) in members.items():
        assert e.steps_done == case.nt, name
    pack.close()
    for name, (e, case) in members.items():
        assert_same(results(e, case), solo[name], name)
        e.close()
    assert np.abs(solo["dlgain"]["fields"][2]).max() > 0 and np.abs(solo["c5gain"]["I0"]).max() > 0


def test_pack_refuses_misuse(tmp_path):
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    a = eng.Engine(cases.mixed(str(tmp_path / "a"), flags=FIXED, nt=40), 64)
    b = eng.Engine(cases.mixed(str(tmp_path / "b"), flags=FIXED, nt=40), 32)
    pack = eng.Pack(64)
    pack.add(a)
    with pytest.raises(eng.EngineError, match="pack member"):
        a.step(1)
    with pytest.raises(eng.EngineError, match="already in a pack"):
        pack.add(a)
    with pytest.raises(eng.EngineError, match="precision"):
        pack.add(b)
    pack.step(5)
    pack.remove(a)
    a.step(1)  # alone again
    assert a.steps_done == 6
    pack.close()
    a.close()
    b.close()
