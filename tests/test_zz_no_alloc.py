"""cfd_port_gather_static (the gather that allocates nothing, staging through
a fixed device array) against cfd_port_gather, through the C ABI: equal bit
for bit on two calls in a row, the second on a smaller port of another
simulation, in both precisions; a port larger than the array is refused."""
import numpy as np
import pytest

import cases
import cpsfdtd_amd as eng
from cpsfdtd_amd.case import FIXED, CaseWriter

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [64, 32])
def test_static_gather_equals_gather(tmp_path, precision):
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    big = eng.Engine(cases.mixed(str(tmp_path / "a"), n=(60, 56, 48), flags=FIXED, dl=True), precision)
    small = eng.Engine(cases.mixed(str(tmp_path / "b"), flags=FIXED), precision)
    print("static gather FP%d" % precision, flush=True)
    big.step(45)
    small.step(30)
    seen = []
    for e, q in ((big, 0), (small, 1), (big, 1), (small, 0)):
        want = e.port_gather(q)
        got = e.port_gather_static(q)
        assert got.shape == want.shape and np.array_equal(got, want), q
        assert np.abs(want).max() > 0
        seen.append(len(want))
    assert seen[0] > seen[1]  # the second call gathers fewer points
    big.close()
    small.close()


def test_static_gather_refuses_a_port_beyond_its_array(tmp_path):
    """A 600 x 300 plane port (180000 points) exceeds the 174762 points the
    array holds: refused with a clear error before any device work."""
    d = 10e-9
    w = CaseWriter(str(tmp_path))
    w.materials([(1, 1, 0, 0)])
    w.grid((601, 301, 4), (d, d, d), [])
    w.source(500e-9, 5e-15, 10e-15)
    w.point_source(10, 10, 2, comp=2, amp=1.0)
    w.plane_port_cells("p.txt", "z", 2, 0, 599, 0, 299, (d, d, d))
    e = eng.Engine(w.run(4, flags=FIXED), 32)
    assert len(e.port_points(0)) == 180000
    with pytest.raises(eng.EngineError, match="static gather holds at most 174762"):
        e.port_gather_static(0)
    assert e.port_gather(0).shape == (180000, 6)
    e.close()
