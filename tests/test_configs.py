"""Workload generators (cpsfdtd_amd/configs.py): the seeded rod placement's
generator against the C++ standard's known answer, and the host model of each
workload against the oracle (ID maps, port points, dispersive/gain cells)."""
import itertools

import numpy as np

from cpsfdtd_amd import configs
import cpsfdtd_amd as eng


def test_mt19937_64_known_answer():
    """[rand.predef]: the 10000th output of a default-constructed
    std::mt19937_64 is 9981545732273789042."""
    g = configs.MT19937_64()
    for _ in range(9999):
        g.next()
    assert g.next() == 9981545732273789042


def test_mt19937_64_seeding():
    """std::mt19937_64(seed) with seed 5489 is the default engine, and the
    uniform draw is the top 53 bits."""
    a, b = configs.MT19937_64(5489), configs.MT19937_64()
    x = [a.next() for _ in range(5)]
    assert x == [b.next() for _ in range(5)]
    assert x[0] == 14514284786278117030
    c = configs.MT19937_64(5489)
    assert c.uniform() == (x[0] >> 11) / 2.0 ** 53


def test_rod_positions():
    pts = configs.rod_positions(400, 192, 192, 3, 7.0, 1234)
    assert len(pts) == 400
    assert pts == configs.rod_positions(400, 192, 192, 3, 7.0, 1234)
    p = np.array(pts)
    assert p.min() >= 4 and p.max() < 188
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1) + np.eye(400) * 1e9
    assert d2.min() >= 49.0


def model_vs_oracle(case, oracle_lib):
    m = eng.Model(case)
    o = oracle_lib.OracleSim(case)
    assert tuple(m.N) == tuple(o.N)
    assert np.array_equal(m.id_map(), o.id_map())
    for q in range(len(case.ports)):
        assert np.array_equal(np.asarray(m.port_points(q)).reshape(-1, 3), o.port_points(q))
    return m, o


def test_c1_model_matches_oracle(tmp_path, oracle_lib):
    case = configs.c1(str(tmp_path), nt=10)
    m, o = model_vs_oracle(case, oracle_lib)
    assert o.N == (128, 128, 128)
    assert (m.id_map() == 1).sum() > 30000  # the sphere


def test_c3_model_matches_oracle(tmp_path, oracle_lib):
    case = configs.c3(str(tmp_path), nt=10)
    m, o = model_vs_oracle(case, oracle_lib)
    assert o.N == (192, 192, 64)
    idm = m.id_map()
    gain = m.cells(True)
    assert len(gain) == (idm[:192, :192, 1:64] == 5).sum() > 500000
    assert (idm == 1).sum() > 200000  # the rods
    assert len(m.port_points(0)) > 10000


def test_c5_request_stream_statistics():
    """The serving stream's documented moments: 2000 requests at 20/s, edges
    log-uniform over 32-320 (mean edge ~125, mean cells ~1.9e6), steps
    uniform over 500-8000, a 40/30/30 mix."""
    rq = configs.c5_requests()
    assert len(rq) == 2000 and rq == configs.c5_requests()
    t = np.array([r[0] for r in rq])
    assert np.all(np.diff(t) > 0) and 90 < t[-1] < 110
    N = np.array([r[1]["N"] for r in rq], dtype=float)
    assert N.min() >= 32 and N.max() <= 320 and 115 < N.mean() < 135
    assert 1.6e6 < N.prod(axis=1).mean() < 2.3e6
    steps = np.array([r[1]["steps"] for r in rq])
    assert steps.min() >= 500 and steps.max() <= 8000 and 4000 < steps.mean() < 4500
    kinds = [r[1]["kind"] for r in rq]
    assert abs(kinds.count("sphere") / 2000 - 0.4) < 0.04 and abs(kinds.count("dl") / 2000 - 0.3) < 0.04


def test_c5_request_model_builds(tmp_path):
    rq = configs.c5_requests(20)
    for k, (_, p) in enumerate(rq[:6]):
        d = tmp_path / ("r%d" % k)
        d.mkdir()
        m = eng.Model(configs.c5(str(d), p))
        assert tuple(m.N) == p["N"] and len(m.port_points(0)) > 0
