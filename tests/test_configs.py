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
    """c3's rods: addition at 7 cells then relaxation to the configured 8-cell
    minimum spacing, seeded and repeatable."""
    pts = configs.rod_positions(400, 192, 192, 3, 8.0, 1234, rsa_spacing=7.0)
    assert len(pts) == 400
    assert pts == configs.rod_positions(400, 192, 192, 3, 8.0, 1234, rsa_spacing=7.0)
    p = np.array(pts)
    assert p.min() >= 4 and p.max() < 188
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1) + np.eye(400) * 1e9
    assert d2.min() >= 64.0
    # relaxation moves rods: the addition-only layout had pairs under 8 cells
    q = np.array(configs.rod_positions(400, 192, 192, 3, 7.0, 1234))
    e2 = ((q[:, None, :] - q[None, :, :]) ** 2).sum(-1) + np.eye(400) * 1e9
    assert 49.0 <= e2.min() < 64.0


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


def test_c5_media_clear_of_cpml(tmp_path):
    """No request of the stream puts a non-vacuum cell inside the 10-cell
    CPML (the reference's CPML E terms assume vacuum, CPML.cpp:524)."""
    rq = configs.c5_requests(40)
    kinds = set()
    for k, (_, p) in enumerate(rq):
        if p["N"][0] * p["N"][1] * p["N"][2] > 3e6:
            continue
        kinds.add(p["kind"])
        d = tmp_path / ("r%d" % k)
        d.mkdir()
        idm = eng.Model(configs.c5(str(d), p)).id_map()
        inner = idm[10:-10, 10:-10, 10:-10].copy()
        assert (idm != 0).sum() == (inner != 0).sum() > 0, (k, p)
    assert kinds == {"sphere", "dl", "gain"}


def _slab_case(directory, inset, nt):
    """A c5 gain request at 40x40x48 with its slab ``inset`` cells from the
    side faces (0: through the side CPML)."""
    from cpsfdtd_amd.case import CaseWriter, Shape, FIXED
    nx, ny, nz, dx = 40, 40, 48, 10e-9
    w = CaseWriter(directory)
    w.materials([(1, 1, 0, 0), (2.25, 1, 0, 0), (1, 1, 0, 0), (1, 1, 0, 0), (1, 1, 0, 0), (2.25, 1, 0, 0)])
    h = nz // 4
    w.grid((nx, ny, nz), (dx, dx, dx),
           [Shape(5, "cuboid", x0=inset * dx, y0=inset * dx, z0=(nz // 2 - h // 2) * dx, a=(nx + 1 - 2 * inset) * dx,
                  b=(ny + 1 - 2 * inset) * dx, c=h * dx)])
    w.cpml(10, m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    w.source(600e-9, 10e-15, 30e-15)
    w.point_source(nx // 2, ny // 2, nz // 2, comp=2, amp=1.0)
    w.gain(600e-9, 1e14, 1e-13, 1e-10, 1e-13, 3.3e24, 1e9)
    return w.run(nt)


def test_dielectric_through_cpml_diverges_on_oracle(tmp_path, oracle_lib):
    """The reference's arithmetic: the eps_r 2.25 gain slab of a c5 request
    stays bounded when it ends 12 cells from the side faces, and grows
    without bound when it runs through the side CPML (vacuum dt/eps0 on the
    CPML E terms, CPML.cpp:524)."""
    peak = {}
    for inset in (12, 0):
        d = tmp_path / ("i%d" % inset)
        d.mkdir()
        o = oracle_lib.OracleSim(_slab_case(str(d), inset, 400))
        o.step(400)
        with np.errstate(invalid="ignore"):
            peak[inset] = max(float(np.nan_to_num(np.abs(o.field(c)), nan=np.inf).max()) for c in range(6))
    assert peak[12] < 1e-4  # the source pulse is still rising (3.4e-6)
    assert peak[0] > 1e6    # non-finite by then
