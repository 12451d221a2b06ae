"""The serving loop (cfd_server_*, cpsfdtd_amd.Server): requests of different
shapes run together in step quanta, each on its own stream or as members of
packs (one launch per pass over all of them); each
request's results must equal a solo run of the same case bit for bit (FP64
and FP32), a failing request must not disturb the others, and the memory
budget must hold requests back.  The first eight requests of the seeded c5
stream run at their real sizes: FP64 against the oracle over their first 200
steps (fixture), and served to the end against solo runs."""
import hashlib
import os

import numpy as np
import pytest

import cpsfdtd_amd as eng
from cpsfdtd_amd import configs

pytestmark = pytest.mark.gpu

SMALL = [
    ({"N": (36, 40, 44), "steps": 150, "kind": "sphere"}, 64),
    ({"N": (48, 33, 40), "steps": 90, "kind": "dl"}, 64),
    ({"N": (40, 40, 52), "steps": 120, "kind": "gain"}, 64),
    ({"N": (33, 45, 38), "steps": 200, "kind": "sphere"}, 32),
    ({"N": (44, 36, 48), "steps": 75, "kind": "dl"}, 32),
]


def make(root, k, params):
    d = os.path.join(str(root), "r%d" % k)
    os.makedirs(d)
    return configs.c5(d, params)


@pytest.mark.parametrize("pack_cells,step_cells", [(1e9, 0), (0, 0), (0, 1)], ids=["packed", "streams", "one-stepping"])
def test_server_results_equal_solo_runs(tmp_path, pack_cells, step_cells):
    srv = eng.Server(max_active=3, quantum=37, pack_cells=pack_cells, step_cells=step_cells)
    cases, ids = [], []
    for k, (p, prec) in enumerate(SMALL):
        case = make(tmp_path, k, p)
        cases.append((case, prec))
        ids.append(srv.submit(case, prec))
    print("server: %d requests submitted" % len(ids), flush=True)
    for rid in ids:
        assert srv.wait(rid, 120.0) == eng.REQ_DONE, srv.error(rid)
    st = srv.stats()
    assert st["peak_active"] <= 3 and st["steps_launched"] == sum(p["steps"] for p, _ in SMALL)
    done = []
    for rid, (case, prec) in zip(ids, cases):
        e = eng.Engine(case, prec)
        e.step(case.nt)
        assert np.array_equal(srv.result(rid, 0, "intensity"), e.port_intensity(0)), rid
        assert np.array_equal(srv.result(rid, 0, "dft"), e.port_dft(0)), rid
        e.close()
        _, (t_sub, t_start, t_done) = srv.poll(rid)
        assert 0 <= t_sub <= t_start <= t_done
        done.append(t_done)
        srv.release(rid)
    srv.close()
    if step_cells:  # one request steps at a time, the earliest admitted first
        assert done == sorted(done), done


def test_server_isolates_a_failing_request(tmp_path):
    srv = eng.Server(max_active=4, quantum=50)
    good = make(tmp_path, 0, SMALL[0][0])
    bad = make(tmp_path, 1, SMALL[3][0])
    os.remove(bad.path("structure.txt"))
    ids = [srv.submit(good, 64), srv.submit(bad, 64), srv.submit(good, 32)]
    states = [srv.wait(r, 120.0) for r in ids]
    assert states == [eng.REQ_DONE, eng.REQ_FAILED, eng.REQ_DONE]
    assert "structure" in srv.error(ids[1])
    with pytest.raises(eng.EngineError, match="failed"):
        srv.result(ids[1], 0)
    assert srv.result(ids[0], 0).shape == (good.nt,)
    srv.close()


def test_server_budget_admits_one_at_a_time(tmp_path):
    srv = eng.Server(hbm_budget=1, max_active=8, quantum=25)
    ids = [srv.submit(make(tmp_path, k, p), prec) for k, (p, prec) in enumerate(SMALL[:3])]
    assert [srv.wait(r, 120.0) for r in ids] == [eng.REQ_DONE] * 3
    assert srv.stats()["peak_active"] == 1
    times = [srv.poll(r)[1] for r in ids]
    for a, b in zip(times, times[1:]):
        assert b[1] >= a[2] - 1e-3  # the next started after the previous finished
    srv.close()


def test_server_watchdog_fails_a_diverging_request(tmp_path):
    """A request whose medium has a large negative conductivity (|aE| > 1,
    growing fields) overflows to inf/NaN; the watchdog fails it with a clear
    message while a healthy request beside it completes."""
    good = make(tmp_path, 0, SMALL[0][0])
    bad = make(tmp_path, 1, {"N": (36, 36, 36), "steps": 400, "kind": "sphere"})
    with open(bad.path("material.txt")) as fh:
        lines = fh.read().splitlines()
    assert lines[2].split()[0] == "2.25"  # the sphere's medium (epsR muR sigma sigmaM)
    lines[2] = "2.25\t1\t-2e6\t0"
    with open(bad.path("material.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    srv = eng.Server(max_active=2, quantum=64)
    ids = [srv.submit(bad, 32), srv.submit(good, 32)]
    assert srv.wait(ids[0], 120.0) == eng.REQ_FAILED
    assert "non-finite" in srv.error(ids[0])
    assert srv.wait(ids[1], 120.0) == eng.REQ_DONE
    srv.close()


@pytest.mark.parametrize("policy", [eng.POLICY_FIFO, eng.POLICY_SHORTEST], ids=["fifo", "shortest"])
def test_server_admission_policy(tmp_path, policy):
    """One request at a time: while a first request holds the slot, a large
    and then a small request queue; first come, first served starts the
    large one next, shortest work first (cells x steps) the small one."""
    import time
    srv = eng.Server(max_active=1, quantum=64, policy=policy)
    first = make(tmp_path, 0, {"N": (96, 96, 96), "steps": 3000, "kind": "sphere"})
    large = make(tmp_path, 1, {"N": (60, 60, 60), "steps": 600, "kind": "sphere"})
    small = make(tmp_path, 2, {"N": (32, 32, 32), "steps": 100, "kind": "sphere"})
    a = srv.submit(first, 32)
    while srv.poll(a)[0] == eng.REQ_QUEUED:
        time.sleep(0.005)
    b = srv.submit(large, 32)
    c = srv.submit(small, 32)
    assert srv.poll(a)[0] == eng.REQ_RUNNING  # both queued behind it
    for rid in (a, b, c):
        assert srv.wait(rid, 120.0) == eng.REQ_DONE, srv.error(rid)
    tb, tc = srv.poll(b)[1][1], srv.poll(c)[1][1]
    assert (tc < tb) if policy == eng.POLICY_SHORTEST else (tb < tc)
    srv.close()


def test_engine_nonfinite_count(tmp_path):
    case = make(tmp_path, 0, SMALL[0][0])
    e = eng.Engine(case, 64)
    e.step(10)
    assert e.nonfinite() == 0
    f = e.field(1)
    f[3, 4, 5] = np.nan
    f[6, 7, 8] = np.inf
    e.set_field(1, f)
    assert e.nonfinite() == 2
    e.close()


def test_server_rejects_a_malformed_request_at_submit(tmp_path):
    """A request whose configuration fails the schema check (ports announced
    without their files) is refused by cfd_server_submit with the reason,
    before anything is copied or queued; the server goes on serving."""
    import ctypes as C
    srv = eng.Server(max_active=2, quantum=50)
    good = make(tmp_path, 0, SMALL[0][0])
    cfg, keep = eng.make_config(good, 64)
    cfg.port_files = C.POINTER(C.c_char_p)()
    rid = C.c_longlong(-1)
    L = eng.library()
    assert L.cfd_server_submit(srv.h, C.byref(cfg), C.byref(rid)) == -1
    msg = L.cfd_server_last_error().decode()
    assert "request rejected" in msg and "port_files" in msg, msg
    assert srv.stats()["queued"] == 0
    r = srv.submit(good, 64)
    assert srv.wait(r, 120.0) == eng.REQ_DONE
    srv.close()


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def test_c5_first_requests_real_size(tmp_path):
    """The first eight requests of the seeded c5 stream at their real grid
    sizes (up to 239 x 313 x 210 cells, Drude-Lorentz, gain and dielectric
    members): the FP64 engine equals the oracle's first 200 steps bit for bit
    (fields, port DFT; |E|^2 sums within 1e-12), and served in FP32 to their
    last step (up to 6954 steps; those under 2e6 cells packed, the others on
    their own streams) each result equals its solo run bit for bit."""
    g = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c5_first8.npz"),
                     allow_pickle=False))
    steps = int(g["steps"][0])
    rq = configs.c5_requests(8)
    cases = []
    for q, (_, p) in enumerate(rq):
        case = make(tmp_path, q, p)
        assert tuple(g["N%d" % q]) == p["N"] and case.nt == int(g["nt%d" % q][0])
        cases.append(case)
        print("c5 request %d %s %s: FP64 %d steps vs oracle" % (q, p["N"], p["kind"], steps), flush=True)
        e = eng.Engine(case, 64)
        e.step(steps)
        for c in range(6):
            assert _sha(e.field(c)) == bytes(g["field_sha256%d" % q][c]).decode(), (q, c)
        ref = g["intensity%d" % q]
        got = e.port_intensity(0)
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref), q
        d = e.port_dft(0)
        assert _sha(np.stack([d.real, d.imag], axis=-1)) == bytes(g["dft_sha256%d" % q]).decode(), q
        e.close()
    print("c5: serving the eight requests in FP32", flush=True)
    srv = eng.Server(pack_cells=2e6)
    ids = [srv.submit(case, 32) for case in cases]
    for rid in ids:
        assert srv.wait(rid, 300.0) == eng.REQ_DONE, srv.error(rid)
    for q, (rid, case) in enumerate(zip(ids, cases)):
        e = eng.Engine(case, 32)
        e.step(case.nt)
        assert np.array_equal(srv.result(rid, 0, "intensity"), e.port_intensity(0)), q
        assert np.array_equal(srv.result(rid, 0, "dft"), e.port_dft(0)), q
        assert np.abs(e.port_intensity(0)).max() > 0
        e.close()
        srv.release(rid)
    srv.close()


@pytest.mark.parametrize("precision", [32, 64])
def test_admission_estimate_covers_the_simulation(tmp_path, precision):
    """The server admits by cfd_model_device_bytes of the built model: for
    cases with recorded port samples, a 64-frequency DFT, CPML on every face,
    Drude-Lorentz and gain cells, and a c5 request, it is at least what the
    simulation then allocates (cfd_device_bytes) and at most 10 % + 2 MB
    above it."""
    import cases
    from cpsfdtd_amd.case import FIXED
    made = [
        cases.mixed(str(tmp_path / "a"), dl=True, gain=True, flags=FIXED),
        cases.mixed(str(tmp_path / "b"), freqs=tuple(4e14 + 1e12 * q for q in range(64)), stride=4, record=False),
        configs.c1(str(tmp_path / "c"), nt=300, record=True),
        make(tmp_path, 9, SMALL[2][0]),
    ]
    for case in made:
        est = eng.Model(case).device_bytes(precision)
        e = eng.Engine(case, precision)
        got = e.device_bytes()
        e.close()
        print("estimate %d, allocated %d" % (est, got), flush=True)
        assert got <= est <= got * 1.10 + (2 << 20), (case.dir, est, got)
