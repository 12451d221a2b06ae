"""Host-side set-up of the engine (parsers, painters, coefficient tables, CPML
grading, TF/SF correction lists, ports, dispersive/gain cell lists) against the
FP64 oracle: integer maps and every FP64 table must be bit-identical.  Runs on
the CPU through cfd_model_* (no device)."""
import ctypes as C

import numpy as np
import pytest

import cases
import cpsfdtd_amd as eng
from cpsfdtd_amd.case import FIXED, REF_ZEROHEAP, SRC_REF_PULSE

COMP_E = {0: (1, 2), 1: (0, 2), 2: (0, 1)}   # face axis -> components of record parts 1, 2
COMP_H = {0: (5, 4), 1: (5, 3), 2: (4, 3)}
SIGN = {0: (1, -1), 1: (-1, 1), 2: (1, -1)}
FACE_ORDER = (1, 1, 0, 0, 2, 2)


def regroup(o, h, flags):
    ijk, tab = o.tfsf_points()
    v = o.tfsf_vec()
    E0, H0, CE, CH = v[0:3], v[3:6], v[6:9], v[9:12]
    amp = {
        0: {0: (CE[0] * H0[2], CE[0] * H0[1]), 1: (CE[1] * H0[2], CE[1] * H0[0]), 2: (CE[2] * H0[1], CE[2] * H0[0])},
        1: {0: (CH[0] * E0[1], CH[0] * E0[2]), 1: (CH[1] * E0[0], CH[1] * E0[2]), 2: (CH[2] * E0[0], CH[2] * E0[1])},
    }[h]
    comps = COMP_H if h else COMP_E
    far = 0 if flags & 2 else 1
    return ijk, tab, amp, comps, far


def build_expected(o, h, flags, face_id):
    ijk, tab, amp, comps, far = regroup(o, h, flags)
    tg = {}
    order = []
    for q in range(len(ijk)):
        f = face_id[q]
        a = FACE_ORDER[f]
        e = f & 1
        flip = -1.0 if e else 1.0
        i, j, k, part = (int(x) for x in ijk[q])
        if h:
            off = far if e else -1
            if a == 0:
                i += off
            elif a == 1:
                j += off
            else:
                k += off
        for c in range(2):
            if not part & (1 << c):
                continue
            s = SIGN[a][c] * flip
            coef = -amp[a][c] if s < 0 else amp[a][c]
            t = tab[q, 4 * c:4 * c + 4]  # this pair's dE pE dH pH
            if h:
                prof = 0.0 if ((flags & REF_ZEROHEAP) and a == 0 and part != 3) else t[3]
                term = (coef, t[2], prof)
            else:
                term = (coef, t[0], t[1])
            key = (comps[a][c], i, j, k)
            if key not in tg:
                tg[key] = []
                order.append(key)
            tg[key].append(term)
    return order, tg


CASES = [
    ("mixed_ref", dict(), None),
    ("mixed_fixed", dict(flags=FIXED), None),
    ("mixed_zs", dict(zs=120e-9, flags=FIXED | SRC_REF_PULSE), None),
    ("mixed_dl_gain", dict(dl=True, gain=True, flags=0), None),
    ("mixed_periodic_x", dict(periodic=1, npml=(0, 5, 6), flags=FIXED), None),
    ("mixed_periodic_xy", dict(periodic=3, npml=(0, 0, 6), flags=FIXED), None),
    ("mixed_normal", dict(theta=0.0, phi=0.0, psi=0.0, flags=FIXED), None),
]


@pytest.mark.parametrize("name,kw,_", CASES, ids=[c[0] for c in CASES])
def test_model_matches_oracle(oracle_lib, tmp_path, name, kw, _):
    kw = dict(kw)
    case = cases.mixed(str(tmp_path), freqs=(5e14, 6e14), stride=2, **kw)
    o = oracle_lib.OracleSim(case)
    m = eng.Model(case)
    assert m.N == o.N and m.d == o.d and m.dt == o.dt and m.nt == o.nt
    assert np.array_equal(m.id_map(), o.id_map())
    for w, nm in enumerate(("aE", "bE", "aH", "bH")):
        assert np.array_equal(m.table(nm), o.coefs()[w]), nm
    for w in range(12):
        assert np.array_equal(m.table("stretch%d" % w), o.stretch(w)), w
    for a, ax in enumerate("xyz"):
        for e in range(2):
            bE, cE = o.cpml(a * 4 + e * 2)
            bH, cH = o.cpml(a * 4 + e * 2 + 1)
            if len(bE) == 0:
                continue
            assert np.array_equal(m.table("pml_Eb%s%d" % (ax, e + 1)), bE)
            assert np.array_equal(m.table("pml_Ec%s%d" % (ax, e + 1)), cE)
            assert np.array_equal(m.table("pml_Hb%s%d" % (ax, e + 1)), bH)
            assert np.array_equal(m.table("pml_Hc%s%d" % (ax, e + 1)), cH)
    e, hw = o.tfsf_wave()
    me, mh = m.table("Einc"), m.table("Hinc")
    n = min(len(e), len(me))
    assert n >= case.nt and np.array_equal(me[:n], e[:n]) and np.array_equal(mh[:n], hw[:n])
    for q in range(2):
        assert np.array_equal(m.port_points(q), o.port_points(q))
    if kw.get("dl"):
        assert np.array_equal(m.table("dl"), o.dl_coefs())
        assert len(m.cells(False)) == o.dl_ncells() > 0
    if kw.get("gain"):
        assert len(m.cells(True)) == len(o.gain_pop()) > 0


def _faces_from_records(ijk, periodic=0):
    """Split the oracle's record sequence (faces y1 y2 x1 x2 z1 z2, each
    contiguous) into faces: a face's records all lie on its plane, and the next
    face's first record does not (the box is at least two cells thick)."""
    n = len(ijk)
    faces = np.empty(n, np.int32)
    f, q = 0, 0
    while q < n:
        a = FACE_ORDER[f]
        if a < 2 and periodic & (1 << a):  # no faces on a periodic axis
            f += 1
            continue
        c = ijk[q, a]
        while q < n and ijk[q, a] == c:
            faces[q] = f
            q += 1
        f += 1
    return faces


@pytest.mark.parametrize("kw", [dict(), dict(flags=FIXED), dict(flags=REF_ZEROHEAP), dict(periodic=3, npml=(0, 0, 6), flags=FIXED)],
                         ids=["ref", "fixed", "zeroheap", "periodic"])
def test_tfsf_correction_lists(oracle_lib, tmp_path, kw):
    case = cases.mixed(str(tmp_path), **kw)
    o = oracle_lib.OracleSim(case)
    m = eng.Model(case)
    ijk, _ = o.tfsf_points()
    faces = _faces_from_records(ijk, case.periodic)
    for h in (0, 1):
        order, expect = build_expected(o, h, case.flags, faces)
        tg, terms = m.tfsf(h)
        assert len(tg) == len(order)
        got = {}
        for t in tg:
            comp, i, j, k, first, count = (int(x) for x in t)
            got[(comp, i, j, k)] = [tuple(terms[q]) for q in range(first, first + count)]
        for key in order:
            assert key in got, key
            assert len(got[key]) == len(expect[key])
            for a, b in zip(got[key], expect[key]):
                assert a[0] == b[0] and a[1] == b[1] and a[2] == b[2], (key, a, b)


def test_shipped_inputs(tmp_path):
    """The reference's own example inputs parse and paint as the reference does."""
    import hashlib
    import os
    import shutil
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    g = dict(np.load(os.path.join(gold, "ref_shipped.npz"), allow_pickle=False))
    for tag, structure, pml in (("structure", "structure.txt", True), ("shi1", "Shi1.txt", False)):
        d = tmp_path / tag
        d.mkdir()
        for f in ("material.txt", structure) + (("CFS_PML.txt",) if pml else ()):
            shutil.copy(os.path.join(gold, "inputs", f), str(d / ("structure.txt" if f == structure else f)))
        w = eng.case.CaseWriter(str(d))
        w.case.pml = pml
        case = w.run(1, flags=REF_ZEROHEAP)
        m = eng.Model(case)
        ids = m.id_map()
        assert hashlib.sha256(ids.astype("<i2").tobytes()).hexdigest() == bytes(g[tag + "_sha256"]).decode()
        st = np.concatenate([m.table("stretch%d" % w) for w in range(12)])
        assert np.array_equal(st, g[tag + "_stretch"])


def test_bad_inputs_fail_loudly(tmp_path):
    case = cases.mixed(str(tmp_path))
    with open(case.path("CFS_PML.txt"), "w") as fh:
        fh.write("40 40\n")
    with pytest.raises(eng.EngineError):
        eng.Model(case)
    case = cases.mixed(str(tmp_path / "b"))
    case.ports.append((case.path("missing_port.txt"), 0))
    with pytest.raises(eng.EngineError, match="cannot open"):
        eng.Model(case)


def _check(cfg):
    """cfd_config_check: (rc, message)."""
    L = eng.library()
    rc = L.cfd_config_check(C.byref(cfg))
    return rc, L.cfd_model_last_error().decode() if rc < 0 else ""


def test_config_schema_rejected_before_any_file_is_read(tmp_path):
    """Each malformed field of a request is refused by cfd_config_check (and
    by the model build, which applies it first) with a message naming it;
    the copies and set-up code never see such a configuration."""
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED)
    cfg, keep = eng.make_config(case)
    assert _check(cfg) == (0, "")
    null_ptr = C.POINTER(C.c_char_p)()
    bad = [("structure", None, "material and structure"),
           ("precision", 16, "precision"),
           ("nt", -1, "nt"),
           ("n_src", -2, "negative"),
           ("port_files", null_ptr, "without port_files"),
           ("src_ijkc", C.POINTER(C.c_int)(), "without src_ijkc"),
           ("periodic", 4, "periodic")]
    for field, value, msg in bad:
        c2, keep2 = eng.make_config(case)
        setattr(c2, field, value)
        rc, text = _check(c2)
        assert rc == -1 and msg in text, (field, text)
    c2, keep2 = eng.make_config(case)
    c2.n_freq, c2.freqs = 2, C.POINTER(C.c_double)()
    assert "without freqs" in _check(c2)[1]
    c2, keep2 = eng.make_config(case)
    c2.port_files[1] = None
    assert "port 1: no file" in _check(c2)[1]
    c2, keep2 = eng.make_config(case)
    kinds = (C.c_int * 2)(0, 7)
    c2.port_kinds = C.cast(kinds, C.POINTER(C.c_int))
    assert "port 1: kind" in _check(c2)[1]
    h = C.c_void_p()
    assert eng.library().cfd_model_build(C.byref(c2), C.byref(h)) == -1  # the build refuses it too


def test_config_optional_arrays(tmp_path):
    """port_kinds, src_amp and src_phase may be null (plane ports, amplitude
    1, phase 0): such a configuration passes the check and builds, with the
    same port point lists as the one that gives those values."""
    case = cases.mixed(str(tmp_path), flags=FIXED)
    case.ports = [p for p in case.ports if p[1] == 0]  # plane ports only
    case.sources = [(s[0], s[1], s[2], s[3], 1.0, 0.0) for s in case.sources]
    assert case.ports and case.sources
    cfg, keep = eng.make_config(case)
    nulls, keep2 = eng.make_config(case)
    nulls.port_kinds = C.POINTER(C.c_int)()
    nulls.src_amp = C.POINTER(C.c_double)()
    nulls.src_phase = C.POINTER(C.c_double)()
    assert _check(nulls) == (0, "")
    L = eng.library()
    ha, hb = C.c_void_p(), C.c_void_p()
    assert L.cfd_model_build(C.byref(cfg), C.byref(ha)) == 0 and L.cfd_model_build(C.byref(nulls), C.byref(hb)) == 0
    try:
        for q in range(len(case.ports)):
            n = L.cfd_model_port(ha, q, None)
            a, b = np.zeros(3 * n, np.int32), np.zeros(3 * n, np.int32)
            L.cfd_model_port(ha, q, a.ctypes.data_as(C.POINTER(C.c_int)))
            L.cfd_model_port(hb, q, b.ctypes.data_as(C.POINTER(C.c_int)))
            assert np.array_equal(a, b)
    finally:
        L.cfd_model_free(ha)
        L.cfd_model_free(hb)
