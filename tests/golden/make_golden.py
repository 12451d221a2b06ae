"""Writes the golden fixtures under tests/golden/ from the reference's own 3-D
core (oracle/_ref/refdrv, built by `make -C oracle ref` from /root/reference).

Run in a container that has /root/reference:
    python tests/golden/make_golden.py [shipped] [mixed] [tiny] [text] [extra] [facade]
(no argument: all of them)

Fixtures (all data, no reference source):
  inputs/               the reference's example inputs used here (resourses/*.txt)
  ref_shipped.npz       ID-map digest, histogram and centre planes for
                        resourses/structure.txt and Shi1.txt, and the CPML tables
                        resourses/CFS_PML.txt produces on structure.txt's mesh
  ref_mixed.npz         setup tables of tests.cases.mixed (every painter, CPML,
                        oblique TF/SF, plane + sphere port), with k0 = 0 and with
                        a nonzero beam reference height (k0 = zs/dz)
  ref_tiny.npz          40 steps of tests.cases.tiny: full fields and port samples
  ref_extra.npz         the continuous-wave TF/SF form (TF_SF.cpp:917-926):
                        Einc/Hinc and incident vectors of tests.cases.mixed(cw=1)
                        over 600 steps (ramp and steady branches), and SHA-256
                        digests of the fields and port samples after 300 steps
                        of tests.cases.tiny(cw=1); media with mu_r != 1 and a
                        magnetic conductivity (Structure.cpp:352-357): the
                        coefficient tables of tests.cases.mixed(magnetic=True)
                        and digests after 100 steps of tests.cases.tiny(
                        magnetic=True)
  ref_facade.npz        TF/SF through the one-face methods (TF_SF.cpp:1090-1410):
                        SHA-256 digests and L2 norms of the fields and port
                        samples after 80 steps of tests.cases.tiny with the
                        faces (0 x1 ... 5 z2) in three orders; Source::Pulse
                        and Source::ContinuousWave (Sources.cpp:60-76) of
                        tests.cases.tiny's source file at ntimestep = q / 2,
                        q < 2000, amplitude 1.7, locations 0 and 2e-7 m;
                        TF_SF::Einc_onPoint and ::Hinc_onPoint (TF_SF.cpp:
                        966-986) of tests.cases.tiny at retarded = q / 4 - 1,
                        q < 4 nt + 4: the pulse (nt = 80) and the continuous
                        wave (nt = 300: its ramp writes past the tables when
                        Nt is shorter, F19)
  ref_port_text.npz     the text the reference's Ports.cpp Export_* methods write
                        over 40 steps of tests.cases.tiny(text_ports=True) (a
                        line, a 2-D plane and a partial sphere port; plane
                        ports also Export_ID_onPort), as UTF-8
                        bytes per file; the entries its loops print from past
                        the last port point (defect F11: Num_of_Port_Points
                        counts one point more than it fills, and the plane
                        exports then read a further row) are cut off here, after
                        checking that their number is exactly that
"""
import hashlib
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import pyoracle as po  # noqa: E402
import cases  # noqa: E402

SHIPPED = ["material.txt", "structure.txt", "Shi1.txt", "CFS_PML.txt"]


def id_digest(ids):
    return hashlib.sha256(np.ascontiguousarray(ids, dtype="<i2").tobytes()).hexdigest()


def tables(prefix, r, out):
    for k, v in r.items():
        out[prefix + k] = v


def _values(line):
    """Tab-terminated values of one exported line."""
    assert line.endswith("\t"), line[-20:]
    return line[:-1].split("\t")


def _tabbed(values):
    return "".join(v + "\t" for v in values)


def strip_f11(texts, q, sphere, nt):
    """One port's exports without the entries read past its P filled points:
    one extra point record (the last line of the point and E lists, one value
    per row of a sphere port) and, for plane ports, one further row of Nw
    values in the value exports.  Asserts those counts."""
    t = {k[len("port%d_" % q):-4]: v for k, v in texts.items() if k.startswith("port%d_" % q)}
    pts = t["points"].split("\n")
    assert pts[-1] == "" and pts[-2] == "0\t0\t0"  # the unfilled record, zero-filled heap
    head = pts[:1] if sphere else []
    body = pts[1:-2] if sphere else pts[:-2]
    P = len(body)
    r = {"points": "\n".join(head + body) + "\n"}
    last = t["E_last"].split("\n")
    assert len(last) == P + 2 and last[-1] == ""
    r["E_last"] = "\n".join(last[:P]) + "\n"
    if sphere:
        extra = 1
        for k in ("Ex_file", "I_file"):  # one line, no newline
            v = _values(t[k])
            assert len(v) == P + extra
            r[k] = _tabbed(v[:P])
    else:
        rows = t["Ex_file"].split("\n")
        nw = len(_values(rows[0]))
        assert P % nw == 0 and len(rows) == P // nw + 2 and rows[-1] == ""
        assert all(len(_values(x)) == nw for x in rows[:-1])
        extra = nw
        r["Ex_file"] = "".join(x + "\n" for x in rows[:P // nw])
        rows = t["ID_file"].split("\n")  # same rows of Nw
        assert len(rows) == P // nw + 2 and rows[-1] == ""
        assert all(len(_values(x)) == nw for x in rows[:-1])
        r["ID_file"] = "".join(x + "\n" for x in rows[:P // nw])
        assert t["I_file"].endswith("\n")
        v = _values(t["I_file"][:-1])
        assert len(v) == P + extra
        r["I_file"] = _tabbed(v[:P]) + "\n"
    for k in ("Ex", "I"):  # one line per step
        lines = t[k].split("\n")
        assert len(lines) == nt + 1 and lines[-1] == ""
        cut = []
        for x in lines[:-1]:
            v = _values(x)
            assert len(v) == P + extra
            cut.append(_tabbed(v[:P]) + "\n")
        r[k] = "".join(cut)
    return {"port%d_%s" % (q, k): v for k, v in r.items()}


def digest(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest().encode(),
                         dtype=np.uint8)


def make_shipped():
    inp = os.path.join(HERE, "inputs")
    os.makedirs(inp, exist_ok=True)
    for f in SHIPPED:
        shutil.copy(os.path.join(cases.REF_RESOURCES, f), os.path.join(inp, f))
    out = {}
    for tag, sfile, pml in [("structure", "structure.txt", True), ("shi1", "Shi1.txt", False)]:
        d = tempfile.mkdtemp()
        case = cases.shipped(d, sfile, with_pml=pml)
        r = po.ref_tables(case)
        dims = r["dims"]
        ids = r["id"].reshape(dims[0] + 1, dims[1] + 1, dims[2] + 1)
        out[tag + "_dims"] = dims
        out[tag + "_d"] = r["d"]
        out[tag + "_sha256"] = np.frombuffer(id_digest(ids).encode(), dtype=np.uint8)
        out[tag + "_hist"] = np.bincount(ids.ravel().astype(np.int64), minlength=16)
        out[tag + "_plane_x"] = ids[dims[0] // 2].copy()
        out[tag + "_plane_y"] = ids[:, dims[1] // 2].copy()
        out[tag + "_plane_z"] = ids[:, :, dims[2] // 2].copy()
        out[tag + "_coef"] = r["coef"]
        out[tag + "_stretch"] = r["stretch"]
        if pml:
            out[tag + "_cpml_len"] = r["cpml_len"]
            out[tag + "_cpml"] = r["cpml"]
            out[tag + "_cpml_c"] = r["cpml_c"]
    np.savez_compressed(os.path.join(HERE, "ref_shipped.npz"), **out)


def make_mixed():
    out = {}
    # k0set: the TF_SF object's unset k0 is given the value the oracle derives,
    # (int)(zs/dz), so both place the beam reference point identically
    for tag, kw, k0 in [("k0zero", {}, None), ("k0set", {"zs": 120e-9}, int(120e-9 / 10e-9))]:
        d = tempfile.mkdtemp()
        case = cases.mixed(d, **kw)
        r = po.ref_tables(case, k0=k0)
        tables(tag + "_", r, out)
    np.savez_compressed(os.path.join(HERE, "ref_mixed.npz"), **out)


def make_tiny():
    d = tempfile.mkdtemp()
    case = cases.tiny(d)
    F, P = po.ref_run(case, case.nt, src=(8, 9, 10, 2.0))
    out = {"nt": np.array([case.nt])}
    for c, nm in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        out[nm] = F[c]
    out["port0"] = P[0]
    np.savez_compressed(os.path.join(HERE, "ref_tiny.npz"), **out)


def make_text():
    d = tempfile.mkdtemp()
    case = cases.tiny(d, text_ports=True)
    texts = po.ref_port_text(case, case.nt, src=(8, 9, 10, 2.0))
    out = {}
    for q, (_, kind) in enumerate(case.ports):
        for name, text in strip_f11(texts, q, kind == 1, case.nt).items():
            out[name] = np.frombuffer(text.encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "ref_port_text.npz"), **out)


def make_extra():
    out = {}
    # nt >= the 5-period ramp (437 steps here): the reference writes its ramp
    # past Einc/Hinc[Nt] otherwise (TF_SF.cpp:917-921)
    r = po.ref_tables(cases.mixed(tempfile.mkdtemp(), cw=1, nt=600))
    for k in ("tfsf_einc", "tfsf_hinc", "tfsf_vec"):
        out["cw_" + k] = r[k]
    r = po.ref_tables(cases.mixed(tempfile.mkdtemp(), magnetic=True))
    out["mag_coef"] = r["coef"]
    for tag, kw, nt in (("cw", {"cw": 1}, 300), ("mag", {"magnetic": True}, 100)):
        case = cases.tiny(tempfile.mkdtemp(), nt=nt, **kw)
        F, P = po.ref_run(case, nt, src=(8, 9, 10, 2.0))
        out[tag + "_nt"] = np.array([nt])
        out[tag + "_field_sha256"] = np.array([digest(F[c]) for c in range(6)])
        out[tag + "_field_l2"] = np.array([float(np.linalg.norm(F[c])) for c in range(6)])
        out[tag + "_port0_sha256"] = digest(P[0])
    np.savez_compressed(os.path.join(HERE, "ref_extra.npz"), **out)


FACE_ORDERS = ([2, 0, 4, 5, 1, 3], [4], [3, 1])
SOURCE_LOCS = (0.0, 2e-7)


def make_facade():
    out = {"face_nt": np.array([80])}
    for q, faces in enumerate(FACE_ORDERS):
        case = cases.tiny(tempfile.mkdtemp(), nt=80)
        F, P = po.ref_run(case, 80, src=(8, 9, 10, 2.0), faces=faces)
        out["faces%d" % q] = np.array(faces)
        out["faces%d_field_sha256" % q] = np.array([digest(F[c]) for c in range(6)])
        out["faces%d_field_l2" % q] = np.array([float(np.linalg.norm(F[c])) for c in range(6)])
        out["faces%d_port0_sha256" % q] = digest(P[0])
    case = cases.tiny(tempfile.mkdtemp(), nt=80)
    for q, loc in enumerate(SOURCE_LOCS):
        r = po.ref_source(case, 2000, 1.7, loc)
        out["pulse%d" % q] = r["pulse"]
        out["cw%d" % q] = r["cw"]
    out["source_locs"] = np.array(SOURCE_LOCS)
    for cw, nt in ((0, 80), (1, 300)):
        out["inc_cw%d" % cw] = po.ref_inc(cases.tiny(tempfile.mkdtemp(), nt=nt, cw=cw), 4 * nt + 4)
        out["inc_cw%d_nt" % cw] = np.array([nt])
    np.savez_compressed(os.path.join(HERE, "ref_facade.npz"), **out)


MAKERS = {"shipped": make_shipped, "mixed": make_mixed, "tiny": make_tiny, "text": make_text, "extra": make_extra,
          "facade": make_facade}


def main():
    po.build()
    if not po.ref_available():
        raise SystemExit("reference driver missing: run make -C oracle ref (needs /root/reference)")
    for name in sys.argv[1:] or list(MAKERS):
        MAKERS[name]()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
