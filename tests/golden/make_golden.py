"""Writes the golden fixtures under tests/golden/ from the reference's own 3-D
core (oracle/_ref/refdrv, built by `make -C oracle ref` from /root/reference).

Run in a container that has /root/reference:
    python tests/golden/make_golden.py

Fixtures (all data, no reference source):
  inputs/               the reference's example inputs used here (resourses/*.txt)
  ref_shipped.npz       ID-map digest, histogram and centre planes for
                        resourses/structure.txt and Shi1.txt, and the CPML tables
                        resourses/CFS_PML.txt produces on structure.txt's mesh
  ref_mixed.npz         setup tables of tests.cases.mixed (every painter, CPML,
                        oblique TF/SF, plane + sphere port), with k0 = 0 and with
                        a nonzero beam reference height (k0 = zs/dz)
  ref_tiny.npz          40 steps of tests.cases.tiny: full fields and port samples
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


def main():
    po.build()
    if not po.ref_available():
        raise SystemExit("reference driver missing: run make -C oracle ref (needs /root/reference)")
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

    out = {}
    # k0set: the TF_SF object's unset k0 is given the value the oracle derives,
    # (int)(zs/dz), so both place the beam reference point identically
    for tag, kw, k0 in [("k0zero", {}, None), ("k0set", {"zs": 120e-9}, int(120e-9 / 10e-9))]:
        d = tempfile.mkdtemp()
        case = cases.mixed(d, **kw)
        r = po.ref_tables(case, k0=k0)
        tables(tag + "_", r, out)
    np.savez_compressed(os.path.join(HERE, "ref_mixed.npz"), **out)

    d = tempfile.mkdtemp()
    case = cases.tiny(d)
    F, P = po.ref_run(case, case.nt, src=(8, 9, 10, 2.0))
    out = {"nt": np.array([case.nt])}
    for c, nm in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        out[nm] = F[c]
    out["port0"] = P[0]
    np.savez_compressed(os.path.join(HERE, "ref_tiny.npz"), **out)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
