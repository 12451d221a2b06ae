"""ctypes binding to the FP64 oracle (oracle/liboracle.so) and the reference
driver (oracle/_ref/refdrv).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORC_LIB selects another build (the sanitizer build, _san/liboracle.so)
LIB_PATH = os.environ.get("ORC_LIB") or os.path.join(HERE, "liboracle.so")
REFDRV = os.path.join(HERE, "_ref", "refdrv")
REFDRV_O0 = os.path.join(HERE, "_ref", "refdrv_O0")  # reference sources at -O0 (timing only)


def build(quiet=True):
    """Compile liboracle.so (and the reference driver when the sources exist)."""
    targets = ["all"]
    if os.path.isdir("/root/reference"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


class _Config(C.Structure):
    _fields_ = [
        ("material", C.c_char_p), ("structure", C.c_char_p), ("pml", C.c_char_p), ("tfsf", C.c_char_p),
        ("dl", C.c_char_p), ("gain", C.c_char_p), ("source", C.c_char_p),
        ("n_ports", C.c_int), ("port_files", C.POINTER(C.c_char_p)), ("port_kinds", C.POINTER(C.c_int)),
        ("n_src", C.c_int), ("src_ijkc", C.POINTER(C.c_int)), ("src_amp", C.POINTER(C.c_double)),
        ("src_phase", C.POINTER(C.c_double)),
        ("nt", C.c_int), ("n_freq", C.c_int), ("freqs", C.POINTER(C.c_double)), ("dft_stride", C.c_int),
        ("record_ports", C.c_int), ("periodic", C.c_int), ("flags", C.c_uint),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(_Config)]
        L.orc_destroy.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        for name, args, res in [
            ("orc_dims", [vp, ip, dp], None), ("orc_nmedia", [vp], C.c_int),
            ("orc_get_id", [vp, C.POINTER(C.c_short)], None), ("orc_get_coef", [vp, C.c_int, dp], None),
            ("orc_get_stretch", [vp, C.c_int, dp], C.c_int), ("orc_get_cpml", [vp, C.c_int, dp, dp], C.c_int),
            ("orc_get_tfsf_points", [vp, ip, dp], C.c_int), ("orc_get_tfsf_wave", [vp, dp, dp], C.c_int),
            ("orc_get_tfsf_vec", [vp, dp], None), ("orc_port_npts", [vp, C.c_int], C.c_int),
            ("orc_port_points", [vp, C.c_int, ip], None), ("orc_port_samples", [vp, C.c_int, dp], None),
            ("orc_port_dft", [vp, C.c_int, dp], None), ("orc_port_intensity", [vp, C.c_int, dp], None),
            ("orc_step", [vp, C.c_int], C.c_int), ("orc_steps_done", [vp], C.c_int),
            ("orc_get_field", [vp, C.c_int, dp], None), ("orc_set_field", [vp, C.c_int, dp], None),
            ("orc_gain_ncells", [vp], C.c_int), ("orc_gain_pop", [vp, dp], None),
            ("orc_dl_ncells", [vp], C.c_int), ("orc_dl_coefs", [vp, dp], C.c_int),
            ("orc_source_wave", [vp, C.c_int, dp], C.c_int),
            ("orc_dl_0d", [vp, C.c_int, dp, dp], C.c_int), ("orc_gain_0d", [vp, C.c_int, dp, dp, dp], C.c_int),
            ("orc_set_tfsf_faces", [vp, ip, C.c_int], C.c_int), ("orc_dl_state", [vp, dp], C.c_int),
            ("orc_gain_state", [vp, dp], C.c_int), ("orc_special_cells", [vp, C.c_int, ip], C.c_int),
            ("orc_dl_poles", [vp, ip], None),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _enc(s):
    return s.encode() if s else None


class OracleSim:
    """FP64 CPU restatement of the reference time loop for one case."""

    def __init__(self, case):
        L = lib()
        f = case.files
        self._keep = []
        cfg = _Config()
        for k in ("material", "structure", "pml", "tfsf", "dl", "gain", "source"):
            setattr(cfg, k, _enc(f[k]))
        np_ = len(case.ports)
        pf = (C.c_char_p * max(np_, 1))(*[_enc(p) for p, _ in case.ports])
        pk = np.array([k for _, k in case.ports] or [0], dtype=np.int32)
        src = np.array([s[:4] for s in case.sources] or [[0, 0, 0, 0]], dtype=np.int32).reshape(-1)
        amp = np.array([s[4] for s in case.sources] or [0.0], dtype=np.float64)
        ph = np.array([s[5] for s in case.sources] or [0.0], dtype=np.float64)
        fr = np.array(case.freqs or [0.0], dtype=np.float64)
        self._keep += [pf, pk, src, amp, ph, fr]
        cfg.n_ports = np_
        cfg.port_files = pf
        cfg.port_kinds = _ip(pk)
        cfg.n_src = len(case.sources)
        cfg.src_ijkc = _ip(src)
        cfg.src_amp = _dp(amp)
        cfg.src_phase = _dp(ph)
        cfg.nt = case.nt
        cfg.n_freq = len(case.freqs)
        cfg.freqs = _dp(fr)
        cfg.dft_stride = case.dft_stride
        cfg.record_ports = int(case.record_ports)
        cfg.periodic = case.periodic
        cfg.flags = case.flags
        self.h = L.orc_create(C.byref(cfg))
        if not self.h:
            raise RuntimeError("oracle: " + L.orc_last_error().decode())
        n = np.zeros(4, np.int32)
        d = np.zeros(4, np.float64)
        L.orc_dims(self.h, _ip(n), _dp(d))
        self.N = tuple(int(x) for x in n[:3])
        self.nt = int(n[3])
        self.d = tuple(float(x) for x in d[:3])
        self.dt = float(d[3])
        self.shape = tuple(x + 1 for x in self.N)
        self.nports = np_
        self.nf = len(case.freqs)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        if lib().orc_step(self.h, n):
            raise RuntimeError(lib().orc_last_error().decode())

    @property
    def steps_done(self):
        return lib().orc_steps_done(self.h)

    def field(self, comp):
        a = np.empty(self.shape, np.float64)
        lib().orc_get_field(self.h, comp, _dp(a))
        return a

    def fields(self):
        return [self.field(c) for c in range(6)]

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        lib().orc_set_field(self.h, comp, _dp(a))

    def id_map(self):
        a = np.empty(self.shape, np.int16)
        lib().orc_get_id(self.h, a.ctypes.data_as(C.POINTER(C.c_short)))
        return a

    def coefs(self):
        n = lib().orc_nmedia(self.h)
        out = []
        for w in range(4):
            a = np.empty(n, np.float64)
            lib().orc_get_coef(self.h, w, _dp(a))
            out.append(a)
        return out

    def stretch(self, which):
        n = lib().orc_get_stretch(self.h, which, None)
        a = np.empty(n, np.float64)
        lib().orc_get_stretch(self.h, which, _dp(a))
        return a

    def cpml(self, which):
        n = lib().orc_get_cpml(self.h, which, None, None)
        b = np.empty(n, np.float64)
        c = np.empty(n, np.float64)
        if n:
            lib().orc_get_cpml(self.h, which, _dp(b), _dp(c))
        return b, c

    def tfsf_points(self):
        n = lib().orc_get_tfsf_points(self.h, None, None)
        ijk = np.empty((max(n, 1), 4), np.int32)
        tab = np.empty((max(n, 1), 8), np.float64)  # per record: pair 1 (dE pE dH pH), pair 2
        lib().orc_get_tfsf_points(self.h, _ip(ijk), _dp(tab))
        return ijk[:n], tab[:n]

    def tfsf_wave(self):
        n = lib().orc_get_tfsf_wave(self.h, None, None)
        e = np.empty(n, np.float64)
        h = np.empty(n, np.float64)
        lib().orc_get_tfsf_wave(self.h, _dp(e), _dp(h))
        return e, h

    def tfsf_vec(self):
        a = np.empty(12, np.float64)
        lib().orc_get_tfsf_vec(self.h, _dp(a))
        return a

    def port_points(self, q):
        n = lib().orc_port_npts(self.h, q)
        a = np.empty((n, 3), np.int32)
        lib().orc_port_points(self.h, q, _ip(a))
        return a

    def port_samples(self, q):
        n = lib().orc_port_npts(self.h, q)
        a = np.zeros((self.steps_done, n, 6), np.float64)
        lib().orc_port_samples(self.h, q, _dp(a))
        return a

    def port_dft(self, q):
        n = lib().orc_port_npts(self.h, q)
        a = np.zeros((n, self.nf, 6, 2), np.float64)
        lib().orc_port_dft(self.h, q, _dp(a))
        return a[..., 0] + 1j * a[..., 1]

    def port_intensity(self, q):
        a = np.zeros(self.steps_done, np.float64)
        lib().orc_port_intensity(self.h, q, _dp(a))
        return a

    def gain_pop(self):
        n = lib().orc_gain_ncells(self.h)
        a = np.zeros((n, 4), np.float64)
        lib().orc_gain_pop(self.h, _dp(a))
        return a

    def source_wave(self, q):
        n = lib().orc_source_wave(self.h, q, None)
        a = np.zeros(n, np.float64)
        if n:
            lib().orc_source_wave(self.h, q, _dp(a))
        return a

    def dl_ncells(self):
        return lib().orc_dl_ncells(self.h)

    def set_tfsf_faces(self, faces):
        """Step with TF_SF's one-face updates in this order (face = 2 * axis
        + high: 0 x1 ... 5 z2) instead of the whole update; [] restores it."""
        f = np.ascontiguousarray(faces, dtype=np.int32)
        if lib().orc_set_tfsf_faces(self.h, _ip(f), len(f)) < 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def dl_state(self):
        """[cell][3][nD + 2 nL + 1]: JD poles, JL poles, JL_1 poles, E_1."""
        poles = np.zeros(2, np.int32)
        lib().orc_dl_poles(self.h, _ip(poles))
        per = int(poles[0]) + 2 * int(poles[1]) + 1
        n = lib().orc_dl_state(self.h, None)
        a = np.zeros(max(n, 1) * 3 * per, np.float64)
        lib().orc_dl_state(self.h, _dp(a))
        return a[:n * 3 * per].reshape(n, 3, per)

    def gain_state(self):
        """[cell][6]: P x y z, previous P x y z."""
        n = lib().orc_gain_state(self.h, None)
        a = np.zeros((max(n, 1), 6), np.float64)
        lib().orc_gain_state(self.h, _dp(a))
        return a[:n]

    def special_cells(self, kind):
        """[cell][3] i j k of the Drude-Lorentz (0) or gain (1) cells."""
        n = lib().orc_special_cells(self.h, kind, None)
        a = np.zeros((max(n, 1), 3), np.int32)
        lib().orc_special_cells(self.h, kind, _ip(a))
        return a[:n]

    def dl_coefs(self):
        a = np.zeros(64, np.float64)
        n = lib().orc_dl_coefs(self.h, _dp(a))
        return a[:n]

    def dl_0d(self, curl):
        curl = np.ascontiguousarray(curl, dtype=np.float64)
        out = np.zeros_like(curl)
        lib().orc_dl_0d(self.h, len(curl), _dp(curl), _dp(out))
        return out

    def gain_0d(self, efield):
        e = np.ascontiguousarray(efield, dtype=np.float64)
        pop = np.zeros((len(e), 4), np.float64)
        p = np.zeros(len(e), np.float64)
        lib().orc_gain_0d(self.h, len(e), _dp(e), _dp(pop), _dp(p))
        return pop, p


# ---------------------------------------------------------------------------
# reference driver (compiled reference sources), this container only
# ---------------------------------------------------------------------------
def ref_available():
    return os.path.exists(REFDRV)


def _port_arg(case):
    if not case.ports:
        return []
    return ["--ports", ",".join(("s:" if k == 1 else "p:") + p for p, k in case.ports)]


def ref_tables(case, k0=None):
    """Run ``refdrv tables`` on the case; returns a dict of arrays."""
    out = tempfile.mkdtemp(prefix="refdrv_")
    args = [REFDRV, "tables", case.dir, out] + _port_arg(case)
    if k0 is not None:
        args += ["--k0", str(k0)]
    if not case.tfsf:
        args += ["--notfsf"]
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    r = {}
    for name in os.listdir(out):
        stem, ext = os.path.splitext(name)
        dt = {".i32": np.int32, ".i16": np.int16, ".f64": np.float64}[ext]
        r[stem] = np.fromfile(os.path.join(out, name), dtype=dt)
    return r


FACE_NAMES = ("x1", "x2", "y1", "y2", "z1", "z2")


def ref_run(case, nsteps, src=None, k0=None, faces=None):
    """Run the reference loop; returns (fields[6], port samples list).
    faces: TF/SF through the one-face methods in this order (0 x1 ... 5 z2)."""
    out = tempfile.mkdtemp(prefix="refdrv_")
    args = [REFDRV, "run", case.dir, out, str(nsteps)] + _port_arg(case)
    if faces:
        args += ["--faces", ",".join(FACE_NAMES[f] for f in faces)]
    if src is not None:
        args += ["--src"] + [str(x) for x in src]
    if k0 is not None:
        args += ["--k0", str(k0)]
    if not case.tfsf:
        args += ["--notfsf"]
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    dims = None
    fields = []
    for nm in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        fields.append(np.fromfile(os.path.join(out, nm + ".f64"), dtype=np.float64))
    ports = []
    for q in range(len(case.ports)):
        ports.append(np.fromfile(os.path.join(out, "port%d_samples.f64" % q), dtype=np.float64).reshape(nsteps, -1, 6))
    return fields, ports


def ref_source(case, nhalf, amp, loc):
    """Source::Pulse and Source::ContinuousWave of the case's source file at
    ntimestep = q / 2 (q < nhalf), amplitude amp, location loc (metres)."""
    out = tempfile.mkdtemp(prefix="refdrv_")
    subprocess.run([REFDRV, "source", case.dir, out, str(nhalf), repr(float(amp)), repr(float(loc))], check=True,
                   stdout=subprocess.DEVNULL)
    r = {k: np.fromfile(os.path.join(out, k + ".f64"), dtype=np.float64) for k in ("pulse", "cw")}
    shutil.rmtree(out, ignore_errors=True)
    return r


def ref_inc(case, nq):
    """TF_SF::Einc_onPoint, ::Hinc_onPoint at retarded = q / 4 - 1 (q < nq),
    [q][E, H]."""
    out = tempfile.mkdtemp(prefix="refdrv_")
    subprocess.run([REFDRV, "inc", case.dir, out, str(nq)], check=True, stdout=subprocess.DEVNULL)
    r = np.fromfile(os.path.join(out, "inc.f64"), dtype=np.float64).reshape(nq, 2)
    shutil.rmtree(out, ignore_errors=True)
    return r


def ref_port_text(case, nsteps, src=None):
    """Run the reference loop with --export; returns {file name: text} of the
    files its own Ports.cpp Export_* methods wrote."""
    out = tempfile.mkdtemp(prefix="refdrv_")
    args = [REFDRV, "run", case.dir, out, str(nsteps), "--export"] + _port_arg(case)
    if src is not None:
        args += ["--src"] + [str(x) for x in src]
    if not case.tfsf:
        args += ["--notfsf"]
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    texts = {}
    for name in sorted(os.listdir(out)):
        if name.endswith(".txt"):
            with open(os.path.join(out, name)) as f:
                texts[name] = f.read()
    shutil.rmtree(out, ignore_errors=True)
    return texts


def ref_bench(case, nsteps, threads=None, opt="O2"):
    """Time the reference loop (sources at -O2, or -O0 with opt="O0")."""
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    drv = REFDRV_O0 if opt == "O0" else REFDRV
    args = [drv, "bench", case.dir, str(nsteps)] + (["--notfsf"] if not case.tfsf else [])
    r = subprocess.run(args, check=True, capture_output=True, text=True, env=env)
    line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
    return json.loads(line)
