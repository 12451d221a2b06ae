"""MI355X-native 3-D FDTD engine (cpsFDTD's time-stepping loop on HIP/gfx950).

Python-side host interface over the C ABI in include/cpsfdtd.h
(libcpsfdtd.so).  The engine itself is C++ and HIP; this module only loads
the library, fills the configuration struct and moves arrays.  There is no
CPU fallback: if the library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .case import Case  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
# CFD_LIB selects another build of the same library (the sanitizer build of
# tests/run_sanitized.sh); the default is the in-tree one
LIB_PATH = os.environ.get("CFD_LIB") or os.path.join(HERE, "lib", "libcpsfdtd.so")

CFD_F32 = 32
CFD_F64 = 64


class EngineError(RuntimeError):
    pass


class _Config(C.Structure):
    _fields_ = [
        ("material", C.c_char_p), ("structure", C.c_char_p), ("pml", C.c_char_p), ("tfsf", C.c_char_p),
        ("dl", C.c_char_p), ("gain", C.c_char_p), ("source", C.c_char_p),
        ("n_ports", C.c_int), ("port_files", C.POINTER(C.c_char_p)), ("port_kinds", C.POINTER(C.c_int)),
        ("n_src", C.c_int), ("src_ijkc", C.POINTER(C.c_int)), ("src_amp", C.POINTER(C.c_double)),
        ("src_phase", C.POINTER(C.c_double)),
        ("nt", C.c_int), ("n_freq", C.c_int), ("freqs", C.POINTER(C.c_double)), ("dft_stride", C.c_int),
        ("record_ports", C.c_int), ("periodic", C.c_int), ("flags", C.c_uint), ("precision", C.c_int),
    ]


# every entry point declared in include/cpsfdtd.h: name -> (argtypes, restype)
_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
SIGNATURES = {
    "cfd_last_error": ([], C.c_char_p),
    "cfd_device_info": ([C.c_char_p, C.c_int, _ip, C.POINTER(C.c_size_t)], C.c_int),
    "cfd_create": ([C.POINTER(_Config), C.POINTER(_vp)], C.c_int),
    "cfd_destroy": ([_vp], C.c_int),
    "cfd_info": ([_vp, _ip, _dp, _ip], C.c_int),
    "cfd_device_bytes": ([_vp], C.c_size_t),
    "cfd_step": ([_vp, C.c_int], C.c_int),
    "cfd_steps_done": ([_vp], C.c_int),
    "cfd_sync": ([_vp], C.c_int),
    "cfd_basic_h": ([_vp], C.c_int),
    "cfd_basic_e": ([_vp], C.c_int),
    "cfd_cpml_h": ([_vp], C.c_int),
    "cfd_cpml_e": ([_vp], C.c_int),
    "cfd_tfsf_h": ([_vp, C.c_int], C.c_int),
    "cfd_tfsf_e": ([_vp, C.c_int], C.c_int),
    "cfd_tfsf_face_h": ([_vp, C.c_int, C.c_int], C.c_int),
    "cfd_tfsf_face_e": ([_vp, C.c_int, C.c_int], C.c_int),
    "cfd_dl_e": ([_vp], C.c_int),
    "cfd_gain_pol": ([_vp], C.c_int),
    "cfd_gain_e": ([_vp], C.c_int),
    "cfd_gain_pop": ([_vp], C.c_int),
    "cfd_sources": ([_vp, C.c_int], C.c_int),
    "cfd_ports_sample": ([_vp, C.c_int], C.c_int),
    "cfd_advance": ([_vp], C.c_int),
    "cfd_read_field": ([_vp, C.c_int, _dp], C.c_int),
    "cfd_write_field": ([_vp, C.c_int, _dp], C.c_int),
    "cfd_read_region": ([_vp, C.c_int, _ip, _dp], C.c_int),
    "cfd_read_id": ([_vp, C.POINTER(C.c_int16)], C.c_int),
    "cfd_port_npts": ([_vp, C.c_int], C.c_int),
    "cfd_port_points": ([_vp, C.c_int, _ip], C.c_int),
    "cfd_port_samples": ([_vp, C.c_int, _dp], C.c_int),
    "cfd_port_gather": ([_vp, C.c_int, _dp], C.c_int),
    "cfd_port_gather_static": ([_vp, C.c_int, _dp], C.c_int),
    "cfd_port_dft": ([_vp, C.c_int, _dp], C.c_int),
    "cfd_port_intensity": ([_vp, C.c_int, _dp], C.c_int),
    "cfd_gain_read": ([_vp, _dp], C.c_int),
    "cfd_dl_state": ([_vp, _dp], C.c_int),
    "cfd_gain_state": ([_vp, _dp], C.c_int),
    "cfd_special_cells": ([_vp, C.c_int, _ip], C.c_int),
    "cfd_tfsf_wave": ([_vp, _dp, _dp], C.c_int),
    "cfd_model_build": ([C.POINTER(_Config), C.POINTER(_vp)], C.c_int),
    "cfd_create_from_model": ([_vp, C.c_int, C.POINTER(_vp)], C.c_int),
    "cfd_model_free": ([_vp], C.c_int),
    "cfd_model_last_error": ([], C.c_char_p),
    "cfd_model_dims": ([_vp, _ip, _dp], C.c_int),
    "cfd_model_id": ([_vp, C.POINTER(C.c_int16)], C.c_int),
    "cfd_model_table": ([_vp, C.c_char_p, _dp], C.c_int),
    "cfd_model_tfsf": ([_vp, C.c_int, _ip, _dp, _ip], C.c_int),
    "cfd_model_port": ([_vp, C.c_int, _ip], C.c_int),
    "cfd_model_cells": ([_vp, C.c_int, C.POINTER(C.c_longlong)], C.c_int),
    "cfd_kernel_times": ([_vp, C.c_void_p, _dp, _ip, C.c_int], C.c_int),
    "cfd_batch_create": ([C.POINTER(_Config), C.c_int, C.POINTER(_vp)], C.c_int),
    "cfd_batch_destroy": ([_vp], C.c_int),
    "cfd_batch_size": ([_vp], C.c_int),
    "cfd_batch_step": ([_vp, C.c_int], C.c_int),
    "cfd_batch_sync": ([_vp], C.c_int),
    "cfd_batch_member": ([_vp, C.c_int, C.POINTER(_vp)], C.c_int),
    "cfd_batch_device_bytes": ([_vp, C.POINTER(C.c_size_t)], C.c_int),
    "cfd_batch_set_timing": ([_vp, C.c_int], C.c_int),
    "cfd_batch_kernel_times": ([_vp, C.c_void_p, _dp, _ip, C.c_int], C.c_int),
    "cfd_set_timing": ([_vp, C.c_int], C.c_int),
    "cfd_query": ([_vp], C.c_int),
    "cfd_check_finite": ([_vp], C.c_int),
    "cfd_server_create": ([C.c_void_p, C.POINTER(_vp)], C.c_int),
    "cfd_server_destroy": ([_vp], C.c_int),
    "cfd_server_submit": ([_vp, C.POINTER(_Config), C.POINTER(C.c_longlong)], C.c_int),
    "cfd_server_poll": ([_vp, C.c_longlong, _dp], C.c_int),
    "cfd_server_wait": ([_vp, C.c_longlong, C.c_double], C.c_int),
    "cfd_server_result": ([_vp, C.c_longlong, C.c_int, C.c_int, _dp, C.POINTER(C.c_size_t)], C.c_int),
    "cfd_server_request_error": ([_vp, C.c_longlong], C.c_char_p),
    "cfd_server_release": ([_vp, C.c_longlong], C.c_int),
    "cfd_server_stats": ([_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_size_t), _ip, _ip], C.c_int),
    "cfd_server_last_error": ([], C.c_char_p),
    "cfd_port_flux": ([C.c_void_p, C.c_int, C.c_int, _dp], C.c_int),
    "cfd_dft_flux": ([_dp, C.c_int, C.c_int, C.c_int, C.c_double, _dp], C.c_int),
    "cfd_tfsf_tables": ([C.c_char_p, C.c_double, C.c_int, _dp, _dp], C.c_int),
    "cfd_model_device_bytes": ([_vp, C.c_int], C.c_size_t),
    "cfd_image_colormap": ([_dp, C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_ubyte)], C.c_int),
    "cfd_image_materials": ([C.POINTER(C.c_short), C.c_int, C.c_int, C.POINTER(C.c_ubyte)], C.c_int),
    "cfd_image_write_png": ([C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_ubyte)], C.c_int),
    "cfd_image_last_error": ([], C.c_char_p),
    "cfd_config_check": ([C.POINTER(_Config)], C.c_int),
    "cfd_pack_create": ([C.c_int, C.POINTER(_vp)], C.c_int),
    "cfd_pack_destroy": ([_vp], C.c_int),
    "cfd_pack_add": ([_vp, _vp], C.c_int),
    "cfd_pack_remove": ([_vp, _vp], C.c_int),
    "cfd_pack_size": ([_vp], C.c_int),
    "cfd_pack_step": ([_vp, C.c_int], C.c_int),
    "cfd_pack_sync": ([_vp], C.c_int),
    "cfd_pack_query": ([_vp], C.c_int),
    "cfd_pack_check_finite": ([_vp, C.POINTER(_vp), C.c_int, _ip], C.c_int),
    "cfd_pack_set_timing": ([_vp, C.c_int], C.c_int),
    "cfd_pack_kernel_times": ([_vp, C.c_void_p, _dp, _ip, C.c_int], C.c_int),
}

_lib = None


def library(path: str = LIB_PATH):
    """Load libcpsfdtd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise EngineError("libcpsfdtd.so not built: run `python -m cpsfdtd_amd.build` (%s)" % path)
        L = C.CDLL(path)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise EngineError(library().cfd_last_error().decode())
    return rc


def _enc(s):
    return s.encode() if s else None


def device_info():
    name = C.create_string_buffer(256)
    ncu = C.c_int(0)
    mem = C.c_size_t(0)
    _check(library().cfd_device_info(name, 256, C.byref(ncu), C.byref(mem)))
    return {"name": name.value.decode(), "cu": ncu.value, "hbm_bytes": mem.value}


def make_config(case, precision=CFD_F64):
    """Fill the C configuration struct from a Case; returns (config, keepalive)."""
    f = case.files
    keep = []
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
    keep += [pf, pk, src, amp, ph, fr]
    cfg.n_ports = np_
    cfg.port_files = pf
    cfg.port_kinds = pk.ctypes.data_as(_ip)
    cfg.n_src = len(case.sources)
    cfg.src_ijkc = src.ctypes.data_as(_ip)
    cfg.src_amp = amp.ctypes.data_as(_dp)
    cfg.src_phase = ph.ctypes.data_as(_dp)
    cfg.nt = case.nt
    cfg.n_freq = len(case.freqs)
    cfg.freqs = fr.ctypes.data_as(_dp)
    cfg.dft_stride = case.dft_stride
    cfg.record_ports = int(case.record_ports)
    cfg.periodic = case.periodic
    cfg.flags = case.flags
    cfg.precision = precision
    return cfg, keep


class Engine:
    """One device-resident simulation (handle of the C ABI)."""

    def __init__(self, case, precision=CFD_F64, _handle=None):
        L = library()
        if _handle is None:
            cfg, self._keep = make_config(case, precision)
            h = C.c_void_p()
            _check(L.cfd_create(C.byref(cfg), C.byref(h)))
            self._owned = True
        else:  # borrowed view of a batch member
            h, self._keep, self._owned = _handle, None, False
        self.h = h
        self.precision = precision
        n = np.zeros(4, np.int32)
        d = np.zeros(4, np.float64)
        cnt = np.zeros(4, np.int32)
        _check(L.cfd_info(self.h, n.ctypes.data_as(_ip), d.ctypes.data_as(_dp), cnt.ctypes.data_as(_ip)))
        self.N = tuple(int(x) for x in n[:3])
        self.nt = int(n[3])
        self.d = tuple(float(x) for x in d[:3])
        self.dt = float(d[3])
        self.shape = tuple(x + 1 for x in self.N)
        self.nports, self.n_dl, self.n_gain, self.n_tfsf = (int(x) for x in cnt)
        self.nf = len(case.freqs)
        self.dl_per = 0  # values per component of a Drude-Lorentz cell: nD + 2 nL + 1
        if case.files.get("dl"):
            with open(case.files["dl"]) as f:
                nD, nL = (int(f.readline().split()[0]) for _ in range(2))
            self.dl_per = nD + 2 * nL + 1

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                library().cfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping -------------------------------------------------------
    def step(self, n=1):
        _check(library().cfd_step(self.h, n))

    def sync(self):
        _check(library().cfd_sync(self.h))

    @property
    def steps_done(self):
        return _check(library().cfd_steps_done(self.h))

    def device_bytes(self):
        return int(library().cfd_device_bytes(self.h))

    def module_step(self, n, tfsf=True):
        """One step as the reference's main.cpp sequences its modules
        (main.cpp:267-343), one device pass per module call."""
        L = library()
        for f in (L.cfd_basic_h,):
            _check(f(self.h))
        if tfsf:
            _check(L.cfd_tfsf_h(self.h, n))
        _check(L.cfd_cpml_h(self.h))
        _check(L.cfd_gain_pol(self.h))
        _check(L.cfd_basic_e(self.h))
        _check(L.cfd_dl_e(self.h))
        _check(L.cfd_gain_e(self.h))
        _check(L.cfd_gain_pop(self.h))
        if tfsf:
            _check(L.cfd_tfsf_e(self.h, n))
        _check(L.cfd_sources(self.h, n))
        _check(L.cfd_cpml_e(self.h))
        if self.nports:
            _check(L.cfd_ports_sample(self.h, n))
        _check(L.cfd_advance(self.h))

    # -- host access ----------------------------------------------------
    def field(self, comp):
        a = np.empty(self.shape, np.float64)
        _check(library().cfd_read_field(self.h, comp, a.ctypes.data_as(_dp)))
        return a

    def fields(self):
        return [self.field(c) for c in range(6)]

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        _check(library().cfd_write_field(self.h, comp, a.ctypes.data_as(_dp)))

    def id_map(self):
        a = np.empty(self.shape, np.int16)
        _check(library().cfd_read_id(self.h, a.ctypes.data_as(C.POINTER(C.c_int16))))
        return a

    def port_points(self, q):
        n = _check(library().cfd_port_npts(self.h, q))
        a = np.empty((n, 3), np.int32)
        _check(library().cfd_port_points(self.h, q, a.ctypes.data_as(_ip)))
        return a

    def port_samples(self, q):
        n = _check(library().cfd_port_npts(self.h, q))
        a = np.zeros((self.steps_done, n, 6), np.float64)
        _check(library().cfd_port_samples(self.h, q, a.ctypes.data_as(_dp)))
        return a

    def region(self, comp, i0, i1, j0, j1, k0, k1):
        """Cells [i0,i1) x [j0,j1) x [k0,k1) of one component, [i][j][k]."""
        box = np.array([i0, i1, j0, j1, k0, k1], np.int32)
        a = np.empty((i1 - i0, j1 - j0, k1 - k0), np.float64)
        _check(library().cfd_read_region(self.h, comp, box.ctypes.data_as(_ip), a.ctypes.data_as(_dp)))
        return a

    def nonfinite(self):
        """Number of NaN / inf values in the six fields now."""
        return _check(library().cfd_check_finite(self.h))

    def port_gather(self, q):
        """Current field values at port q's points, [points][6]."""
        n = _check(library().cfd_port_npts(self.h, q))
        a = np.zeros((n, 6), np.float64)
        _check(library().cfd_port_gather(self.h, q, a.ctypes.data_as(_dp)))
        return a

    def port_gather_static(self, q):
        """port_gather through cfd_port_gather_static (no allocation)."""
        n = _check(library().cfd_port_npts(self.h, q))
        a = np.zeros((n, 6), np.float64)
        _check(library().cfd_port_gather_static(self.h, q, a.ctypes.data_as(_dp)))
        return a

    def port_dft(self, q):
        n = _check(library().cfd_port_npts(self.h, q))
        a = np.zeros((n, self.nf, 6, 2), np.float64)
        _check(library().cfd_port_dft(self.h, q, a.ctypes.data_as(_dp)))
        return a[..., 0] + 1j * a[..., 1]

    def port_flux(self, q, axis):
        """Spectral energy through plane port q (normal ``axis``, 0 x 1 y 2 z),
        J/Hz per DFT frequency (cfd_port_flux)."""
        a = np.zeros(self.nf, np.float64)
        _check(library().cfd_port_flux(self.h, q, axis, a.ctypes.data_as(_dp)))
        return a

    def port_intensity(self, q):
        a = np.zeros(self.steps_done, np.float64)
        _check(library().cfd_port_intensity(self.h, q, a.ctypes.data_as(_dp)))
        return a

    def gain_pop(self):
        a = np.zeros((self.n_gain, 4), np.float64)
        _check(library().cfd_gain_read(self.h, a.ctypes.data_as(_dp)))
        return a

    def gain_state(self):
        """[cell][6]: polarisation x y z, then its previous values."""
        n = _check(library().cfd_gain_state(self.h, None))
        a = np.zeros((max(n, 1), 6), np.float64)
        _check(library().cfd_gain_state(self.h, a.ctypes.data_as(_dp)))
        return a[:n]

    def dl_state(self):
        """[cell][3][nD + 2 nL + 1]: Drude currents, Lorentz currents, their
        previous values, the previous field, per component."""
        n = _check(library().cfd_dl_state(self.h, None))
        a = np.zeros(max(n, 1) * 3 * self.dl_per, np.float64)
        _check(library().cfd_dl_state(self.h, a.ctypes.data_as(_dp)))
        return a[:n * 3 * self.dl_per].reshape(n, 3, self.dl_per)

    def special_cells(self, kind):
        """[cell][3] i j k of the Drude-Lorentz (0) or gain (1) cells."""
        n = _check(library().cfd_special_cells(self.h, kind, None))
        a = np.zeros((max(n, 1), 3), np.int32)
        _check(library().cfd_special_cells(self.h, kind, a.ctypes.data_as(_ip)))
        return a[:n]

    def tfsf_wave(self):
        """The incident tables Einc, Hinc."""
        n = _check(library().cfd_tfsf_wave(self.h, None, None))
        e, h = np.zeros(n), np.zeros(n)
        _check(library().cfd_tfsf_wave(self.h, e.ctypes.data_as(_dp), h.ctypes.data_as(_dp)))
        return e, h

    # -- timing ---------------------------------------------------------
    def set_timing(self, on=True):
        _check(library().cfd_set_timing(self.h, 1 if on else 0))

    def kernel_times(self):
        return _kernel_times(library().cfd_kernel_times, self.h)


def _kernel_times(fn, h):
    """{pass name: (total ms, launches)} from cfd_kernel_times / cfd_batch_kernel_times."""
    cap = 64
    names = (C.c_char * 32 * cap)()
    ms = np.zeros(cap, np.float64)
    cnt = np.zeros(cap, np.int32)
    n = _check(fn(h, C.cast(names, C.c_void_p), ms.ctypes.data_as(_dp), cnt.ctypes.data_as(_ip), cap))
    return {bytes(names[i]).split(b"\0")[0].decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}


class Pack:
    """Simulations of any grid size and physics stepped together, each pass of
    a step one launch over all members (cfd_pack_*).  Members are Engines of
    the pack's precision; add() at step 0 (or a multiple of 16 DFT strides),
    remove() before stepping or freeing a member alone."""

    def __init__(self, precision=CFD_F64):
        h = C.c_void_p()
        _check(library().cfd_pack_create(precision, C.byref(h)))
        self.h = h
        self.precision = precision
        self.members = []

    def close(self):
        if getattr(self, "h", None):
            library().cfd_pack_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return _check(library().cfd_pack_size(self.h))

    def add(self, engine):
        _check(library().cfd_pack_add(self.h, engine.h))
        self.members.append(engine)

    def remove(self, engine):
        _check(library().cfd_pack_remove(self.h, engine.h))
        self.members.remove(engine)

    def step(self, n=1):
        _check(library().cfd_pack_step(self.h, n))

    def sync(self):
        _check(library().cfd_pack_sync(self.h))

    def set_timing(self, on):
        _check(library().cfd_pack_set_timing(self.h, int(on)))

    def kernel_times(self):
        return _kernel_times(library().cfd_pack_kernel_times, self.h)

    def check_finite(self, engines=None):
        """Non-finite field values of each given member (default: all)."""
        engines = list(self.members if engines is None else engines)
        sims = (_vp * len(engines))(*[e.h for e in engines])
        counts = np.zeros(len(engines), np.int32)
        _check(library().cfd_pack_check_finite(self.h, sims, len(engines), counts.ctypes.data_as(_ip)))
        return counts.tolist()


class Batch:
    """Simulations of one shape stepped together: each pass of a step is one
    launch over all instances (cfd_batch_*).  member(i) is a read-only Engine
    view of instance i (fields, ports, gain populations)."""

    def __init__(self, cases, precision=CFD_F64):
        L = library()
        cfgs = (_Config * len(cases))()
        self._keep = []
        for i, case in enumerate(cases):
            cfg, keep = make_config(case, precision)
            cfgs[i] = cfg
            self._keep.append((cfg, keep))
        h = C.c_void_p()
        _check(L.cfd_batch_create(cfgs, len(cases), C.byref(h)))
        self.h = h
        self.cases = list(cases)
        self.precision = precision
        self._members = {}

    def close(self):
        if getattr(self, "h", None):
            for m in self._members.values():
                m.h = None
            library().cfd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return _check(library().cfd_batch_size(self.h))

    def step(self, n=1):
        _check(library().cfd_batch_step(self.h, n))

    def sync(self):
        _check(library().cfd_batch_sync(self.h))

    def member(self, i):
        if i not in self._members:
            h = C.c_void_p()
            _check(library().cfd_batch_member(self.h, i, C.byref(h)))
            self._members[i] = Engine(self.cases[i], self.precision, _handle=h)
        return self._members[i]

    def device_bytes(self):
        v = C.c_size_t()
        _check(library().cfd_batch_device_bytes(self.h, C.byref(v)))
        return int(v.value)

    def set_timing(self, on):
        _check(library().cfd_batch_set_timing(self.h, int(on)))

    def kernel_times(self):
        return _kernel_times(library().cfd_batch_kernel_times, self.h)


def _image_check(rc):
    if rc < 0:
        raise EngineError(library().cfd_image_last_error().decode())


def image_colormap(values, cmin, cmax, rgb=None):
    """cfd_image_colormap (OutImage::plot_save's colour map) of a map
    values[x][y]; returns pixels [y][x][3] (uint8).  Updates rgb in place when
    given (NaN values leave their pixels as they were)."""
    v = np.ascontiguousarray(values, np.float64)
    nx, ny = v.shape
    if rgb is None:
        rgb = np.zeros((ny, nx, 3), np.uint8)
    if rgb.shape != (ny, nx, 3) or rgb.dtype != np.uint8 or not rgb.flags.c_contiguous:
        raise ValueError("rgb must be a contiguous uint8 array [ny][nx][3]")
    _image_check(library().cfd_image_colormap(v.ctypes.data_as(_dp), nx, ny, float(cmin), float(cmax),
                                              rgb.ctypes.data_as(C.POINTER(C.c_ubyte))))
    return rgb


def image_materials(ids):
    """cfd_image_materials (OutImage::structure_plot_save) of ids[x][y];
    returns pixels [y][x][3] (uint8)."""
    v = np.ascontiguousarray(ids, np.int16)
    nx, ny = v.shape
    rgb = np.zeros((ny, nx, 3), np.uint8)
    _image_check(library().cfd_image_materials(v.ctypes.data_as(C.POINTER(C.c_short)), nx, ny,
                                               rgb.ctypes.data_as(C.POINTER(C.c_ubyte))))
    return rgb


def write_png(path, rgb):
    """cfd_image_write_png: pixels [height][width][3] (uint8) to an RGB PNG."""
    p = np.ascontiguousarray(rgb, np.uint8)
    if p.ndim != 3 or p.shape[2] != 3:
        raise ValueError("rgb must be [height][width][3]")
    _image_check(library().cfd_image_write_png(os.fsencode(path), p.shape[1], p.shape[0],
                                               p.ctypes.data_as(C.POINTER(C.c_ubyte))))


def dft_flux(dft, axis, scale):
    """cfd_dft_flux on complex DFT accumulators [points][freqs][6] held on the
    host (a served request's results): scale * Re sum (E_a H_b* - E_b H_a*)."""
    dft = np.asarray(dft)
    a = np.ascontiguousarray(np.stack([dft.real, dft.imag], axis=-1), np.float64)
    out = np.zeros(dft.shape[1], np.float64)
    if library().cfd_dft_flux(a.ctypes.data_as(_dp), dft.shape[0], dft.shape[1], axis, scale, out.ctypes.data_as(_dp)) < 0:
        raise EngineError("cfd_dft_flux: bad arguments")
    return out


class _ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int), ("quantum", C.c_int), ("policy", C.c_int),
                ("pack_cells", C.c_double), ("step_cells", C.c_double)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3
POLICY_FIFO, POLICY_SHORTEST = 0, 1


class Server:
    """Serving loop (cfd_server_*): submit cases, then poll or wait; results
    are the port |E|^2 series and DFTs, on the host once a request is done."""

    def __init__(self, hbm_budget=0, max_active=0, quantum=0, policy=0, pack_cells=0, step_cells=0):
        o = _ServerOpts(hbm_budget, max_active, quantum, policy, float(pack_cells), float(step_cells))
        h = C.c_void_p()
        if library().cfd_server_create(C.byref(o), C.byref(h)) < 0:
            raise EngineError(library().cfd_server_last_error().decode())
        self.h = h
        self.cases = {}

    def _chk(self, rc):
        if rc < 0:
            raise EngineError(library().cfd_server_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            library().cfd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, case, precision=CFD_F64):
        cfg, keep = make_config(case, precision)
        rid = C.c_longlong()
        self._chk(library().cfd_server_submit(self.h, C.byref(cfg), C.byref(rid)))
        self.cases[rid.value] = case
        return rid.value

    def poll(self, rid):
        t = np.zeros(3, np.float64)
        st = self._chk(library().cfd_server_poll(self.h, rid, t.ctypes.data_as(_dp)))
        return st, tuple(float(x) for x in t)

    def wait(self, rid, timeout=600.0):
        return self._chk(library().cfd_server_wait(self.h, rid, timeout))

    def error(self, rid):
        return library().cfd_server_request_error(self.h, rid).decode()

    def result(self, rid, port, what="intensity"):
        w = 1 if what == "dft" else 0
        n = C.c_size_t()
        self._chk(library().cfd_server_result(self.h, rid, port, w, None, C.byref(n)))
        a = np.zeros(n.value, np.float64)
        self._chk(library().cfd_server_result(self.h, rid, port, w, a.ctypes.data_as(_dp), C.byref(n)))
        if w:
            nf = len(self.cases[rid].freqs)
            a = a.reshape(-1, nf, 6, 2)
            return a[..., 0] + 1j * a[..., 1]
        return a

    def release(self, rid):
        self._chk(library().cfd_server_release(self.h, rid))
        self.cases.pop(rid, None)

    def stats(self):
        st, pb, pa, q = C.c_longlong(), C.c_size_t(), C.c_int(), C.c_int()
        self._chk(library().cfd_server_stats(self.h, C.byref(st), C.byref(pb), C.byref(pa), C.byref(q)))
        return {"steps_launched": st.value, "peak_bytes": pb.value, "peak_active": pa.value, "queued": q.value}


class Model:
    """Host-side set-up of a case (no device): the tables the engine uploads."""

    def __init__(self, case, precision=CFD_F64):
        L = library()
        cfg, keep = make_config(case, precision)
        h = C.c_void_p()
        if L.cfd_model_build(C.byref(cfg), C.byref(h)) < 0:
            raise EngineError(L.cfd_model_last_error().decode())
        self.h = h
        n = np.zeros(4, np.int32)
        d = np.zeros(4, np.float64)
        L.cfd_model_dims(self.h, n.ctypes.data_as(_ip), d.ctypes.data_as(_dp))
        self.N = tuple(int(x) for x in n[:3])
        self.nt = int(n[3])
        self.d = tuple(float(x) for x in d[:3])
        self.dt = float(d[3])
        self.shape = tuple(x + 1 for x in self.N)

    def __del__(self):
        try:
            if self.h:
                library().cfd_model_free(self.h)
                self.h = None
        except Exception:
            pass

    def id_map(self):
        a = np.empty(self.shape, np.int16)
        library().cfd_model_id(self.h, a.ctypes.data_as(C.POINTER(C.c_int16)))
        return a

    def device_bytes(self, precision=CFD_F64):
        """Device bytes a simulation of this model allocates (cfd_model_device_bytes)."""
        return int(library().cfd_model_device_bytes(self.h, precision))

    def table(self, name):
        L = library()
        n = L.cfd_model_table(self.h, name.encode(), None)
        if n < 0:
            raise EngineError("no table " + name)
        a = np.empty(n, np.float64)
        L.cfd_model_table(self.h, name.encode(), a.ctypes.data_as(_dp))
        return a

    def tfsf(self, h):
        L = library()
        nterm = C.c_int(0)
        nt = L.cfd_model_tfsf(self.h, h, None, None, C.byref(nterm))
        tg = np.empty((max(nt, 1), 6), np.int32)
        tm = np.empty((max(nterm.value, 1), 3), np.float64)
        L.cfd_model_tfsf(self.h, h, tg.ctypes.data_as(_ip), tm.ctypes.data_as(_dp), C.byref(nterm))
        return tg[:nt], tm[:nterm.value]

    def port_points(self, q):
        L = library()
        n = L.cfd_model_port(self.h, q, None)
        a = np.empty((max(n, 0), 3), np.int32)
        L.cfd_model_port(self.h, q, a.ctypes.data_as(_ip))
        return a

    def cells(self, gain=False):
        L = library()
        n = L.cfd_model_cells(self.h, 1 if gain else 0, None)
        a = np.empty(max(n, 1), np.int64)
        L.cfd_model_cells(self.h, 1 if gain else 0, a.ctypes.data_as(C.POINTER(C.c_longlong)))
        return a[:n]
