"""Case descriptions and writers for cpsFDTD's on-disk input formats.

A :class:`Case` is a directory holding the reference's text inputs
(``material.txt``, ``structure.txt``, ``CFS_PML.txt``, ``TF_SF.txt``,
``DrudeLorentz.txt``, port files, ``source.txt``) plus the run options the
reference hard-codes in ``main.cpp`` (point sources, which ports, step count).
Formats follow the parsers cited on each writer.  Numbers are written with
17 significant digits so the parsed doubles are exactly the ones intended.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

C0 = 2.99792458E8
EPS0 = 8.8541878176E-12
MU0 = 1.2566370614E-6

# behaviour switches (same bit values as the engine's and the oracle's)
FIX_CPML_HSIGN = 1     # CPML.cpp:643-682 H-side y/z psi sign -> Roden-Gedney
FIX_TFSF_FARFACE = 2   # TF_SF.cpp:1222,1255,1288 far-face H index
SRC_REF_PULSE = 4      # Sources.cpp:72-76 Pulse() waveform for point sources
REF_ZEROHEAP = 8       # reproduce the reference on zero-filled heap memory
FIX_TFSF_YEE = 16      # TF_SF.cpp:155-887 incident values at each component's Yee position
FIXED = FIX_CPML_HSIGN | FIX_TFSF_FARFACE | FIX_TFSF_YEE

PORT_PLANE = 0
PORT_SPHERE = 1


def _f(x: float) -> str:
    return repr(float(x)) if isinstance(x, float) else str(x)


def _line(*vals) -> str:
    return "\t".join(_f(v) if isinstance(v, float) else str(v) for v in vals) + "\n"


@dataclass
class Shape:
    """One ``structure.txt`` shape line (Structure.cpp:97-118)."""
    material: int
    kind: str
    theta: float = 0.0
    phi: float = 0.0
    psi: float = 0.0
    x0: float = 0.0
    y0: float = 0.0
    z0: float = 0.0
    a: float = 0.0
    b: float = 0.0
    c: float = 0.0


@dataclass
class Case:
    """A run: input files in ``dir`` plus the options main.cpp hard-codes."""
    dir: str
    nt: int = 0
    pml: bool = False
    tfsf: bool = False
    dl: bool = False
    gain: bool = False
    source: bool = False
    ports: List[Tuple[str, int]] = field(default_factory=list)
    sources: List[Tuple[int, int, int, int, float, float]] = field(default_factory=list)
    freqs: List[float] = field(default_factory=list)
    dft_stride: int = 1
    record_ports: bool = False
    periodic: int = 0
    flags: int = 0

    def path(self, name: str) -> str:
        return os.path.join(self.dir, name)

    @property
    def files(self) -> dict:
        return {
            "material": self.path("material.txt"),
            "structure": self.path("structure.txt"),
            "pml": self.path("CFS_PML.txt") if self.pml else None,
            "tfsf": self.path("TF_SF.txt") if self.tfsf else None,
            "dl": self.path("DrudeLorentz.txt") if self.dl else None,
            "gain": self.path("gain.txt") if self.gain else None,
            "source": self.path("source.txt") if self.source else None,
        }


class CaseWriter:
    """Builds a case directory file by file."""

    def __init__(self, directory: str):
        os.makedirs(directory, exist_ok=True)
        self.case = Case(dir=os.path.abspath(directory))

    def _write(self, name: str, text: str) -> str:
        p = self.case.path(name)
        with open(p, "w") as fh:
            fh.write(text)
        return p

    # Structure.cpp:26-52
    def materials(self, media: Sequence[Tuple[float, float, float, float]]):
        text = _line(len(media)) + "".join(_line(*(float(v) for v in m)) for m in media)
        self._write("material.txt", text)
        return self

    # Structure.cpp:68-119.  Lengths are written as (N + 1/2) d so that the
    # reference's N = (int)(L/d) truncation lands on N.
    def grid(self, n: Tuple[int, int, int], d: Tuple[float, float, float], shapes: Sequence[Shape] = (),
             freespace: bool = False):
        L = [(n[a] + 0.5) * d[a] for a in range(3)]
        text = _line(*L) + _line(*(float(x) for x in d)) + _line(1 if freespace else 0) + _line(len(shapes))
        for idx, s in enumerate(shapes):
            text += _line(idx, s.material, s.kind, float(s.theta), float(s.phi), float(s.psi),
                          float(s.x0), float(s.y0), float(s.z0), float(s.a), float(s.b), float(s.c))
        self._write("structure.txt", text)
        return self

    def structure_text(self, text: str):
        self._write("structure.txt", text)
        return self

    # CPML.cpp:29-62
    def cpml(self, n, m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05):
        if isinstance(n, int):
            n = ((n, n), (n, n), (n, n))
        text = "".join(_line(a, b) for a, b in n)
        text += _line(float(m), float(malp)) + _line(float(kapmax), float(sig_ratio)) + _line(float(alpmax))
        self._write("CFS_PML.txt", text)
        self.case.pml = True
        return self

    # TF_SF.cpp:35-89, 898-907
    def tfsf(self, lam, dlam, w0, z0, xs, ys, zs, theta, phi, E0, psi, box, nt, cw=0):
        text = _line(float(lam)) + _line(float(dlam)) + _line(float(w0), float(z0))
        text += _line(float(xs), float(ys), float(zs)) + _line(float(theta), float(phi))
        text += _line(float(E0), float(psi))
        text += "".join(_line(a, b) for a, b in box)
        text += _line(int(nt)) + _line(int(cw))
        self._write("TF_SF.txt", text)
        self.case.tfsf = True
        return self

    # 3-D Drude-Lorentz layout (line 3 first token = eps_inf; rad/s units)
    def drude_lorentz(self, eps_inf, drude: Sequence[Tuple[float, float]], lorentz: Sequence[Tuple[float, float, float]]):
        text = _line(len(drude)) + _line(len(lorentz)) + _line(float(eps_inf))
        text += "".join(_line(float(w), float(g)) for w, g in drude)
        text += "".join(_line(float(w), float(dl), float(de)) for w, dl, de in lorentz)
        self._write("DrudeLorentz.txt", text)
        self.case.dl = True
        return self

    # FourLevelLaser.cpp:37-93 layout plus line 15 = pump rate
    def gain(self, lambda_a, gamma_a, t32, t21, t10, ntot, pump, lambda_b=0.0, gamma_b=0.0, t30=1.0):
        vals = ["ID", "0", "0\t0", "0\t0", _f(float(lambda_b)), _f(float(lambda_a)), _f(float(gamma_a)),
                _f(float(gamma_b)), _f(float(t30)), _f(float(t32)), _f(float(t21)), _f(float(t10)),
                _f(float(ntot)), "0.0", _f(float(pump))]
        self._write("gain.txt", "\n".join(vals) + "\n")
        self.case.gain = True
        return self

    # Sources.cpp:36-46
    def source(self, lam, fwhm, delay):
        self._write("source.txt", _line(float(lam)) + _line(float(fwhm)) + _line(float(delay)))
        self.case.source = True
        return self

    def point_source(self, i, j, k, comp=2, amp=1.0, phase=0.0):
        self.case.sources.append((int(i), int(j), int(k), int(comp), float(amp), float(phase)))
        return self

    # Ports.cpp:213-229
    def plane_port(self, name, normal, position, w1, w2, h1, h2):
        self._write(name, _line(normal) + _line(float(position)) + _line(float(w1), float(w2)) + _line(float(h1), float(h2)))
        self.case.ports.append((self.case.path(name), PORT_PLANE))
        return self

    def plane_port_cells(self, name, normal, c, a1, a2, b1, b2, d):
        """Plane port covering cell indices [a1,a2] x [b1,b2] at index c (half-cell offsets
        keep the reference's truncating conversions exact)."""
        dn, dw, dh = {"x": (d[0], d[1], d[2]), "y": (d[1], d[0], d[2]), "z": (d[2], d[0], d[1])}[normal]
        return self.plane_port(name, normal, (c + 0.5) * dn, (a1 + 0.5) * dw, (a2 + 0.5) * dw,
                               (b1 + 0.5) * dh, (b2 + 0.5) * dh)

    # Ports.cpp:40-62
    def sphere_port(self, name, r, xc, yc, zc, t1, t2, p1, p2):
        self._write(name, _line(float(r)) + _line(float(xc), float(yc), float(zc)) + _line(float(t1), float(t2))
                    + _line(float(p1), float(p2)))
        self.case.ports.append((self.case.path(name), PORT_SPHERE))
        return self

    def run(self, nt, freqs=(), dft_stride=1, record_ports=False, periodic=0, flags=0):
        c = self.case
        c.nt, c.freqs, c.dft_stride, c.record_ports, c.periodic, c.flags = (
            int(nt), [float(f) for f in freqs], int(dft_stride), bool(record_ports), int(periodic), int(flags))
        return c


def dlambda_for_fwhm(lam: float, tau: float) -> float:
    """dlambda that gives TF_SF.cpp:911-914's envelope an amplitude FWHM of ``tau``.

    The envelope is exp(-sig t^2) with sig = dw^2/(16 ln 2), so its amplitude
    FWHM is 8 ln2 / dw.  dnu = C/(lam - dl/2) - C/(lam + dl/2) is solved for dl.
    """
    dw = 8.0 * math.log(2.0) / tau
    dnu = dw / (2.0 * math.pi)
    return (-C0 + math.sqrt(C0 * C0 + dnu * dnu * lam * lam)) / (dnu / 2.0)


def courant_dt(d: Tuple[float, float, float]) -> float:
    """Structure.cpp:124."""
    return 0.99 / (C0 * math.sqrt(1.0 / (d[0] * d[0]) + 1.0 / (d[1] * d[1]) + 1.0 / (d[2] * d[2])))
