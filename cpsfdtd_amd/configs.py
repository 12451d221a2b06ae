"""Generators for the benchmark workloads, written in the reference's input
formats (so the reference's own parsers and the engine read the same files).

  c1  plain FP64 Yee, 128^3, dielectric sphere, Ez point source, line port
  c2  256^3, 4x4 Drude-Lorentz spheres, 12-cell CPML, oblique Gaussian-beam
      TF/SF, two plane ports with a 64-frequency running DFT
  c3  192x192x64 disordered rod slab with four-level gain, periodic x/y
  c4  batched periodic unit cells (see cpsfdtd_amd.batch)

``scale`` < 1 shrinks a workload isotropically for tests (cell counts, PML and
box insets scale; physics parameters stay).
"""
from __future__ import annotations

import math

import numpy as np

from .case import (C0, CaseWriter, Shape, dlambda_for_fwhm, FIXED)


def _s(n, scale, lo=1):
    return max(lo, int(round(n * scale)))


def c1(directory, scale=1.0, nt=1000, flags=FIXED, record=True):
    """128^3, dx=10 nm, 10-cell CPML (m=3, sigma_max=0.8(m+1)/(eta0 dx),
    kappa_max=5, alpha_max=0.05), eps_r=2.25 sphere of radius 20 cells at the
    centre, Ez source at (40,64,64) (500 nm, 5 fs, 15 fs delay), line port along
    x through the centre recording every step."""
    N = _s(128, scale, 16)
    d = (10e-9, 10e-9, 10e-9)
    c = N // 2
    r = _s(20, scale, 2)
    w = CaseWriter(directory)
    w.materials([(1, 1, 0, 0), (2.25, 1, 0, 0)])
    w.grid((N, N, N), d, [Shape(1, "ellipsoid", x0=c * d[0], y0=c * d[1], z0=c * d[2], a=r * d[0], b=r * d[1], c=r * d[2])])
    w.cpml(_s(10, scale, 4), m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    w.source(500e-9, 5e-15, 15e-15)
    w.point_source(_s(40, scale, 5), c, c, comp=2, amp=1.0)
    w.plane_port_cells("lineport.txt", "z", c, 0, N - 1, c, c, d)
    return w.run(nt, record_ports=record, flags=flags)


C2_FREQS = [C0 / lam for lam in np.linspace(400e-9, 800e-9, 64)]


def c2(directory, scale=1.0, nt=4000, flags=FIXED, nfreq=64, stride=8, record=False):
    """256^3, dx=5 nm, 12-cell CPML; TF/SF box 20 cells inside the PML driving
    an oblique Gaussian beam (theta 30, phi 45, p-polarised psi_E = 90 deg,
    waist 400 nm, 600 nm, 8 fs amplitude FWHM); 4x4 Drude-Lorentz spheres
    (radius 15 cells, pitch 50 cells; eps_inf 3.7, wp 1.37e16, gamma 2.7e13,
    Lorentz (d_eps, w0, delta) = (1.0, 4.2e15, 1e15), (0.5, 6.1e15, 2e15));
    reflected and transmitted plane ports with a running DFT at ``nfreq``
    frequencies evenly spaced in wavelength over 400-800 nm."""
    N = _s(256, scale, 32)
    d = (5e-9, 5e-9, 5e-9)
    npml = _s(12, scale, 4)
    inset = npml + _s(20, scale, 3)
    c = N // 2
    r = _s(15, scale, 2)
    pitch = _s(50, scale, 2 * r + 2)
    first = c - (3 * pitch) // 2
    w = CaseWriter(directory)
    w.materials([(1, 1, 0, 0), (1, 1, 0, 0), (1, 1, 0, 0)])
    sphere = Shape(2, "ellipsoid", x0=first * d[0], y0=first * d[1], z0=c * d[2], a=r * d[0], b=r * d[1], c=r * d[2])
    grid = Shape(0, "multi", theta=4, phi=4, psi=1, x0=first * d[0], y0=first * d[1], z0=c * d[2],
                 a=pitch * d[0], b=pitch * d[1], c=0.0)
    w.grid((N, N, N), d, [sphere, grid])
    w.cpml(npml, m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    lam = 600e-9
    w.tfsf(lam, dlambda_for_fwhm(lam, 8e-15), 400e-9, 0.0, (c + 0.5) * d[0], (c + 0.5) * d[1], (c + 0.5) * d[2],
           30.0, 45.0, 1.0, 90.0, ((inset, inset), (inset, inset), (inset, inset)), nt)
    w.drude_lorentz(3.7, [(1.37e16, 2.7e13)], [(4.2e15, 1e15, 1.0), (6.1e15, 2e15, 0.5)])
    lo, hi = npml, N - npml
    refl = npml + (inset - npml) // 2
    w.plane_port_cells("port_r.txt", "z", refl, lo, hi, lo, hi, d)
    w.plane_port_cells("port_t.txt", "z", N - inset - (inset - npml) // 2, lo, hi, lo, hi, d)
    freqs = C2_FREQS if nfreq == 64 else [C0 / l for l in np.linspace(400e-9, 800e-9, nfreq)]
    return w.run(nt, freqs=freqs, dft_stride=stride, record_ports=record, flags=flags)


def c2_reference_core(directory, scale=1.0, nt=4000):
    """The part of c2 the reference can run (its 3-D build has no
    Drude-Lorentz class): same grid, spheres (ID 2 cells are then skipped by
    BasicFDTD.cpp:38), CPML and TF/SF, no ports."""
    case = c2(directory, scale=scale, nt=nt, nfreq=0)
    case.dl = False
    case.ports = []
    case.freqs = []
    return case
