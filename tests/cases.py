"""Small input cases shared by the tests and the golden-fixture script.

Every case is written in the reference's own text formats by
cpsfdtd_amd.case.CaseWriter.  Sizes are chosen so the FP64 oracle finishes
each in well under a second.
"""
from __future__ import annotations

import os
import shutil

from cpsfdtd_amd.case import (CaseWriter, Shape, dlambda_for_fwhm, REF_ZEROHEAP, SRC_REF_PULSE)

REF_RESOURCES = "/root/reference/resourses"


def mixed(directory, n=(40, 36, 32), npml=(6, 5, 6), box=9, nt=60, flags=REF_ZEROHEAP | SRC_REF_PULSE,
          zs=0.0, z0=0.0, theta=30.0, phi=45.0, psi=90.0, src=(12, 14, 16), dl=False, gain=False, dl_medium=None,
          record=True, freqs=(), stride=1, periodic=0, cw=0, magnetic=False, extra_media=0):
    """Every shape painter, CPML on all faces, an oblique Gaussian-beam TF/SF
    (cw=1: the continuous-wave form), a soft Ez source, a plane and a sphere
    port.  magnetic: media 1 and 3 get mu_r 2 and 1.5, medium 3 a magnetic
    conductivity of 1e9 ohm/m (sigma_M dt / 2 mu0 mu_r about 0.5 %).
    extra_media: that many more dielectrics after the six, the last painted
    as a small block (more than eight media: the engine's coefficient tables
    instead of its selects)."""
    d = (10e-9, 10e-9, 10e-9)
    w = CaseWriter(directory)
    w.materials([(1, 1, 0, 0), (2.25, 2.0 if magnetic else 1, 0, 0), (1, 1, 0, 0),
                 (4.0, 1.5 if magnetic else 1, 0.01, 1e9 if magnetic else 0), (1, 1, 0, 0), (1.5, 1, 0, 0)] +
                [(1.2 + 0.3 * q, 1, 0.002 * q, 0) for q in range(extra_media)])
    sx, sy, sz = n[0] / 40.0, n[1] / 36.0, n[2] / 32.0
    w.grid(n, d, [
        Shape(1, "ellipsoid", x0=200e-9 * sx, y0=180e-9 * sy, z0=160e-9 * sz, a=80e-9, b=70e-9, c=60e-9),
        Shape(3, "cuboid", x0=50e-9 * sx + 20e-9, y0=60e-9 * sy + 20e-9, z0=70e-9 * sz, a=60e-9, b=50e-9, c=40e-9),
        Shape(4, "cuboid", theta=0.3, phi=0.2, psi=0.1, x0=200e-9 * sx, y0=150e-9 * sy, z0=100e-9 * sz,
              a=60e-9, b=50e-9, c=40e-9),
        Shape(1, "cylinder", x0=300e-9 * sx, y0=100e-9 * sy, z0=50e-9 * sz, a=30e-9, b=40e-9, c=150e-9 * sz),
        Shape(2 if dl else 1, "ellipsoid", x0=120e-9 * sx, y0=250e-9 * sy, z0=150e-9 * sz, a=50e-9, b=40e-9, c=45e-9),
        Shape(5 if gain else 1, "cuboid", x0=250e-9 * sx, y0=230e-9 * sy, z0=200e-9 * sz, a=70e-9, b=60e-9, c=50e-9),
        Shape(4, "multi", theta=2, phi=2, psi=1, x0=100e-9 * sx, y0=100e-9 * sy, z0=200e-9 * sz, a=150e-9 * sx,
              b=120e-9 * sy, c=0.0),
    ] + ([  # dispersive / gain blocks reaching into CPML corners
        Shape(2, "cuboid", x0=40e-9, y0=180e-9 * sy, z0=40e-9, a=60e-9, b=50e-9, c=60e-9)] if dl else []) + ([
        Shape(5, "cuboid", x0=(n[0] - 4) * 10e-9, y0=40e-9, z0=(n[2] - 4) * 10e-9, a=60e-9, b=60e-9, c=60e-9)]
        if gain else []) + ([
        Shape(5 + extra_media, "cuboid", x0=80e-9 * sx, y0=200e-9 * sy, z0=60e-9 * sz, a=70e-9, b=60e-9, c=80e-9)]
        if extra_media else []))
    if npml:
        w.cpml(((npml[0], npml[0]), (npml[1], npml[1]), (npml[2], npml[2])), m=3.0, malp=1.0, kapmax=5.0,
               sig_ratio=1.0, alpmax=0.05)
    if box:
        if zs != 0.0:
            z0 = zs / d[2]  # numerically the value the reference leaves in z0 (TF_SF.cpp:57)
        w.tfsf(500e-9, dlambda_for_fwhm(500e-9, 5e-15), 300e-9, z0, n[0] * 5e-9, n[1] * 5e-9, zs,
               theta, phi, 1.0, psi, ((box, box), (box, box), (box, box)), nt, cw=cw)
    w.source(500e-9, 5e-15, 10e-15)
    if src:
        w.point_source(*src, comp=2, amp=3.0)
    if dl:  # dl_medium: (eps_inf, [(wp, gamma)], [(w0, delta, deps)]), default 1 Drude + 2 Lorentz poles
        w.drude_lorentz(*(dl_medium or (3.7, [(1.37e16, 2.7e13)], [(4.2e15, 1e15, 1.0), (6.1e15, 2e15, 0.5)])))
    if gain:
        w.gain(600e-9, 1e14, 1e-13, 1e-10, 1e-13, 3.3e24, 1e9)
    w.plane_port_cells("pp.txt", "z", n[2] // 2, 2, n[0] - 3, 3, n[1] - 6, d)
    w.sphere_port("sp.txt", 100e-9, n[0] * 5e-9, n[1] * 5e-9, n[2] * 5e-9, 30.0, 150.0, 0.0, 300.0)
    return w.run(nt, freqs=freqs, dft_stride=stride, record_ports=record, periodic=periodic, flags=flags)


def tiny(directory, nt=40, flags=REF_ZEROHEAP | SRC_REF_PULSE, text_ports=False, cw=0, magnetic=False):
    """24x22x20 grid with every module the reference has; small enough that
    full field snapshots fit in a fixture.  text_ports adds a 2-D x-normal
    plane port (6 x 4 points) and a partial sphere port, for the text
    exports.  cw=1: continuous-wave TF/SF.  magnetic: media 1 and 3 get mu_r
    2 and 1.5, medium 3 a magnetic conductivity of 1e9 ohm/m."""
    d = (10e-9, 10e-9, 10e-9)
    n = (24, 22, 20)
    w = CaseWriter(directory)
    w.materials([(1, 1, 0, 0), (2.25, 2.0 if magnetic else 1, 0, 0), (1, 1, 0, 0),
                 (4.0, 1.5 if magnetic else 1, 0.02, 1e9 if magnetic else 0), (1, 1, 0, 0)])
    w.grid(n, d, [
        Shape(1, "ellipsoid", x0=120e-9, y0=110e-9, z0=100e-9, a=40e-9, b=45e-9, c=35e-9),
        Shape(3, "cuboid", x0=60e-9, y0=70e-9, z0=60e-9, a=30e-9, b=40e-9, c=30e-9),
        Shape(4, "cuboid", x0=150e-9, y0=140e-9, z0=120e-9, a=20e-9, b=20e-9, c=20e-9),
    ])
    w.cpml(((4, 4), (4, 4), (4, 4)), m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    w.tfsf(300e-9, dlambda_for_fwhm(300e-9, 3e-15), 200e-9, 0.0, 120e-9, 110e-9, 0.0, 25.0, 35.0, 1.0, 70.0,
           ((6, 6), (6, 6), (6, 6)), nt, cw=cw)
    w.source(300e-9, 3e-15, 5e-15)
    w.point_source(8, 9, 10, comp=2, amp=2.0)
    w.plane_port_cells("line.txt", "z", 10, 1, 22, 11, 11, d)
    if text_ports:
        w.plane_port_cells("face.txt", "x", 16, 4, 9, 5, 8, d)
        w.sphere_port("cap.txt", 40e-9, 120e-9, 110e-9, 100e-9, 60.0, 120.0, 0.0, 180.0)
    return w.run(nt, record_ports=True, flags=flags)


def shipped(directory, structure="structure.txt", with_pml=True):
    """The reference's own example inputs (resourses/material.txt and a
    structure file, optionally resourses/CFS_PML.txt)."""
    os.makedirs(directory, exist_ok=True)
    shutil.copy(os.path.join(REF_RESOURCES, "material.txt"), os.path.join(directory, "material.txt"))
    shutil.copy(os.path.join(REF_RESOURCES, structure), os.path.join(directory, "structure.txt"))
    w = CaseWriter(directory)
    if with_pml:
        shutil.copy(os.path.join(REF_RESOURCES, "CFS_PML.txt"), os.path.join(directory, "CFS_PML.txt"))
        w.case.pml = True
    return w.run(1)
