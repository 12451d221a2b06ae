"""Pins of the oracle's functions that the reference's own runs leave open:
the Roden-Gedney CPML H-side sign (CFD_FIX_CPML_HSIGN, defect F6 of
CPML.cpp:643-682) by what it is for, absorbing an outgoing pulse; the
Gaussian point-source waveform (Sources.cpp:29-76 reads FWHM and delay but
ignores them, so this waveform is the engine's own formula) by its envelope
and spectrum; and the engine's host copy of that waveform against the
oracle's, bit for bit."""
import math

import numpy as np
import pytest

from cpsfdtd_amd.case import C0, EPS0, MU0, CaseWriter, FIX_CPML_HSIGN, SRC_REF_PULSE
import cpsfdtd_amd as eng

FLIP_HX = 1 << 30  # oracle-only planted mistake: x-slab CPML H terms with the opposite sign


def pulse_box(directory, n=60, npml=10, nt=900):
    """A vacuum box of n^3 cells (10 nm) with npml-cell CPML on every face and
    a short Ez pulse (500 nm, 2 fs FWHM, 4 fs delay) at the centre."""
    d = 10e-9
    w = CaseWriter(directory)
    w.materials([(1, 1, 0, 0)])
    w.grid((n, n, n), (d, d, d), [])
    w.cpml(npml, m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    w.source(500e-9, 2e-15, 4e-15)
    w.point_source(n // 2, n // 2, n // 2, comp=2, amp=1.0)
    return w.run(nt)


def interior_energy(o, npml):
    """Discrete field energy (J per unit cell volume) inside the CPML box."""
    s = slice(npml, -npml)
    e = sum(float(np.sum(o.field(c)[s, s, s] ** 2)) for c in range(3))
    h = sum(float(np.sum(o.field(c)[s, s, s] ** 2)) for c in range(3, 6))
    return 0.5 * (EPS0 * e + MU0 * h)


def residual(oracle_lib, directory, flags, n=60, npml=10, last=600, every=50):
    """Largest interior energy over the run (sampled every ``every`` steps)
    and the energy left at step ``last``."""
    case = pulse_box(directory, n, npml, last + 1)
    case.flags = flags
    o = oracle_lib.OracleSim(case)
    es = []
    while o.steps_done < last:
        o.step(every)
        es.append(interior_energy(o, npml))
    return max(es), es[-1]


def test_fixed_cpml_hsign_absorbs_better(oracle_lib, tmp_path):
    """The pulse (about 2e-10 J per unit cell volume at its peak, step 200)
    has left the box by step 400.  With the Roden-Gedney H-side signs the
    CPML leaves under 1e-10 of the peak energy at step 600 (measured 9e-13,
    the round-off floor), and at least 10 times less than with the
    reference's y/z-slab signs (measured 3e-7 of the peak: the reference's
    layer still reflects).  The same check fails when the x-slab sign is
    flipped instead (a planted mistake in the oracle only; measured 9e-8)."""
    res = {}
    for name, flags in (("fixed", FIX_CPML_HSIGN), ("ref", 0), ("flip", FIX_CPML_HSIGN | FLIP_HX)):
        (tmp_path / name).mkdir()
        peak, left = residual(oracle_lib, str(tmp_path / name), flags)
        res[name] = left / peak
    print("residual / peak energy at step 600: %s" % res)

    def absorbs(name):
        return res[name] < 1e-10 and res["ref"] > 10 * res[name]
    assert absorbs("fixed")
    assert not absorbs("ref")
    assert not absorbs("flip")


def source_case(directory, lam, fwhm, delay, nt, flags=0):
    d = 10e-9
    w = CaseWriter(directory)
    w.materials([(1, 1, 0, 0)])
    w.grid((8, 8, 8), (d, d, d), [])
    w.source(lam, fwhm, delay)
    w.point_source(4, 4, 4, comp=2, amp=1.0)
    w.point_source(3, 4, 4, comp=1, amp=-2.5, phase=0.7)
    return w.run(nt, flags=flags)


def envelope(x):
    from scipy.signal import hilbert
    return np.abs(hilbert(x))


def waveform_ok(w, dt, lam, fwhm, delay):
    """Envelope peak at the delay (within a step) with unit amplitude (0.1 %),
    amplitude FWHM equal to fwhm (0.5 %), spectrum peak at c / lambda (0.5 %)."""
    t = np.arange(len(w)) * dt
    env = envelope(w)
    k = int(np.argmax(env))
    above = np.nonzero(env >= env[k] / 2)[0]
    width = (above[-1] - above[0]) * dt
    f = np.linspace(0.5, 1.5, 4001) * C0 / lam
    spec = np.abs(np.exp(-2j * np.pi * np.outer(f, t)) @ w)
    return (abs(t[k] - delay) <= dt and abs(env[k] - 1.0) < 1e-3 and abs(width / fwhm - 1) < 5e-3
            and abs(f[int(np.argmax(spec))] / (C0 / lam) - 1) < 5e-3)


@pytest.mark.parametrize("lam,fwhm,delay", [(500e-9, 5e-15, 15e-15), (600e-9, 20e-15, 60e-15)])
def test_point_source_waveform(oracle_lib, tmp_path, lam, fwhm, delay):
    """The Gaussian point source of the c1 and c3 workloads: its amplitude
    envelope peaks at the file's delay with the source's amplitude, has the
    file's FWHM as its full width at half maximum, and its spectrum peaks at
    c / lambda.  The same check rejects an intensity FWHM, a missing ln 2, a
    reversed delay and an angular frequency used as an ordinary one."""
    nt = 6000
    o = oracle_lib.OracleSim(source_case(str(tmp_path), lam, fwhm, delay, nt))
    dt = o.dt
    w = o.source_wave(0)
    assert len(w) == nt + 1
    assert waveform_ok(w, dt, lam, fwhm, delay)
    t = np.arange(nt + 1) * dt
    nu = C0 / lam
    wrong = {
        "intensity FWHM": np.exp(-2 * math.log(2) * (t - delay) ** 2 / fwhm ** 2) * np.sin(2 * math.pi * nu * (t - delay)),
        "no ln 2": np.exp(-4 * (t - delay) ** 2 / fwhm ** 2) * np.sin(2 * math.pi * nu * (t - delay)),
        "reversed delay": np.exp(-4 * math.log(2) * (t + delay - 2 * delay * 0.9) ** 2 / fwhm ** 2) *
        np.sin(2 * math.pi * nu * (t - delay)),
        "angular carrier": np.exp(-4 * math.log(2) * (t - delay) ** 2 / fwhm ** 2) * np.sin(nu * (t - delay)),
    }
    for name, x in wrong.items():
        assert not waveform_ok(x, dt, lam, fwhm, delay), name
    # the second source: amplitude -2.5 and its phase on the same envelope
    w2 = o.source_wave(1)
    g = np.exp(-4 * math.log(2) * (t - delay) ** 2 / fwhm ** 2)
    assert np.allclose(w2, -2.5 * g * np.sin(2 * math.pi * nu * (t - delay) + 0.7), rtol=0, atol=1e-12)


@pytest.mark.parametrize("flags", [0, SRC_REF_PULSE])
def test_engine_source_wave_equals_oracle(oracle_lib, tmp_path, flags):
    """The engine's host model computes the waveform table that its source
    kernel adds, in both forms (the Gaussian above and the reference's
    Source::Pulse), bit for bit as the oracle does."""
    nt = 1200
    case = source_case(str(tmp_path), 500e-9, 5e-15, 15e-15, nt, flags)
    o = oracle_lib.OracleSim(case)
    m = eng.Model(case)
    tab = m.table("src_wave")
    assert len(tab) == 2 * (nt + 1)
    for q in range(2):
        assert np.array_equal(tab[q * (nt + 1):(q + 1) * (nt + 1)], o.source_wave(q)), q
    assert np.abs(tab).max() > 0.5
