"""Closed forms and invariants that pin the oracle where the reference has no
runnable code (its Drude-Lorentz and gain classes are 2-D only, PBC is a stub,
ports have no DFT), and the Yee update's own invariants (cavity modes on the
discrete dispersion relation, conserved discrete energy).  The GPU parity tests carry these pins to the engine bit
for bit.  Each check is chosen so that a plausible mistake (a flipped sign, a
dropped time step, a wrong index) fails it."""
import math

import numpy as np
import pytest
import scipy.linalg

import cases
from cpsfdtd_amd.case import C0, EPS0, MU0, CaseWriter, Shape, FIXED

DL_POLES = dict(eps_inf=3.7, drude=[(1.37e16, 8.0e14)], lorentz=[(4.2e15, 1e15, 1.0), (6.1e15, 2e15, 0.5)])


def dl_case(tmp_path, d=10e-9, **kw):
    p = dict(DL_POLES)
    p.update(kw)
    w = CaseWriter(str(tmp_path))
    w.materials([(1, 1, 0, 0), (1, 1, 0, 0), (1, 1, 0, 0)])
    w.grid((8, 8, 8), (d, d, d), [])
    w.drude_lorentz(p["eps_inf"], p["drude"], p["lorentz"])
    return w.run(10), p


def h_discrete(z, dt, p):
    """E(z)/C(z) of the ADE as written down from its difference equations:
    trapezoidal Ampere at n+1/2, trapezoidal Drude J' + g J = e0 wp^2 E,
    central-difference Lorentz J'' + 2 d J' + w0^2 J = e0 De w0^2 E'."""
    amp = EPS0 * p["eps_inf"] * (z - 1) / dt
    chi = 0
    for wp, g in p["drude"]:
        chi = chi + EPS0 * wp * wp * (z + 1) / 2 / ((z - 1) / dt + g * (z + 1) / 2)
    for w0, dl, de in p["lorentz"]:
        lhs = (z - 2 + 1 / z) / dt ** 2 + dl * (z - 1 / z) / dt + w0 * w0
        chi = chi + EPS0 * de * w0 * w0 * (z - 1 / z) / (2 * dt) / lhs
    return 1.0 / (amp + (z + 1) / 2 * chi)


def h_continuous(w, p):
    """1 / (i w e0 eps(w)) with eps = eps_inf + Drude + Lorentz (e^{+iwt})."""
    iw = 1j * w
    s = iw * EPS0 * p["eps_inf"]
    for wp, g in p["drude"]:
        s = s + EPS0 * wp * wp / (g + iw)
    for w0, dl, de in p["lorentz"]:
        s = s + EPS0 * de * w0 * w0 * iw / (w0 * w0 - w * w + 2j * dl * w)
    return 1.0 / s


def test_drude_lorentz_single_cell_transfer_function(oracle_lib, tmp_path):
    """The oracle's ADE (coefficients C1..C6, kappa, beta, alpha, xi, gamma of
    Drude_Lorentz.cpp:74-108 with F9/F10 fixed) driven by a windowed curl
    sequence reproduces the z-domain transfer function of its defining
    difference equations."""
    case, p = dl_case(tmp_path)
    o = oracle_lib.OracleSim(case)
    dt = o.dt
    n = np.arange(60000)
    w1 = 2 * math.pi * C0 / 600e-9
    c = np.exp(-((n - 1200) / 300.0) ** 2) * np.sin(w1 * n * dt)
    y = o.dl_0d(c)
    assert abs(y[-100:]).max() < 1e-12 * abs(y).max()  # decayed inside the window
    for lam in (450e-9, 550e-9, 600e-9, 700e-9):
        w = 2 * math.pi * C0 / lam
        z = np.exp(1j * w * dt)
        Y = np.sum(y * z ** (-n))
        Cz = np.sum(c * z ** (-n))
        got = Y / (z * Cz)
        want = h_discrete(z, dt, p)
        assert abs(got - want) / abs(want) < 1e-8, (lam, got, want)


def dl_recurrence(c, dt, C1, C2, C3, C4, C5, C6, kap, beta, alpha, xi, gam):
    """Single-cell ADE recurrence with given coefficients (the form of
    Drude_Lorentz.cpp:172-183), for checking that the transfer-function test
    rejects a wrong coefficient."""
    nD, nL = len(C4), len(C5)
    JD, JL, JL1 = np.zeros(nD), np.zeros(nL), np.zeros(nL)
    E = E1 = 0.0
    out = np.empty(len(c))
    for n in range(len(c)):
        t = E
        En = C1 * E + C2 * E1 + C3 * c[n] - np.dot(C4, JD) - np.dot(C5, JL) - np.dot(C6, JL1)
        JD = kap * JD + beta * (En + t)
        JL, JL1 = alpha * JL + xi * JL1 + gam * (En - E1) / 2.0 / dt, JL
        E1, E = t, En
        out[n] = E
    return out


def test_transfer_function_rejects_reference_c1_and_c5(oracle_lib, tmp_path):
    """The reference's C1 (sign of the loss terms, F9) and C5 (no dt, F10) give
    a response that leaves the transfer function by orders of magnitude more
    than the tolerance used above, so that check does discriminate."""
    case, p = dl_case(tmp_path)
    o = oracle_lib.OracleSim(case)
    dt = o.dt
    nD, nL = len(p["drude"]), len(p["lorentz"])
    co = o.dl_coefs()
    C1, C2, C3 = co[0:3]
    k = 3
    C4 = co[k:k + nD]; k += nD
    C5 = co[k:k + nL]; k += nL
    C6 = co[k:k + nL]; k += nL
    kap = co[k:k + nD]; k += nD
    beta = co[k:k + nD]; k += nD
    alpha = co[k:k + nL]; k += nL
    xi = co[k:k + nL]; k += nL
    gam = co[k:k + nL]
    n = np.arange(12000)
    w1 = 2 * math.pi * C0 / 600e-9
    c = np.exp(-((n - 1200) / 300.0) ** 2) * np.sin(w1 * n * dt)
    base = dl_recurrence(c, dt, C1, C2, C3, C4, C5, C6, kap, beta, alpha, xi, gam)
    np.testing.assert_allclose(base, o.dl_0d(c), rtol=1e-12, atol=1e-12 * np.abs(base).max())
    den = 2 * dt / C3
    wrong_c1 = (2 * EPS0 * p["eps_inf"] + dt * beta.sum()) / den
    wrong_c5 = (1.0 + alpha) / den
    w = 2 * math.pi * C0 / 600e-9
    z = np.exp(1j * w * dt)
    want = h_discrete(z, dt, p)
    for bad in (dl_recurrence(c, dt, wrong_c1, C2, C3, C4, C5, C6, kap, beta, alpha, xi, gam),
                dl_recurrence(c, dt, C1, C2, C3, C4, wrong_c5, C6, kap, beta, alpha, xi, gam)):
        Y = np.sum(bad * z ** (-n))
        with np.errstate(invalid="ignore", over="ignore"):
            got = Y / (z * np.sum(c * z ** (-n)))
            rel = abs(got - want) / abs(want)
        assert not rel <= 1e-3, rel  # an unstable (nan/inf) response counts as rejected


def test_drude_lorentz_continuous_limit(oracle_lib, tmp_path):
    """As dt -> 0 the discrete response converges to the continuous
    Drude-Lorentz medium at second order: pins the physical meaning of the
    file's parameters (rad/s, delta as the half linewidth, Delta-eps).  The
    curl sample c[n] sits at (n+1/2) dt and E^{n} at n dt, so the continuous
    ratio carries exp(i w dt / 2); without it the error is first order."""
    case, p = dl_case(tmp_path)
    dt = oracle_lib.OracleSim(case).dt
    for lam in (500e-9, 800e-9, 1200e-9):
        w = 2 * math.pi * C0 / lam
        hc = h_continuous(w, p)
        err = [abs(h_discrete(np.exp(1j * w * h), h, p) * np.exp(0.5j * w * h) - hc) / abs(hc)
               for h in (dt, dt / 2, dt / 4, dt / 8)]
        assert err[0] < 2e-3, (lam, err)
        for a, b in zip(err[1:], err[2:]):
            assert 3.5 < a / b < 4.5, (lam, err)


def gain_case(tmp_path, pump=1e13, t32=1e-13, t21=1e-12, t10=1e-13, ntot=3.3e24, d=10e-9):
    w = CaseWriter(str(tmp_path))
    w.materials([(1, 1, 0, 0)] * 6)
    w.grid((8, 8, 8), (d, d, d), [])
    w.gain(600e-9, 1e14, t32, t21, t10, ntot, pump)
    return w.run(10)


def test_gain_relaxation_matches_rate_equations(oracle_lib, tmp_path):
    """With no field the four-level populations follow the linear rate
    equations dN3 = Wp N0 - N3/t32, dN2 = N3/t32 - N2/t21, dN1 = N2/t21 - N1/t10,
    N0 = Ntot - N1 - N2 - N3 (matrix exponential).  The decays are
    Crank-Nicolson and the pump term explicit (the oracle's 3-D restatement of
    the FourLevelLaser.cpp:296-316 population step),
    so the error at a fixed physical time halves with dt (dt follows the cell
    size through the Courant limit)."""
    pump, t32, t21, t10, ntot = 1e13, 1e-13, 1e-12, 1e-13, 3.3e24
    A = np.array([[-pump, 1 / t10, 0, 0],
                  [0, -1 / t10, 1 / t21, 0],
                  [0, 0, -1 / t21, 1 / t32],
                  [pump, 0, 0, -1 / t32]])
    N0 = np.array([ntot, 0, 0, 0])
    err = []
    for k, d in enumerate((10e-9, 5e-9, 2.5e-9)):
        sub = tmp_path / ("d%d" % k)
        sub.mkdir()
        o = oracle_lib.OracleSim(gain_case(sub, pump, t32, t21, t10, ntot, d=d))
        nsteps = 6000 << k
        pop, P = o.gain_0d(np.zeros(nsteps))
        assert np.all(P == 0.0)
        np.testing.assert_allclose(pop.sum(axis=1), ntot, rtol=1e-13)
        want = scipy.linalg.expm(A * nsteps * o.dt) @ N0
        assert want[3] > 0.01 * ntot  # the pump really moved population
        err.append(np.abs(pop[-1] - want).max() / np.abs(want[1:]).max())
    assert err[0] < 1e-3, err
    for a, b in zip(err, err[1:]):
        assert 1.8 < a / b < 2.2, err


def test_gain_population_conserved_under_drive(oracle_lib, tmp_path):
    o = oracle_lib.OracleSim(gain_case(tmp_path))
    n = np.arange(8000)
    w = 2 * math.pi * C0 / 600e-9
    pop, P = o.gain_0d(1e8 * np.sin(w * n * o.dt))
    np.testing.assert_allclose(pop.sum(axis=1), 3.3e24, rtol=1e-13)
    assert np.abs(P).max() > 0


def test_gain_drive_couples_to_populations(oracle_lib, tmp_path):
    """A pumped medium driven at its emission line exchanges energy with the
    field through the N1-N2 transition: the drive changes N2 relative to the
    undriven run, and without a drive the polarisation stays exactly zero."""
    o = oracle_lib.OracleSim(gain_case(tmp_path, pump=1e13))
    n = np.arange(30000)
    w = 2 * math.pi * C0 / 600e-9
    drive = 5e8 * np.sin(w * n * o.dt)
    popd, Pd = o.gain_0d(drive)
    pop0, P0 = o.gain_0d(np.zeros_like(drive))
    assert np.all(P0 == 0.0) and np.abs(Pd).max() > 0
    assert abs(popd[-1, 2] - pop0[-1, 2]) > 1e-9 * pop0[-1, 2]


def pbc_case(tmp_path, nx=12, ny=10, nz=48):
    d = (10e-9, 10e-9, 10e-9)
    w = CaseWriter(str(tmp_path))
    w.materials([(1, 1, 0, 0), (2.25, 1, 0, 0), (4.0, 1, 0.05, 0)])
    w.grid((nx, ny, nz), d, [Shape(1, "cuboid", x0=0.0, y0=0.0, z0=20 * d[2], a=(nx + 1) * d[0], b=(ny + 1) * d[1], c=6 * d[2]),
                            Shape(2, "cuboid", x0=0.0, y0=0.0, z0=30 * d[2], a=(nx + 1) * d[0], b=(ny + 1) * d[1], c=3 * d[2])])
    w.cpml(((0, 0), (0, 0), (8, 8)), m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    w.tfsf(400e-9, 200e-9, 1e3, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, ((0, 0), (0, 0), (12, 12)), 300)
    return w.run(300, periodic=3, flags=FIXED)


def test_periodic_plane_wave_translation_invariant(oracle_lib, tmp_path):
    """Normal incidence on an x/y-uniform layered stack in a periodic cell:
    every column carries the same field, exactly (a wrong wrap index breaks it)."""
    o = oracle_lib.OracleSim(pbc_case(tmp_path))
    o.step(300)
    fields = [o.field(c)[:o.N[0], :o.N[1], :] for c in range(6)]
    assert np.abs(fields[1]).max() > 1e-3 and np.abs(fields[3]).max() > 0  # Ey, Hx carry the wave
    for c, f in enumerate(fields):
        assert np.array_equal(f, np.broadcast_to(f[:1, :1, :], f.shape)), c


def test_port_dft_equals_numpy(oracle_lib, tmp_path):
    """The running DFT at the ports equals numpy's DFT of the recorded samples
    at the same stride: X(f) = sum x_n exp(-2 pi i f t_n), t_n = (n+1) dt for E
    and (n+1/2) dt for H."""
    freqs = (4.0e14, 5.5e14, 7.5e14)
    case = cases.mixed(str(tmp_path), freqs=freqs, stride=3, flags=FIXED, nt=90)
    o = oracle_lib.OracleSim(case)
    o.step(90)
    for q in range(2):
        x = o.port_samples(q)
        got = o.port_dft(q)
        n = np.arange(0, 90, 3)
        for fi, f in enumerate(freqs):
            tE = (n + 1.0) * o.dt
            tH = (n + 0.5) * o.dt
            for c in range(6):
                t = tE if c < 3 else tH
                want = np.sum(x[n, :, c] * np.exp(-2j * np.pi * f * t)[:, None], axis=0)
                scale = np.abs(x[:, :, c]).max() * len(n) + 1e-300
                assert np.abs(got[:, fi, c] - want).max() <= 1e-12 * scale


def test_gain_polarisation_is_a_lorentzian_at_the_emission_line(oracle_lib, tmp_path):
    """The four-level medium's polarisation responds to the field as a damped
    oscillator at the file's emission line.  With a slow lower level the
    pumped steady state is absorbing (N1 > N2; an inverted single cell with
    the field's back-action is a laser, not a linear system), and a weak
    broadband pulse's response P(w)/E(w) peaks at 2 pi c / lambda_a (0.5 %;
    a 610 nm line is off by 1.7 %) with a power FWHM equal to the file's
    linewidth (5 %; measured 1 %).  Pins the polarisation coefficients'
    frequency and damping terms by behaviour."""
    lam, gamma = 600e-9, 1e14
    o = oracle_lib.OracleSim(gain_case(tmp_path, pump=1e13, t32=1e-13, t21=1e-13, t10=2e-12))
    dt = o.dt
    wa = 2 * math.pi * C0 / lam
    n_pump, n_probe, width = 800000, 12000, 60.0
    n = np.arange(n_pump + n_probe)
    drive = np.where(n >= n_pump, 1e3 * np.exp(-(((n - n_pump - 300) / width) ** 2)) * np.cos(wa * n * dt), 0.0)
    pop, P = o.gain_0d(drive)
    dN = pop[n_pump, 1] - pop[n_pump, 2]
    assert dN > 0 and abs(pop[-1, 1] - pop[-1, 2] - dN) < 1e-3 * dN  # absorbing, steady over the probe
    bE0 = o.coefs()[1][0]
    E = drive - bE0 / dt * np.diff(P, prepend=0.0)  # the field the medium sees (back-action included)
    Pw, Ew = P[n_pump:], np.concatenate(([0.0], E[n_pump:-1]))  # P[n] responds to E[n-1]
    assert abs(Pw[-50:]).max() < 1e-3 * abs(Pw).max()  # rung down inside the window
    w = wa + gamma * np.linspace(-3, 3, 601)
    ph = np.exp(-1j * np.outer(w, np.arange(n_probe) * dt))
    H2 = np.abs((ph @ Pw) / (ph @ Ew)) ** 2
    k = int(np.argmax(H2))
    assert abs(w[k] - wa) < 0.005 * wa, (w[k], wa)
    above = np.nonzero(H2 >= H2[k] / 2)[0]
    fwhm = w[above[-1]] - w[above[0]]
    assert abs(fwhm - gamma) < 0.05 * gamma, (fwhm, gamma)


def _tfsf_vacuum(tmp_path, flags, theta, phi, psi, n=40, npml=8, box=12, nt=400):
    """Vacuum, CPML, a TF/SF box with a plane wave (1 mm waist) of a 5 fs
    pulse at 500 nm (50 cells per wavelength)."""
    from cpsfdtd_amd.case import dlambda_for_fwhm
    d = 10e-9
    w = CaseWriter(str(tmp_path))
    w.materials([(1, 1, 0, 0)])
    w.grid((n, n, n), (d, d, d), [])
    w.cpml(npml, m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    w.tfsf(500e-9, dlambda_for_fwhm(500e-9, 5e-15), 1e-3, 0.0, n * d / 2, n * d / 2, 0.0, theta, phi, 1.0, psi,
           ((box, box), (box, box), (box, box)), nt)
    w.source(500e-9, 5e-15, 10e-15)
    return w.run(nt, flags=flags)


def _leakage(oracle_lib, case, n=40, npml=8, box=12):
    """Peak |E| in the scattered-field region (between the CPML and the box,
    one cell clear of the faces) over peak |E| inside the box, over the run."""
    o = oracle_lib.OracleSim(case)
    sf = np.zeros((n + 1,) * 3, bool)
    sf[npml + 1:n - npml, npml + 1:n - npml, npml + 1:n - npml] = True
    sf[box - 1:n - box + 2, box - 1:n - box + 2, box - 1:n - box + 2] = False
    tf = np.zeros_like(sf)
    tf[box + 1:n - box, box + 1:n - box, box + 1:n - box] = True
    a = b = 0.0
    for _ in range(case.nt // 20 - 1):
        o.step(20)
        F = np.sqrt(sum(o.field(c) ** 2 for c in range(3)))
        a, b = max(a, F[tf].max()), max(b, F[sf].max())
    return b / a


@pytest.mark.parametrize("theta,phi,psi", [(0.0, 0.0, 0.0), (30.0, 45.0, 90.0)], ids=["normal", "oblique"])
def test_tfsf_vacuum_leakage(oracle_lib, tmp_path, theta, phi, psi):
    """A total-field/scattered-field box in vacuum leaves the scattered-field
    region dark: the corrections cancel the incident wave at the faces to the
    grid's numerical dispersion.  With every incident value at its component's
    Yee position the leak is below 1e-3 of the incident peak; taken at the
    cell corner, as TF_SF.cpp:155-887 does, the half-cell timing error leaks
    3-6 % (and the reference's far-face and edge defects add more)."""
    from cpsfdtd_amd.case import FIX_TFSF_YEE, REF_ZEROHEAP, SRC_REF_PULSE
    fixed = _leakage(oracle_lib, _tfsf_vacuum(tmp_path / "a", FIXED, theta, phi, psi))
    corner = _leakage(oracle_lib, _tfsf_vacuum(tmp_path / "b", FIXED & ~FIX_TFSF_YEE, theta, phi, psi))
    ref = _leakage(oracle_lib, _tfsf_vacuum(tmp_path / "c", REF_ZEROHEAP | SRC_REF_PULSE, theta, phi, psi))
    assert fixed < 1e-3
    assert corner > 1e-2
    assert ref > corner


def test_port_flux_is_the_plane_wave_poynting_spectrum(oracle_lib, tmp_path):
    """cfd_dft_flux on a plane port's running DFT: (1) equals the same sum
    computed from numpy transforms of the recorded samples; (2) for a plane
    wave in vacuum (TF/SF, normal incidence) it is the incident spectral
    energy 2 |E_inc(f)|^2 / eta0 x port area at two planes inside the box
    within 1 %; (3) a plane in the scattered-field region sees < 1e-6 of it.
    A swapped tangential pair, a dropped conjugate or a wrong dt/area scale
    fails (1) or (2)."""
    from cpsfdtd_amd.case import dlambda_for_fwhm
    import cpsfdtd_amd as eng
    n, npml, box, d, nt = 32, 6, 10, 10e-9, 1000
    freqs = (5.0e14, 6.0e14, 7.0e14)
    w = CaseWriter(str(tmp_path))
    w.materials([(1, 1, 0, 0)])
    w.grid((n, n, n), (d, d, d), [])
    w.cpml(npml, m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    w.tfsf(500e-9, dlambda_for_fwhm(500e-9, 5e-15), 1e-3, 0.0, n * d / 2, n * d / 2, 0.0, 0.0, 0.0, 1.0, 0.0,
           ((box, box), (box, box), (box, box)), nt)
    w.source(500e-9, 5e-15, 10e-15)
    for q, k in enumerate((13, 18, 8)):  # two planes inside the box, one in the scattered field
        w.plane_port_cells("p%d.txt" % q, "z", k, 12, 20, 12, 20, (d, d, d))
    case = w.run(nt, freqs=freqs, dft_stride=1, record_ports=True, flags=FIXED)
    o = oracle_lib.OracleSim(case)
    o.step(nt)
    dt = o.dt
    scale = 2 * dt * dt * d * d
    fl = [eng.dft_flux(o.port_dft(q), 2, scale) for q in range(3)]
    # (1) the same sum from the recorded samples
    x = o.port_samples(0)
    t = np.arange(nt)
    for fi, f in enumerate(freqs):
        E = [np.sum(x[:, :, c] * np.exp(-2j * np.pi * f * (t + 1.0) * dt)[:, None], axis=0) for c in range(2)]
        H = [np.sum(x[:, :, 3 + c] * np.exp(-2j * np.pi * f * (t + 0.5) * dt)[:, None], axis=0) for c in range(2)]
        want = scale * np.real(np.sum(E[0] * np.conj(H[1]) - E[1] * np.conj(H[0])))
        assert abs(fl[0][fi] - want) <= 1e-12 * abs(want)
    # (2) the incident wave's spectral energy through the port's area
    einc = eng.Model(case).table("Einc")[:nt]
    eta0 = math.sqrt(MU0 / EPS0)
    npts = len(o.port_points(0))
    inc = np.array([dt * abs(np.sum(einc * np.exp(-2j * np.pi * f * t * dt))) for f in freqs])
    expect = 2 * inc ** 2 / eta0 * npts * d * d
    for q in (0, 1):
        assert np.all(np.abs(fl[q] / expect - 1) < 1e-2), (q, fl[q] / expect)
    # (3) the scattered-field region is dark
    assert np.all(np.abs(fl[2]) < 1e-6 * expect)


@pytest.mark.parametrize("eps_r", [1.0, 2.25])
def test_yee_cavity_modes_and_energy(oracle_lib, tmp_path, eps_r):
    """The Yee update itself (BasicFDTD.cpp:25-81) against its closed forms.
    A cell periodic in x and y with a uniform medium carries x-polarised waves
    uniform in x and y between the z walls, where Ex (on the E nodes k = 0 and
    k = Nz) stays zero: its standing modes Ex ~ sin(p pi k / Nz) are exact
    eigenmodes of the scheme, with
        sin(w dt / 2) = (c dt / (n dz)) sin(p pi / (2 Nz)),   n^2 = eps_r,
    and the discrete energy eps|E^n|^2/2 + mu H^(n-1/2).H^(n+1/2)/2 is
    conserved.  Started from a Gaussian Ex profile, the measured mode
    frequencies (Hann-windowed DFT peaks) match the discrete relation within
    1e-5 and miss the continuous one, c p / (2 n Nz dz), by over 1 % at p = 15;
    a wrong time step, cell size, medium coefficient or curl sign fails."""
    Nz, d, nt = 40, 10e-9, 6000
    w = CaseWriter(str(tmp_path))
    w.materials([(eps_r, 1, 0, 0)])
    w.grid((4, 4, Nz), (d, d, d), [])
    case = w.run(nt, periodic=3)
    o = oracle_lib.OracleSim(case)
    ex = o.field(0)
    k = np.arange(ex.shape[2])
    prof = np.exp(-((k - 13.0) / 3.0) ** 2)
    prof[0] = prof[-1] = 0.0
    ex[:, :, :] = prof[None, None, :]
    o.set_field(0, ex)
    dt = o.dt
    e_prev, h_prev = o.field(0)[0, 0].copy(), o.field(4)[0, 0].copy()
    probe, energy = np.empty(nt), np.empty(nt)
    for n in range(nt):
        o.step(1)
        e, h = o.field(0)[0, 0].copy(), o.field(4)[0, 0].copy()
        energy[n] = 0.5 * EPS0 * eps_r * np.sum(e_prev ** 2) + 0.5 * MU0 * np.sum(h_prev[:Nz] * h[:Nz])
        e_prev, h_prev = e, h
        probe[n] = e[7]
    assert (energy.max() - energy.min()) / energy.mean() < 1e-12
    spec = np.abs(np.fft.rfft(probe * np.hanning(nt), 8 * nt))
    df = 1.0 / (8 * nt * dt)
    nidx = math.sqrt(eps_r)
    for p in (1, 5, 10, 15):
        f_disc = 2.0 / dt * math.asin(C0 * dt / (nidx * d) * math.sin(p * math.pi / (2 * Nz))) / (2 * math.pi)
        f_cont = C0 * p / (2 * nidx * Nz * d)
        i0 = int(round(f_disc / df))
        i = i0 - 40 + int(np.argmax(spec[i0 - 40:i0 + 40]))
        a, b, c = np.log(spec[i - 1]), np.log(spec[i]), np.log(spec[i + 1])
        f_meas = (i + 0.5 * (a - c) / (a - 2 * b + c)) * df
        assert abs(f_meas / f_disc - 1) < 1e-5, (p, f_meas / f_disc - 1)
        if p == 15:
            assert abs(f_meas / f_cont - 1) > 1e-2, (p, f_meas / f_cont - 1)


def _block_case(directory, dispersive):
    """A 6x5x4-cell block of eps_r 2.25 in a CPML-bounded grid with a soft
    source: as a plain medium (material 1) or as the Drude-Lorentz medium
    (ID 2) with eps_inf = 2.25 and no poles."""
    w = CaseWriter(str(directory))
    w.materials([(1, 1, 0, 0), (2.25, 1, 0, 0), (1, 1, 0, 0)])
    w.grid((24, 22, 20), (10e-9, 10e-9, 10e-9),
           [Shape(2 if dispersive else 1, "cuboid", x0=120e-9, y0=110e-9, z0=100e-9, a=60e-9, b=50e-9, c=40e-9)])
    if dispersive:
        w.drude_lorentz(2.25, [], [])
    w.cpml(((4, 4), (4, 4), (4, 4)), m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    w.source(300e-9, 3e-15, 5e-15)
    w.point_source(8, 9, 10, comp=2, amp=2.0)
    return w.run(300, flags=FIXED)


def test_drude_lorentz_without_poles_is_the_dielectric(oracle_lib, tmp_path):
    """With no Drude or Lorentz pole the ADE update is the basic update of a
    lossless dielectric of permittivity eps_inf (C1 = 1, C2 = 0, C3 = bE):
    the block as ADE cells and as a plain eps_r 2.25 medium give the same
    fields bit for bit after 300 steps."""
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    plain = oracle_lib.OracleSim(_block_case(tmp_path / "a", False))
    ade = oracle_lib.OracleSim(_block_case(tmp_path / "b", True))
    assert ade.dl_ncells() > 0
    plain.step(300)
    ade.step(300)
    for c in range(6):
        assert np.array_equal(plain.field(c), ade.field(c)), c
    assert np.abs(plain.field(2)).max() > 1e-6
