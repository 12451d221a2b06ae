"""HIP engine (through the C ABI) against the FP64 oracle and the reference's
golden fixtures.

FP64: fields, port samples and DFT accumulators are compared bit for bit (the
engine keeps the reference's operand order and is built without FP
contraction); reductions whose order differs (port intensity sums) within a
relative 1e-12.  FP32: port time series within a relative L2 of 1e-4 of the
FP64 oracle on these short runs.
"""
import os

import numpy as np
import pytest

import cases
import cpsfdtd_amd as eng
from cpsfdtd_amd.case import FIXED, REF_ZEROHEAP, SRC_REF_PULSE

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)


def assert_fields_equal(e, o, what=""):
    for c in range(6):
        a, b = e.field(c), o.field(c)
        assert np.array_equal(a, b), "%s comp %d max|d|=%g rel=%g" % (what, c, np.abs(a - b).max(), rel_l2(a, b))


def run_both(case, steps, module=False, precision=64):
    o = cases_oracle(case)
    o.step(steps)
    e = eng.Engine(case, precision)
    if module:
        for n in range(steps):
            e.module_step(n, tfsf=case.tfsf)
    else:
        e.step(steps)
    e.sync()
    return e, o


def cases_oracle(case):
    import pyoracle
    return pyoracle.OracleSim(case)


def test_tiny_matches_reference_fixture(tmp_path):
    """40 steps of the reference's own core (fixture) reproduced on the GPU."""
    g = dict(np.load(os.path.join(GOLD, "ref_tiny.npz"), allow_pickle=False))
    case = cases.tiny(str(tmp_path))
    e = eng.Engine(case, 64)
    e.step(int(g["nt"][0]))
    for c, nm in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        assert np.array_equal(e.field(c).ravel(), g[nm]), nm
    assert np.array_equal(e.port_samples(0), g["port0"])


@pytest.mark.parametrize("module", [False, True], ids=["fused", "module"])
@pytest.mark.parametrize("flags", [REF_ZEROHEAP | SRC_REF_PULSE, FIXED], ids=["ref", "fixed"])
def test_mixed_f64_bit_exact(tmp_path, module, flags):
    case = cases.mixed(str(tmp_path), flags=flags)
    e, o = run_both(case, 60, module=module)
    assert_fields_equal(e, o)
    for q in range(2):
        assert np.array_equal(e.port_samples(q), o.port_samples(q))
        assert rel_l2(e.port_intensity(q), o.port_intensity(q)) < 1e-12


@pytest.mark.parametrize("form", ["merged", "split", "grid"])
def test_update_forms_bit_exact(tmp_path, monkeypatch, form):
    """Every fused-step form: merged (one launch per half-step, lean or CPML
    path per tile), split (lean box, shell boxes, special cells) and the single
    full-physics grid kernel per half-step."""
    monkeypatch.setenv("CFD_UPDATE_KERNEL", form)
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED)
    e, o = run_both(case, 40)
    assert_fields_equal(e, o, form)
    assert np.array_equal(e.gain_pop(), o.gain_pop())


@pytest.mark.parametrize("form", ["merged", "split"])
def test_coefficient_tables_beyond_eight_media_f64_bit_exact(tmp_path, monkeypatch, form):
    """Ten media: the update kernels read each medium's coefficients from
    their tables (up to eight they select them from the kernel arguments)."""
    monkeypatch.setenv("CFD_UPDATE_KERNEL", form)
    case = cases.mixed(str(tmp_path), dl=True, flags=FIXED, extra_media=4)
    e, o = run_both(case, 40)
    assert_fields_equal(e, o, form)
    assert np.array_equal(e.port_samples(0), o.port_samples(0))


def test_dl_gain_f64_bit_exact(tmp_path):
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED)
    e, o = run_both(case, 50)
    assert e.n_dl > 0 and e.n_gain > 0
    assert_fields_equal(e, o, "dl+gain")
    assert np.array_equal(e.gain_pop(), o.gain_pop())


@pytest.mark.parametrize("poles", ["2d1l", "0d1l"])
def test_dl_other_pole_counts_f64_bit_exact(tmp_path, poles):
    """Pole counts without a fixed-count special-pass instantiation (two
    Drude and one Lorentz pole; one Lorentz pole alone) take the generic
    cell update; bit-exact against the oracle with dispersive and gain cells
    in the CPML corners."""
    medium = {"2d1l": (2.5, [(1.37e16, 2.7e13), (4.0e15, 5e13)], [(5.0e15, 1e15, 0.8)]),
              "0d1l": (2.0, [], [(4.2e15, 1e15, 1.5)])}[poles]
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED, dl_medium=medium)
    print("dl poles", poles, flush=True)
    e, o = run_both(case, 40)
    assert e.n_dl > 0
    assert_fields_equal(e, o, "dl " + poles)
    assert np.array_equal(e.gain_pop(), o.gain_pop())


def test_dl_without_poles_is_the_dielectric_f64(tmp_path, oracle_lib):
    """Zero Drude and Lorentz poles take the generic special-pass update; it
    must reduce to the plain dielectric bit for bit (as on the oracle)."""
    from test_closed_forms import _block_case
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    o = oracle_lib.OracleSim(_block_case(tmp_path / "a", False))
    o.step(300)
    print("dl without poles", flush=True)
    e = eng.Engine(_block_case(tmp_path / "b", True), 64)
    assert e.n_dl > 0
    e.step(300)
    assert_fields_equal(e, o, "ade without poles")
    e.close()


def test_dl_gain_module_f64_bit_exact(tmp_path):
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED)
    e, o = run_both(case, 30, module=True)
    assert_fields_equal(e, o, "dl+gain module")
    assert np.array_equal(e.gain_pop(), o.gain_pop())


@pytest.mark.parametrize("form", ["split", "merged"])
@pytest.mark.parametrize("periodic,npml", [(1, (0, 5, 6)), (3, (0, 0, 6))], ids=["x", "xy"])
def test_periodic_f64_bit_exact(tmp_path, monkeypatch, periodic, npml, form):
    """Periodic x (and y) planes, in the split form and in the merged form
    (whose epilogue applies the TF/SF terms on the periodic E planes too)."""
    monkeypatch.setenv("CFD_UPDATE_KERNEL", form)
    case = cases.mixed(str(tmp_path), periodic=periodic, npml=npml, flags=FIXED, theta=0.0, phi=0.0, psi=0.0)
    e, o = run_both(case, 50)
    assert_fields_equal(e, o, "periodic")


def test_port_dft_f64(tmp_path):
    case = cases.mixed(str(tmp_path), freqs=(4.0e14, 5.5e14, 7.5e14), stride=3, flags=FIXED)
    e, o = run_both(case, 60)
    for q in range(2):
        assert np.array_equal(e.port_dft(q), o.port_dft(q))
    # the plane port's flux spectrum: the device run's, the same sum on the
    # oracle's accumulators (stride 3, 10 nm cells)
    scale = 2 * (3 * o.dt) ** 2 * 10e-9 * 10e-9
    assert np.array_equal(e.port_flux(0, 2), eng.dft_flux(o.port_dft(0), 2, scale))


def test_fp32_close_to_fp64_oracle(tmp_path):
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED)
    e, o = run_both(case, 60, precision=32)
    for q in range(2):
        err = rel_l2(e.port_samples(q), o.port_samples(q))
        assert err < 1e-4, err
    for c in range(6):
        assert rel_l2(e.field(c), o.field(c)) < 1e-4


def test_bad_input_raises(tmp_path):
    case = cases.mixed(str(tmp_path))
    with open(case.path("CFS_PML.txt"), "w") as fh:
        fh.write("100 100\n")
    with pytest.raises(eng.EngineError):
        eng.Engine(case, 64)


def test_run_length_enforced(tmp_path):
    case = cases.mixed(str(tmp_path), nt=5)
    e = eng.Engine(case, 32)
    e.step(5)
    with pytest.raises(eng.EngineError, match="run finished"):
        e.step(1)


def test_graph_replayed_blocks_bit_exact(tmp_path):
    """Step blocks replayed from a captured HIP graph (aligned 128-step blocks)
    against the oracle: eager steps, a first replay, host reads, a second
    replay, eager steps again; fields, port samples, DFT and gain populations
    compared after every stage."""
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED, nt=400, freqs=(4.0e14, 6.5e14), stride=2)
    o = cases_oracle(case)
    e = eng.Engine(case, 64)
    done = 0
    for k in (3, 253, 128, 16):  # eager to 128, replay 128-256, replay 256-384, eager to 400
        e.step(k)
        o.step(k)
        done += k
        assert e.steps_done == done
        assert_fields_equal(e, o, "after %d steps" % done)
        for q in range(2):
            assert np.array_equal(e.port_samples(q), o.port_samples(q)), (done, q)
        assert np.array_equal(e.gain_pop(), o.gain_pop()), done
    for q in range(2):
        assert np.array_equal(e.port_dft(q), o.port_dft(q))


def test_graph_replay_equals_eager_fp32(tmp_path, monkeypatch):
    case = cases.mixed(str(tmp_path), dl=True, gain=True, flags=FIXED, nt=300, freqs=(5.0e14,), stride=1)
    runs = []
    for graphs in ("1", "0"):
        monkeypatch.setenv("CFD_GRAPHS", graphs)
        e = eng.Engine(case, 32)
        e.step(300)
        runs.append(([e.field(c) for c in range(6)], [e.port_dft(q) for q in range(2)]))
        e.close()
    for a, b in zip(runs[0][0] + runs[0][1], runs[1][0] + runs[1][1]):
        assert np.array_equal(a, b)


def _odd_case(directory, n, npml, dl):
    """A grid of ragged shape: a dielectric block, an optional Drude-Lorentz
    block, a soft Ez source and CPML of the given thickness per axis."""
    from cpsfdtd_amd.case import CaseWriter, Shape
    d = 10e-9
    w = CaseWriter(directory)
    w.materials([(1, 1, 0, 0), (2.25, 1, 0, 0), (1, 1, 0, 0)])
    shapes = [Shape(1, "cuboid", x0=d, y0=d, z0=d, a=(n[0] // 2) * d, b=(n[1] // 2) * d, c=(n[2] // 2) * d)]
    if dl:
        shapes.append(Shape(2, "cuboid", x0=(n[0] // 2) * d, y0=0.0, z0=0.0, a=(n[0] + 1) * d, b=(n[1] + 1) * d,
                            c=(n[2] // 3 + 1) * d))
    w.grid(n, (d, d, d), shapes)
    if any(npml):
        w.cpml(((npml[0], npml[0]), (npml[1], npml[1]), (npml[2], npml[2])), m=3.0, malp=1.0, kapmax=5.0,
               sig_ratio=1.0, alpmax=0.05)
    w.source(500e-9, 5e-15, 10e-15)
    w.point_source(n[0] // 2, n[1] // 2, n[2] // 2, comp=2, amp=1.0)
    if dl:
        w.drude_lorentz(3.7, [(1.37e16, 2.7e13)], [(4.2e15, 1e15, 1.0)])
    return w.run(30, flags=FIXED)


@pytest.mark.parametrize("form", ["merged", "split", "grid"])
@pytest.mark.parametrize("n,npml,dl", [
    ((2, 2, 2), (0, 0, 0), False),        # the smallest grid the model accepts
    ((3, 70, 5), (0, 4, 0), False),       # one ragged x-tile, y spans 18 tile rows
    ((65, 3, 4), (4, 0, 0), True),        # x just past one tile width, CPML in x
    ((130, 9, 7), (5, 2, 2), True),       # three x-tiles, the last with 2 columns; thin CPML on every axis
], ids=["2x2x2", "3x70x5", "65x3x4", "130x9x7"])
def test_ragged_shapes_f64_bit_exact(tmp_path, monkeypatch, form, n, npml, dl):
    """Tile edges in every form: partial x-tiles and tile rows, axes shorter
    than a tile, slabs of a few cells, Drude-Lorentz cells in a CPML slab."""
    monkeypatch.setenv("CFD_UPDATE_KERNEL", form)
    case = _odd_case(str(tmp_path), n, npml, dl)
    e, o = run_both(case, 30)
    assert_fields_equal(e, o, "%s %s" % (form, n))


def test_empty_case_steps_and_stays_zero(tmp_path):
    """A grid with nothing in it (no sources, TF/SF, CPML or ports) steps in
    both precisions, with and without graph replay, and stays exactly zero."""
    from cpsfdtd_amd.case import CaseWriter
    w = CaseWriter(str(tmp_path))
    w.materials([(1, 1, 0, 0)])
    w.grid((33, 17, 9), (10e-9, 10e-9, 10e-9), [])
    case = w.run(300)
    for prec in (32, 64):
        e = eng.Engine(case, prec)
        e.step(7)      # eager
        e.step(256)    # includes graph-replayed blocks
        e.sync()
        for c in range(6):
            assert not np.any(e.field(c)), (prec, c)
        assert e.nonfinite() == 0
        e.close()


@pytest.mark.parametrize("tag,kw", [("cw", {"cw": 1}), ("mag", {"magnetic": True})])
def test_tiny_variants_match_reference_fixture(tmp_path, tag, kw):
    """The reference's own loop with continuous-wave TF/SF and with magnetic
    media (fixture digests of its fields and port samples), reproduced on the
    GPU bit for bit."""
    import hashlib
    g = dict(np.load(os.path.join(GOLD, "ref_extra.npz"), allow_pickle=False))
    nt = int(g[tag + "_nt"][0])
    case = cases.tiny(str(tmp_path), nt=nt, **kw)
    print("tiny", tag, nt, "steps", flush=True)
    e = eng.Engine(case, 64)
    e.step(nt)

    def sha(a):
        return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
    for c in range(6):
        assert sha(e.field(c).ravel()) == bytes(g[tag + "_field_sha256"][c]).decode(), c
    assert sha(e.port_samples(0)) == bytes(g[tag + "_port0_sha256"]).decode()


@pytest.mark.parametrize("kw", [{"cw": 1, "nt": 600}, {"magnetic": True, "dl": True, "gain": True}],
                         ids=["cw", "magnetic"])
def test_mixed_variants_f64_bit_exact(tmp_path, kw):
    """Continuous-wave TF/SF over its ramp and into the steady sine (600
    steps), and magnetic media beside Drude-Lorentz and gain cells, in the
    fixed physics: engine = oracle bit for bit, graph-replayed blocks
    included."""
    nt = kw.pop("nt", 200)
    case = cases.mixed(str(tmp_path), flags=FIXED, nt=nt, **kw)
    print("mixed", sorted(kw), nt, "steps", flush=True)
    e, o = run_both(case, nt)
    assert_fields_equal(e, o)
    for q in range(2):
        assert np.array_equal(e.port_samples(q), o.port_samples(q))


def test_sources_sharing_a_cell_f64_bit_exact(tmp_path):
    """Three soft sources on one (cell, component), one on another component
    of the same cell and one elsewhere: added one after another in source
    order, as the reference does, so the engine equals the oracle bit for
    bit (no lost contribution, no other rounding)."""
    from cpsfdtd_amd.case import CaseWriter
    d = 10e-9
    w = CaseWriter(str(tmp_path))
    w.materials([(1, 1, 0, 0), (2.25, 1, 0, 0)])
    w.grid((24, 20, 18), (d, d, d), [])
    w.cpml(4, m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    w.source(400e-9, 3e-15, 6e-15)
    for amp, ph, comp, cell in ((1.0, 0.0, 2, (12, 10, 9)), (-0.37, 0.4, 2, (12, 10, 9)), (0.9, 1.1, 1, (12, 10, 9)),
                                (2.5, 2.0, 2, (12, 10, 9)), (0.5, 0.0, 0, (7, 6, 5))):
        w.point_source(*cell, comp=comp, amp=amp, phase=ph)
    case = w.run(300, flags=FIXED)
    e, o = run_both(case, 300)
    assert_fields_equal(e, o, "shared source cell")
    assert np.abs(o.field(2)).max() > 0


def _post_case(directory, src, port_z, sphere_x):
    """40 x 36 x 32 cells, 5-cell CPML, a normal-incidence TF/SF box 8 cells in
    from the faces, a Drude-Lorentz sphere of radius 4 cells at (sphere_x, 18,
    16), a soft Ez source at src and a plane port across the grid at
    port_z."""
    from cpsfdtd_amd.case import CaseWriter, Shape, dlambda_for_fwhm
    d = 10e-9
    n = (40, 36, 32)
    w = CaseWriter(directory)
    w.materials([(1, 1, 0, 0), (1, 1, 0, 0), (1, 1, 0, 0)])
    w.grid(n, (d, d, d), [Shape(2, "ellipsoid", x0=sphere_x * d, y0=18 * d, z0=16 * d, a=4 * d, b=4 * d, c=4 * d)])
    w.cpml(5, m=3.0, malp=1.0, kapmax=5.0, sig_ratio=1.0, alpmax=0.05)
    w.tfsf(500e-9, dlambda_for_fwhm(500e-9, 5e-15), 1e-3, 0.0, 20 * d, 18 * d, 0.0, 0.0, 0.0, 1.0, 90.0,
           ((8, 8), (8, 8), (8, 8)), 60)
    w.drude_lorentz(3.7, [(1.37e16, 2.7e13)], [(4.2e15, 1e15, 1.0), (6.1e15, 2e15, 0.5)])
    w.source(500e-9, 5e-15, 10e-15)
    w.point_source(*src, comp=2, amp=2.0)
    w.plane_port_cells("pp.txt", "z", port_z, 2, n[0] - 3, 2, n[1] - 3, (d, d, d))
    return w.run(60, record_ports=True, flags=FIXED)


@pytest.mark.parametrize("kind", ["all", "ports_apart", "sources_apart", "none"])
def test_post_e_launch_f64_bit_exact(tmp_path, kind):
    """The passes after the E update (special cells, TF/SF-E, sources, port
    sampling) share one launch as far as their cells allow: every outcome of
    the host's test, with the launches it is expected to choose, equals the
    oracle bit for bit.  all: nothing shared; ports_apart: the port plane
    crosses the TF/SF x and y faces; sources_apart: the source sits in the
    Drude-Lorentz sphere; none: the sphere crosses a TF/SF face."""
    src, port_z, sphere_x = {"all": ((12, 12, 12), 4, 20), "ports_apart": ((12, 12, 12), 16, 20),
                             "sources_apart": ((20, 18, 16), 4, 20), "none": ((12, 12, 12), 4, 8)}[kind]
    case = _post_case(str(tmp_path), src, port_z, sphere_x)
    e, o = run_both(case, 60)
    assert_fields_equal(e, o, kind)
    assert np.array_equal(e.port_samples(0), o.port_samples(0))
    assert np.abs(o.field(2)).max() > 0
    t = eng.Engine(case, eng.CFD_F64)
    t.set_timing(True)
    t.step(3)
    names = set(t.kernel_times())
    t.close()
    expect = {"all": {"post_e"}, "ports_apart": {"post_e", "ports"}, "sources_apart": {"post_e", "sources", "ports"},
              "none": {"special_e", "tfsf_e", "sources", "ports"}}[kind]
    launched = names & {"post_e", "special_e", "tfsf_e", "sources", "ports"}
    assert launched == expect, (kind, sorted(names))
