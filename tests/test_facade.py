"""The C++ facade (include/cpsfdtd.hpp): the reference's class names and call
sequence (main.cpp:115-343) over the C ABI.

tests/facade/main_loop.cpp is a driver in the shape of the reference's main:
Structure, Field, CPML, TF_SF, Drude_Lorentz, FourLevelLaser, PlainPort and
SphericalPort objects, then the per-step module calls in the reference's
order.  Its outputs (final fields, port values read through port_point, the
text written by Export_Ex_onPort / Export_E_onPort / Export_Port) are checked
against the FP64 oracle: bit for bit for values, character for character for
text (default stream formatting of the same doubles).
"""
import os
import subprocess

import numpy as np
import pytest

import cases
import cpsfdtd_amd as eng
from port_text import g6
from cpsfdtd_amd import build as engine_build
from cpsfdtd_amd.case import FIXED

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = engine_build.FACADE_DRIVER


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        if not os.path.exists(engine_build.LIB):
            pytest.skip("libcpsfdtd.so is not built")
        engine_build.build_facade_driver()
    return DRIVER


def driver_args(case, steps, out):
    f = case.files
    args = [DRIVER, "--material", f["material"], "--structure", f["structure"], "--steps", str(steps),
            "--flags", str(case.flags), "--out", str(out)]
    for key in ("pml", "tfsf", "dl", "gain", "source"):
        if f[key]:
            args += ["--" + key, f[key]]
    planes = [p for p, k in case.ports if k == 0]
    spheres = [p for p, k in case.ports if k == 1]
    assert [p for p, _ in case.ports] == planes + spheres, "driver orders plane ports first"
    for p in planes:
        args += ["--plane", p]
    for p in spheres:
        args += ["--sphere", p]
    for s in case.sources:
        assert s[5] == 0.0
        args += ["--src"] + [str(v) for v in s[:4]] + [repr(float(s[4]))]
    return args


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "cpsfdtd.hpp"\nint main() { return 0; }\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                    "-I" + os.path.join(ROOT, "include"), str(src)], check=True)


def test_port_exports_before_the_run_follow_main_cpp_order(tmp_path):
    """main.cpp:144-160 exports the sphere ports' point lists right after
    constructing them, then constructs the plane ports and saves their
    structure images, all before the first update.  Through the facade this
    runs host-side only (no simulation is started, so later constructors
    still register), and writes what the reference's Ports.cpp wrote for the
    same ports (tests/golden/ref_port_text.npz) and the OutImage structure map
    of those IDs."""
    import port_text
    from test_image import decode_png
    if not os.path.exists(engine_build.LIB):
        pytest.skip("libcpsfdtd.so is not built")
    case = cases.tiny(str(tmp_path / "case"), text_ports=True)
    src = tmp_path / "order.cpp"
    src.write_text(r'''
#include "cpsfdtd.hpp"
using namespace cpsfdtd;
int main(int, char **argv)
{
    const std::string d = argv[1], out = argv[2];
    Structure S((d + "/material.txt").c_str());
    S.generate_mesh((d + "/structure.txt").c_str());
    S.InitializeBasicFDTDCoefficients();
    SphericalPort cap(S, (d + "/cap.txt").c_str());
    cap.Export_Port((out + "/cap_points.txt").c_str());
    PlainPort line(S, (d + "/line.txt").c_str());
    PlainPort face(S, (d + "/face.txt").c_str());
    face.Export_Port((out + "/face_points.txt").c_str());
    line.Export_ID_onPort((out + "/line_ID.txt").c_str());
    OutImage img(line.Nw, line.Nh, 1, 0, 0.02, (out + "/line").c_str());
    img.structure_plot_save(line.Export_ID());
    return S.ctx->sim ? 3 : 0;
}
''')
    exe = tmp_path / "order"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L" + engine_build.LIBDIR, "-lcpsfdtd", "-Wl,-rpath," + engine_build.LIBDIR], check=True)
    out = tmp_path / "out"
    out.mkdir()
    subprocess.run([str(exe), case.dir, str(out)], check=True, timeout=60)
    g = dict(np.load(os.path.join(ROOT, "tests", "golden", "ref_port_text.npz"), allow_pickle=False))
    line, face, cap = (port_text.fixture(g, q) for q in range(3))
    assert (out / "cap_points.txt").read_text() == cap["points"]
    assert (out / "face_points.txt").read_text() == face["points"]
    assert (out / "line_ID.txt").read_text() == line["ID_file"]
    ids = np.array([[int(v) for v in line["ID_file"].split()]], np.int16).T  # [Nw][Nh = 1]
    assert np.array_equal(decode_png(str(out / "line_structure.png")), eng.image_materials(ids))


def test_tfsf_public_parameters(tmp_path):
    """The facade TF_SF exposes the reference's public parameters (main.cpp
    prints tfsf.lambda and tfsf.nu and loops n < tfsf.Nt), read and derived
    as TF_SF.cpp does: angles to radians, nu = c / lambda, T = 1 / nu,
    nT = (int)(T / dt), i0 = (int)(xs / dx)."""
    if not os.path.exists(engine_build.LIB):
        pytest.skip("libcpsfdtd.so is not built")
    case = cases.tiny(str(tmp_path / "case"))
    src = tmp_path / "tf.cpp"
    src.write_text(r'''
#include <cstdio>
#include "cpsfdtd.hpp"
using namespace cpsfdtd;
int main(int, char **argv)
{
    const std::string d = argv[1];
    Structure S((d + "/material.txt").c_str());
    S.generate_mesh((d + "/structure.txt").c_str());
    TF_SF tfsf(S, (d + "/TF_SF.txt").c_str());
    int steps = 0;
    for (int n = 1; n < tfsf.Nt; ++n) ++steps;  // main.cpp:267
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %d %d %d %d %d %d %d %d %.17g %.17g\n", tfsf.lambda,
                tfsf.nu, tfsf.T, tfsf.theta_inc, tfsf.phi_inc, tfsf.psi_E, tfsf.E0, tfsf.Nt, tfsf.nT, tfsf.i0,
                tfsf.j0, tfsf.nTFSFx1, tfsf.nTFSFz2, steps, (int)tfsf.cwmode, tfsf.w0, tfsf.kz);
    return 0;
}
''')
    exe = tmp_path / "tf"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L" + engine_build.LIBDIR, "-lcpsfdtd", "-Wl,-rpath," + engine_build.LIBDIR], check=True)
    v = subprocess.run([str(exe), case.dir], check=True, capture_output=True, text=True, timeout=60).stdout.split()
    lines = open(case.files["tfsf"]).read().splitlines()
    lam = float(lines[0].split()[0])
    xs = float(lines[3].split()[0])
    th, ph = (float(x) for x in lines[4].split()[:2])
    e0, psi = (float(x) for x in lines[5].split()[:2])
    nt = int(lines[9].split()[0])
    dt = float(subprocess.run([DRIVER] + ["--material", case.files["material"], "--structure",
                                          case.files["structure"], "--mesh-only", "--out", str(tmp_path)],
                              check=True, capture_output=True, text=True).stdout.split()[6]) \
        if os.path.exists(DRIVER) else None
    pi, c = 3.1415926535897932384626433832795, 2.99792458E8
    nu = c / lam
    assert float(v[0]) == lam and float(v[1]) == nu and float(v[2]) == 1.0 / nu
    assert float(v[3]) == th * pi / 180.0 and float(v[4]) == ph * pi / 180.0 and float(v[5]) == psi * pi / 180.0
    assert float(v[6]) == e0 and int(v[7]) == nt == 40 and int(v[13]) == nt - 1
    if dt is not None:
        assert int(v[8]) == int((1.0 / nu) / dt)
    assert int(v[9]) == int(xs / 10e-9) and int(v[11]) == 6 and int(v[12]) == 6
    assert float(v[16]) == np.cos(th * pi / 180.0)


def test_structure_mesh_matches_oracle(tmp_path, driver, oracle_lib):
    """Structure(material) + generate_mesh(structure) through the facade, no
    device: dimensions, time step and the ID map equal the oracle's."""
    case = cases.mixed(str(tmp_path / "case"))
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run(driver_args(case, 1, out) + ["--mesh-only"], check=True, capture_output=True, text=True)
    n = r.stdout.split()
    o = oracle_lib.OracleSim(case)
    assert [int(v) for v in n[:3]] == list(o.N)
    assert float(n[6]) == o.dt
    want = o.id_map()
    got = np.fromfile(out / "id.i16", dtype=np.int16).reshape(want.shape)
    assert np.array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["ref", "dl_gain"])
def test_main_loop_matches_oracle(tmp_path, driver, oracle_lib, variant):
    if variant == "ref":
        case = cases.mixed(str(tmp_path / "case"))
    else:
        case = cases.mixed(str(tmp_path / "case"), dl=True, gain=True, flags=FIXED)
    steps = 40
    out = tmp_path / "out"
    out.mkdir()
    print("facade main_loop", variant, flush=True)
    subprocess.run(driver_args(case, steps, out), check=True, timeout=120)
    o = oracle_lib.OracleSim(case)
    o.step(steps)
    for c, nm in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        want = o.field(c)
        got = np.fromfile(out / ("field_%s.f64" % nm), dtype=np.float64).reshape(want.shape)
        assert np.array_equal(got, want), nm
    for q in range(len(case.ports)):
        want = o.port_samples(q)
        got = np.fromfile(out / ("port%d_samples.f64" % q), dtype=np.float64).reshape(want.shape)
        assert np.array_equal(got, want), q
        rows = (out / ("port%d_Ex.txt" % q)).read_text().split("\n")
        assert rows[-1] == "" and len(rows) == steps + 1
        for n in (0, steps // 2, steps - 1):
            assert rows[n] == "".join(g6(v) + "\t" for v in want[n, :, 0]), (q, n)
        last = (out / ("port%d_E_last.txt" % q)).read_text()
        assert last == "".join("%s\t%s\t%s\n" % (g6(a), g6(b), g6(c)) for a, b, c in want[-1, :, :3])
        pts = o.port_points(q)
        lines = (out / ("port%d_points.txt" % q)).read_text().splitlines()
        if case.ports[q][1] == 1:
            lines = lines[1:]  # sphere ports begin with the centre in cells
        assert lines == ["%d\t%d\t%d" % tuple(p) for p in pts]


@pytest.mark.gpu
def test_main_loop_port_text_matches_reference_exports(tmp_path, driver, oracle_lib):
    """Every port text file the facade's Export_* methods write (a line, a 2-D
    plane and a sphere port over 40 steps) equals, byte for byte, what the
    reference's own Ports.cpp wrote for the same run
    (tests/golden/ref_port_text.npz, without the values the reference prints
    from past the last point, defect F11).  The OutImage PNGs of the plane
    ports (structure_plot_save(Export_ID()), plot_save(Export_EIntensity()))
    hold the colour maps of the oracle's ID map and last-step intensities."""
    import port_text
    from test_image import decode_png
    g = dict(np.load(os.path.join(ROOT, "tests", "golden", "ref_port_text.npz"), allow_pickle=False))
    case = cases.tiny(str(tmp_path / "case"), text_ports=True)
    out = tmp_path / "out"
    out.mkdir()
    cmax = 4e-6
    print("facade port text", flush=True)
    subprocess.run(driver_args(case, case.nt, out) + ["--images", repr(cmax)], check=True, timeout=120)
    for q in range(len(case.ports)):
        want = port_text.fixture(g, q)
        for k in want:
            assert (out / ("port%d_%s.txt" % (q, k))).read_text() == want[k], (q, k)
    o = oracle_lib.OracleSim(case)
    o.step(case.nt)
    ids3 = o.id_map()
    for q, (path, kind) in enumerate(case.ports):
        if kind != 0:
            continue
        nw = port_text.plane_nw(path, o.d)
        pts = o.port_points(q)
        x = o.port_samples(q)[-1]
        inten = x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1] + x[:, 2] * x[:, 2]
        as_map = lambda v: np.asarray(v).reshape(-1, nw).T  # point n -> [n % Nw][n // Nw]
        got = decode_png(str(out / ("port%d_img%06d.png" % (q, case.nt))))
        assert np.array_equal(got, eng.image_colormap(as_map(inten), 0.0, cmax)), q
        got = decode_png(str(out / ("port%d_img_structure.png" % q)))
        assert np.array_equal(got, eng.image_materials(as_map(ids3[pts[:, 0], pts[:, 1], pts[:, 2]]))), q


def _compile(tmp_path, name, code):
    if not os.path.exists(engine_build.LIB):
        pytest.skip("libcpsfdtd.so is not built")
    src = tmp_path / (name + ".cpp")
    src.write_text(code)
    exe = tmp_path / name
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L" + engine_build.LIBDIR, "-lcpsfdtd", "-Wl,-rpath," + engine_build.LIBDIR], check=True)
    return str(exe)


FACADE_GOLD = os.path.join(ROOT, "tests", "golden", "ref_facade.npz")


def test_source_class_matches_reference(tmp_path):
    """cpsfdtd::Source (Sources.h:24-40) on the source file of the tiny case:
    Pulse and ContinuousWave at ntimestep = q / 2 (q < 2000, through the
    ramp and into the steady sine), amplitude 1.7, locations 0 and 2e-7 m,
    equal bit for bit to the reference's own Source (fixture written by
    oracle/_ref/refdrv from Sources.cpp); the file's lambda, FWHM and delay
    and nu = c / lambda, T = 1 / nu are the public members."""
    g = dict(np.load(FACADE_GOLD, allow_pickle=False))
    case = cases.tiny(str(tmp_path / "case"), nt=80)
    exe = _compile(tmp_path, "src", r'''
#include <cstdio>
#include "cpsfdtd.hpp"
using namespace cpsfdtd;
int main(int, char **argv)
{
    const std::string d = argv[1], out = argv[2];
    Structure S((d + "/material.txt").c_str());
    S.generate_mesh((d + "/structure.txt").c_str());
    Source src(S.dt, (d + "/source.txt").c_str());
    const double locs[2] = {0.0, 2e-7};
    FILE *f = std::fopen(out.c_str(), "wb");
    for (double loc : locs)
        for (int q = 0; q < 2000; q++) {
            const double v[2] = {src.Pulse(1.7, q / 2.0, loc), src.ContinuousWave(1.7, q / 2.0, loc)};
            std::fwrite(v, sizeof v, 1, f);
        }
    std::fclose(f);
    std::printf("%.17g %.17g %.17g %.17g %.17g\n", src.lambda, src.FWHM, src.delay, src.nu, src.T);
    return S.ctx->sim ? 3 : 0;
}
''')
    out = tmp_path / "wave.f64"
    v = subprocess.run([exe, case.dir, str(out)], check=True, capture_output=True, text=True, timeout=60).stdout.split()
    w = np.fromfile(out, dtype=np.float64).reshape(2, 2000, 2)
    for q in range(2):
        assert np.array_equal(w[q, :, 0], g["pulse%d" % q]), q
        assert np.array_equal(w[q, :, 1], g["cw%d" % q]), q
    assert np.abs(g["cw0"]).max() > 1.6 and np.abs(g["pulse1"]).max() > 1.0  # both forms reach their amplitude
    lines = open(os.path.join(case.dir, "source.txt")).read().splitlines()
    lam, fwhm, delay = (float(lines[k].split()[0]) for k in range(3))
    assert [float(x) for x in v[:3]] == [lam, fwhm, delay]
    assert float(v[3]) == 2.99792458E8 / lam and float(v[4]) == 1.0 / float(v[3])


@pytest.mark.parametrize("cw", [0, 1], ids=["pulse", "cw"])
def test_tfsf_inc_on_point_matches_reference(tmp_path, cw):
    """TF_SF::Einc_onPoint and ::Hinc_onPoint (TF_SF.cpp:966-986) through
    the facade, host only (no simulation starts), at retarded = q / 4 - 1 up
    to the table's end (0 before the wave starts, linear between entries),
    equal bit for bit to the reference's own TF_SF in both modes; the public
    Einc / Hinc tables extend past Nt on demand with the same values."""
    g = dict(np.load(FACADE_GOLD, allow_pickle=False))
    nt = int(g["inc_cw%d_nt" % cw][0])
    want = g["inc_cw%d" % cw]
    case = cases.tiny(str(tmp_path / "case"), nt=nt, cw=cw)
    exe = _compile(tmp_path, "inc", r'''
#include <cstdio>
#include <cstdlib>
#include "cpsfdtd.hpp"
using namespace cpsfdtd;
int main(int, char **argv)
{
    const std::string d = argv[1], out = argv[2];
    const int nq = std::atoi(argv[3]);
    Structure S((d + "/material.txt").c_str());
    S.generate_mesh((d + "/structure.txt").c_str());
    TF_SF tf(S, (d + "/TF_SF.txt").c_str());
    const size_t len0 = tf.Einc.size();
    FILE *f = std::fopen(out.c_str(), "wb");
    for (int q = 0; q < nq; q++) {
        const double v[2] = {tf.Einc_onPoint(q / 4.0 - 1.0), tf.Hinc_onPoint(q / 4.0 - 1.0)};
        std::fwrite(v, sizeof v, 1, f);
    }
    std::fclose(f);
    const double late = tf.Einc_onPoint(len0 + 10.5);  // past the table: it grows
    std::printf("%zu %zu %.17g %.17g\n", len0, tf.Einc.size(), late, tf.Einc[len0 - 1]);
    return S.ctx->sim ? 3 : 0;
}
''')
    out = tmp_path / "inc.f64"
    v = subprocess.run([exe, case.dir, str(out), str(len(want))], check=True, capture_output=True, text=True,
                       timeout=60).stdout.split()
    got = np.fromfile(out, dtype=np.float64).reshape(-1, 2)
    assert np.array_equal(got, want)
    assert np.abs(want).max() > (0.5 if cw else 1e-3)
    assert int(v[0]) == nt + 3 and int(v[1]) > int(v[0]) + 10 and np.isfinite(float(v[2]))


def test_module_public_parameters(tmp_path):
    """The public parameters of CFS_PML (CFS_PML.h:37-39), Drude_Lorentz
    (Drude_Lorentz.h:33-35) and FourLevelLaser (FourLevelLaser.h:31-38) are
    the values of their files, line for line, and omega_a = 2 pi c /
    lambda_a; constructing them starts no simulation."""
    case = cases.mixed(str(tmp_path / "case"), dl=True, gain=True)
    exe = _compile(tmp_path, "params", r'''
#include <cstdio>
#include "cpsfdtd.hpp"
using namespace cpsfdtd;
int main(int, char **argv)
{
    const std::string d = argv[1];
    Structure S((d + "/material.txt").c_str());
    S.generate_mesh((d + "/structure.txt").c_str());
    CFS_PML pml(S, argv[2]);
    Drude_Lorentz DL(S, argv[3]);
    FourLevelLaser FL(S, argv[4]);
    std::printf("%d %d %d %d %d %d %.17g %.17g %.17g %.17g %.17g\n", pml.Npmlx1, pml.Npmlx2, pml.Npmly1, pml.Npmly2,
                pml.Npmlz1, pml.Npmlz2, pml.p_order, pml.malp, pml.kapmax, pml.sig_ratio, pml.alpmax);
    std::printf("%d %d %.17g", DL.num_of_D_poles, DL.num_of_L_poles, DL.epsInf);
    for (size_t q = 0; q < DL.omega_D.size(); q++) std::printf(" %.17g %.17g", DL.omega_D[q], DL.gamma_D[q]);
    for (size_t q = 0; q < DL.omega_L.size(); q++)
        std::printf(" %.17g %.17g %.17g", DL.omega_L[q], DL.delta_L[q], DL.delta_eps_L[q]);
    std::printf("\n%s %d %d %d %d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
                FL.distri, FL.NumSources, FL.i1, FL.i2, FL.j1, FL.j2, FL.lambda_b, FL.lambda_a, FL.gamma_a, FL.gamma_b,
                FL.t30, FL.t32, FL.t21, FL.t10, FL.N_0, FL.N_1, FL.pump_rate, FL.omega_a);
    return S.ctx->sim ? 3 : 0;
}
''')
    f = case.files
    rows = subprocess.run([exe, case.dir, f["pml"], f["dl"], f["gain"]], check=True, capture_output=True, text=True,
                          timeout=60).stdout.splitlines()

    def numbers(path, skip=0):
        return [float(x) for line in open(path).read().splitlines()[skip:] for x in line.split()]
    pml = [float(x) for x in rows[0].split()]
    assert pml == numbers(case.files["pml"])[:11]
    dl = [float(x) for x in rows[1].split()]
    assert dl == numbers(case.files["dl"])[:len(dl)] and len(dl) == 3 + 2 * dl[0] + 3 * dl[1]
    fl = rows[2].split()
    lines = open(case.files["gain"]).read().splitlines()
    assert fl[0] == lines[0].split()[0]
    file_vals = [float(x) for line in lines[1:15] for x in line.split()[:2 if line in lines[2:4] else 1]]
    assert [float(x) for x in fl[1:17]] == file_vals
    assert float(fl[17]) == 2.0 * 3.1415926535897932384626433832795 * 2.99792458E8 / float(fl[7])


@pytest.mark.gpu
@pytest.mark.parametrize("order", [0, 1, 2])
def test_main_loop_tfsf_faces_match_reference(tmp_path, driver, order):
    """The main loop with the TF/SF correction through the facade's one-face
    methods (x1UpdateHField ... z2UpdateEField, TF_SF.cpp:1090-1410) in the
    fixture's orders, including z1 alone (the reference's z1 forms drop the
    profile factor): fields and port samples after 80 steps equal the
    reference's own run bit for bit (tests/golden/ref_facade.npz)."""
    import hashlib
    g = dict(np.load(FACADE_GOLD, allow_pickle=False))
    faces = [int(f) for f in g["faces%d" % order]]
    case = cases.tiny(str(tmp_path / "case"), nt=80)
    out = tmp_path / "out"
    out.mkdir()
    names = ("x1", "x2", "y1", "y2", "z1", "z2")
    print("facade faces", [names[f] for f in faces], flush=True)
    subprocess.run(driver_args(case, 80, out) + ["--faces", ",".join(names[f] for f in faces)], check=True, timeout=120)

    def sha(a):
        return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
    for c, nm in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        got = np.fromfile(out / ("field_%s.f64" % nm), dtype=np.float64)
        assert sha(got) == bytes(g["faces%d_field_sha256" % order][c]).decode(), nm
    assert sha(np.fromfile(out / "port0_samples.f64", dtype=np.float64)) == \
        bytes(g["faces%d_port0_sha256" % order]).decode()


@pytest.mark.gpu
def test_main_loop_cfs_pml_and_state_match_oracle(tmp_path, driver, oracle_lib):
    """The main loop driven through CFS_PML (Initialize_PML, UpdatePMLHFields,
    UpdatePMLEFields; CFS_PML.h:31-35) on the Drude-Lorentz + gain case equals
    the oracle bit for bit (fields, port samples), and the state read back
    through Drude_Lorentz::DLM() and FourLevelLaser (location, N0..N3, Pax
    Pay Paz) equals the oracle's cell lists, currents, previous fields,
    populations and polarisations bit for bit."""
    case = cases.mixed(str(tmp_path / "case"), dl=True, gain=True, flags=FIXED)
    steps = 40
    out = tmp_path / "out"
    out.mkdir()
    print("facade CFS_PML + state", flush=True)
    subprocess.run(driver_args(case, steps, out) + ["--cfs", "--state"], check=True, timeout=120)
    o = oracle_lib.OracleSim(case)
    o.step(steps)
    for c, nm in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        want = o.field(c)
        got = np.fromfile(out / ("field_%s.f64" % nm), dtype=np.float64).reshape(want.shape)
        assert np.array_equal(got, want), nm
    for q in range(len(case.ports)):
        want = o.port_samples(q)
        assert np.array_equal(np.fromfile(out / ("port%d_samples.f64" % q), dtype=np.float64).reshape(want.shape), want)
    st = o.dl_state()
    n, _, per = st.shape
    dl = np.fromfile(out / "dl_state.f64", dtype=np.float64).reshape(n, 3 + 3 * per)
    assert n > 0 and np.array_equal(dl[:, :3].astype(np.int32), o.special_cells(0))
    assert np.array_equal(dl[:, 3:].reshape(n, 3, per), st)
    assert np.abs(st).max() > 0
    pop, pol = o.gain_pop(), o.gain_state()
    gs = np.fromfile(out / "gain_state.f64", dtype=np.float64).reshape(len(pop), 10)
    assert len(pop) > 0 and np.array_equal(gs[:, :3].astype(np.int32), o.special_cells(1))
    assert np.array_equal(gs[:, 3:7], pop) and np.array_equal(gs[:, 7:10], pol[:, :3])
    assert np.abs(pol[:, :3]).max() > 0
