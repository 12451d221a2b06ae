"""Image export (cfd_image_*, the reference's OutImage.cpp): colour maps of a
plane port's maps and the PNG files they are saved in.  CPU only.

The reference draws with CImg and saves with CImg::save_png; neither CImg nor
libpng is in this image, so the reference cannot be run here.  The colour map
is pinned by values worked out by hand from OutImage.cpp:140-202 (including
its thresholds (cmax + cmin) / 3 and (cmax + third) / 2, which differ from
equal thirds of [cmin, cmax] when cmin != 0, and the truncating int
conversion); the PNG files are decoded independently here (chunk CRCs, IHDR,
zlib stream, filter bytes)."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import cpsfdtd_amd as eng
from cpsfdtd_amd import build as engine_build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", autouse=True)
def lib():
    if not os.path.exists(engine_build.LIB):
        pytest.skip("libcpsfdtd.so is not built")


def decode_png(path):
    """8-bit RGB PNG -> pixels [height][width][3]; checks every chunk CRC."""
    b = open(path, "rb").read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(b):
        n, = struct.unpack(">I", b[pos:pos + 4])
        kind, data = b[pos + 4:pos + 8], b[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", b[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(kind + data) == crc, kind
        chunks.append((kind, data))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1] == (b"IEND", b"")
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(b"".join(d for k, d in chunks if k == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(rows[:, 0] == 0)  # filter type 0 on every row
    return rows[:, 1:].reshape(h, w, 3).copy()


def test_colormap_known_values():
    # cmin 0, cmax 0.02 (main.cpp:149): thresholds 0.00667 and 0.01333
    v = [-1.0, 0.0, 0.001, 0.006, 0.007, 0.013, 0.014, 0.0199, 0.02, 1.0]
    want = [(0, 0, 0), (0, 0, 0), (12, 0, 0), (76, 0, 0), (254, 88, 0), (254, 165, 0),
            (254, 254, 177), (254, 254, 252), (254, 254, 254), (254, 254, 254)]
    rgb = eng.image_colormap(np.array(v)[:, None], 0.0, 0.02)
    assert rgb.shape == (1, len(v), 3)
    assert [tuple(int(c) for c in p) for p in rgb[0]] == want
    # cmin -1, cmax 2: the reference's thresholds are 1/3 and 7/6 (equal
    # thirds would be 0 and 1, and give (254, 101, 0) and (254, 254, 177))
    rgb = eng.image_colormap(np.array([[0.2], [1.1], [1.5]]), -1.0, 2.0)
    assert [tuple(int(c) for c in p) for p in rgb[0]] == [(101, 0, 0), (254, 177, 0), (254, 254, 211)]


def test_colormap_layout_and_nan():
    """map[x][y] -> pixel row y, column x; a NaN leaves its pixel as it was."""
    m = np.zeros((3, 2))
    m[2, 0] = 0.001  # x = 2, y = 0
    m[0, 1] = 1.0    # x = 0, y = 1
    rgb = eng.image_colormap(m, 0.0, 0.02)
    assert rgb.shape == (2, 3, 3)
    assert tuple(rgb[0, 2]) == (12, 0, 0) and tuple(rgb[1, 0]) == (254, 254, 254)
    m[0, 1] = np.nan
    m[2, 0] = 0.0
    eng.image_colormap(m, 0.0, 0.02, rgb)
    assert tuple(rgb[1, 0]) == (254, 254, 254) and tuple(rgb[0, 2]) == (0, 0, 0)


def test_material_colours():
    ids = np.arange(10, dtype=np.int16).reshape(5, 2)  # x 0..4, y 0..1
    rgb = eng.image_materials(ids)
    table = [(254, 254, 254), (254, 0, 0), (0, 0, 254), (0, 254, 0), (0, 0, 0),
             (254, 254, 0), (0, 254, 254), (254, 0, 254), (127, 0, 254), (0, 127, 254)]
    for x in range(5):
        for y in range(2):
            assert tuple(rgb[y, x]) == table[ids[x, y]]
    for bad in (10, -1):
        with pytest.raises(eng.EngineError, match="no colour"):
            eng.image_materials(np.array([[0, bad]], np.int16))


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    for h, w in ((1, 1), (7, 13), (64, 300)):
        rgb = rng.integers(0, 255, size=(h, w, 3), dtype=np.uint8)
        p = str(tmp_path / ("r%d_%d.png" % (h, w)))
        eng.write_png(p, rgb)
        assert np.array_equal(decode_png(p), rgb)
    with pytest.raises(eng.EngineError, match="cannot open"):
        eng.write_png(str(tmp_path / "no" / "such" / "dir.png"), rgb)
    with pytest.raises(eng.EngineError, match="bad image size"):  # a side above 16384
        eng.write_png(str(tmp_path / "wide.png"), np.zeros((1, 16385, 3), np.uint8))


def test_outimage_facade(tmp_path):
    """cpsfdtd::OutImage (include/cpsfdtd.hpp) on the CPU: plot_save writes
    <base>NNNNNN.png only on steps that are multiples of the save modulus, with
    NNNNNN = step / modulus; structure_plot_save writes <base>_structure.png."""
    src = tmp_path / "oi.cpp"
    src.write_text(r'''
#include "cpsfdtd.hpp"
int main(int, char **argv)
{
    double a[3][2] = {{0.0, 1.0}, {0.007, 0.001}, {0.014, -1.0}};
    double *rows[3] = {a[0], a[1], a[2]};
    short s[3][2] = {{0, 1}, {2, 3}, {4, 9}};
    short *srows[3] = {s[0], s[1], s[2]};
    cpsfdtd::OutImage img(3, 2, 5, 0.0, 0.02, (std::string(argv[1]) + "/f_").c_str());
    img.plot_save(rows, 3);
    img.plot_save(rows, 10);
    img.structure_plot_save(srows);
    return 0;
}
''')
    exe = tmp_path / "oi"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L" + engine_build.LIBDIR, "-lcpsfdtd", "-Wl,-rpath," + engine_build.LIBDIR], check=True)
    out = tmp_path / "o"
    out.mkdir()
    subprocess.run([str(exe), str(out)], check=True, timeout=60)
    assert sorted(os.listdir(out)) == ["f_000002.png", "f__structure.png"]
    want = eng.image_colormap(np.array([[0.0, 1.0], [0.007, 0.001], [0.014, -1.0]]), 0.0, 0.02)
    assert np.array_equal(decode_png(str(out / "f_000002.png")), want)
    assert np.array_equal(decode_png(str(out / "f__structure.png")),
                          eng.image_materials(np.array([[0, 1], [2, 3], [4, 9]], np.int16)))
