"""Text layout of the reference's port exports (Ports.cpp), written from port
samples for the tests.  std::ostream prints a double with its default format,
%g at precision 6; every exported value is followed by a tab.

Layouts (P points, samples [steps][P][6]):
  points   Export_Port: "i\\tj\\tk" per point (Ports.cpp:355-362); a sphere
           port first prints its centre in cells (Ports.cpp:101-109)
  E_last   Export_E_onPort(file): "Ex\\tEy\\tEz" per point (Ports.cpp:111-119, 363-371)
  Ex, I    Export_Ex_onPort / Export_EIntensity_onPort(ofstream&), one line
           per call (Ports.cpp:128-135, 180-186, 378-388, 501-510)
  Ex_file  Export_Ex_onPort(file): sphere ports one line without a newline
           (Ports.cpp:120-127), plane ports Nw values per line (Ports.cpp:366-377)
  I_file   Export_EIntensity_onPort(file): sphere one line without a newline
           (Ports.cpp:172-179), plane one line with a newline (Ports.cpp:491-500)
  ID_file  plane ports, Export_ID_onPort: the material IDs at the points,
           Nw per line (Ports.cpp:308-320)
The committed fixture tests/golden/ref_port_text.npz holds what the
reference's own Ports.cpp writes; these functions are checked against it.
"""
import numpy as np

KEYS = ("points", "E_last", "Ex", "I", "Ex_file", "I_file")
PLANE_KEYS = KEYS + ("ID_file",)


def g6(v):
    return "%g" % v


def _tabbed(values):
    return "".join(g6(v) + "\t" for v in values)


def plane_nw(path, d):
    """Points along a plane port's first in-plane axis, from its file as
    Ports.cpp:232-283 computes them: (int)(w2/dw) - (int)(w1/dw) + 1."""
    lines = open(path).read().splitlines()
    normal = lines[0].split()[0]
    w1, w2 = (float(x) for x in lines[2].split()[:2])
    dw = d[1] if normal == "x" else d[0]
    return int(w2 / dw) - int(w1 / dw) + 1


def sphere_centre(path, d):
    """A sphere port's centre in cells, xc/dx yc/dy zc/dz (Ports.cpp:104)."""
    xc, yc, zc = (float(x) for x in open(path).read().splitlines()[1].split()[:3])
    return xc / d[0], yc / d[1], zc / d[2]


def exports(samples, points, sphere, nw=None, centre=None, ids=None):
    """{key: text} for one port after len(samples) steps; ids: the material
    ID at each point (plane ports)."""
    s = np.asarray(samples, np.float64)
    ex = s[:, :, 0]
    inten = s[:, :, 0] * s[:, :, 0] + s[:, :, 1] * s[:, :, 1] + s[:, :, 2] * s[:, :, 2]
    head = ("%s\t%s\t%s\n" % tuple(g6(c) for c in centre)) if sphere else ""
    r = {"points": head + "".join("%d\t%d\t%d\n" % tuple(p) for p in points),
         "E_last": "".join("%s\t%s\t%s\n" % (g6(a), g6(b), g6(c)) for a, b, c in s[-1, :, :3]),
         "Ex": "".join(_tabbed(row) + "\n" for row in ex),
         "I": "".join(_tabbed(row) + "\n" for row in inten)}
    if sphere:
        r["Ex_file"] = _tabbed(ex[-1])
        r["I_file"] = _tabbed(inten[-1])
    else:
        n = s.shape[1]
        r["Ex_file"] = "".join(_tabbed(ex[-1][i:i + nw]) + "\n" for i in range(0, n, nw))
        r["I_file"] = _tabbed(inten[-1]) + "\n"
        if ids is not None:
            r["ID_file"] = "".join("".join("%d\t" % v for v in ids[i:i + nw]) + "\n" for i in range(0, n, nw))
    return r


def fixture(g, q):
    """{key: text} of port q from the loaded ref_port_text.npz."""
    return {k: bytes(g["port%d_%s" % (q, k)]).decode() for k in PLANE_KEYS if "port%d_%s" % (q, k) in g}
