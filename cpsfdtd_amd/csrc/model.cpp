// model.cpp -- builds cfd::Model from the reference's input files.
//
// Each section names the reference code it follows.  Expressions keep the
// reference's operand order and are compiled without FP contraction so that
// the material-ID map is bit-identical and every FP64 table equals the
// reference's.  Where the reference reads memory it never wrote, the engine
// takes the decisions recorded for the defect register (F1, F4, F11, F16);
// CFD_REF_ZEROHEAP instead reproduces what the reference computes on
// zero-filled memory.
#include "model.hpp"

#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <unordered_map>
#include <sstream>

namespace cfd {
namespace {

// getline()-per-line reader with istringstream >> semantics (CRLF tolerated:
// '\r' is whitespace to operator>>).
class LineReader {
public:
    explicit LineReader(const char *path) : path_(path ? path : "") {
        std::ifstream f(path_);
        if (!f) throw Error("cannot open " + path_);
        std::string s;
        while (std::getline(f, s)) lines_.push_back(s);
    }
    std::istringstream next() {
        ++n_;
        return std::istringstream(n_ <= (int)lines_.size() ? lines_[n_ - 1] : std::string());
    }
    [[noreturn]] void bad(const std::string &what) const {
        throw Error(path_ + ": line " + std::to_string(n_) + ": cannot read " + what);
    }
    template <class... T>
    void read(const std::string &what, T &...v) {
        std::istringstream is = next();
        if (!(is >> ... >> v)) bad(what);
    }

private:
    std::string path_;
    std::vector<std::string> lines_;
    int n_ = 0;
};

struct ShapeRec {
    int index = 0, material = 0;
    std::string type;
    double theta = 0, phi = 0, psi = 0, x0 = 0, y0 = 0, z0 = 0, a = 0, b = 0, c = 0;
    int ix0 = 0, jy0 = 0, kz0 = 0, ia = 0, jb = 0, kc = 0;
};

class Painter {
public:
    explicit Painter(Model &m) : m_(m) {}
    void set(int i, int j, int k, int mat) {
        if (i < 0 || j < 0 || k < 0 || i > m_.Nx || j > m_.Ny || k > m_.Nz) return;  // clipped
        m_.id[m_.flat(i, j, k)] = (int16_t)mat;
    }
    // Structure.cpp:159-203
    void cuboid(const ShapeRec &s) {
        if (s.theta == 0 && s.phi == 0 && s.psi == 0) {
            for (int i = 0; i < s.ia; ++i)
                for (int j = 0; j < s.jb; ++j)
                    for (int k = 0; k < s.kc; ++k) set(i + s.ix0, j + s.jy0, k + s.kz0, s.material);
            return;
        }
        const double dx = m_.dx, dy = m_.dy, dz = m_.dz;
        for (double x = 0; x < s.a; x = x + (dx / 2.0))
            for (double y = 0; y < s.b; y = y + (dy / 2.0))
                for (double z = 0; z < s.c; z = z + (dz / 2.0)) {
                    double x1 = x;
                    double y1 = std::cos(s.theta) * (y) - std::sin(s.theta) * (z);
                    double z1 = std::sin(s.theta) * (y) + std::cos(s.theta) * (z);
                    double x2 = std::cos(s.phi) * (x1) + std::sin(s.phi) * (z1);
                    double y2 = y1;
                    double z2 = -std::sin(s.phi) * (x1) + std::cos(s.phi) * (z1);
                    double x3 = std::cos(s.psi) * (x2) - std::sin(s.psi) * (y2);
                    double y3 = std::sin(s.psi) * (x2) + std::cos(s.psi) * (y2);
                    double z3 = z2;
                    int i3 = (int)((x3 + s.x0) / dx), j3 = (int)((y3 + s.y0) / dy), k3 = (int)((z3 + s.z0) / dz);
                    if (i3 >= 0 && i3 <= m_.Nx && j3 >= 0 && j3 <= m_.Ny && k3 >= 0 && k3 <= m_.Nz)
                        m_.id[m_.flat(i3, j3, k3)] = (int16_t)s.material;
                }
    }
    // Structure.cpp:205-221.  The three terms are non-negative (or NaN), and a
    // rounded sum of non-negative terms is never below either term, so a row
    // (i) or column (i, j) whose partial sum already exceeds 1 (or is NaN)
    // holds no painted cell: skipping it leaves the painted set unchanged and
    // scans a sphere's bounding columns instead of the whole grid.
    void ellipsoid(const ShapeRec &s) {
        const double dx = m_.dx, dy = m_.dy, dz = m_.dz;
        for (int i = 0; i < m_.Nx; i++) {
            double tx = ((double)i * dx - s.x0) * ((double)i * dx - s.x0) / (s.a * s.a);
            if (!(tx <= 1.0)) continue;
            for (int j = 0; j < m_.Ny; j++) {
                double ty = ((double)j * dy - s.y0) * ((double)j * dy - s.y0) / (s.b * s.b);
                double txy = tx + ty;
                if (!(txy <= 1.0)) continue;
                for (int k = 0; k < m_.Nz; k++) {
                    double t = txy + ((double)k * dz - s.z0) * ((double)k * dz - s.z0) / (s.c * s.c);
                    if (t <= 1.0) m_.id[m_.flat(i, j, k)] = (int16_t)s.material;
                }
            }
        }
    }
    // Structure.cpp:223-236 (rows whose x term alone exceeds 1 skipped, as
    // for the ellipsoid)
    void cylinder(const ShapeRec &s) {
        const double dx = m_.dx, dy = m_.dy;
        for (int i = 0; i < m_.Nx; i++) {
            const double tx = ((double)i * dx - s.x0) * ((double)i * dx - s.x0) / (s.a * s.a);
            if (!(tx <= 1.0)) continue;
            for (int j = 0; j < m_.Ny; j++) {
                double t = tx + ((double)j * dy - s.y0) * ((double)j * dy - s.y0) / (s.b * s.b);
                if (t <= 1.0)
                    for (int k = 0; k < s.kc; k++) set(i, j, k + s.kz0, s.material);
            }
        }
    }
    // Structure.cpp:238-306
    void multi(const ShapeRec &multi, const ShapeRec &single) {
        ShapeRec my = single;
        my.ia = (int)(my.a / m_.dx);
        my.jb = (int)(my.b / m_.dy);
        my.kc = (int)(my.c / m_.dz);
        for (int i = 0; i < multi.theta; ++i) {
            my.x0 = multi.x0 + i * multi.a;
            my.ix0 = (int)(my.x0 / m_.dx);
            for (int j = 0; j < multi.phi; ++j) {
                my.y0 = multi.y0 + j * multi.b;
                my.jy0 = (int)(my.y0 / m_.dy);
                for (int k = 0; k < multi.psi; ++k) {
                    my.z0 = multi.z0 + k * multi.c;
                    my.kz0 = (int)(my.z0 / m_.dz);
                    paint(my, nullptr);
                }
            }
        }
    }
    void paint(const ShapeRec &s, const std::vector<ShapeRec> *all) {
        if (s.type == "cuboid") cuboid(s);
        else if (s.type == "ellipsoid") ellipsoid(s);
        else if (s.type == "cylinder") cylinder(s);
        else if (s.type == "multi" && all) {
            if (s.material < 0 || s.material >= (int)all->size())
                throw Error("multi shape refers to shape line " + std::to_string(s.material) + ", which does not exist");
            multi(s, (*all)[s.material]);
        }
    }

private:
    Model &m_;
};

// ellipsoid's k loop is innermost with (i,j) terms hoisted; the hoisted sum
// (tx + ty) + tz is the reference's left-to-right order, so values match.

// Structure.cpp:26-52, 68-157, 309-359
void load_structure(Model &m, const char *matf, const char *meshf) {
    {
        LineReader L(matf);
        L.read("NoMedia", m.NoMedia);
        if (m.NoMedia <= 0 || m.NoMedia > 255) throw Error(std::string(matf) + ": NoMedia must be 1..255");
        m.epsR.resize(m.NoMedia); m.muR.resize(m.NoMedia); m.sig.resize(m.NoMedia); m.sigM.resize(m.NoMedia);
        for (int i = 0; i < m.NoMedia; i++) L.read("epsR muR sig sigM", m.epsR[i], m.muR[i], m.sig[i], m.sigM[i]);
    }
    std::vector<ShapeRec> shapes;
    int freespace = 0;
    {
        LineReader L(meshf);
        L.read("Lx Ly Lz", m.Lx, m.Ly, m.Lz);
        L.read("dx dy dz", m.dx, m.dy, m.dz);
        {
            std::istringstream is = L.next();
            is >> freespace;
        }
        int nsh = 0;
        {
            std::istringstream is = L.next();
            is >> nsh;
        }
        for (int i = 0; i < nsh; i++) {
            ShapeRec s;
            std::istringstream is = L.next();
            is >> s.index >> s.material >> s.type >> s.theta >> s.phi >> s.psi >> s.x0 >> s.y0 >> s.z0 >> s.a >> s.b >> s.c;
            s.ix0 = (int)(s.x0 / m.dx); s.jy0 = (int)(s.y0 / m.dy); s.kz0 = (int)(s.z0 / m.dz);
            s.ia = (int)(s.a / m.dx); s.jb = (int)(s.b / m.dy); s.kc = (int)(s.c / m.dz);
            shapes.push_back(s);
        }
    }
    m.Nx = (int)(m.Lx / m.dx);
    m.Ny = (int)(m.Ly / m.dy);
    m.Nz = (int)(m.Lz / m.dz);
    if (m.Nx < 2 || m.Ny < 2 || m.Nz < 2) throw Error("grid must have at least 2 cells per axis");
    // cell indices are 64-bit throughout; beyond 2^36 cells no device memory holds the fields anyway
    if ((double)(m.Nx + 1) * (m.Ny + 1) * (m.Nz + 1) > 68719476736.0) throw Error("grid too large for one simulation");
    m.dt = 0.99 / (kC * std::sqrt(1.0 / (m.dx * m.dx) + 1.0 / (m.dy * m.dy) + 1.0 / (m.dz * m.dz)));
    m.id.assign(m.cells(), 0);
    if (!freespace) {
        Painter p(m);
        for (const ShapeRec &s : shapes) p.paint(s, &shapes);
    }
    {
        const int none = -100000;  // no 16-bit ID has this value
        std::vector<int> bad(8, none);  // per thread: an undefined material it found
        const size_t n = m.id.size();
        host_parallel([&](int t, int nth) {
            const size_t a = n * t / nth, b = n * (t + 1) / nth;
            for (size_t q = a; q < b; q++)
                if (m.id[q] < 0 || m.id[q] >= m.NoMedia) {
                    bad[t] = m.id[q];
                    break;
                }
        });
        for (int v : bad)
            if (v != none) throw Error("structure uses material " + std::to_string(v) + ", not defined in the material file");
    }
    // Structure.cpp:309-359
    const int len[kNumStretch] = {m.Nz, m.Nz, m.Ny, m.Ny, m.Nx, m.Nx, m.Nz, m.Nz, m.Ny, m.Ny, m.Nx, m.Nx};
    const double inv[kNumStretch] = {1.0 / m.dz, 1.0 / m.dz, 1.0 / m.dy, 1.0 / m.dy, 1.0 / m.dx, 1.0 / m.dx,
                                     1.0 / m.dz, 1.0 / m.dz, 1.0 / m.dy, 1.0 / m.dy, 1.0 / m.dx, 1.0 / m.dx};
    for (int t = 0; t < kNumStretch; t++) m.stretch[t].assign(len[t], inv[t]);
    m.aE.resize(m.NoMedia); m.bE.resize(m.NoMedia); m.aH.resize(m.NoMedia); m.bH.resize(m.NoMedia);
    const double dt = m.dt;
    for (int i = 0; i < m.NoMedia; ++i) {
        m.aE[i] = (2.0 * kEps0 * m.epsR[i] - m.sig[i] * dt) / (2.0 * kEps0 * m.epsR[i] + m.sig[i] * dt);
        m.bE[i] = (2.0 * dt) / (2.0 * kEps0 * m.epsR[i] + m.sig[i] * dt);
        m.aH[i] = (2.0 * kMu0 * m.muR[i] - m.sigM[i] * dt) / (2.0 * kMu0 * m.muR[i] + m.sigM[i] * dt);
        m.bH[i] = (2.0 * dt) / (2.0 * kMu0 * m.muR[i] + m.sigM[i] * dt);
    }
}

// CPML.cpp:139-229 grading: E side on integer cells, H side on half cells
struct Grading {
    std::vector<double> kappa, b, c;
};
Grading grade(bool efield, bool firstend, int npml, double dt, double del, double m, double ma, double sigmax,
              double kapmax, double alpmax) {
    const int n = efield ? npml + 1 : npml;
    std::vector<double> sigma(n), alpha(n);
    Grading g;
    g.kappa.resize(n);
    g.b.assign(n, 0.0);
    g.c.assign(n, 0.0);
    for (int i = 1; i <= n; ++i) {
        double psk, pa;
        if (efield) {
            if (firstend) {
                psk = std::pow((((double)npml + 1.0 - i) / ((double)npml)), m);
                pa = std::pow((((double)i - 1.0) / ((double)npml)), ma);
            } else {
                psk = std::pow((((double)i - 1.0) / npml), m);
                pa = std::pow((((double)npml + 1.0 - i) / npml), ma);
            }
        } else {
            if (firstend) {
                psk = std::pow(((npml - i + 0.5) / npml), m);
                pa = std::pow(((i - 0.5) / npml), ma);
            } else {
                psk = std::pow(((i - 0.5) / npml), m);
                pa = std::pow(((npml - i + 0.5) / npml), ma);
            }
        }
        sigma[i - 1] = sigmax * psk;
        g.kappa[i - 1] = 1.0 + (kapmax - 1.0) * psk;
        alpha[i - 1] = alpmax * pa;
    }
    for (int i = 0; i < n; ++i) {  // H side: all npml entries (F1)
        g.b[i] = std::exp(-(sigma[i] / g.kappa[i] + alpha[i]) * dt / kEps0);
        if ((sigma[i] == 0.0) && (alpha[i] == 0.0)) g.c[i] = 0.0;
        else g.c[i] = sigma[i] * (g.b[i] - 1.0) / (sigma[i] + g.kappa[i] * alpha[i]) / g.kappa[i] / del;
    }
    return g;
}

// CPML.cpp:29-68, 231-527
void load_cpml(Model &m, const char *path) {
    LineReader L(path);
    for (int a = 0; a < 3; a++) L.read("PML thickness", m.npml[a][0], m.npml[a][1]);
    L.read("m malp", m.pm_m, m.pm_malp);
    L.read("kapmax sig_ratio", m.pm_kapmax, m.pm_sig_ratio);
    L.read("alpmax", m.pm_alpmax);
    if (m.pbx) m.npml[0][0] = m.npml[0][1] = 0;
    if (m.pby) m.npml[1][0] = m.npml[1][1] = 0;
    const int N[3] = {m.Nx, m.Ny, m.Nz};
    const double d[3] = {m.dx, m.dy, m.dz};
    for (int a = 0; a < 3; a++)
        if (m.npml[a][0] < 0 || m.npml[a][1] < 0 || m.npml[a][0] + m.npml[a][1] >= N[a])
            throw Error(std::string(path) + ": PML thickness does not fit the grid");
    const double eta0 = std::sqrt(kMu0 / kEps0);
    const int foldE[3][2] = {{cEzy, cEyz}, {cEzx, cExz}, {cExy, cEyx}};
    const int foldH[3][2] = {{cHzy, cHyz}, {cHzx, cHxz}, {cHxy, cHyx}};
    for (int a = 0; a < 3; a++) {
        const double sigmax = m.pm_sig_ratio * (0.8 * (m.pm_m + 1) / (d[a] * eta0));
        for (int e = 0; e < 2; e++) {
            const int n = m.npml[a][e];
            if (n == 0) continue;
            Grading gE = grade(true, e == 0, n, m.dt, d[a], m.pm_m, m.pm_malp, sigmax, m.pm_kapmax, m.pm_alpmax);
            Grading gH = grade(false, e == 0, n, m.dt, d[a], m.pm_m, m.pm_malp, sigmax, m.pm_kapmax, m.pm_alpmax);
            if (m.flags & CFD_REF_ZEROHEAP) { gH.b[n - 1] = 0.0; gH.c[n - 1] = 0.0; }
            // kappa folded into the stretch tables (CPML.cpp:361-519; second end bound Npml?2, F2)
            for (int i = 0; i < (e == 0 ? n + 1 : n); ++i) {
                const int t = e == 0 ? i : N[a] - n + i;
                m.stretch[foldE[a][0]][t] /= gE.kappa[i];
                m.stretch[foldE[a][1]][t] /= gE.kappa[i];
            }
            for (int i = 0; i < n; ++i) {
                const int t = e == 0 ? i : N[a] - n + i;
                m.stretch[foldH[a][0]][t] /= gH.kappa[i];
                m.stretch[foldH[a][1]][t] /= gH.kappa[i];
            }
            m.pbE[a][e] = gE.b; m.pcE[a][e] = gE.c; m.pbH[a][e] = gH.b; m.pcH[a][e] = gH.c;
        }
    }
    m.cpmlE = m.dt / kEps0;
    m.cpmlH = m.bH[0];
    m.has_pml = true;
}

// TF_SF.cpp:966-986 table extension and :898-955 waveforms
void tfsf_waveforms(Model &m, double lambda, double dlambda, int cwmode, int upto) {
    m.Einc.assign(upto + 3, 0.0);
    m.Hinc.assign(upto + 3, 0.0);
    const double dt = m.dt, nu = kC / lambda, T = 1.0 / nu;
    const int nT = (int)(T / dt), nd = 5 * nT;
    const double dnu = kC / (lambda - dlambda / 2.0) - kC / (lambda + dlambda / 2.0);
    const double domega = 2.0 * kPi * dnu;
    const double sig = domega * domega / (16 * std::log(2));
    for (int n = 0; n <= upto + 2; ++n) {
        if (cwmode) {
            if (n < nd) {
                m.Einc[n] = std::exp(-((n * dt) - 5.0 * T) * ((n * dt) - 5 * T) / (4.0 * T * T)) * std::sin(2.0 * kPi * nu * (n * dt));
                m.Hinc[n] = std::exp(-((n * dt + dt / 2.0) - 5.0 * T) * ((n * dt + dt / 2.0) - 5 * T) / (4.0 * T * T)) *
                            std::sin(2.0 * kPi * nu * (n * dt + dt / 2.0));
            } else {
                m.Einc[n] = std::sin(2.0 * kPi * nu * (n * dt));
                m.Hinc[n] = std::sin(2.0 * kPi * nu * (n * dt + dt / 2.0));
            }
        } else {
            const double t0 = 5 * T;
            m.Einc[n] = std::exp(-sig * ((double)n * dt - t0) * ((double)n * dt - t0)) * std::sin(2.0 * kPi * nu * ((double)n * dt - t0));
            m.Hinc[n] = std::exp(-sig * ((double)n * dt + dt / 2.0 - t0) * ((double)n * dt + dt / 2.0 - t0)) *
                        std::sin(2.0 * kPi * nu * ((double)n * dt + dt / 2.0 - t0));
        }
    }
}

// the incident tables of a TF/SF file alone (TF_SF.cpp:33-89 lines 1, 2, 10,
// 11 and :898-955): entries 0 .. max(Nt, nt) + 2
void tfsf_file_tables(const char *path, double dt, int nt, std::vector<double> &einc, std::vector<double> &hinc) {
    if (!(dt > 0)) throw Error("TF/SF tables need a time step > 0");
    LineReader L(path);
    double lambda, dlambda;
    int Nt = 0, cw = 0;
    L.read("lambda", lambda);
    L.read("dlambda", dlambda);
    for (int q = 0; q < 7; q++) L.next();
    L.read("Nt", Nt);
    {
        std::istringstream is = L.next();
        is >> cw;
    }
    Model t;
    t.dt = dt;
    tfsf_waveforms(t, lambda, dlambda, cw, std::max(Nt, nt));
    einc = std::move(t.Einc);
    hinc = std::move(t.Hinc);
}

// TF_SF.cpp:29-959 (geometry, tables) and :988-1301 (which component each face
// record corrects, with which sign, in which order)
void load_tfsf(Model &m, const char *path, int nt_req) {
    LineReader L(path);
    double lambda, dlambda, w0, z0, xs, ys, zs, theta, phi, E0, psiE;
    int nb[3][2], Nt = 0, cw = 0;
    L.read("lambda", lambda);
    L.read("dlambda", dlambda);
    L.read("w0 z0", w0, z0);
    const double zR = kPi * w0 * w0 / lambda;
    L.read("xs ys zs", xs, ys, zs);
    const int i0 = (int)(xs / m.dx), j0 = (int)(ys / m.dy), k0 = (int)(zs / m.dz);  // F4
    L.read("theta phi", theta, phi);
    theta = theta * kPi / 180.0;
    phi = phi * kPi / 180.0;
    L.read("E0 psi", E0, psiE);
    psiE = psiE * kPi / 180.0;
    for (int a = 0; a < 3; a++) L.read("TF/SF box", nb[a][0], nb[a][1]);
    L.read("Nt", Nt);
    {
        std::istringstream is = L.next();
        is >> cw;
    }
    if (m.pbx) nb[0][0] = nb[0][1] = 0;
    if (m.pby) nb[1][0] = nb[1][1] = 0;
    const int i1 = nb[0][0], i2 = m.Nx - nb[0][1], j1 = nb[1][0], j2 = m.Ny - nb[1][1], k1 = nb[2][0], k2 = m.Nz - nb[2][1];
    if ((!m.pbx && (i1 < 1 || i2 >= m.Nx)) || (!m.pby && (j1 < 1 || j2 >= m.Ny)) || k1 < 1 || k2 >= m.Nz ||
        i2 <= i1 || j2 <= j1 || k2 <= k1)
        throw Error(std::string(path) + ": TF/SF box must lie strictly inside the grid");
    m.tf_box[0][0] = i1; m.tf_box[0][1] = i2; m.tf_box[1][0] = j1; m.tf_box[1][1] = j2; m.tf_box[2][0] = k1; m.tf_box[2][1] = k2;
    const double eta0 = std::sqrt(kMu0 / kEps0);
    // TF_SF.cpp:112-117 (Taflove)
    const double E0x = E0 * (std::cos(psiE) * std::sin(phi) - std::sin(psiE) * std::cos(theta) * std::cos(phi));
    const double E0y = E0 * (-std::cos(psiE) * std::cos(phi) - std::sin(psiE) * std::cos(theta) * std::sin(phi));
    const double E0z = E0 * std::sin(psiE) * std::sin(theta);
    const double H0x = E0 / eta0 * (std::sin(psiE) * std::sin(phi) + std::cos(psiE) * std::cos(theta) * std::cos(phi));
    const double H0y = E0 / eta0 * (-std::sin(psiE) * std::cos(phi) + std::cos(psiE) * std::cos(theta) * std::sin(phi));
    const double H0z = -E0 / eta0 * std::cos(psiE) * std::sin(theta);
    const double kx = std::sin(theta) * std::cos(phi), ky = std::sin(theta) * std::sin(phi), kz = std::cos(theta);
    const double kabs = std::sqrt(kx * kx + ky * ky + kz * kz);
    // TF_SF.cpp:124-138
    const int ci[2] = {i1, i2}, cj[2] = {j1, j2}, ck[2] = {k1, k2};
    double l0 = 0;
    for (int q = 0; q < 8; q++) {
        const int a = (q >> 2) & 1, b = (q >> 1) & 1, c = q & 1;
        const double v = (kx * (ci[a] - i0) * m.dx + ky * (cj[b] - j0) * m.dy + kz * (ck[c] - k0) * m.dz) / kC;
        if (q == 0 || l0 > v) l0 = v;
    }
    // TF_SF.cpp:155-166
    auto at = [&](double x, double y, double z, double &delay, double &prof) {
        const double r2 = x * x + y * y + z * z;
        double rdk = (x * kx + y * ky + z * kz) / kabs;
        const double rho2 = r2 - rdk * rdk;
        rdk = rdk - z0;
        const double wz = w0 * std::sqrt(1 + rdk * rdk / (zR * zR));
        prof = std::exp(-rho2 / wz / wz);
        delay = ((x * kx + y * ky + z * kz) / kC - l0) / m.dt;
    };
    const double CE[3] = {m.dt / (kEps0 * m.dx), m.dt / (kEps0 * m.dy), m.dt / (kEps0 * m.dz)};
    const double CH[3] = {m.dt / (kMu0 * m.dx), m.dt / (kMu0 * m.dy), m.dt / (kMu0 * m.dz)};
    // per face axis: components corrected by record parts 1 and 2, their C*amplitude and low-face sign
    const int compE[3][2] = {{1, 2}, {0, 2}, {0, 1}};  // x: Ey Ez, y: Ex Ez, z: Ex Ey
    const double ampE[3][2] = {{CE[0] * H0z, CE[0] * H0y}, {CE[1] * H0z, CE[1] * H0x}, {CE[2] * H0y, CE[2] * H0x}};
    const int compH[3][2] = {{5, 4}, {5, 3}, {4, 3}};  // x: Hz Hy, y: Hz Hx, z: Hy Hx
    const double ampH[3][2] = {{CH[0] * E0y, CH[0] * E0z}, {CH[1] * E0x, CH[1] * E0z}, {CH[2] * E0x, CH[2] * E0y}};
    const double sgn[3][2] = {{+1, -1}, {-1, +1}, {+1, -1}};
    const int far = (m.flags & CFD_FIX_TFSF_FARFACE) ? 0 : 1;

    // targets in order of first appearance; terms kept flat with their
    // target's number, grouped per target (stable) once every face is done.
    // Face (a, e) corrects cells whose coordinate a is that face's plane (E:
    // the box face; H: one cell outside the low face, `far` past the high
    // one), each (cell, component) once.  A target can therefore come from a
    // second face only if a second coordinate lies on its axis's planes: only
    // those targets are looked up (key: cell * 8 + component); every other is
    // new where it appears.
    std::unordered_map<unsigned long long, int> eidx, hidx;
    std::vector<int> eown, hown;
    std::vector<TfsfTerm> eterms, hterms;
    {
        const size_t nrec = 2 * ((size_t)(i2 - i1 + 1) * (k2 - k1 + 1) + (size_t)(j2 - j1 + 1) * (k2 - k1 + 1) +
                                 (size_t)(i2 - i1 + 1) * (j2 - j1 + 1));
        m.tfE.reserve(2 * nrec); m.tfH.reserve(2 * nrec);
        eown.reserve(2 * nrec); hown.reserve(2 * nrec);
        eterms.reserve(2 * nrec); hterms.reserve(2 * nrec);
        m.tf_face_terms.reserve(4 * nrec);
    }
    const int planeE[3][2] = {{i1, i2}, {j1, j2}, {k1, k2}};
    const int planeH[3][2] = {{i1 - 1, i2 + far}, {j1 - 1, j2 + far}, {k1 - 1, k2 + far}};
    auto add = [&](bool h, int comp, int i, int j, int k, double coef, double delay, double prof) {
        auto &tg = h ? m.tfH : m.tfE;
        const int(&P)[3][2] = h ? planeH : planeE;
        const int on = (i == P[0][0] || i == P[0][1]) + (j == P[1][0] || j == P[1][1]) + (k == P[2][0] || k == P[2][1]);
        int t = (int)tg.size();
        if (on >= 2) {
            const unsigned long long key = (unsigned long long)m.flat(i, j, k) * 8u + (unsigned)comp;
            const auto ins = (h ? hidx : eidx).emplace(key, t);
            t = ins.first->second;
            if (ins.second) tg.push_back({comp, i, j, k, 0, 0});
        } else
            tg.push_back({comp, i, j, k, 0, 0});
        (h ? hown : eown).push_back(t);
        (h ? hterms : eterms).push_back({coef, delay, prof});
    };
    // a face alone: its records' corrections, each target once; the z1
    // variants leave out the transverse profile (TF_SF.cpp:1126-1143 and
    // :1375-1392 comment the factor out), the other faces keep it
    auto add_face = [&](bool h, int face, int comp, int i, int j, int k, double coef, double delay, double prof) {
        m.tfFace[face][h].push_back({comp, i, j, k, (int)m.tf_face_terms.size(), 1});
        m.tf_face_terms.push_back({coef, delay, face == 4 ? 1.0 : prof});
    };
    const int order[3] = {1, 0, 2};  // y faces, x faces, z faces
    for (int oa = 0; oa < 3; oa++) {
        const int a = order[oa];
        if ((a == 0 && m.pbx) || (a == 1 && m.pby)) continue;
        for (int e = 0; e < 2; e++) {
            const double flip = e ? -1.0 : 1.0;
            const double sh = e ? 0.5 : -0.5;
            const int c0 = a == 0 ? (e ? i2 : i1) : a == 1 ? (e ? j2 : j1) : (e ? k2 : k1);
            // record list of this face: main loop, then edge A (part 1), edge B (part 2)
            struct Rec { int i, j, k, part; };
            std::vector<Rec> recs;
            if (a == 1) {
                for (int i = i1; i < i2; ++i) for (int k = k1; k < k2; ++k) recs.push_back({i, c0, k, 3});
                for (int i = i1; i < i2; ++i) recs.push_back({i, c0, k2, 1});
                if (!m.pbx) for (int k = k1; k < k2; ++k) recs.push_back({i2, c0, k, 2});
            } else if (a == 0) {
                for (int j = j1; j < j2; ++j) for (int k = k1; k < k2; ++k) recs.push_back({c0, j, k, 3});
                for (int j = j1; j < j2; ++j) recs.push_back({c0, j, k2, 1});
                for (int k = k1; k < k2; ++k) recs.push_back({c0, j2, k, 2});
            } else {
                for (int i = i1; i < i2; ++i) for (int j = j1; j < j2; ++j) recs.push_back({i, j, c0, 3});
                if (!m.pby) for (int i = i1; i < i2; ++i) recs.push_back({i, j2, c0, 1});
                if (!m.pbx) for (int j = j1; j < j2; ++j) recs.push_back({i2, j, c0, 2});
            }
            // incident values of each record's component pairs: the H one (for
            // the E correction) half a cell across the face, the E one on it;
            // with CFD_FIX_TFSF_YEE both also half a cell along the E
            // component's own axis, their Yee position (the reference: the
            // cell corner).  Independent per record: computed on host threads.
            struct Inc { double dE, pE, dH, pH; };
            std::vector<Inc> inc(2 * recs.size());
            host_parallel([&](int t, int nth) {
                const size_t r0 = recs.size() * t / nth, r1 = recs.size() * (t + 1) / nth;
                for (size_t q = r0; q < r1; q++) {
                    const Rec &r = recs[q];
                    for (int c = 0; c < 2; c++) {
                        if (!(r.part & (1 << c))) continue;
                        double o[3] = {0, 0, 0};
                        if (m.flags & CFD_FIX_TFSF_YEE) o[compE[a][c]] = 0.5;
                        const double x = m.dx * (r.i - i0 + o[0]), y = m.dy * (r.j - j0 + o[1]), z = m.dz * (r.k - k0 + o[2]);
                        double xe = x, ye = y, ze = z;
                        if (a == 0) xe = m.dx * (r.i - i0 + sh);
                        if (a == 1) ye = m.dy * (r.j - j0 + sh);
                        if (a == 2) ze = m.dz * (r.k - k0 + sh);
                        Inc &v = inc[2 * q + c];
                        at(xe, ye, ze, v.dE, v.pE);
                        at(x, y, z, v.dH, v.pH);
                    }
                }
            });
            for (size_t q = 0; q < recs.size(); q++) {
                const Rec &r = recs[q];
                int hi = r.i, hj = r.j, hk = r.k;
                const int off = e ? far : -1;
                if (a == 0) hi += off; else if (a == 1) hj += off; else hk += off;
                for (int c = 0; c < 2; c++) {
                    if (!(r.part & (1 << c))) continue;
                    const double dE = inc[2 * q + c].dE, pE = inc[2 * q + c].pE;
                    const double dH = inc[2 * q + c].dH, pH = inc[2 * q + c].pH;
                    const double s = sgn[a][c] * flip;
                    const double cE = s < 0 ? -ampE[a][c] : ampE[a][c], cH = s < 0 ? -ampH[a][c] : ampH[a][c];
                    add(false, compE[a][c], r.i, r.j, r.k, cE, dE, pE);
                    const double ph = ((m.flags & CFD_REF_ZEROHEAP) && a == 0 && r.part != 3) ? 0.0 : pH;
                    add(true, compH[a][c], hi, hj, hk, cH, dH, ph);
                    add_face(false, 2 * a + e, compE[a][c], r.i, r.j, r.k, cE, dE, pE);
                    add_face(true, 2 * a + e, compH[a][c], hi, hj, hk, cH, dH, ph);
                }
            }
        }
    }
    m.tf_terms.resize(eterms.size() + hterms.size());
    size_t base = 0;
    for (int h = 0; h < 2; h++) {
        auto &tg = h ? m.tfH : m.tfE;
        const auto &terms = h ? hterms : eterms;
        const auto &own = h ? hown : eown;
        for (int t : own) tg[t].count++;
        size_t o = base;
        for (auto &g : tg) {
            g.first = (int)o;
            o += g.count;
            g.count = 0;
        }
        for (size_t q = 0; q < terms.size(); q++) {
            TfsfTarget &g = tg[own[q]];
            m.tf_terms[g.first + g.count++] = terms[q];
        }
        base = o;
    }
    if (nt_req <= 0) nt_req = Nt;
    tfsf_waveforms(m, lambda, dlambda, cw, std::max(Nt, nt_req));
    if (m.nt <= 0) m.nt = Nt;
    m.has_tfsf = true;
}

// 3-D Drude-Lorentz (Drude_Lorentz.cpp:30-108 with F9, F10 and F12 decisions)
void load_dl(Model &m, const char *path) {
    LineReader L(path);
    L.read("number of Drude poles", m.nD);
    L.read("number of Lorentz poles", m.nL);
    if (m.nD < 0 || m.nD > 8 || m.nL < 0 || m.nL > 8) throw Error(std::string(path) + ": 0..8 poles of each kind");
    L.read("eps_inf", m.dl_epsInf);
    const double dt = m.dt, sigma = 0.0;
    double t1 = 0, t2 = 0;
    for (int q = 0; q < m.nD; q++) {
        double w, g;
        L.read("Drude pole", w, g);
        m.dlKap.push_back((1.0 - g * dt / 2.0) / (1.0 + g * dt / 2.0));
        m.dlBeta.push_back((w * w * kEps0 * dt / 2.0) / (1.0 + g * dt / 2.0));
        t1 = t1 + m.dlBeta.back();
    }
    for (int q = 0; q < m.nL; q++) {
        double w, dl, de;
        L.read("Lorentz pole", w, dl, de);
        m.dlAlpha.push_back((2.0 - w * w * dt * dt) / (1.0 + dl * dt));
        m.dlXi.push_back((dl * dt - 1.0) / (1.0 + dl * dt));
        m.dlGam.push_back((kEps0 * de * w * w * dt * dt) / (1.0 + dl * dt));
        t2 = t2 + m.dlGam.back();
    }
    const double den = 2.0 * kEps0 * m.dl_epsInf + sigma * dt + dt * t1 + t2 / 2.0;
    m.dlC1 = (2.0 * kEps0 * m.dl_epsInf - sigma * dt - dt * t1) / den;
    m.dlC2 = t2 / 2.0 / den;
    m.dlC3 = 2.0 * dt / den;
    for (int q = 0; q < m.nD; q++) m.dlC4.push_back((1.0 + m.dlKap[q]) * dt / den);
    for (int q = 0; q < m.nL; q++) {
        m.dlC5.push_back((1.0 + m.dlAlpha[q]) * dt / den);
        m.dlC6.push_back(m.dlXi[q] * dt / den);
    }
    m.has_dl = true;
}

// Four-level gain, 3-D single-electron model with pump rate; file layout of
// FourLevelLaser.cpp:37-93 plus line 15 = pump rate (F14)
void load_gain(Model &m, const char *path) {
    LineReader L(path);
    double v[15] = {0};
    for (int q = 0; q < 15; q++) {
        std::istringstream is = L.next();
        if (q < 4) continue;
        if (!(is >> v[q])) L.bad("gain parameter");
    }
    m.g_lambda = v[5]; m.g_gamma = v[6]; m.g_t32 = v[9]; m.g_t21 = v[10]; m.g_t10 = v[11]; m.g_Ntot = v[12]; m.g_Wp = v[14];
    if (m.g_t32 <= 0 || m.g_t21 <= 0 || m.g_t10 <= 0 || m.g_lambda <= 0) throw Error(std::string(path) + ": lifetimes and wavelength must be positive");
    const double dt = m.dt;
    m.g_omega = 2.0 * kPi * kC / m.g_lambda;
    m.g_kappa = 6.0 * kPi * kEps0 * kC * kC * kC / (m.g_omega * m.g_omega * m.g_t21);
    // FourLevelLaser.cpp:100-102
    m.gCPa1 = (2.0 - m.g_omega * m.g_omega * dt * dt) / (1.0 + m.g_gamma * dt / 2.0);
    m.gCPa2 = (m.g_gamma * dt / 2.0 - 1.0) / (1.0 + m.g_gamma * dt / 2.0);
    m.gCPa3 = dt * dt * m.g_kappa / (1.0 + m.g_gamma * dt / 2.0);
    m.has_gain = true;
}

// Sources.cpp:29-76
void load_sources(Model &m, const cfd_config &cfg) {
    if (cfg.n_src <= 0) return;
    if (!cfg.source) throw Error("point sources need a source file");
    LineReader L(cfg.source);
    double lambda, fwhm, delay;
    L.read("lambda", lambda);
    L.read("FWHM", fwhm);
    L.read("delay", delay);
    const double dt = m.dt, nu = kC / lambda, T = 1.0 / nu;
    const int nT = (int)(T / dt), nw = 2 * nT, nd = 4 * nw;
    const double pinudt = 2.0 * kPi * nu * dt;
    m.src_wave.assign((size_t)cfg.n_src * (m.nt + 1), 0.0);
    for (int q = 0; q < cfg.n_src; q++) {
        const int *s = cfg.src_ijkc + 4 * q;
        if (s[0] < 0 || s[1] < 0 || s[2] < 0 || s[0] > m.Nx || s[1] > m.Ny || s[2] > m.Nz || s[3] < 0 || s[3] > 2)
            throw Error("point source " + std::to_string(q) + " lies outside the grid");
        m.src.insert(m.src.end(), s, s + 4);
        const double A = cfg.src_amp ? cfg.src_amp[q] : 1.0, ph = cfg.src_phase ? cfg.src_phase[q] : 0.0;
        for (int n = 0; n <= m.nt; n++) {
            double w;
            if (m.flags & CFD_SRC_REF_PULSE) {
                const double retard = n - 0.0 / kC / dt;
                w = (A * std::exp(-(double)(retard - nd) * (retard - nd) / (nw * nw)) * std::sin(pinudt * (retard - nd)));
            } else {
                const double t = (double)n * dt - delay;
                w = A * std::exp(-4.0 * std::log(2.0) * t * t / (fwhm * fwhm)) * std::sin(2.0 * kPi * nu * t + ph);
            }
            m.src_wave[(size_t)q * (m.nt + 1) + n] = w;
        }
    }
}

// Ports.cpp:28-86 (sphere) and :201-306 (plane); the reference's extra unset
// record (F11) is not created
Port load_port(const Model &m, const char *path, int kind) {
    LineReader L(path);
    Port P;
    P.kind = kind;
    const double dx = m.dx, dy = m.dy, dz = m.dz;
    if (kind == CFD_PORT_SPHERE) {
        double r, xc, yc, zc, t1, t2, p1, p2;
        L.read("radius", r);
        L.read("centre", xc, yc, zc);
        L.read("theta range", t1, t2);
        L.read("phi range", p1, p2);
        t1 = t1 * kPi / 180.0; t2 = t2 * kPi / 180.0; p1 = p1 * kPi / 180.0; p2 = p2 * kPi / 180.0;
        const double dphi = dx / r, dtheta = dz / r;
        for (double phi = p1; phi <= p2; phi = phi + dphi)
            for (double th = t1; th <= t2; th = th + dtheta) {
                P.ijk.push_back((int)((xc + r * std::sin(th) * std::cos(phi)) / dx));
                P.ijk.push_back((int)((yc + r * std::sin(th) * std::sin(phi)) / dy));
                P.ijk.push_back((int)((zc + r * std::cos(th)) / dz));
            }
    } else {
        char nd = 0;
        double pos, w1, w2, h1, h2;
        L.read("normal direction", nd);
        L.read("position", pos);
        L.read("w1 w2", w1, w2);
        L.read("h1 h2", h1, h2);
        int a1, a2, b1, b2, c1;
        if (nd == 'x') { c1 = (int)(pos / dx); a1 = (int)(w1 / dy); a2 = (int)(w2 / dy); b1 = (int)(h1 / dz); b2 = (int)(h2 / dz); }
        else if (nd == 'y') { c1 = (int)(pos / dy); a1 = (int)(w1 / dx); a2 = (int)(w2 / dx); b1 = (int)(h1 / dz); b2 = (int)(h2 / dz); }
        else if (nd == 'z') { c1 = (int)(pos / dz); a1 = (int)(w1 / dx); a2 = (int)(w2 / dx); b1 = (int)(h1 / dy); b2 = (int)(h2 / dy); }
        else throw Error(std::string(path) + ": normal direction must be x, y or z");
        if (a2 < a1 || b2 < b1) throw Error(std::string(path) + ": empty port");
        for (int b = b1; b <= b2; b++)
            for (int a = a1; a <= a2; a++) {
                if (nd == 'x') P.ijk.insert(P.ijk.end(), {c1, a, b});
                else if (nd == 'y') P.ijk.insert(P.ijk.end(), {a, c1, b});
                else P.ijk.insert(P.ijk.end(), {a, b, c1});
            }
    }
    for (size_t q = 0; q < P.ijk.size(); q += 3) {
        if (m.pbx && P.ijk[q] == m.Nx) P.ijk[q] = 0;
        if (m.pby && P.ijk[q + 1] == m.Ny) P.ijk[q + 1] = 0;
        if (P.ijk[q] < 0 || P.ijk[q + 1] < 0 || P.ijk[q + 2] < 0 || P.ijk[q] > m.Nx || P.ijk[q + 1] > m.Ny || P.ijk[q + 2] > m.Nz)
            throw Error(std::string(path) + ": port point " + std::to_string(q / 3) + " lies outside the grid");
    }
    return P;
}

// dispersive and gain cells in (i, j, k) order: ranges of i scanned on host
// threads, their lists joined in i order
void cell_lists(Model &m) {
    if (!m.has_dl && !m.has_gain) return;
    const int ie0 = m.pbx ? 0 : 1, je0 = m.pby ? 0 : 1;
    std::vector<std::vector<long long>> dl(8), gain(8);
    host_parallel([&](int t, int nth) {
        const int a = ie0 + (int)((long long)(m.Nx - ie0) * t / nth), b = ie0 + (int)((long long)(m.Nx - ie0) * (t + 1) / nth);
        for (int i = a; i < b; i++)
            for (int j = je0; j < m.Ny; j++) {
                const size_t p0 = m.flat(i, j, 0);
                for (int k = 1; k < m.Nz; k++) {
                    const int id = m.id[p0 + k];
                    if (m.has_dl && id == 2) dl[t].push_back((long long)(p0 + k));
                    if (m.has_gain && id == 5) gain[t].push_back((long long)(p0 + k));
                }
            }
    });
    for (int t = 0; t < 8; t++) {
        m.dl_cells.insert(m.dl_cells.end(), dl[t].begin(), dl[t].end());
        m.gain_cells.insert(m.gain_cells.end(), gain[t].begin(), gain[t].end());
    }
}

}  // namespace

void tfsf_tables(const char *path, double dt, int nt, std::vector<double> &einc, std::vector<double> &hinc) {
    tfsf_file_tables(path, dt, nt, einc, hinc);
}

void check_config(const cfd_config &cfg) {
    if (!cfg.material || !cfg.structure) throw Error("material and structure files are required");
    if (cfg.precision != CFD_F32 && cfg.precision != CFD_F64) throw Error("precision must be 32 or 64");
    if (cfg.nt < 0) throw Error("nt must not be negative");
    if (cfg.n_ports < 0 || cfg.n_src < 0 || cfg.n_freq < 0) throw Error("negative port, source or frequency count");
    if (cfg.n_ports > 0) {
        if (!cfg.port_files) throw Error("n_ports > 0 without port_files");
        for (int q = 0; q < cfg.n_ports; q++) {
            if (!cfg.port_files[q]) throw Error("port " + std::to_string(q) + ": no file");
            if (cfg.port_kinds && cfg.port_kinds[q] != CFD_PORT_PLANE && cfg.port_kinds[q] != CFD_PORT_SPHERE)
                throw Error("port " + std::to_string(q) + ": kind must be CFD_PORT_PLANE or CFD_PORT_SPHERE");
        }
    }
    if (cfg.n_src > 0 && !cfg.src_ijkc) throw Error("n_src > 0 without src_ijkc");
    if (cfg.n_freq > 0 && !cfg.freqs) throw Error("n_freq > 0 without freqs");
    if (cfg.periodic & ~3) throw Error("periodic takes bits 1 (x) and 2 (y) only");
}

Model build_model(const cfd_config &cfg) {
    check_config(cfg);
    Model m;
    m.flags = cfg.flags;
    m.pbx = cfg.periodic & 1;
    m.pby = (cfg.periodic >> 1) & 1;
    load_structure(m, cfg.material, cfg.structure);
    if (cfg.pml) load_cpml(m, cfg.pml);
    m.nt = cfg.nt;
    if (cfg.tfsf) load_tfsf(m, cfg.tfsf, cfg.nt);
    if (m.nt <= 0) throw Error("number of time steps not given (nt <= 0 and no TF/SF file)");
    if (cfg.dl) load_dl(m, cfg.dl);
    if (cfg.gain) load_gain(m, cfg.gain);
    cell_lists(m);
    load_sources(m, cfg);
    m.dft_stride = cfg.dft_stride > 0 ? cfg.dft_stride : 1;
    m.record = cfg.record_ports != 0;
    if (cfg.n_freq > 0) {
        m.freqs.assign(cfg.freqs, cfg.freqs + cfg.n_freq);
        const int nf = cfg.n_freq;
        m.twE.resize((size_t)nf * m.nt * 2);
        m.twH.resize((size_t)nf * m.nt * 2);
        host_parallel([&](int t, int nth) {  // frequencies shared among host threads
            for (int f = t; f < nf; f += nth)
                for (int n = 0; n < m.nt; n++) {
                    const double tE = ((double)n + 1.0) * m.dt, tH = ((double)n + 0.5) * m.dt;
                    const double aE = 2.0 * kPi * m.freqs[f] * tE, aH = 2.0 * kPi * m.freqs[f] * tH;
                    const size_t o = ((size_t)f * m.nt + n) * 2;
                    m.twE[o] = std::cos(aE); m.twE[o + 1] = std::sin(aE);
                    m.twH[o] = std::cos(aH); m.twH[o + 1] = std::sin(aH);
                }
        });
    }
    for (int q = 0; q < cfg.n_ports; q++) m.ports.push_back(load_port(m, cfg.port_files[q], cfg.port_kinds ? cfg.port_kinds[q] : 0));
    return m;
}

}  // namespace cfd
