/*
 * cpsfdtd.hpp -- C++ facade with the reference's class names and call
 * signatures, over the C ABI of libcpsfdtd.so (cpsfdtd.h).
 *
 * A driver written against the reference's classes (main.cpp:115-343) keeps
 * its shape:
 *
 *     cpsfdtd::Structure S("material.txt");          // Structure.h:35
 *     S.generate_mesh("structure.txt");              // Structure.h:37
 *     S.InitializeBasicFDTDCoefficients();           // Structure.h:38
 *     cpsfdtd::Field F(S, "3D");                     // Fields.h:26
 *     cpsfdtd::CPML pml(S, "CFS_PML.txt");           // CPML.h:25
 *     pml.InitializeCPML(S);
 *     cpsfdtd::TF_SF tfsf(S, "TF_SF.txt");           // TF_SF.h:32
 *     cpsfdtd::PlainPort port(S, "port.txt");        // Ports.h:75
 *     for (int n = 0; n < nt; ++n) {                 // main.cpp:267-343
 *         cpsfdtd::BasicHFieldUpdate(F, S);
 *         tfsf.UpdateHField(n, F, S);
 *         pml.UpdateCPMLHFields(F);
 *         cpsfdtd::BasicEFieldUpdate(F, S);
 *         tfsf.UpdateEField(n, F, S);
 *         pml.UpdateCPMLEFields(F);
 *         port.Update_Fields_on_Port(F);
 *     }
 *
 * Differences a caller sees:
 *  - Objects register their input file; the device simulation is created by
 *    the first update call, from every object constructed before it.  A
 *    module constructed after that throws.
 *  - Time steps are counted at H half-steps: every BasicHFieldUpdate after
 *    the first one starts a new step (the reference's loop order).
 *  - F.Ex[i][j][k] reads and writes a host mirror of the device field,
 *    refreshed after device updates and written back before the next one
 *    (the reference's public double*** arrays, Fields.h:29-34).
 *  - Errors throw cpsfdtd::Error with the library's message (the reference
 *    checks nothing).
 *  - Ports have no trailing unset point (the reference's Num_of_Port_Points
 *    counts one point it never sets, Ports.cpp:74,238,259,280); their text
 *    exports are otherwise the reference's, byte for byte.  Point lists,
 *    Export_Port and Export_ID work before the first update without
 *    starting the simulation.
 *  - OutImage writes its PNG files but opens no display window.
 *  - Options the reference hard-codes or lacks (precision, number of steps,
 *    defect switches, periodic axes, point sources, port DFT, keeping every
 *    port sample) are set on the Structure before the first update.
 */
#ifndef CPSFDTD_HPP
#define CPSFDTD_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cpsfdtd.h"

namespace cpsfdtd {

class Error : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class Field;

namespace detail {

inline void check(int rc, const char *what)
{
    if (rc < 0) throw Error(std::string(what) + ": " + cfd_last_error());
}

// a parameter file read line by line, one istringstream per line, as the
// reference's constructors read theirs (getline + istringstream)
class Lines {
public:
    Lines(const char *who, const char *fileName)
    {
        std::ifstream in(fileName);
        if (!in) throw Error(std::string(who) + ": cannot open " + fileName);
        std::string l;
        while (std::getline(in, l)) lines_.push_back(l);
    }
    std::istringstream next() { return std::istringstream(n_ < lines_.size() ? lines_[n_++] : std::string()); }
private:
    std::vector<std::string> lines_;
    size_t n_ = 0;
};

// the reference's Constants.h values
constexpr double kPi = 3.1415926535897932384626433832795, kC = 2.99792458E8;

// One run: the files the objects registered, the options, and the device
// simulation once it exists.
struct Context {
    std::string material, structure, pml, tfsf, dl, gain, source;
    std::vector<std::string> port_files;
    std::vector<int> port_kinds;
    std::vector<int> src_ijkc;
    std::vector<double> src_amp, src_phase, freqs;
    int dft_stride = 1, record_ports = 0, nt = 0, periodic = 0, precision = CFD_F64;
    unsigned flags = 0;
    cfd_sim *sim = nullptr;
    Field *field = nullptr;
    long generation = 0;   // bumped by every call that may change device fields
    int step = 0;          // steps begun so far minus one (index of the current step)
    bool h_started = false;
    int sampled_step = -1; // step whose port samples were taken

    Context() = default;
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    ~Context()
    {
        if (sim) cfd_destroy(sim);
    }
    void require_unstarted(const char *what) const
    {
        if (sim) throw Error(std::string(what) + " registered after the simulation started");
    }
    cfd_sim *get()
    {
        if (sim) return sim;
        std::vector<const char *> pf;
        for (const std::string &p : port_files) pf.push_back(p.c_str());
        cfd_config c{};
        auto opt = [](const std::string &s) { return s.empty() ? nullptr : s.c_str(); };
        c.material = material.c_str();
        c.structure = structure.c_str();
        c.pml = opt(pml);
        c.tfsf = opt(tfsf);
        c.dl = opt(dl);
        c.gain = opt(gain);
        c.source = opt(source);
        c.n_ports = (int)pf.size();
        c.port_files = pf.empty() ? nullptr : pf.data();
        c.port_kinds = port_kinds.empty() ? nullptr : port_kinds.data();
        c.n_src = (int)src_amp.size();
        c.src_ijkc = src_ijkc.empty() ? nullptr : src_ijkc.data();
        c.src_amp = src_amp.empty() ? nullptr : src_amp.data();
        c.src_phase = src_phase.empty() ? nullptr : src_phase.data();
        c.nt = nt;
        c.n_freq = (int)freqs.size();
        c.freqs = freqs.empty() ? nullptr : freqs.data();
        c.dft_stride = dft_stride;
        c.record_ports = record_ports;
        c.periodic = periodic;
        c.flags = flags;
        c.precision = precision;
        check(cfd_create(&c, &sim), "cfd_create");
        return sim;
    }
    template <class Fn>
    void device(Fn &&fn, const char *what);  // flush host writes, run, bump generation
};

}  // namespace detail

// ---------------------------------------------------------------------------
// Structure (Structure.h:33-60): media, mesh, dimensions, ID map
// ---------------------------------------------------------------------------
class Structure {
public:
    explicit Structure(const char *fileName) : ctx(std::make_shared<detail::Context>()) { ctx->material = fileName; }
    Structure(const Structure &) = delete;
    Structure &operator=(const Structure &) = delete;

    void generate_mesh(const char *fileName)
    {
        ctx->require_unstarted("Structure::generate_mesh");
        ctx->structure = fileName;
        cfd_config c{};
        c.material = ctx->material.c_str();
        c.structure = ctx->structure.c_str();
        c.nt = 1;
        c.precision = CFD_F64;
        cfd_model *m = nullptr;
        if (cfd_model_build(&c, &m) < 0) throw Error(std::string("generate_mesh: ") + cfd_model_last_error());
        int n[4];
        double d[4];
        cfd_model_dims(m, n, d);
        Nx = n[0]; Ny = n[1]; Nz = n[2];
        dx = d[0]; dy = d[1]; dz = d[2]; dt = d[3];
        id_.assign((size_t)(Nx + 1) * (Ny + 1) * (Nz + 1), 0);
        cfd_model_id(m, id_.data());
        cfd_model_free(m);
    }
    void InitializeBasicFDTDCoefficients() {}  // the coefficients are built with the simulation

    // ID[i][j][k] (Structure.h:43), read-only
    struct IdRow2 {
        const Structure *s; int i, j;
        short operator[](int k) const { return s->id_[((size_t)i * (s->Ny + 1) + j) * (s->Nz + 1) + k]; }
    };
    struct IdRow1 {
        const Structure *s; int i;
        IdRow2 operator[](int j) const { return IdRow2{s, i, j}; }
    };
    struct IdMap {
        const Structure *s;
        IdRow1 operator[](int i) const { return IdRow1{s, i}; }
    };
    IdMap ID{this};

    double dx = 0, dy = 0, dz = 0, dt = 0;
    int Nx = 0, Ny = 0, Nz = 0;

    // options the reference hard-codes or lacks; set before the first update
    void set_precision(int bits) { ctx->require_unstarted("precision"); ctx->precision = bits; }
    void set_steps(int nt) { ctx->require_unstarted("steps"); ctx->nt = nt; }
    void set_flags(unsigned flags) { ctx->require_unstarted("flags"); ctx->flags = flags; }
    void set_periodic(int axes) { ctx->require_unstarted("periodic axes"); ctx->periodic = axes; }
    void set_source_file(const char *fileName) { ctx->require_unstarted("source"); ctx->source = fileName; }
    void add_point_source(int i, int j, int k, int comp, double amp, double phase = 0.0)
    {
        ctx->require_unstarted("point source");
        ctx->src_ijkc.insert(ctx->src_ijkc.end(), {i, j, k, comp});
        ctx->src_amp.push_back(amp);
        ctx->src_phase.push_back(phase);
    }
    void set_port_dft(const std::vector<double> &freqs_hz, int stride)
    {
        ctx->require_unstarted("port DFT");
        ctx->freqs = freqs_hz;
        ctx->dft_stride = stride;
    }
    void record_port_samples(bool on) { ctx->require_unstarted("port recording"); ctx->record_ports = on ? 1 : 0; }

    cfd_sim *simulation() { return ctx->get(); }  // the device simulation, created on first use
    int step() const { return ctx->step; }

    std::shared_ptr<detail::Context> ctx;

private:
    std::vector<int16_t> id_;
};

// ---------------------------------------------------------------------------
// Field (Fields.h:23-45): Ex..Hz with [i][j][k] access through host mirrors
// ---------------------------------------------------------------------------
class Field {
public:
    Field(Structure &sS, const char *smode = "3D") : ctx(sS.ctx), mode(smode ? smode : "")
    {
        if (ctx->field) throw Error("one Field per Structure");
        ctx->field = this;
        ncell = (size_t)(sS.Nx + 1) * (sS.Ny + 1) * (sS.Nz + 1);
        ny1 = sS.Ny + 1;
        nz1 = sS.Nz + 1;
        if (!ncell || !sS.Nx) throw Error("Field: generate_mesh has not run");
    }
    ~Field()
    {
        if (ctx->field == this) ctx->field = nullptr;
    }
    Field(const Field &) = delete;
    Field &operator=(const Field &) = delete;

    class Ref {
    public:
        operator double() const { return f->host(c)[p]; }
        Ref &operator=(double v)
        {
            f->host(c)[p] = v;
            f->dirty[c] = true;
            return *this;
        }
        Ref &operator+=(double v) { return *this = (double)*this + v; }
        Ref(Field *f_, int c_, size_t p_) : f(f_), c(c_), p(p_) {}
    private:
        Field *f; int c; size_t p;
    };
    struct Row2 {
        Field *f; int c; size_t base;
        Ref operator[](int k) const { return Ref(f, c, base + k); }
    };
    struct Row1 {
        Field *f; int c; int i;
        Row2 operator[](int j) const { return Row2{f, c, ((size_t)i * f->ny1 + j) * f->nz1}; }
    };
    struct Component {
        Field *f; int c;
        Row1 operator[](int i) const { return Row1{f, c, i}; }
    };
    Component Ex{this, 0}, Ey{this, 1}, Ez{this, 2}, Hx{this, 3}, Hy{this, 4}, Hz{this, 5};

    // whole component, [i][j][k] flat, k fastest
    const std::vector<double> &values(int comp) { return host(comp); }

    // write host changes back to the device (also done before every update)
    void flush()
    {
        if (!ctx->sim) return;
        for (int c = 0; c < 6; c++)
            if (dirty[c]) {
                detail::check(cfd_write_field(ctx->sim, c, mirror[c].data()), "cfd_write_field");
                dirty[c] = false;
            }
    }

private:
    std::vector<double> &host(int c)
    {
        std::vector<double> &v = mirror[c];
        if (v.size() != ncell) {
            v.assign(ncell, 0.0);
            valid[c] = -1;
        }
        if (ctx->sim && valid[c] != ctx->generation && !dirty[c]) {
            detail::check(cfd_read_field(ctx->sim, c, v.data()), "cfd_read_field");
            valid[c] = ctx->generation;
        }
        return v;
    }

    std::shared_ptr<detail::Context> ctx;
    std::string mode;
    size_t ncell = 0;
    int ny1 = 0, nz1 = 0;
    std::vector<double> mirror[6];
    long valid[6] = {-1, -1, -1, -1, -1, -1};
    bool dirty[6] = {false, false, false, false, false, false};
};

template <class Fn>
void detail::Context::device(Fn &&fn, const char *what)
{
    cfd_sim *s = get();
    if (field) field->flush();
    check(fn(s), what);
    generation++;
}

// ---------------------------------------------------------------------------
// BasicFDTD.h:23-24
// ---------------------------------------------------------------------------
inline void BasicHFieldUpdate(Field &, Structure &sS)
{
    detail::Context &c = *sS.ctx;
    if (c.h_started) {  // a new time step begins with its H half-step
        c.device([](cfd_sim *s) { return cfd_advance(s); }, "cfd_advance");
        c.step++;
    }
    c.h_started = true;
    c.device([](cfd_sim *s) { return cfd_basic_h(s); }, "BasicHFieldUpdate");
}
inline void BasicEFieldUpdate(Field &, Structure &sS)
{
    sS.ctx->device([](cfd_sim *s) { return cfd_basic_e(s); }, "BasicEFieldUpdate");
}

// ---------------------------------------------------------------------------
// CPML (CPML.h:22-30) and CFS_PML (CFS_PML.h:28-37): same physics
// ---------------------------------------------------------------------------
class CPML {
public:
    CPML(Structure &sS, const char *fileName) : ctx(sS.ctx)
    {
        ctx->require_unstarted("CPML");
        ctx->pml = fileName;
    }
    void InitializeCPML(Structure &) {}
    void UpdateCPMLEFields(Field &) { ctx->device([](cfd_sim *s) { return cfd_cpml_e(s); }, "UpdateCPMLEFields"); }
    void UpdateCPMLHFields(Field &) { ctx->device([](cfd_sim *s) { return cfd_cpml_h(s); }, "UpdateCPMLHFields"); }
private:
    std::shared_ptr<detail::Context> ctx;
};

class CFS_PML {
public:
    CFS_PML(Structure &sS, const char *fileName) : ctx(sS.ctx)
    {
        ctx->require_unstarted("CFS_PML");
        ctx->pml = fileName;
        // the public parameters as CFS_PML.cpp:30-71 reads them
        detail::Lines L("CFS_PML", fileName);
        L.next() >> Npmlx1 >> Npmlx2;
        L.next() >> Npmly1 >> Npmly2;
        L.next() >> Npmlz1 >> Npmlz2;
        L.next() >> p_order >> malp;
        L.next() >> kapmax >> sig_ratio;
        L.next() >> alpmax;
    }
    // CFS_PML.h:37-39.  sigmax is per axis here (0.8 (m + 1) / (d eta0) x
    // sig_ratio, CPML.cpp:361-519) and reported for none; alpmin is never read
    // by the reference.  Num_of_PML_points has no counterpart: psi lives in
    // per-slab arrays, not a point list.
    double p_order = 0, kapmax = 0, alpmin = 0, alpmax = 0, malp = 0, sig_ratio = 0, sigmax = 0;
    int Npmlx1 = 0, Npmly1 = 0, Npmlz1 = 0, Npmlx2 = 0, Npmly2 = 0, Npmlz2 = 0;
    void Initialize_PML(Structure &) {}
    void UpdatePMLEFields(Field &) { ctx->device([](cfd_sim *s) { return cfd_cpml_e(s); }, "UpdatePMLEFields"); }
    void UpdatePMLHFields(Field &) { ctx->device([](cfd_sim *s) { return cfd_cpml_h(s); }, "UpdatePMLHFields"); }
private:
    std::shared_ptr<detail::Context> ctx;
};

// ---------------------------------------------------------------------------
// TF_SF (TF_SF.h:30-58): face corrections at time step n
// ---------------------------------------------------------------------------
class TF_SF {
public:
    TF_SF(Structure &sS, const char *fileName) : ctx(sS.ctx)
    {
        ctx->require_unstarted("TF_SF");
        ctx->tfsf = fileName;
        // the reference's public parameters (TF_SF.h:54-66), read as
        // TF_SF.cpp:33-89 and 899-910 read them (constants of Constants.h)
        const double pi = 3.1415926535897932384626433832795, C = 2.99792458E8;
        std::ifstream in(fileName);
        if (!in) throw Error(std::string("TF_SF: cannot open ") + fileName);
        std::string line;
        auto next = [&]() {
            std::getline(in, line);
            return std::istringstream(line);
        };
        next() >> lambda;
        next() >> dlambda;
        {
            std::istringstream is = next();
            is >> w0 >> z0;
        }
        {
            std::istringstream is = next();
            is >> xs >> ys >> zs;
        }
        {
            std::istringstream is = next();
            is >> theta_inc >> phi_inc;
        }
        {
            std::istringstream is = next();
            is >> E0 >> psi_E;
        }
        {
            std::istringstream is = next();
            is >> nTFSFx1 >> nTFSFx2;
        }
        {
            std::istringstream is = next();
            is >> nTFSFy1 >> nTFSFy2;
        }
        {
            std::istringstream is = next();
            is >> nTFSFz1 >> nTFSFz2;
        }
        next() >> Nt;
        next() >> cwmode;
        theta_inc = theta_inc * pi / 180.0;
        phi_inc = phi_inc * pi / 180.0;
        psi_E = psi_E * pi / 180.0;
        k_num = 2.0 * pi / lambda;
        zR = pi * w0 * w0 / lambda;
        if (sS.dx > 0 && sS.dy > 0 && sS.dz > 0) {
            i0 = (int)(xs / sS.dx);
            j0 = (int)(ys / sS.dy);
            k0 = (int)(zs / sS.dz);  // left unset by the reference (F4); the engine uses zs/dz
        }
        kx = std::sin(theta_inc) * std::cos(phi_inc);
        ky = std::sin(theta_inc) * std::sin(phi_inc);
        kz = std::cos(theta_inc);
        kabs = std::sqrt(kx * kx + ky * ky + kz * kz);
        nu = C / lambda;
        T = 1.0 / nu;
        if (sS.dt > 0) nT = (int)(T / sS.dt);
        file = fileName;
        dt = sS.dt;
        if (dt > 0) tables(std::max(Nt, ctx->nt));
    }

    // the reference's public parameters; angles in radians.  z0 is line 3's
    // beam-waist position (the reference overwrites it with zs/dz, F4)
    double theta_inc = 0, phi_inc = 0, psi_E = 0, E0 = 0;
    int nTFSFx1 = 0, nTFSFx2 = 0, nTFSFy1 = 0, nTFSFy2 = 0, nTFSFz1 = 0, nTFSFz2 = 0;
    double kx = 0, ky = 0, kz = 0, kabs = 0;
    double lambda = 0, dlambda = 0, nu = 0, T = 0;
    int nT = 0, Nt = 0;
    double k_num = 0, w0 = 0, z0 = 0, zR = 0;
    double xs = 0, ys = 0, zs = 0;
    int i0 = 0, j0 = 0, k0 = 0;
    bool cwmode = false;
    void UpdateEField(int timestep, Field &, Structure &)
    {
        ctx->device([timestep](cfd_sim *s) { return cfd_tfsf_e(s, timestep); }, "TF_SF::UpdateEField");
    }
    void UpdateHField(int timestep, Field &, Structure &)
    {
        ctx->device([timestep](cfd_sim *s) { return cfd_tfsf_h(s, timestep); }, "TF_SF::UpdateHField");
    }
    // one face of the box alone (TF_SF.cpp:1090-1199, 1302-1410); the z1
    // forms leave the transverse profile out, as the reference's do
    void x1UpdateEField(int n, Field &, Structure &) { face(false, CFD_FACE_X1, n); }
    void x2UpdateEField(int n, Field &, Structure &) { face(false, CFD_FACE_X2, n); }
    void y1UpdateEField(int n, Field &, Structure &) { face(false, CFD_FACE_Y1, n); }
    void y2UpdateEField(int n, Field &, Structure &) { face(false, CFD_FACE_Y2, n); }
    void z1UpdateEField(int n, Field &, Structure &) { face(false, CFD_FACE_Z1, n); }
    void z2UpdateEField(int n, Field &, Structure &) { face(false, CFD_FACE_Z2, n); }
    void x1UpdateHField(int n, Field &, Structure &) { face(true, CFD_FACE_X1, n); }
    void x2UpdateHField(int n, Field &, Structure &) { face(true, CFD_FACE_X2, n); }
    void y1UpdateHField(int n, Field &, Structure &) { face(true, CFD_FACE_Y1, n); }
    void y2UpdateHField(int n, Field &, Structure &) { face(true, CFD_FACE_Y2, n); }
    void z1UpdateHField(int n, Field &, Structure &) { face(true, CFD_FACE_Z1, n); }
    void z2UpdateHField(int n, Field &, Structure &) { face(true, CFD_FACE_Z2, n); }

    // TF_SF.cpp:966-986: the incident E (H) wave at a retarded time index,
    // linear between table entries, 0 before the wave starts.  The tables
    // (Einc, Hinc, TF_SF.h:57) are computed on the host from the file and the
    // Structure's dt and grow past Nt on demand.
    double Einc_onPoint(double retarded) { return on_point(Einc, retarded); }
    double Hinc_onPoint(double retarded) { return on_point(Hinc, retarded); }
    std::vector<double> Einc, Hinc;
private:
    void face(bool h, int f, int n)
    {
        ctx->device([h, f, n](cfd_sim *s) { return h ? cfd_tfsf_face_h(s, f, n) : cfd_tfsf_face_e(s, f, n); },
                    h ? "TF_SF::*UpdateHField (one face)" : "TF_SF::*UpdateEField (one face)");
    }
    void tables(int upto)
    {
        const int len = cfd_tfsf_tables(file.c_str(), dt, upto, nullptr, nullptr);
        if (len < 0) throw Error(std::string("TF_SF tables: ") + cfd_model_last_error());
        Einc.assign((size_t)len, 0.0);
        Hinc.assign((size_t)len, 0.0);
        cfd_tfsf_tables(file.c_str(), dt, upto, Einc.data(), Hinc.data());
    }
    double on_point(std::vector<double> &tab, double retarded)
    {
        if (retarded <= 0) return 0.0;
        const int ntemp = (int)std::floor(retarded);
        const double resid = retarded - (double)ntemp;
        if ((size_t)ntemp + 1 >= tab.size()) {
            if (!(dt > 0)) throw Error("TF_SF::*inc_onPoint: construct TF_SF after Structure::generate_mesh");
            tables(std::max(ntemp + 1, ctx->nt));
        }
        return ((1.0 - resid) * tab[ntemp] + resid * tab[ntemp + 1]);
    }
    std::string file;
    double dt = 0;
    std::shared_ptr<detail::Context> ctx;
};

// ---------------------------------------------------------------------------
// Source (Sources.h:24-40): the point-source waveforms of a source file
// (line 1 lambda, 2 FWHM, 3 delay; Sources.cpp:29-57), host only
// ---------------------------------------------------------------------------
class Source {
public:
    Source(double deltat, const char *fileName) : dt(deltat)
    {
        detail::Lines L("Source", fileName);
        L.next() >> lambda;
        L.next() >> FWHM;
        L.next() >> delay;
        nu = detail::kC / lambda;
        T = 1.0 / nu;
        nT = (int)(T / dt);
        pinudt = 2.0 * detail::kPi * nu * dt;
        nw = 2 * nT;
        nd = 4 * nw;
        nTrans = nd + nw;
    }
    // Sources.cpp:71-76: a Gaussian-windowed sine centred nd steps after the
    // retarded time origin (ntimestep - location / (C dt)); FWHM and delay
    // are read but not used, as in the reference
    double Pulse(double Amplitude, double ntimestep, double location)
    {
        retard = ntimestep - location / detail::kC / dt;
        const double x = retard - nd;
        return Amplitude * std::exp(-x * x / (double)(nw * nw)) * std::sin(pinudt * x);
    }
    // Sources.cpp:60-69: the same window as a ramp up to retarded time nd,
    // then the steady sine
    double ContinuousWave(double Amplitude, double ntimestep, double location)
    {
        retard = ntimestep - location / detail::kC / dt;
        if (retard < nd) {
            const double x = retard - nd;
            return Amplitude * std::exp(-x * x / (double)(nw * nw)) * std::sin(pinudt * retard);
        }
        return Amplitude * std::sin(pinudt * retard);
    }
    double lambda = 0, nu = 0, T = 0, delay = 0, FWHM = 0;
private:
    double dt = 0, retard = 0, pinudt = 0;
    int nd = 0, nw = 0, nT = 0, nTrans = 0;
};

// ---------------------------------------------------------------------------
// Drude_Lorentz (Drude_Lorentz.h:28-45), 3-D, on ID 2 cells
// ---------------------------------------------------------------------------
// one Drude-Lorentz cell's state (DL_material, Drude_Lorentz.h:22-27, in
// 3-D): per component the Drude currents and the Lorentz currents with their
// previous values, one value per pole, and the previous field
struct DL_material {
    int ix = 0, jy = 0, kz = 0;
    std::vector<double> JDx, JDy, JDz, JLx, JLy, JLz, JLx_1, JLy_1, JLz_1;
    double Ex_1 = 0, Ey_1 = 0, Ez_1 = 0;
};

class Drude_Lorentz {
public:
    Drude_Lorentz(Structure &sS, const char *fileName) : ctx(sS.ctx)
    {
        ctx->require_unstarted("Drude_Lorentz");
        ctx->dl = fileName;
        // the public parameters (Drude_Lorentz.h:33-35): pole counts,
        // eps_inf, then one line per Drude pole (omega, gamma) and per
        // Lorentz pole (omega, delta, delta eps); sigma is 0 (F10)
        detail::Lines L("Drude_Lorentz", fileName);
        L.next() >> num_of_D_poles;
        L.next() >> num_of_L_poles;
        L.next() >> epsInf;
        for (int q = 0; q < num_of_D_poles; q++) {
            double w = 0, g = 0;
            L.next() >> w >> g;
            omega_D.push_back(w);
            gamma_D.push_back(g);
        }
        for (int q = 0; q < num_of_L_poles; q++) {
            double w = 0, d = 0, de = 0;
            L.next() >> w >> d >> de;
            omega_L.push_back(w);
            delta_L.push_back(d);
            delta_eps_L.push_back(de);
        }
    }
    void EFieldUpdate(Field &, Structure &) { ctx->device([](cfd_sim *s) { return cfd_dl_e(s); }, "Drude_Lorentz::EFieldUpdate"); }

    double sigma = 0, epsInf = 0;
    std::vector<double> omega_D, gamma_D, delta_eps_L, omega_L, delta_L;
    int num_of_D_poles = 0, num_of_L_poles = 0;

    // device state, read back (the simulation starts if it has not)
    int num_of_DLM_gridcells()
    {
        const int n = cfd_special_cells(ctx->get(), 0, nullptr);
        detail::check(n, "cfd_special_cells");
        return n;
    }
    std::vector<DL_material> DLM()
    {
        cfd_sim *s = ctx->get();
        const int n = num_of_DLM_gridcells();
        std::vector<int> ijk((size_t)n * 3);
        detail::check(cfd_special_cells(s, 0, ijk.data()), "cfd_special_cells");
        const int nD = num_of_D_poles, nL = num_of_L_poles, per = nD + 2 * nL + 1;
        std::vector<double> v((size_t)n * 3 * per);
        detail::check(cfd_dl_state(s, v.data()), "cfd_dl_state");
        std::vector<DL_material> out((size_t)n);
        for (int r = 0; r < n; r++) {
            DL_material &M = out[(size_t)r];
            M.ix = ijk[3 * r]; M.jy = ijk[3 * r + 1]; M.kz = ijk[3 * r + 2];
            std::vector<double> *JD[3] = {&M.JDx, &M.JDy, &M.JDz}, *JL[3] = {&M.JLx, &M.JLy, &M.JLz},
                                *JL1[3] = {&M.JLx_1, &M.JLy_1, &M.JLz_1};
            double *E1[3] = {&M.Ex_1, &M.Ey_1, &M.Ez_1};
            for (int c = 0; c < 3; c++) {
                const double *w = v.data() + ((size_t)r * 3 + c) * per;
                JD[c]->assign(w, w + nD);
                JL[c]->assign(w + nD, w + nD + nL);
                JL1[c]->assign(w + nD + nL, w + nD + 2 * nL);
                *E1[c] = w[nD + 2 * nL];
            }
        }
        return out;
    }
private:
    std::shared_ptr<detail::Context> ctx;
};

// ---------------------------------------------------------------------------
// FourLevelLaser (FourLevelLaser.h:20-33), 3-D, on ID 5 cells
// ---------------------------------------------------------------------------
class FourLevelLaser {
public:
    FourLevelLaser(Structure &sS, const char *fileName) : ctx(sS.ctx)
    {
        ctx->require_unstarted("FourLevelLaser");
        ctx->gain = fileName;
        // the public parameters as FourLevelLaser.cpp:37-97 reads them (line
        // 15, the pump rate, is this engine's; F14)
        detail::Lines L("FourLevelLaser", fileName);
        std::string d;
        L.next() >> d;
        std::snprintf(distri, sizeof distri, "%s", d.c_str());
        L.next() >> NumSources;
        L.next() >> i1 >> i2;
        L.next() >> j1 >> j2;
        L.next() >> lambda_b;
        L.next() >> lambda_a;
        L.next() >> gamma_a;
        L.next() >> gamma_b;
        L.next() >> t30;
        L.next() >> t32;
        L.next() >> t21;
        L.next() >> t10;
        L.next() >> N_0;
        L.next() >> N_1;
        L.next() >> pump_rate;
        omega_a = 2.0 * detail::kPi * detail::kC / lambda_a;
        omega_b = 2.0 * detail::kPi * detail::kC / lambda_b;
        idif = i2 - i1;
        jdif = j2 - j1;
    }
    // FourLevelLaser.h:31-38.  NumSources, i1..jdif and distri are the file's
    // values; the gain cells themselves are the ID 5 cells (location()).
    int NumSources = 0;
    int i1 = 0, i2 = 0, j1 = 0, j2 = 0, idif = 0, jdif = 0;
    double N_0 = 0, N_1 = 0;
    double gamma_a = 0, gamma_b = 0, lambda_a = 0, lambda_b = 0, omega_a = 0, omega_b = 0;
    double t30 = 0, t32 = 0, t21 = 0, t10 = 0, pump_rate = 0;
    char distri[40] = {0};

    // device state, read back (the simulation starts if it has not):
    // i j k of each gain cell (location), populations N0..N3 and the
    // polarisation Pax Pay Paz (one transition in this 3-D model)
    std::vector<int> location()
    {
        cfd_sim *s = ctx->get();
        const int n = cfd_special_cells(s, 1, nullptr);
        detail::check(n, "cfd_special_cells");
        std::vector<int> ijk((size_t)n * 3);
        detail::check(cfd_special_cells(s, 1, ijk.data()), "cfd_special_cells");
        return ijk;
    }
    std::vector<double> N0() { return level(0); }
    std::vector<double> N1() { return level(1); }
    std::vector<double> N2() { return level(2); }
    std::vector<double> N3() { return level(3); }
    std::vector<double> Pax() { return pol(0); }
    std::vector<double> Pay() { return pol(1); }
    std::vector<double> Paz() { return pol(2); }
    // [cell][6]: polarisation x y z, then its previous values
    std::vector<double> polarization()
    {
        cfd_sim *s = ctx->get();
        const int n = cfd_gain_state(s, nullptr);
        detail::check(n, "cfd_gain_state");
        std::vector<double> v((size_t)n * 6);
        detail::check(cfd_gain_state(s, v.data()), "cfd_gain_state");
        return v;
    }
    void UpdatePolarizationDensities(Field &) { ctx->device([](cfd_sim *s) { return cfd_gain_pol(s); }, "UpdatePolarizationDensities"); }
    void UpdateElectricField(Field &) { ctx->device([](cfd_sim *s) { return cfd_gain_e(s); }, "UpdateElectricField"); }
    void UpdateStatesPopulations(Field &) { ctx->device([](cfd_sim *s) { return cfd_gain_pop(s); }, "UpdateStatesPopulations"); }
    // populations of every gain cell, [cell][N0..N3]
    std::vector<double> populations()
    {
        int counts[4];
        detail::check(cfd_info(ctx->get(), nullptr, nullptr, counts), "cfd_info");
        std::vector<double> v((size_t)counts[2] * 4);
        detail::check(cfd_gain_read(ctx->sim, v.data()), "cfd_gain_read");
        return v;
    }
private:
    std::vector<double> level(int l)
    {
        const std::vector<double> p = populations();
        std::vector<double> v;
        for (size_t r = 0; r < p.size() / 4; r++) v.push_back(p[r * 4 + l]);
        return v;
    }
    std::vector<double> pol(int c)
    {
        const std::vector<double> p = polarization();
        std::vector<double> v;
        for (size_t r = 0; r < p.size() / 6; r++) v.push_back(p[r * 6 + c]);
        return v;
    }
    std::shared_ptr<detail::Context> ctx;
};

// soft point sources of the Structure's options at step n (main.cpp:284-288)
inline void UpdateSources(Structure &sS, int timestep)
{
    sS.ctx->device([timestep](cfd_sim *s) { return cfd_sources(s, timestep); }, "sources");
}

// ---------------------------------------------------------------------------
// Ports (Ports.h:20-92)
// ---------------------------------------------------------------------------
struct Port_Points {
    int i, j, k;
    double Ex, Ey, Ez, Hx, Hy, Hz;
};

namespace detail {

class PortBase {
public:
    std::vector<Port_Points> port_point;
    int Num_of_Port_Points = 0;

    // Ports.cpp:88-99 / 342-353: this step's six values at every point (the
    // device also records them, and feeds the running DFT when configured)
    void Update_Fields_on_Port(Field &)
    {
        Context &c = *ctx;
        cfd_sim *s = c.get();
        if (c.field) c.field->flush();
        if (c.sampled_step != c.step) {
            check(cfd_ports_sample(s, c.step), "cfd_ports_sample");
            c.sampled_step = c.step;
        }
        points();
        std::vector<double> v(port_point.size() * 6);
        if (!v.empty()) check(cfd_port_gather(s, index, v.data()), "cfd_port_gather");
        for (size_t t = 0; t < port_point.size(); t++) {
            Port_Points &q = port_point[t];
            q.Ex = v[t * 6]; q.Ey = v[t * 6 + 1]; q.Ez = v[t * 6 + 2];
            q.Hx = v[t * 6 + 3]; q.Hy = v[t * 6 + 4]; q.Hz = v[t * 6 + 5];
        }
    }
    // Ports.cpp:112-119: one line per point, Ex Ey Ez
    void Export_E_onPort(const char *fileName)
    {
        std::ofstream out(fileName);
        Export_E_onPort(out);
    }
    void Export_E_onPort(std::ofstream &out)
    {
        for (const Port_Points &q : port_point) out << q.Ex << "\t" << q.Ey << "\t" << q.Ez << std::endl;
    }
    // the port's point list, i j k per line (Ports.cpp:356-362)
    void points_to(std::ostream &out)
    {
        points();
        for (const Port_Points &q : port_point) out << q.i << "\t" << q.j << "\t" << q.k << std::endl;
    }
    int port_index() const { return index; }

protected:
    PortBase(Structure &sS, const char *fileName, int kind) : ctx(sS.ctx)
    {
        ctx->require_unstarted("port");
        index = (int)ctx->port_files.size();
        ctx->port_files.push_back(fileName);
        ctx->port_kinds.push_back(kind);
    }
    // The point list depends on the mesh and the port file only.  Before the
    // device simulation exists (main.cpp:144-151 exports the sphere ports'
    // lists right after constructing them, then constructs the plane ports), it comes
    // from the host model of this port alone, so asking for it does not start
    // the simulation; afterwards from the simulation.  Both are the same
    // computation.
    void points()
    {
        if (!port_point.empty()) return;
        std::vector<int> ijk;
        int n = 0;
        if (ctx->sim) {
            n = cfd_port_npts(ctx->sim, index);
            check(n, "cfd_port_npts");
            ijk.resize((size_t)n * 3);
            if (n) check(cfd_port_points(ctx->sim, index, ijk.data()), "cfd_port_points");
        } else {
            cfd_config c{};
            const char *file = ctx->port_files[index].c_str();
            int kind = ctx->port_kinds[index];
            c.material = ctx->material.c_str();
            c.structure = ctx->structure.c_str();
            c.n_ports = 1;
            c.port_files = &file;
            c.port_kinds = &kind;
            c.nt = 1;
            c.precision = CFD_F64;
            cfd_model *m = nullptr;
            if (cfd_model_build(&c, &m) < 0) throw Error(std::string("port ") + file + ": " + cfd_model_last_error());
            n = cfd_model_port(m, 0, nullptr);
            ijk.resize((size_t)std::max(n, 0) * 3);
            if (n > 0) cfd_model_port(m, 0, ijk.data());
            cfd_model_free(m);
            if (n < 0) throw Error(std::string("port ") + file + ": no point list");
        }
        port_point.resize(n);
        for (int t = 0; t < n; t++) port_point[t] = Port_Points{ijk[3 * t], ijk[3 * t + 1], ijk[3 * t + 2], 0, 0, 0, 0, 0, 0};
        Num_of_Port_Points = n;
    }
    template <class Get>
    void row(std::ostream &out, Get get, int nw, bool per_row_newline)
    {
        // Ports.cpp:377-388 (plane ports: rows of Nw values) and 121-143
        // (sphere ports: one row); values separated and ended by a tab
        int ii = 0;
        const int n = (int)port_point.size();
        while (ii < n) {
            for (int i = 0; i < nw && ii < n; ++i, ++ii) out << get(port_point[ii]) << "\t";
            if (per_row_newline) out << std::endl;
        }
    }
    std::shared_ptr<Context> ctx;
    int index = 0;
};

}  // namespace detail

class SphericalPort : public detail::PortBase {
public:
    SphericalPort(Structure &sS, const char *fileName) : PortBase(sS, fileName, CFD_PORT_SPHERE), dx(sS.dx), dy(sS.dy), dz(sS.dz)
    {
        std::ifstream in(fileName);
        std::string line;
        std::getline(in, line);  // r  (Ports.cpp:40-43)
        std::istringstream(line) >> r;
        std::getline(in, line);  // xc yc zc
        std::istringstream(line) >> xc >> yc >> zc;
    }
    void Export_Port(const char *fileName)  // Ports.cpp:101-109: centre in cells, then the points
    {
        std::ofstream out(fileName);
        out << xc / dx << "\t" << yc / dy << "\t" << zc / dz << std::endl;
        points_to(out);
    }
    void Export_Ex_onPort(const char *f) { file_row(f, [](const Port_Points &q) { return q.Ex; }); }
    void Export_Ey_onPort(const char *f) { file_row(f, [](const Port_Points &q) { return q.Ey; }); }
    void Export_Ez_onPort(const char *f) { file_row(f, [](const Port_Points &q) { return q.Ez; }); }
    void Export_EIntensity_onPort(const char *f) { file_row(f, intensity); }
    void Export_Ex_onPort(std::ofstream &o) { step_row(o, [](const Port_Points &q) { return q.Ex; }); }
    void Export_Ey_onPort(std::ofstream &o) { step_row(o, [](const Port_Points &q) { return q.Ey; }); }
    void Export_Ez_onPort(std::ofstream &o) { step_row(o, [](const Port_Points &q) { return q.Ez; }); }
    void Export_EIntensity_onPort(std::ofstream &o) { step_row(o, intensity); }
    using PortBase::Export_E_onPort;

    double r = 0, xc = 0, yc = 0, zc = 0;

private:
    static double intensity(const Port_Points &q) { return q.Ex * q.Ex + q.Ey * q.Ey + q.Ez * q.Ez; }
    template <class Get>
    void file_row(const char *f, Get g)  // Ports.cpp:130-137: values, no newline
    {
        std::ofstream out(f);
        row(out, g, (int)port_point.size(), false);
    }
    template <class Get>
    void step_row(std::ofstream &o, Get g)  // Ports.cpp:138-144: one line per call
    {
        row(o, g, (int)port_point.size(), false);
        o << std::endl;
    }
    double dx, dy, dz;
};

class PlainPort : public detail::PortBase {
public:
    // The Structure must outlive the port (Export_ID reads its ID map).
    PlainPort(Structure &sS, const char *fileName) : PortBase(sS, fileName, CFD_PORT_PLANE), S(&sS)
    {
        // Ports.cpp:211-283: normal, position, w1 w2, h1 h2; the in-plane
        // indices run over (int)(w1/dw)..(int)(w2/dw) and (int)(h1/dh)..
        // (int)(h2/dh), with (dw, dh) = (dy, dz) for an x-normal port,
        // (dx, dz) for y, (dx, dy) for z
        std::ifstream in(fileName);
        std::string line;
        std::getline(in, line);
        std::istringstream(line) >> normal_direction;
        std::getline(in, line);
        std::istringstream(line) >> position;
        std::getline(in, line);
        std::istringstream(line) >> w1 >> w2;
        std::getline(in, line);
        std::istringstream(line) >> h1 >> h2;
        const double dw = normal_direction == 'x' ? sS.dy : sS.dx;
        const double dh = normal_direction == 'z' ? sS.dy : sS.dz;
        if (!(dw > 0) || !(dh > 0)) throw Error("PlainPort: construct it after Structure::generate_mesh");
        Nw = (int)(w2 / dw) - (int)(w1 / dw) + 1;
        Nh = (int)(h2 / dh) - (int)(h1 / dh) + 1;
        if (Nw < 1 || Nh < 1) throw Error(std::string("PlainPort: empty port in ") + fileName);
        out_store.assign((size_t)Nw * Nh, 0.0);  // Ports.cpp:296-303: port_out [Nw][Nh], zeroed
        out_rows.resize(Nw);
        for (int i = 0; i < Nw; ++i) out_rows[i] = out_store.data() + (size_t)i * Nh;
        port_out = out_rows.data();
    }
    PlainPort(const PlainPort &) = delete;
    PlainPort &operator=(const PlainPort &) = delete;

    void Export_Port(const char *fileName)  // Ports.cpp:355-362
    {
        std::ofstream out(fileName);
        points_to(out);
    }
    void Export_Ex_onPort(const char *f) { file_rows(f, [](const Port_Points &q) { return q.Ex; }); }
    void Export_Ey_onPort(const char *f) { file_rows(f, [](const Port_Points &q) { return q.Ey; }); }
    void Export_Ez_onPort(const char *f) { file_rows(f, [](const Port_Points &q) { return q.Ez; }); }
    void Export_EIntensity_onPort(const char *f) { file_line(f, intensity); }
    void Export_Ex_onPort(std::ofstream &o) { step_row(o, [](const Port_Points &q) { return q.Ex; }); }
    void Export_Ey_onPort(std::ofstream &o) { step_row(o, [](const Port_Points &q) { return q.Ey; }); }
    void Export_Ez_onPort(std::ofstream &o) { step_row(o, [](const Port_Points &q) { return q.Ez; }); }
    void Export_EIntensity_onPort(std::ofstream &o) { step_row(o, intensity); }
    using PortBase::Export_E_onPort;

    // Ports.cpp:308-320: the material IDs, Nw per line
    void Export_ID_onPort(const char *fileName)
    {
        std::ofstream out(fileName);
        short **m = Export_ID();
        for (int j = 0; j < Nh; ++j) {
            for (int i = 0; i < Nw; ++i) out << m[i][j] << "\t";
            out << std::endl;
        }
    }
    // Ports.cpp:322-340: [Nw][Nh] material IDs; the array belongs to the port
    // and holds until the next Export_ID call
    short **Export_ID()
    {
        points();
        check_shape();
        id_store.resize(port_point.size());
        id_rows.resize(Nw);
        for (int i = 0; i < Nw; ++i) id_rows[i] = id_store.data() + (size_t)i * Nh;
        for (size_t n = 0; n < port_point.size(); ++n) {
            const Port_Points &q = port_point[n];
            id_rows[n % Nw][n / Nw] = S->ID[q.i][q.j][q.k];
        }
        return id_rows.data();
    }
    // Ports.cpp:405-415, 442-452, 479-489, 512-522: fill port_out [Nw][Nh]
    // from the last Update_Fields_on_Port and return it
    double **Export_Ex() { return map([](const Port_Points &q) { return q.Ex; }); }
    double **Export_Ey() { return map([](const Port_Points &q) { return q.Ey; }); }
    double **Export_Ez() { return map([](const Port_Points &q) { return q.Ez; }); }
    double **Export_EIntensity() { return map(intensity); }

    // the reference's public members (Ports.h:94-98); point n of port_point
    // sits at (n % Nw, n / Nw) of the in-plane grid
    char normal_direction = 0;
    double position = 0, w1 = 0, w2 = 0, h1 = 0, h2 = 0;
    double **port_out = nullptr;
    int Nw = 0, Nh = 0;

private:
    void check_shape() const
    {
        if (port_point.size() != (size_t)Nw * Nh)
            throw Error("PlainPort: " + std::to_string(port_point.size()) + " points, not Nw x Nh = " +
                        std::to_string(Nw) + " x " + std::to_string(Nh));
    }
    template <class Get>
    double **map(Get g)
    {
        points();
        check_shape();
        for (size_t n = 0; n < port_point.size(); ++n) out_rows[n % Nw][n / Nw] = g(port_point[n]);
        return port_out;
    }
    static double intensity(const Port_Points &q) { return q.Ex * q.Ex + q.Ey * q.Ey + q.Ez * q.Ez; }
    template <class Get>
    void file_rows(const char *f, Get g)  // Ports.cpp:377-388: Nw values per line
    {
        std::ofstream out(f);
        points();
        row(out, g, Nw, true);
    }
    template <class Get>
    void file_line(const char *f, Get g)  // Ports.cpp:491-500: every value on one line
    {
        std::ofstream out(f);
        points();
        row(out, g, (int)port_point.size(), false);
        out << std::endl;
    }
    template <class Get>
    void step_row(std::ofstream &o, Get g)  // Ports.cpp:389-400: the whole port on one line
    {
        row(o, g, (int)port_point.size(), false);
        o << std::endl;
    }
    const Structure *S;
    std::vector<double> out_store;
    std::vector<double *> out_rows;
    std::vector<short> id_store;
    std::vector<short *> id_rows;
};

// ---------------------------------------------------------------------------
// OutImage (OutImage.h:21-41): colour maps of [Nx][Ny] maps (a plane port's
// Export_EIntensity / Export_ID) written as PNG files.  There is no display
// window: plot() only updates the image; plot_save() also writes
// "<base>NNNNNN.png" (NNNNNN = time_step / save_modulus, OutImage.cpp:198),
// structure_plot_save() writes "<base>_structure.png" (OutImage.cpp:214).
// ---------------------------------------------------------------------------
class OutImage {
public:
    OutImage(int N_x, int N_y, int save_modulus, double cmap_min, double cmap_max, const char *file_name_base)
        : savemodulus(save_modulus), Nx(N_x), Ny(N_y), cmapmin(cmap_min), cmapmax(cmap_max),
          filenamebase(file_name_base ? file_name_base : "")
    {
        if (Nx <= 0 || Ny <= 0 || savemodulus <= 0) throw Error("OutImage: sizes and save modulus must be positive");
        rgb.assign((size_t)Nx * Ny * 3, 0);
    }
    void plot(double **dD, int time_step)
    {
        if (time_step % savemodulus == 0) draw(dD);
    }
    void plot_save(double **dD, int time_step)
    {
        if (time_step % savemodulus != 0) return;
        draw(dD);
        char num[16];
        std::snprintf(num, sizeof num, "%06d", time_step / savemodulus);
        save(filenamebase + num + ".png");
    }
    void structure_plot_save(short **dD)
    {
        std::vector<short> v((size_t)Nx * Ny);
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j) v[(size_t)i * Ny + j] = dD[i][j];
        image_check(cfd_image_materials(v.data(), Nx, Ny, rgb.data()));
        save(filenamebase + "_structure.png");
    }

    int savemodulus, Nx, Ny;
    double cmapmin, cmapmax;
    std::string filenamebase;
    std::vector<unsigned char> rgb;  // [Ny][Nx][3], the image's current pixels

private:
    static void image_check(int rc)
    {
        if (rc < 0) throw Error(cfd_image_last_error());
    }
    void draw(double **dD)
    {
        std::vector<double> v((size_t)Nx * Ny);
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j) v[(size_t)i * Ny + j] = dD[i][j];
        image_check(cfd_image_colormap(v.data(), Nx, Ny, cmapmin, cmapmax, rgb.data()));
    }
    void save(const std::string &path) { image_check(cfd_image_write_png(path.c_str(), Nx, Ny, rgb.data())); }
};

}  // namespace cpsfdtd

#endif  // CPSFDTD_HPP
