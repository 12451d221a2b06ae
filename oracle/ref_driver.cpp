// ref_driver.cpp -- drives the reference's own 3-D core (Structure.cpp,
// Fields.cpp, BasicFDTD.cpp, CPML.cpp, TF_SF.cpp, Ports.cpp, Sources.cpp,
// stdafx.cpp, compiled unmodified from /root/reference by oracle/build_ref.sh)
// so that tests can pin the FP64 oracle against it and bench.py can time it.
//
// TEST INFRASTRUCTURE ONLY.  The binary lands in oracle/_ref/ and is never
// part of the product.
//
// Heap memory is zero-filled (global operator new below).  The reference reads
// memory it never writes (CPML.cpp:210 last H coefficient, TF_SF.cpp:1244-1267
// x-face edge records, TF_SF.cpp:125 k0); with zero-filled memory those reads
// are deterministic, and the oracle reproduces them under ORC_REF_ZEROHEAP.
//
// Usage (inputs are read from <case> with the names main.cpp:67-84 uses):
//   refdrv tables <case> <out>                        ID map and setup tables
//   refdrv run <case> <out> <nsteps> [opts]           time loop, fields + ports
//   refdrv bench <case> <nsteps> [opts]               time the loop, print JSON
//   refdrv source <case> <out> <nhalf> <amp> <loc>    Source::Pulse and
//                         ::ContinuousWave at ntimestep = q / 2, q < nhalf
//                         (out/pulse.f64, out/cw.f64)
//   refdrv inc <case> <out> <nq>                      TF_SF::Einc_onPoint and
//                         ::Hinc_onPoint at retarded = q / 4 - 1, q < nq
//                         (out/inc.f64, E and H interleaved)
// opts: --src i j k amp   soft Ez source, amplitude * Source::Pulse(n)
//       --k0 v            value the TF_SF object's unset k0 holds
//       --ports f1,f2     plane port files (p:name) or sphere (s:name)
//       --notfsf          skip TF/SF even when TF_SF.txt exists
//       --faces f1,f2     TF/SF through the one-face methods (x1 ... z2,
//                         x1UpdateHField ... then x1UpdateEField ...) in
//                         this order instead of UpdateHField/UpdateEField
//       --export          (run) also write each port's text through the
//                         reference's own Export_* methods: port{q}_Ex.txt and
//                         port{q}_I.txt (Export_Ex/EIntensity_onPort(ofstream&)
//                         after every step), and after the last step
//                         port{q}_E_last.txt, port{q}_Ex_file.txt,
//                         port{q}_I_file.txt (the const char* forms),
//                         port{q}_points.txt (Export_Port) and, for plane
//                         ports, port{q}_ID_file.txt (Export_ID_onPort)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>
#include <chrono>
#include <unistd.h>
#include <omp.h>

#include "Structure.h"
#include "Fields.h"
#include "BasicFDTD.h"
#define private public
#define protected public
#include "CPML.h"
#include "TF_SF.h"
#include "Ports.h"
#include "Sources.h"
#undef private
#undef protected

void *operator new(std::size_t n)
{
    void *p = std::calloc(1, n ? n : 1);
    if (!p) throw std::bad_alloc();
    return p;
}
void *operator new[](std::size_t n) { return ::operator new(n); }
void operator delete(void *p) noexcept { std::free(p); }
void operator delete[](void *p) noexcept { std::free(p); }
void operator delete(void *p, std::size_t) noexcept { std::free(p); }
void operator delete[](void *p, std::size_t) noexcept { std::free(p); }

static void wr(const std::string &path, const void *data, size_t bytes)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    if (bytes) std::fwrite(data, 1, bytes, f);
    std::fclose(f);
}
static bool exists(const std::string &p) { return access(p.c_str(), R_OK) == 0; }

struct Opts {
    int src = 0, si = 0, sj = 0, sk = 0;
    double samp = 0;
    int k0set = 0, k0 = 0;
    int notfsf = 0;
    int exports = 0;
    std::vector<std::string> ports;
    std::vector<int> faces;  // 0 x1, 1 x2, 2 y1, 3 y2, 4 z1, 5 z2
};

static Opts parse_opts(int argc, char **argv, int first)
{
    Opts o;
    for (int a = first; a < argc; a++) {
        std::string s = argv[a];
        if (s == "--src" && a + 4 < argc) {
            o.src = 1; o.si = atoi(argv[a + 1]); o.sj = atoi(argv[a + 2]); o.sk = atoi(argv[a + 3]); o.samp = atof(argv[a + 4]);
            a += 4;
        } else if (s == "--k0" && a + 1 < argc) {
            o.k0set = 1; o.k0 = atoi(argv[++a]);
        } else if (s == "--notfsf") {
            o.notfsf = 1;
        } else if (s == "--export") {
            o.exports = 1;
        } else if (s == "--faces" && a + 1 < argc) {
            std::string l = argv[++a];
            for (size_t p0 = 0; p0 + 1 < l.size() + 1; p0 += 3) {
                const std::string f = l.substr(p0, 2);
                if (f.size() != 2 || f[0] < 'x' || f[0] > 'z' || (f[1] != '1' && f[1] != '2')) {
                    std::fprintf(stderr, "bad face %s\n", f.c_str());
                    std::exit(2);
                }
                o.faces.push_back(2 * (f[0] - 'x') + (f[1] - '1'));
            }
        } else if (s == "--ports" && a + 1 < argc) {
            std::string l = argv[++a];
            size_t p0 = 0;
            while (p0 <= l.size()) {
                size_t p1 = l.find(',', p0);
                if (p1 == std::string::npos) p1 = l.size();
                if (p1 > p0) o.ports.push_back(l.substr(p0, p1 - p0));
                p0 = p1 + 1;
            }
        }
    }
    return o;
}

struct Sim {
    Structure *S = nullptr;
    Field *F = nullptr;
    CPML *pml = nullptr;
    TF_SF *tf = nullptr;
    Source *src = nullptr;
    std::vector<PlainPort *> pp;
    std::vector<SphericalPort *> sp;
    std::vector<int> kinds;
};

static Sim setup(const Opts &o)
{
    Sim m;
    m.S = new Structure("material.txt");
    m.S->generate_mesh("structure.txt");
    m.S->InitializeBasicFDTDCoefficients();
    m.F = new Field(*m.S, (char *)"TEM");
    if (exists("CFS_PML.txt")) {
        m.pml = new CPML(*m.S, (char *)"CFS_PML.txt");
        m.pml->InitializeCPML(*m.S);
    }
    if (!o.notfsf && exists("TF_SF.txt")) {
        void *mem = ::operator new(sizeof(TF_SF));
        if (o.k0set) static_cast<TF_SF *>(mem)->k0 = o.k0;
        m.tf = new (mem) TF_SF(*m.S, (char *)"TF_SF.txt");
    }
    if (o.src) m.src = new Source(m.S->dt, (char *)"source.txt");
    for (const std::string &p : o.ports) {
        if (p.size() > 2 && p[0] == 's') { m.sp.push_back(new SphericalPort(*m.S, p.c_str() + 2)); m.kinds.push_back(1); }
        else { m.pp.push_back(new PlainPort(*m.S, p.c_str() + 2)); m.kinds.push_back(0); }
    }
    return m;
}

typedef void (TF_SF::*FaceFn)(int, Field &, Structure &);
static const FaceFn kFaceH[6] = {&TF_SF::x1UpdateHField, &TF_SF::x2UpdateHField, &TF_SF::y1UpdateHField,
                                 &TF_SF::y2UpdateHField, &TF_SF::z1UpdateHField, &TF_SF::z2UpdateHField};
static const FaceFn kFaceE[6] = {&TF_SF::x1UpdateEField, &TF_SF::x2UpdateEField, &TF_SF::y1UpdateEField,
                                 &TF_SF::y2UpdateEField, &TF_SF::z1UpdateEField, &TF_SF::z2UpdateEField};

// main.cpp:267-343 order: H, TF/SF-H, CPML-H, E, TF/SF-E, source, CPML-E
static void step(Sim &m, const Opts &o, int n)
{
    BasicHFieldUpdate(*m.F, *m.S);
    if (m.tf && o.faces.empty()) m.tf->UpdateHField(n, *m.F, *m.S);
    if (m.tf) for (int f : o.faces) (m.tf->*kFaceH[f])(n, *m.F, *m.S);
    if (m.pml) m.pml->UpdateCPMLHFields(*m.F);
    BasicEFieldUpdate(*m.F, *m.S);
    if (m.tf && o.faces.empty()) m.tf->UpdateEField(n, *m.F, *m.S);
    if (m.tf) for (int f : o.faces) (m.tf->*kFaceE[f])(n, *m.F, *m.S);
    if (m.src) m.F->Ez[o.si][o.sj][o.sk] = m.F->Ez[o.si][o.sj][o.sk] + m.src->Pulse(o.samp, (double)n, 0.0);
    if (m.pml) m.pml->UpdateCPMLEFields(*m.F);
}

static void dump_field(const Sim &m, double ***A, std::vector<double> &out)
{
    int Nx = m.S->Nx, Ny = m.S->Ny, Nz = m.S->Nz;
    out.resize((size_t)(Nx + 1) * (Ny + 1) * (Nz + 1));
    size_t c = 0;
    for (int i = 0; i <= Nx; i++)
        for (int j = 0; j <= Ny; j++)
            for (int k = 0; k <= Nz; k++) out[c++] = A[i][j][k];
}

static int cmd_tables(const std::string &out, const Opts &o)
{
    Sim m = setup(o);
    Structure &S = *m.S;
    int dims[4] = {S.Nx, S.Ny, S.Nz, S.NoMedia};
    double dd[4] = {S.dx, S.dy, S.dz, S.dt};
    wr(out + "/dims.i32", dims, sizeof dims);
    wr(out + "/d.f64", dd, sizeof dd);
    std::vector<short> id((size_t)(S.Nx + 1) * (S.Ny + 1) * (S.Nz + 1));
    size_t c = 0;
    for (int i = 0; i <= S.Nx; i++) for (int j = 0; j <= S.Ny; j++) for (int k = 0; k <= S.Nz; k++) id[c++] = S.ID[i][j][k];
    wr(out + "/id.i16", id.data(), id.size() * 2);
    std::vector<double> co;
    for (double *t : {S.aE, S.bE, S.aH, S.bH}) co.insert(co.end(), t, t + S.NoMedia);
    wr(out + "/coef.f64", co.data(), co.size() * 8);
    double *st[12] = {S.cExy, S.cEyx, S.cEzx, S.cExz, S.cEzy, S.cEyz, S.cHxy, S.cHyx, S.cHzx, S.cHxz, S.cHzy, S.cHyz};
    int len[12] = {S.Nz, S.Nz, S.Ny, S.Ny, S.Nx, S.Nx, S.Nz, S.Nz, S.Ny, S.Ny, S.Nx, S.Nx};
    std::vector<double> sv;
    for (int t = 0; t < 12; t++) sv.insert(sv.end(), st[t], st[t] + len[t]);
    wr(out + "/stretch.f64", sv.data(), sv.size() * 8);
    if (m.pml) {
        CPML &P = *m.pml;
        // axis*4 + side*2 + (E,H): b then c
        int n[3][2] = {{P.Npmlx1, P.Npmlx2}, {P.Npmly1, P.Npmly2}, {P.Npmlz1, P.Npmlz2}};
        double *tb[12] = {P.bxE1, P.bxH1, P.bxE2, P.bxH2, P.byE1, P.byH1, P.byE2, P.byH2, P.bzE1, P.bzH1, P.bzE2, P.bzH2};
        double *tc[12] = {P.cxE1, P.cxH1, P.cxE2, P.cxH2, P.cyE1, P.cyH1, P.cyE2, P.cyH2, P.czE1, P.czH1, P.czE2, P.czH2};
        std::vector<double> pv;
        std::vector<int> pl;
        for (int w = 0; w < 12; w++) {
            int a = w / 4, e = (w / 2) & 1, h = w & 1;
            int l = h ? n[a][e] : n[a][e] + 1;
            pl.push_back(l);
            pv.insert(pv.end(), tb[w], tb[w] + l);
            pv.insert(pv.end(), tc[w], tc[w] + l);
        }
        wr(out + "/cpml_len.i32", pl.data(), pl.size() * 4);
        wr(out + "/cpml.f64", pv.data(), pv.size() * 8);
        double cp[2] = {P.cpmlE, P.cpmlH};
        wr(out + "/cpml_c.f64", cp, sizeof cp);
    }
    if (m.tf) {
        TF_SF &T = *m.tf;
        int np = T.mxy2_2;  // last written record (TF_SF.cpp:887)
        int seg[19] = {np, T.mxz1, T.mxz1_1, T.mxz1_2, T.mxz2, T.mxz2_1, T.mxz2_2, T.mzy1, T.mzy1_1, T.mzy1_2,
                       T.mzy2, T.mzy2_1, T.mzy2_2, T.mxy1, T.mxy1_1, T.mxy1_2, T.mxy2, T.mxy2_1, T.mxy2_2};
        wr(out + "/tfsf_seg.i32", seg, sizeof seg);
        std::vector<int> ijk;
        std::vector<double> tab;
        for (int q = 0; q < np; q++) {
            const TFSF_point &p = T.TFSFp[q];
            ijk.push_back(p.i); ijk.push_back(p.j); ijk.push_back(p.k);
            for (double v : {p.k_dot_r_E1, p.profile_E1, p.k_dot_r_H1, p.profile_H1, p.k_dot_r_E2, p.profile_E2, p.k_dot_r_H2, p.profile_H2})
                tab.push_back(v);
        }
        wr(out + "/tfsf_ijk.i32", ijk.data(), ijk.size() * 4);
        wr(out + "/tfsf_tab.f64", tab.data(), tab.size() * 8);
        wr(out + "/tfsf_einc.f64", T.Einc, (size_t)(T.Nt + 1) * 8);
        wr(out + "/tfsf_hinc.f64", T.Hinc, (size_t)(T.Nt + 1) * 8);
        double v[14] = {T.E0x, T.E0y, T.E0z, T.H0x, T.H0y, T.H0z, T.CEx, T.CEy, T.CEz, T.CHx, T.CHy, T.CHz, T.l0, (double)T.Nt};
        wr(out + "/tfsf_vec.f64", v, sizeof v);
    }
    size_t ip = 0, is = 0;
    for (size_t q = 0; q < m.kinds.size(); q++) {
        std::vector<int> pts;
        if (m.kinds[q] == 0) {
            PlainPort *P = m.pp[ip++];
            for (int t = 0; t < P->Num_of_Port_Points - 1; t++) {  // F11: last record is never written
                pts.push_back(P->port_point[t].i); pts.push_back(P->port_point[t].j); pts.push_back(P->port_point[t].k);
            }
        } else {
            SphericalPort *P = m.sp[is++];
            for (int t = 0; t < P->Num_of_Port_Points - 1; t++) {
                pts.push_back(P->port_point[t].i); pts.push_back(P->port_point[t].j); pts.push_back(P->port_point[t].k);
            }
        }
        wr(out + "/port" + std::to_string(q) + ".i32", pts.data(), pts.size() * 4);
    }
    return 0;
}

static int cmd_run(const std::string &out, int nsteps, const Opts &o)
{
    Sim m = setup(o);
    std::vector<std::vector<double>> ps(m.kinds.size());
    std::vector<std::ofstream *> exs, ins;
    auto name = [&](size_t q, const char *what) { return out + "/port" + std::to_string(q) + what; };
    if (o.exports)
        for (size_t q = 0; q < m.kinds.size(); q++) {
            exs.push_back(new std::ofstream(name(q, "_Ex.txt")));
            ins.push_back(new std::ofstream(name(q, "_I.txt")));
        }
    for (int n = 0; n < nsteps; n++) {
        step(m, o, n);
        size_t ip = 0, is = 0;
        for (size_t q = 0; q < m.kinds.size(); q++) {
            Port_Points *pt;
            int np;
            if (m.kinds[q] == 0) {
                PlainPort *P = m.pp[ip++];
                P->Update_Fields_on_Port(*m.F); pt = P->port_point; np = P->Num_of_Port_Points - 1;
                if (o.exports) { P->Export_Ex_onPort(*exs[q]); P->Export_EIntensity_onPort(*ins[q]); }
            } else {
                SphericalPort *P = m.sp[is++];
                P->Update_Fields_on_Port(*m.F); pt = P->port_point; np = P->Num_of_Port_Points - 1;
                if (o.exports) { P->Export_Ex_onPort(*exs[q]); P->Export_EIntensity_onPort(*ins[q]); }
            }
            for (int t = 0; t < np; t++)
                for (double v : {pt[t].Ex, pt[t].Ey, pt[t].Ez, pt[t].Hx, pt[t].Hy, pt[t].Hz}) ps[q].push_back(v);
        }
    }
    if (o.exports) {
        size_t ip = 0, is = 0;
        for (size_t q = 0; q < m.kinds.size(); q++) {
            exs[q]->close();
            ins[q]->close();
            if (m.kinds[q] == 0) {
                PlainPort *P = m.pp[ip++];
                P->Export_E_onPort(name(q, "_E_last.txt").c_str());
                P->Export_Ex_onPort(name(q, "_Ex_file.txt").c_str());
                P->Export_EIntensity_onPort(name(q, "_I_file.txt").c_str());
                P->Export_Port(name(q, "_points.txt").c_str());
                P->Export_ID_onPort(name(q, "_ID_file.txt").c_str());
            } else {
                SphericalPort *P = m.sp[is++];
                P->Export_E_onPort(name(q, "_E_last.txt").c_str());
                P->Export_Ex_onPort(name(q, "_Ex_file.txt").c_str());
                P->Export_EIntensity_onPort(name(q, "_I_file.txt").c_str());
                P->Export_Port(name(q, "_points.txt").c_str());
            }
        }
    }
    double ***A[6] = {m.F->Ex, m.F->Ey, m.F->Ez, m.F->Hx, m.F->Hy, m.F->Hz};
    const char *nm[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
    std::vector<double> buf;
    for (int c = 0; c < 6; c++) {
        dump_field(m, A[c], buf);
        wr(out + "/" + nm[c] + ".f64", buf.data(), buf.size() * 8);
    }
    for (size_t q = 0; q < ps.size(); q++) wr(out + "/port" + std::to_string(q) + "_samples.f64", ps[q].data(), ps[q].size() * 8);
    return 0;
}

static int cmd_bench(int nsteps, const Opts &o)
{
    auto t0 = std::chrono::steady_clock::now();
    Sim m = setup(o);
    auto t1 = std::chrono::steady_clock::now();
    step(m, o, 0);  // warm-up, untimed (first-touch of the field pages)
    auto t2 = std::chrono::steady_clock::now();
    for (int n = 1; n <= nsteps; n++) step(m, o, n);
    auto t3 = std::chrono::steady_clock::now();
    double sec = std::chrono::duration<double>(t3 - t2).count();
    double cells = (double)m.S->Nx * m.S->Ny * m.S->Nz;
    std::printf("{\"nx\": %d, \"ny\": %d, \"nz\": %d, \"steps\": %d, \"seconds\": %.6f, \"mcells_per_s\": %.4f, "
                "\"threads\": %d, \"setup_s\": %.3f, \"pml\": %d, \"tfsf\": %d}\n",
                m.S->Nx, m.S->Ny, m.S->Nz, nsteps, sec, cells * nsteps / sec / 1e6, omp_get_max_threads(),
                std::chrono::duration<double>(t1 - t0).count(), m.pml ? 1 : 0, m.tf ? 1 : 0);
    return 0;
}

static int cmd_source(const std::string &out, int nhalf, double amp, double loc)
{
    Structure S("material.txt");
    S.generate_mesh("structure.txt");
    S.InitializeBasicFDTDCoefficients();
    Source src(S.dt, (char *)"source.txt");
    std::vector<double> pulse, cw;
    for (int q = 0; q < nhalf; q++) {
        pulse.push_back(src.Pulse(amp, q / 2.0, loc));
        cw.push_back(src.ContinuousWave(amp, q / 2.0, loc));
    }
    wr(out + "/pulse.f64", pulse.data(), pulse.size() * 8);
    wr(out + "/cw.f64", cw.data(), cw.size() * 8);
    return 0;
}

static int cmd_inc(const std::string &out, int nq)
{
    Opts o;
    Sim m = setup(o);
    if (!m.tf) { std::fprintf(stderr, "no TF_SF.txt\n"); return 2; }
    std::vector<double> v;
    for (int q = 0; q < nq; q++) {
        v.push_back(m.tf->Einc_onPoint(q / 4.0 - 1.0));
        v.push_back(m.tf->Hinc_onPoint(q / 4.0 - 1.0));
    }
    wr(out + "/inc.f64", v.data(), v.size() * 8);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: refdrv tables|run|bench <case> ...\n");
        return 2;
    }
    std::string cmd = argv[1];
    std::string cs = argv[2];
    // std::cout banners from the reference go to stdout; keep stdout for JSON
    if (cmd == "bench") { if (!std::freopen("/dev/null", "w", stderr)) return 2; }
    if (chdir(cs.c_str())) { std::perror(cs.c_str()); return 2; }
    if (cmd == "tables" && argc >= 4) return cmd_tables(argv[3], parse_opts(argc, argv, 4));
    if (cmd == "run" && argc >= 5) return cmd_run(argv[3], atoi(argv[4]), parse_opts(argc, argv, 5));
    if (cmd == "bench" && argc >= 4) return cmd_bench(atoi(argv[3]), parse_opts(argc, argv, 4));
    if (cmd == "inc" && argc >= 5) return cmd_inc(argv[3], atoi(argv[4]));
    if (cmd == "source" && argc >= 7) return cmd_source(argv[3], atoi(argv[4]), atof(argv[5]), atof(argv[6]));
    std::fprintf(stderr, "bad arguments\n");
    return 2;
}
