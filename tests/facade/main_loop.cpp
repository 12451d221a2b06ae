// main_loop.cpp -- a driver in the shape of the reference's main.cpp time
// loop (main.cpp:115-343), written against the C++ facade (cpsfdtd.hpp).
// The tests run it and compare what it writes with the FP64 oracle.
//
//   main_loop --material F --structure F [--pml F] [--tfsf F] [--dl F]
//             [--gain F] [--source F] [--plane F]... [--sphere F]...
//             [--src i j k comp amp] [--steps N] [--flags U] [--out DIR]
//             [--mesh-only] [--images CMAX] [--cfs] [--faces x1,z1,...]
//             [--state]
//
// --cfs drives the CPML through the CFS_PML class (Initialize_PML,
// UpdatePMLHFields, UpdatePMLEFields) instead of CPML.  --faces applies the
// TF/SF correction through the one-face methods (x1UpdateHField ... then
// x1UpdateEField ...) in the given order instead of UpdateHField /
// UpdateEField.  --state, after the last step, writes DIR/dl_state.f64
// (per Drude-Lorentz cell from Drude_Lorentz::DLM(): ix jy kz, then per
// component JD[], JL[], JL_1[] and E_1), DIR/gain_state.f64 (per gain cell
// from FourLevelLaser: i j k, N0..N3, Pax Pay Paz) and DIR/inc.f64
// (TF_SF::Einc_onPoint and Hinc_onPoint at retarded = q / 4 - 1,
// q < 4 steps + 4).
// --mesh-only stops after Structure::generate_mesh (no device): writes
// DIR/id.i16 (ID map, [i][j][k]) and prints Nx Ny Nz dx dy dz dt.
// Writes DIR/field_{Ex..Hz}.f64 (final fields, [i][j][k]), per port q:
// DIR/port{q}_points.txt (Export_Port), DIR/port{q}_Ex.txt and
// DIR/port{q}_I.txt (one line per step, Export_Ex_onPort(ofstream&) and
// Export_EIntensity_onPort(ofstream&)), after the last step
// DIR/port{q}_E_last.txt, DIR/port{q}_Ex_file.txt, DIR/port{q}_I_file.txt
// (Export_E_onPort, Export_Ex_onPort, Export_EIntensity_onPort to a file),
// and DIR/port{q}_samples.f64 ([step][point][6] from port_point after
// Update_Fields_on_Port); plane ports also DIR/port{q}_ID_file.txt
// (Export_ID_onPort).  --images CMAX: for each plane port, OutImage over its
// Nw x Nh map with colour range [0, CMAX] as main.cpp:149-160, 228 uses it:
// DIR/port{q}_img_structure.png (structure_plot_save(Export_ID())) and
// DIR/port{q}_img{steps:06d}.png (plot_save(Export_EIntensity(), steps)).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "cpsfdtd.hpp"

int main(int argc, char **argv)
{
    std::string material, structure, pml, tfsf, dl, gain, source, out = ".";
    std::vector<std::string> planes, spheres;
    std::vector<int> src_ijkc;
    std::vector<double> src_amp;
    int steps = 10;
    unsigned flags = 0;
    bool mesh_only = false, cfs = false, state = false;
    std::vector<int> faces;
    double images_cmax = -1;
    for (int a = 1; a < argc; a++) {
        std::string k = argv[a];
        auto next = [&]() -> std::string {
            if (a + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", k.c_str()); std::exit(2); }
            return argv[++a];
        };
        if (k == "--material") material = next();
        else if (k == "--structure") structure = next();
        else if (k == "--pml") pml = next();
        else if (k == "--tfsf") tfsf = next();
        else if (k == "--dl") dl = next();
        else if (k == "--gain") gain = next();
        else if (k == "--source") source = next();
        else if (k == "--plane") planes.push_back(next());
        else if (k == "--sphere") spheres.push_back(next());
        else if (k == "--steps") steps = std::atoi(next().c_str());
        else if (k == "--flags") flags = (unsigned)std::strtoul(next().c_str(), nullptr, 0);
        else if (k == "--out") out = next();
        else if (k == "--mesh-only") mesh_only = true;
        else if (k == "--images") images_cmax = std::atof(next().c_str());
        else if (k == "--cfs") cfs = true;
        else if (k == "--state") state = true;
        else if (k == "--faces") {
            const std::string l = next();
            for (size_t p = 0; p < l.size(); p += 3) {
                if (l.size() < p + 2 || l[p] < 'x' || l[p] > 'z' || (l[p + 1] != '1' && l[p + 1] != '2')) {
                    std::fprintf(stderr, "bad face list %s\n", l.c_str());
                    return 2;
                }
                faces.push_back(2 * (l[p] - 'x') + (l[p + 1] - '1'));
            }
        }
        else if (k == "--src") {
            for (int q = 0; q < 4; q++) src_ijkc.push_back(std::atoi(next().c_str()));
            src_amp.push_back(std::atof(next().c_str()));
        } else { std::fprintf(stderr, "unknown option %s\n", k.c_str()); return 2; }
    }
    try {
        cpsfdtd::Structure S(material.c_str());
        S.generate_mesh(structure.c_str());
        S.InitializeBasicFDTDCoefficients();
        if (mesh_only) {
            std::vector<int16_t> id;
            for (int i = 0; i <= S.Nx; i++)
                for (int j = 0; j <= S.Ny; j++)
                    for (int k = 0; k <= S.Nz; k++) id.push_back(S.ID[i][j][k]);
            FILE *fh = std::fopen((out + "/id.i16").c_str(), "wb");
            std::fwrite(id.data(), sizeof(int16_t), id.size(), fh);
            std::fclose(fh);
            std::printf("%d %d %d %.17g %.17g %.17g %.17g\n", S.Nx, S.Ny, S.Nz, S.dx, S.dy, S.dz, S.dt);
            return 0;
        }
        S.set_steps(steps);
        S.set_flags(flags);
        if (!source.empty()) S.set_source_file(source.c_str());
        for (size_t q = 0; q < src_amp.size(); q++)
            S.add_point_source(src_ijkc[4 * q], src_ijkc[4 * q + 1], src_ijkc[4 * q + 2], src_ijkc[4 * q + 3], src_amp[q]);
        cpsfdtd::Field F(S, "3D");
        std::unique_ptr<cpsfdtd::CPML> cpml;
        std::unique_ptr<cpsfdtd::CFS_PML> cfspml;
        if (!pml.empty() && cfs) { cfspml.reset(new cpsfdtd::CFS_PML(S, pml.c_str())); cfspml->Initialize_PML(S); }
        else if (!pml.empty()) { cpml.reset(new cpsfdtd::CPML(S, pml.c_str())); cpml->InitializeCPML(S); }
        std::unique_ptr<cpsfdtd::TF_SF> tf;
        if (!tfsf.empty()) tf.reset(new cpsfdtd::TF_SF(S, tfsf.c_str()));
        std::unique_ptr<cpsfdtd::Drude_Lorentz> DL;
        if (!dl.empty()) DL.reset(new cpsfdtd::Drude_Lorentz(S, dl.c_str()));
        std::unique_ptr<cpsfdtd::FourLevelLaser> laser;
        if (!gain.empty()) laser.reset(new cpsfdtd::FourLevelLaser(S, gain.c_str()));
        std::vector<std::unique_ptr<cpsfdtd::PlainPort>> pp;
        std::vector<std::unique_ptr<cpsfdtd::SphericalPort>> sp;
        for (const std::string &f : planes) pp.emplace_back(new cpsfdtd::PlainPort(S, f.c_str()));
        for (const std::string &f : spheres) sp.emplace_back(new cpsfdtd::SphericalPort(S, f.c_str()));

        const size_t nports = pp.size() + sp.size();
        std::vector<std::unique_ptr<std::ofstream>> exOut, inOut;
        std::vector<std::vector<double>> samples(nports);
        auto name = [&](size_t q, const char *what) { return out + "/port" + std::to_string(q) + what; };
        for (size_t q = 0; q < nports; q++) {
            exOut.emplace_back(new std::ofstream(name(q, "_Ex.txt")));
            inOut.emplace_back(new std::ofstream(name(q, "_I.txt")));
        }
        for (size_t q = 0; q < pp.size(); q++) pp[q]->Export_Port((out + "/port" + std::to_string(q) + "_points.txt").c_str());
        for (size_t q = 0; q < sp.size(); q++)
            sp[q]->Export_Port((out + "/port" + std::to_string(pp.size() + q) + "_points.txt").c_str());

        using FaceFn = void (cpsfdtd::TF_SF::*)(int, cpsfdtd::Field &, cpsfdtd::Structure &);
        const FaceFn faceH[6] = {&cpsfdtd::TF_SF::x1UpdateHField, &cpsfdtd::TF_SF::x2UpdateHField,
                                 &cpsfdtd::TF_SF::y1UpdateHField, &cpsfdtd::TF_SF::y2UpdateHField,
                                 &cpsfdtd::TF_SF::z1UpdateHField, &cpsfdtd::TF_SF::z2UpdateHField};
        const FaceFn faceE[6] = {&cpsfdtd::TF_SF::x1UpdateEField, &cpsfdtd::TF_SF::x2UpdateEField,
                                 &cpsfdtd::TF_SF::y1UpdateEField, &cpsfdtd::TF_SF::y2UpdateEField,
                                 &cpsfdtd::TF_SF::z1UpdateEField, &cpsfdtd::TF_SF::z2UpdateEField};
        for (int n = 0; n < steps; ++n) {
            cpsfdtd::BasicHFieldUpdate(F, S);
            if (tf && faces.empty()) tf->UpdateHField(n, F, S);
            if (tf) for (int f : faces) ((*tf).*faceH[f])(n, F, S);
            if (cpml) cpml->UpdateCPMLHFields(F);
            if (cfspml) cfspml->UpdatePMLHFields(F);
            if (laser) laser->UpdatePolarizationDensities(F);
            cpsfdtd::BasicEFieldUpdate(F, S);
            if (DL) DL->EFieldUpdate(F, S);
            if (laser) { laser->UpdateElectricField(F); laser->UpdateStatesPopulations(F); }
            if (tf && faces.empty()) tf->UpdateEField(n, F, S);
            if (tf) for (int f : faces) ((*tf).*faceE[f])(n, F, S);
            if (!src_amp.empty()) cpsfdtd::UpdateSources(S, n);
            if (cpml) cpml->UpdateCPMLEFields(F);
            if (cfspml) cfspml->UpdatePMLEFields(F);
            for (size_t q = 0; q < nports; q++) {
                cpsfdtd::detail::PortBase &P = q < pp.size() ? (cpsfdtd::detail::PortBase &)*pp[q]
                                                              : (cpsfdtd::detail::PortBase &)*sp[q - pp.size()];
                P.Update_Fields_on_Port(F);
                if (q < pp.size()) {
                    pp[q]->Export_Ex_onPort(*exOut[q]);
                    pp[q]->Export_EIntensity_onPort(*inOut[q]);
                } else {
                    sp[q - pp.size()]->Export_Ex_onPort(*exOut[q]);
                    sp[q - pp.size()]->Export_EIntensity_onPort(*inOut[q]);
                }
                for (const cpsfdtd::Port_Points &t : P.port_point)
                    samples[q].insert(samples[q].end(), {t.Ex, t.Ey, t.Ez, t.Hx, t.Hy, t.Hz});
            }
        }
        for (size_t q = 0; q < nports; q++) {
            exOut[q]->close();
            inOut[q]->close();
            cpsfdtd::detail::PortBase &P = q < pp.size() ? (cpsfdtd::detail::PortBase &)*pp[q]
                                                          : (cpsfdtd::detail::PortBase &)*sp[q - pp.size()];
            P.Export_E_onPort(name(q, "_E_last.txt").c_str());
            if (q < pp.size()) {
                pp[q]->Export_Ex_onPort(name(q, "_Ex_file.txt").c_str());
                pp[q]->Export_EIntensity_onPort(name(q, "_I_file.txt").c_str());
                pp[q]->Export_ID_onPort(name(q, "_ID_file.txt").c_str());
                if (images_cmax > 0) {
                    cpsfdtd::OutImage img(pp[q]->Nw, pp[q]->Nh, 1, 0, images_cmax, name(q, "_img").c_str());
                    img.structure_plot_save(pp[q]->Export_ID());
                    img.plot_save(pp[q]->Export_EIntensity(), steps);
                }
            } else {
                sp[q - pp.size()]->Export_Ex_onPort(name(q, "_Ex_file.txt").c_str());
                sp[q - pp.size()]->Export_EIntensity_onPort(name(q, "_I_file.txt").c_str());
            }
            FILE *fh = std::fopen((out + "/port" + std::to_string(q) + "_samples.f64").c_str(), "wb");
            std::fwrite(samples[q].data(), sizeof(double), samples[q].size(), fh);
            std::fclose(fh);
        }
        auto dump = [&](const char *file, const std::vector<double> &v) {
            FILE *fh = std::fopen((out + "/" + file).c_str(), "wb");
            std::fwrite(v.data(), sizeof(double), v.size(), fh);
            std::fclose(fh);
        };
        if (state && DL) {
            std::vector<double> v;
            for (const cpsfdtd::DL_material &M : DL->DLM()) {
                v.insert(v.end(), {(double)M.ix, (double)M.jy, (double)M.kz});
                const std::vector<double> *JD[3] = {&M.JDx, &M.JDy, &M.JDz}, *JL[3] = {&M.JLx, &M.JLy, &M.JLz},
                                          *JL1[3] = {&M.JLx_1, &M.JLy_1, &M.JLz_1};
                const double E1[3] = {M.Ex_1, M.Ey_1, M.Ez_1};
                for (int c = 0; c < 3; c++) {
                    v.insert(v.end(), JD[c]->begin(), JD[c]->end());
                    v.insert(v.end(), JL[c]->begin(), JL[c]->end());
                    v.insert(v.end(), JL1[c]->begin(), JL1[c]->end());
                    v.push_back(E1[c]);
                }
            }
            dump("dl_state.f64", v);
        }
        if (state && laser) {
            const std::vector<int> loc = laser->location();
            const std::vector<double> n0 = laser->N0(), n1 = laser->N1(), n2 = laser->N2(), n3 = laser->N3();
            const std::vector<double> px = laser->Pax(), py = laser->Pay(), pz = laser->Paz();
            std::vector<double> v;
            for (size_t r = 0; r < n0.size(); r++)
                v.insert(v.end(), {(double)loc[3 * r], (double)loc[3 * r + 1], (double)loc[3 * r + 2], n0[r], n1[r], n2[r], n3[r],
                                   px[r], py[r], pz[r]});
            dump("gain_state.f64", v);
        }
        if (state && tf) {
            std::vector<double> v;
            for (int q = 0; q < 4 * steps + 4; q++) {
                v.push_back(tf->Einc_onPoint(q / 4.0 - 1.0));
                v.push_back(tf->Hinc_onPoint(q / 4.0 - 1.0));
            }
            dump("inc.f64", v);
        }
        static const char *names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
        for (int c = 0; c < 6; c++) {
            const std::vector<double> &v = F.values(c);
            FILE *fh = std::fopen((out + "/field_" + names[c] + ".f64").c_str(), "wb");
            std::fwrite(v.data(), sizeof(double), v.size(), fh);
            std::fclose(fh);
        }
        // a read through the reference's array syntax, to exercise the proxies
        std::printf("steps %d Ez[%d][%d][%d] = %.17g\n", S.step() + 1, S.Nx / 2, S.Ny / 2, S.Nz / 2,
                    (double)F.Ez[S.Nx / 2][S.Ny / 2][S.Nz / 2]);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "main_loop: %s\n", e.what());
        return 1;
    }
    return 0;
}
