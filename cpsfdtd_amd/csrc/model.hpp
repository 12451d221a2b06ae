// model.hpp -- host-side model of one simulation: inputs parsed from the
// reference's text formats, integer material-ID map, per-material and per-axis
// coefficient tables, CPML grading, TF/SF correction lists, dispersive and gain
// cell lists, sources and ports.  Everything here is computed once on the host
// in FP64 with the reference's arithmetic (so integer maps and FP64 tables are
// bit-identical to the reference's) and then uploaded by engine.hip.
#pragma once

#include <algorithm>
#include <cstdint>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "cpsfdtd.h"

namespace cfd {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// f(t, nth) on nth host threads (at most 8; t = 0 runs on the caller's), for
// set-up loops whose parts are independent; f splits its own work by t.  An
// exception thrown by any part is rethrown here once every part has ended.
template <class F>
void host_parallel(F &&f) {
    const int nth = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<std::exception_ptr> err(nth);
    auto part = [&](int t) {
        try {
            f(t, nth);
        } catch (...) {
            err[t] = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; t++) {
        try {
            pool.emplace_back(part, t);
        } catch (...) {  // no thread to be had: this part on the caller's
            part(t);
        }
    }
    part(0);
    for (std::thread &th : pool) th.join();
    for (const std::exception_ptr &e : err)
        if (e) std::rethrow_exception(e);
}

// Constants.h:19-26
constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kC = 2.99792458E8;
constexpr double kMu0 = 1.2566370614E-6;
constexpr double kEps0 = 8.8541878176E-12;
constexpr double kHPlanck = 6.62607004081E-34;

// per-axis stretch tables, in Structure.h:53-54 order
enum Stretch { cExy, cEyx, cEzx, cExz, cEzy, cEyz, cHxy, cHyx, cHzx, cHxz, cHzy, cHyz, kNumStretch };

// one TF/SF correction target: field component comp at cell (i,j,k), with its
// terms applied in the reference's face order
struct TfsfTarget {
    int comp, i, j, k;
    int first, count;  // range in Model::tf_terms
};
struct TfsfTerm {
    double coef;   // +-C*{E0,H0} component, sign folded in
    double delay;  // retarded-time offset in steps (TF_SF.cpp:166)
    double prof;   // transverse Gaussian profile (TF_SF.cpp:165)
};

struct Port {
    int kind = CFD_PORT_PLANE;
    std::vector<int> ijk;  // 3 per point
    int npts() const { return (int)(ijk.size() / 3); }
};

struct Model {
    // Structure
    double Lx = 0, Ly = 0, Lz = 0, dx = 0, dy = 0, dz = 0, dt = 0;
    int Nx = 0, Ny = 0, Nz = 0, NoMedia = 0, nt = 0;
    bool pbx = false, pby = false;
    unsigned flags = 0;
    std::vector<double> epsR, muR, sig, sigM, aE, bE, aH, bH;
    std::vector<double> stretch[kNumStretch];
    std::vector<int16_t> id;  // [i][j][k], k fastest

    // CPML
    bool has_pml = false;
    int npml[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    double pm_m = 0, pm_malp = 0, pm_kapmax = 0, pm_sig_ratio = 0, pm_alpmax = 0;
    std::vector<double> pbE[3][2], pcE[3][2], pbH[3][2], pcH[3][2];
    double cpmlE = 0, cpmlH = 0;

    // TF/SF
    bool has_tfsf = false;
    int tf_box[3][2] = {{0, 0}, {0, 0}, {0, 0}};  // i1 i2 / j1 j2 / k1 k2
    std::vector<TfsfTarget> tfE, tfH;
    std::vector<TfsfTerm> tf_terms;
    std::vector<double> Einc, Hinc;  // nt + 3 entries
    // one face alone (TF_SF::x1UpdateEField ... z2UpdateHField), [face][h],
    // face = 2 * axis + (high face): one target per corrected value, terms
    // in tf_face_terms
    std::vector<TfsfTarget> tfFace[6][2];
    std::vector<TfsfTerm> tf_face_terms;

    // Drude-Lorentz (cells with ID == 2)
    bool has_dl = false;
    int nD = 0, nL = 0;
    double dlC1 = 0, dlC2 = 0, dlC3 = 0, dl_epsInf = 0;
    std::vector<double> dlC4, dlC5, dlC6, dlKap, dlBeta, dlAlpha, dlXi, dlGam;
    std::vector<long long> dl_cells;  // flat [i][j][k] index

    // four-level gain (cells with ID == 5)
    bool has_gain = false;
    double g_lambda = 0, g_gamma = 0, g_t32 = 0, g_t21 = 0, g_t10 = 0, g_Ntot = 0, g_Wp = 0;
    double g_omega = 0, g_kappa = 0, gCPa1 = 0, gCPa2 = 0, gCPa3 = 0;
    std::vector<long long> gain_cells;

    // point sources
    std::vector<int> src;          // i j k comp
    std::vector<double> src_wave;  // [nsrc][nt+1]

    // ports
    std::vector<Port> ports;
    std::vector<double> freqs;
    int dft_stride = 1;
    bool record = false;
    std::vector<double> twE, twH;  // [nf][nt][2] cos, sin

    size_t cells() const { return (size_t)(Nx + 1) * (Ny + 1) * (Nz + 1); }
    size_t flat(int i, int j, int k) const { return ((size_t)i * (Ny + 1) + j) * (Nz + 1) + k; }
};

// Builds the model from a configuration; throws cfd::Error on bad input.
// The fields of a configuration read before any file is opened: required
// paths, each count against its array, port kinds, precision, periodic axes.
// Throws Error naming the first problem.
void check_config(const cfd_config &cfg);
Model build_model(const cfd_config &cfg);
// a TF/SF file's incident tables Einc, Hinc (entries 0 .. max(Nt, nt) + 2)
void tfsf_tables(const char *path, double dt, int nt, std::vector<double> &einc, std::vector<double> &hinc);

}  // namespace cfd

// the C ABI's opaque model handle (cfd_model_build, cfd_create_from_model)
struct cfd_model {
    cfd::Model m;
};
