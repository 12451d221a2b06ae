// model_capi.cpp -- host-only inspection of the set-up model (no device).
#include <cstring>
#include <memory>
#include <string>

#include "cpsfdtd.h"
#include "model.hpp"


namespace {
thread_local std::string g_model_error;
const std::vector<double> *find_table(const cfd::Model &m, const std::string &n, std::vector<double> &scratch)
{
    if (n == "aE") return &m.aE;
    if (n == "bE") return &m.bE;
    if (n == "aH") return &m.aH;
    if (n == "bH") return &m.bH;
    if (n == "Einc") return &m.Einc;
    if (n == "Hinc") return &m.Hinc;
    if (n == "src_wave") return &m.src_wave;  // [source][nt + 1]
    if (n.rfind("stretch", 0) == 0) {
        const int t = std::atoi(n.c_str() + 7);
        return (t >= 0 && t < cfd::kNumStretch) ? &m.stretch[t] : nullptr;
    }
    if (n.size() == 8 && n.rfind("pml_", 0) == 0) {  // pml_Ebx1
        const bool h = n[4] == 'H';
        const bool c = n[5] == 'c';
        const int a = n[6] - 'x', e = n[7] - '1';
        if (a < 0 || a > 2 || e < 0 || e > 1) return nullptr;
        return h ? (c ? &m.pcH[a][e] : &m.pbH[a][e]) : (c ? &m.pcE[a][e] : &m.pbE[a][e]);
    }
    if (n == "dl") {
        scratch = {m.dlC1, m.dlC2, m.dlC3};
        for (const auto *v : {&m.dlC4, &m.dlC5, &m.dlC6, &m.dlKap, &m.dlBeta, &m.dlAlpha, &m.dlXi, &m.dlGam})
            scratch.insert(scratch.end(), v->begin(), v->end());
        return &scratch;
    }
    return nullptr;
}
}  // namespace

extern "C" {

int cfd_model_build(const cfd_config *cfg, cfd_model **out)
{
    if (!cfg || !out) return -1;
    *out = nullptr;
    try {
        auto p = std::make_unique<cfd_model>();
        p->m = cfd::build_model(*cfg);
        *out = p.release();
        return 0;
    } catch (const std::exception &e) {
        g_model_error = e.what();
    }
    return -1;
}

const char *cfd_model_last_error(void) { return g_model_error.c_str(); }

int cfd_config_check(const cfd_config *cfg)
{
    if (!cfg) { g_model_error = "null configuration"; return -1; }
    try {
        cfd::check_config(*cfg);
        return 0;
    } catch (const std::exception &e) {
        g_model_error = e.what();
    }
    return -1;
}

int cfd_model_free(cfd_model *m)
{
    delete m;
    return 0;
}

int cfd_model_dims(cfd_model *p, int *n, double *d)
{
    if (!p) return -1;
    const cfd::Model &m = p->m;
    if (n) { n[0] = m.Nx; n[1] = m.Ny; n[2] = m.Nz; n[3] = m.nt; }
    if (d) { d[0] = m.dx; d[1] = m.dy; d[2] = m.dz; d[3] = m.dt; }
    return 0;
}

int cfd_model_id(cfd_model *p, int16_t *host)
{
    if (!p || !host) return -1;
    std::memcpy(host, p->m.id.data(), p->m.id.size() * sizeof(int16_t));
    return 0;
}

int cfd_model_table(cfd_model *p, const char *name, double *out)
{
    if (!p || !name) return -1;
    std::vector<double> scratch;
    const std::vector<double> *v = find_table(p->m, name, scratch);
    if (!v) return -1;
    if (out && !v->empty()) std::memcpy(out, v->data(), v->size() * sizeof(double));
    return (int)v->size();
}

int cfd_model_tfsf(cfd_model *p, int h, int *targets, double *terms, int *nterms)
{
    if (!p) return -1;
    const auto &tg = h ? p->m.tfH : p->m.tfE;
    if (nterms) *nterms = (int)p->m.tf_terms.size();
    if (targets)
        for (size_t t = 0; t < tg.size(); t++) {
            const cfd::TfsfTarget &g = tg[t];
            int *o = targets + 6 * t;
            o[0] = g.comp; o[1] = g.i; o[2] = g.j; o[3] = g.k; o[4] = g.first; o[5] = g.count;
        }
    if (terms)
        for (size_t t = 0; t < p->m.tf_terms.size(); t++) {
            terms[3 * t] = p->m.tf_terms[t].coef;
            terms[3 * t + 1] = p->m.tf_terms[t].delay;
            terms[3 * t + 2] = p->m.tf_terms[t].prof;
        }
    return (int)tg.size();
}

int cfd_model_port(cfd_model *p, int port, int *ijk)
{
    if (!p || port < 0 || port >= (int)p->m.ports.size()) return -1;
    const auto &v = p->m.ports[port].ijk;
    if (ijk) std::memcpy(ijk, v.data(), v.size() * sizeof(int));
    return (int)(v.size() / 3);
}

int cfd_model_cells(cfd_model *p, int gain, long long *flat)
{
    if (!p) return -1;
    const auto &v = gain ? p->m.gain_cells : p->m.dl_cells;
    if (flat) std::memcpy(flat, v.data(), v.size() * sizeof(long long));
    return (int)v.size();
}

int cfd_tfsf_tables(const char *tfsf, double dt, int nt, double *einc, double *hinc)
{
    try {
        std::vector<double> e, h;
        cfd::tfsf_tables(tfsf, dt, nt, e, h);
        if (einc) std::memcpy(einc, e.data(), e.size() * sizeof(double));
        if (hinc) std::memcpy(hinc, h.data(), h.size() * sizeof(double));
        return (int)e.size();
    } catch (const std::exception &ex) {
        g_model_error = ex.what();
        return -1;
    }
}

int cfd_dft_flux(const double *dft, int npts, int nf, int axis, double scale, double *out)
{
    if (!dft || !out || npts < 0 || nf < 0 || axis < 0 || axis > 2) {
        g_model_error = "cfd_dft_flux: bad argument";
        return -1;
    }
    // tangential components (a, b) with (a, b, axis) right-handed: S_axis = E_a H_b - E_b H_a
    const int a = (axis + 1) % 3, b = (axis + 2) % 3;
    for (int f = 0; f < nf; f++) {
        double acc = 0.0;
        for (int p = 0; p < npts; p++) {
            const double *v = dft + ((size_t)p * nf + f) * 12;  // [6][re, im]
            const double ear = v[2 * a], eai = v[2 * a + 1], ebr = v[2 * b], ebi = v[2 * b + 1];
            const double har = v[6 + 2 * a], hai = v[6 + 2 * a + 1], hbr = v[6 + 2 * b], hbi = v[6 + 2 * b + 1];
            // Re(E_a conj(H_b)) - Re(E_b conj(H_a))
            acc += (ear * hbr + eai * hbi) - (ebr * har + ebi * hai);
        }
        out[f] = scale * acc;
    }
    return 0;
}

}  // extern "C"
