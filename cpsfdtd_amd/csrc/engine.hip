// engine.hip -- device state of one simulation, the step sequence, and the C
// ABI declared in include/cpsfdtd.h.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "cpsfdtd.h"
#include "engine.hpp"
#include "kernels.hip"
#include "model.hpp"

namespace cfd {

static thread_local std::string g_error;

#define HIP_OK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) throw ::cfd::Error(std::string(#x) + ": " + hipGetErrorString(e_));          \
    } while (0)

// ----------------------------------------------------------------------------
// per-kernel timing with HIP events on the simulation's stream
// ----------------------------------------------------------------------------
// While the kernel timer brackets a launch, the first kernel it launches
// records the timer's start and stop events in its own dispatch
// (hipExtLaunchKernelGGL), so the time is the kernel's execution alone, not
// the queue's event packets around it; every kernel goes through CFD_LAUNCH.
struct LaunchEvents {
    hipEvent_t a = nullptr, b = nullptr;
    int n = 0;  // kernels launched while bracketed
};
inline thread_local LaunchEvents g_launch_ev;
#define CFD_LAUNCH(K, G, BL, SH, ST, ...)                                                     \
    do {                                                                                     \
        ::cfd::LaunchEvents &ev_ = ::cfd::g_launch_ev;                                       \
        if (ev_.a && ev_.n++ == 0) hipExtLaunchKernelGGL(K, G, BL, SH, ST, ev_.a, ev_.b, 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(K, G, BL, SH, ST, __VA_ARGS__);                              \
    } while (0)

struct KernelTimer {
    bool on = false;
    struct Pending { std::string name; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, int>> acc;

    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        return e;
    }
    template <class F>
    void run(hipStream_t s, const char *name, F &&launch) {
        if (!on) { launch(); return; }
        Pending p{name, get(), get()};
        HIP_OK(hipEventRecord(p.a, s));
        const LaunchEvents outer = g_launch_ev;
        g_launch_ev = LaunchEvents{p.a, p.b, 0};
        try {
            launch();
        } catch (...) {
            g_launch_ev = outer;
            throw;
        }
        const int n = g_launch_ev.n;
        g_launch_ev = outer;
        // one kernel: its dispatch recorded both events; several: from the
        // first one's start to after the last
        if (n != 1) HIP_OK(hipEventRecord(p.b, s));
        pending.push_back(p);
    }
    void collect() {
        for (auto &p : pending) {
            HIP_OK(hipEventSynchronize(p.b));
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, p.a, p.b));
            auto &a = acc[p.name];
            a.first += ms;
            a.second += 1;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    ~KernelTimer() {
        for (auto &p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// ----------------------------------------------------------------------------
// Held while a step block is captured into a graph and while a simulation
// allocates or releases device memory: stream-ordered allocations are refused
// while any stream of the process is capturing, and simulations are created
// and released on other threads (a serving loop) while others capture.
inline std::mutex &capture_mutex() {
    static std::mutex m;
    return m;
}

// Two pinned host staging buffers of kStageBytes each, allocated once per
// process (hipHostMalloc) and shared by every simulation's read-backs under
// stage_mutex(); pageable destinations copy from device memory at a few GB/s,
// pinned ones at the link's rate.
constexpr size_t kStageBytes = size_t(32) << 20;
inline std::mutex &stage_mutex() {
    static std::mutex m;
    return m;
}
inline double **pinned_stage() {  // call with stage_mutex() held
    static double *buf[2] = {nullptr, nullptr};
    for (double *&b : buf)
        if (!b && hipHostMalloc((void **)&b, kStageBytes, hipHostMallocDefault) != hipSuccess) {
            b = nullptr;
            throw Error("cannot allocate pinned staging memory");
        }
    return buf;
}

// FNV-1a-style signature over raw bytes, eight at a time (the tail byte by
// byte): a batch checks that its instances share a geometry
inline void sig_add(uint64_t &h, const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        h = (h ^ w) * 1099511628211ull;
    }
    for (; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
}
template <class T>
inline void sig_add(uint64_t &h, const std::vector<T> &v) {
    const size_t n = v.size();
    sig_add(h, &n, sizeof n);
    if (n) sig_add(h, v.data(), n * sizeof(T));
}

// Fields larger than this are streamed (non-temporal own-cell loads and
// stores in the update kernels); smaller field sets stay in the MI355X's
// 256 MiB Infinity Cache between passes, where plain accesses are faster.
constexpr size_t kStreamFieldBytes = size_t(256) << 20;
inline bool stream_fields(size_t field_bytes) {
    if (const char *v = std::getenv("CFD_STREAM_FIELDS")) return std::string(v) != "0";  // A/B override
    return field_bytes > kStreamFieldBytes;
}

// Merged half-step launch (k_full_h / k_full_e): cs = interior stretch as
// scalars, nt = streamed fields; B = batched instances (blockIdx.y).
template <class R, bool B>
void launch_full(hipStream_t s, bool h, bool cs, bool nt, unsigned ninst, const Dev<R> &d0, const Dev<R> *devs,
                 const MGrid &g, const Box &L, const LeanStretch<R> &st)
{
    const dim3 grid((unsigned)g.per_plane, (unsigned)g.planes, ninst), blk(256);
#define CFD_FULL(K, CS, NT) CFD_LAUNCH((K<R, CS, NT, B>), grid, blk, 0, s, d0, devs, g, L, st)
    if (h) {
        if (cs) { if (nt) CFD_FULL(k_full_h, true, true); else CFD_FULL(k_full_h, true, false); }
        else { if (nt) CFD_FULL(k_full_h, false, true); else CFD_FULL(k_full_h, false, false); }
    } else {
        if (cs) { if (nt) CFD_FULL(k_full_e, true, true); else CFD_FULL(k_full_e, true, false); }
        else { if (nt) CFD_FULL(k_full_e, false, true); else CFD_FULL(k_full_e, false, false); }
    }
#undef CFD_FULL
}

// Pole counts the special pass has a fixed-count instantiation for (the
// workloads' media: one Drude and two Lorentz poles, one Drude pole, or no
// Drude-Lorentz cells at all); any other counts take the generic form.
// poles = {nD, nL} of every instance with Drude-Lorentz cells, or {0, 0}
// when none has any; {-1, -1} when they differ.
template <class R, bool B>
void launch_special(hipStream_t s, dim3 grid, const Dev<R> &d0, const Dev<R> *devs, const long long *cells, int n,
                    const SpecialList *lists, const int (&poles)[2])
{
#define CFD_SPECIAL(ND, NL) CFD_LAUNCH((k_special_e<R, B, ND, NL>), grid, dim3(256), 0, s, d0, devs, cells, n, lists)
    if (poles[0] == 1 && poles[1] == 2) CFD_SPECIAL(1, 2);
    else if (poles[0] == 1 && poles[1] == 0) CFD_SPECIAL(1, 0);
    else if (poles[0] == 0 && poles[1] == 0) CFD_SPECIAL(0, 0);
    else CFD_SPECIAL(-1, -1);
#undef CFD_SPECIAL
}

// fold one instance into the pole counts above (start from {0, 0, 0}: the
// third entry records whether any instance had Drude-Lorentz cells)
inline void merge_poles(int (&poles)[3], long long ndl, int nD, int nL)
{
    if (!ndl || poles[0] < 0) return;
    if (nD + nL == 0) { poles[0] = poles[1] = -1; return; }  // no instantiation skips the ADE for them
    if (!poles[2]) { poles[0] = nD; poles[1] = nL; poles[2] = 1; }
    else if (poles[0] != nD || poles[1] != nL) poles[0] = poles[1] = -1;
}

class SimBase {
public:
    virtual ~SimBase() = default;
    Model m;
    int precision = 64;
    bool member = false;     // instance of a batch: stepped by the batch only
    bool packed = false;     // member of a pack: stepped by the pack only
    bool own_stream = true;
    uint64_t tf_sig = 14695981039346656037ull;    // TF/SF target lists, terms and waveforms
    uint64_t port_sig = 14695981039346656037ull;  // port cells, frequencies, recording
    hipStream_t stream = nullptr;
    int done = 0;
    size_t dev_bytes = 0;
    KernelTimer timer;

    virtual void step(int count) = 0;
    virtual void basic_h() = 0;
    virtual void basic_e() = 0;
    virtual void cpml_h() = 0;
    virtual void cpml_e() = 0;
    virtual void tfsf(bool h, int n) = 0;
    virtual void tfsf_face(bool h, int face, int n) = 0;
    virtual void e_parts(unsigned parts) = 0;
    virtual void sources(int n) = 0;
    virtual void ports(int n) = 0;
    virtual void read_field(int comp, double *host) = 0;
    virtual void write_field(int comp, const double *host) = 0;
    virtual void read_region(int comp, const int *box, double *host) = 0;
    virtual void port_samples(int q, double *host) = 0;
    virtual void port_gather(int q, double *host) = 0;
    // the same values staged through caller-given device memory (out: npts * 6
    // doubles); allocates nothing
    virtual void port_gather_via(int q, double *out, double *host) = 0;
    virtual int port_npts(int q) const = 0;
    virtual void port_dft(int q, double *host) = 0;
    virtual void port_intensity(int q, double *host) = 0;
    virtual void gain_read(double *host) = 0;
    virtual int dl_state(double *host) = 0;
    virtual int gain_state(double *host) = 0;
    virtual long long nonfinite() = 0;
};

template <class R>
class Sim : public SimBase {
public:
    Dev<R> d{};
    std::vector<void *> allocs;
    // TF/SF
    TfTarget *tfE = nullptr, *tfH = nullptr;
    int ntfE = 0, ntfH = 0, tablen = 0;
    double *tcoef = nullptr, *tdelay = nullptr, *tprof = nullptr, *einc = nullptr, *hinc = nullptr;
    // FP32: the same lists in compact form (tf32), used by every fused path
    bool tf32 = false;
    TfTarget32 *tfE32 = nullptr, *tfH32 = nullptr;
    TfTerm32 *tterm32 = nullptr;
    float *einc32 = nullptr, *hinc32 = nullptr;
    // one face alone, uploaded at its first use: targets [face][h], terms
    TfTarget *tfFaceT[6][2] = {};
    int ntfFace[6][2] = {};
    double *fcoef = nullptr, *fdelay = nullptr, *fprof = nullptr;
    // sources, grouped by (cell, component)
    SourceGroup *sgroup = nullptr;
    int nsgroup = 0;
    int *sidx = nullptr;
    double *swave = nullptr;
    // ports
    std::vector<int> dl_slot, gain_slot;  // reference-order cell -> device slot
    std::vector<PortDesc<R>> dports;  // host copy of the port table
    R *gather_buf = nullptr;          // [point][6] for cfd_port_gather
    int gather_max = 0;
    PortDesc<R> *port_tab = nullptr;
    int port_blocks = 0;
    static constexpr int kDftBatch = 16;  // sampled steps buffered per DFT pass
    // replayed step blocks
    bool use_graphs = true;
    int *nbase_dev = nullptr;         // block start read by the captured kernels
    int *count_dev = nullptr;         // the non-finite watchdog's counter
    const int *nbase_cur = nullptr;   // nbase_dev while capturing, else null (absolute steps)
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
    int dft_pending = 0, dft_first = 0;
    double *twE = nullptr, *twH = nullptr;

    // Stream-ordered allocations from the device's default pool, which keeps
    // freed memory for reuse: creating and releasing simulations (a serving
    // loop does it all the time) never synchronises the whole device.
    static void keep_pool_memory() {
        static std::once_flag once;
        std::call_once(once, [] {
            int dev = 0;
            hipMemPool_t pool;
            if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
                uint64_t keep = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
        });
    }
    template <class T>
    T *alloc(size_t count) {
        if (!count) count = 1;
        void *p = nullptr;
        keep_pool_memory();
        HIP_OK(hipMallocAsync(&p, count * sizeof(T), stream));
        HIP_OK(hipMemsetAsync(p, 0, count * sizeof(T), stream));
        allocs.push_back(p);
        dev_bytes += count * sizeof(T);
        return static_cast<T *>(p);
    }
    template <class T, class S>
    T *upload(const std::vector<S> &v) {
        std::vector<T> tmp(v.begin(), v.end());
        T *p = alloc<T>(tmp.size());
        if (!tmp.empty()) HIP_OK(hipMemcpyAsync(p, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return p;
    }
    // the FP32 compact lists (TfTarget32 / TfTerm32) from the exact ones, when
    // every index fits their fields
    void compact_tfsf(const std::vector<TfTarget> &te, const std::vector<TfTarget> &th, const std::vector<double> &c,
                      const std::vector<double> &dl, const std::vector<double> &pr) {
        if ((long long)d.slab * (m.Nz + 1) >= (1LL << 32) || c.size() >= (1u << 27)) return;
        std::vector<TfTarget32> e32, h32;
        for (const std::vector<TfTarget> *tv : {&te, &th})
            for (const TfTarget &t : *tv) {
                if (t.count > 3 || t.comp < 0 || t.comp > 5) return;
                (tv == &te ? e32 : h32).push_back({(unsigned)t.cell, (unsigned)t.comp << 29 | (unsigned)t.count << 27 | (unsigned)t.first});
            }
        std::vector<TfTerm32> terms;
        for (size_t q = 0; q < c.size(); q++) {
            const double base = std::ceil(dl[q]);
            terms.push_back({(float)(c[q] * pr[q]), (int)base, (float)(base - dl[q])});
        }
        tfE32 = upload<TfTarget32>(e32);
        tfH32 = upload<TfTarget32>(h32);
        tterm32 = upload<TfTerm32>(terms);
        einc32 = upload<float>(m.Einc);
        hinc32 = upload<float>(m.Hinc);
        tf32 = true;
    }
    long long dev_index(int i, int j, int k) const {
        return (long long)i + (long long)d.px * ((long long)j + (long long)d.py * k);
    }

    explicit Sim(Model &&model, int prec, hipStream_t shared = nullptr) {
        m = std::move(model);
        precision = prec;
        if (shared) {
            stream = shared;
            own_stream = false;
        } else {
            HIP_OK(hipStreamCreate(&stream));
        }
        try {
            std::lock_guard<std::mutex> lk(capture_mutex());
            setup();
        } catch (...) {
            release();
            throw;
        }
    }
    ~Sim() override { release(); }
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        std::lock_guard<std::mutex> lk(capture_mutex());
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        for (void *p : allocs) (void)hipFreeAsync(p, stream);
        allocs.clear();
        if (stream) (void)hipStreamSynchronize(stream);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    void setup() {
        // measured on MI355X (256^3, 12-cell CPML): merged is faster in both
        // precisions (FP64: 25.8 split -> 27.7 Gcell-updates/s merged, with
        // the tile flags and CPML pairs; split won before them)
        form = kFormMerged;
        if (const char *v = std::getenv("CFD_UPDATE_KERNEL")) {
            const std::string f(v);
            form = f == "grid" ? kFormGrid : f == "merged" ? kFormMerged : kFormSplit;
        }
        d.Nx = m.Nx; d.Ny = m.Ny; d.Nz = m.Nz;
        d.px = ((m.Nx + 1) + 31) / 32 * 32;
        d.py = m.Ny + 1;
        d.slab = (long long)d.px * d.py;
        d.pbx = m.pbx; d.pby = m.pby;
        d.ie0 = m.pbx ? 0 : 1;
        d.je0 = m.pby ? 0 : 1;
        const size_t ncell = (size_t)d.slab * (m.Nz + 1);
        for (int c = 0; c < 6; c++) d.F[c] = alloc<R>(ncell);
        nt_fields = stream_fields(6 * ncell * sizeof(R));
        // material IDs in device order: the reference's [i][j][k] order (k
        // fastest) transposed in 16 x 16 (i, k) blocks, planes of k blocks
        // shared among host threads
        {
            std::vector<uint8_t> id(ncell, 0);
            const int nk = m.Nz + 1, nkb = (nk + 15) / 16;
            host_parallel([&](int t, int nth) {
                for (int kb = t; kb < nkb; kb += nth)
                    for (int j = 0; j <= m.Ny; j++)
                        for (int i0 = 0; i0 <= m.Nx; i0 += 16)
                            for (int i = i0; i <= std::min(m.Nx, i0 + 15); i++)
                                for (int k = 16 * kb; k < std::min(nk, 16 * kb + 16); k++)
                                    id[dev_index(i, j, k)] = (uint8_t)m.id[m.flat(i, j, k)];
            });
            d.id = upload<uint8_t>(id);
        }
        d.aE = upload<R>(m.aE); d.bE = upload<R>(m.bE); d.aH = upload<R>(m.aH); d.bH = upload<R>(m.bH);
        d.nsel = m.NoMedia <= kSelMat ? 1 : 0;
        for (int q = 0; q < kSelMat; q++) {
            const int s = q < m.NoMedia ? q : 0;  // IDs >= NoMedia never occur
            d.selE[0][q] = (R)m.aE[s]; d.selE[1][q] = (R)m.bE[s];
            d.selH[0][q] = (R)m.aH[s]; d.selH[1][q] = (R)m.bH[s];
        }
        for (int t = 0; t < kNumStretch; t++) d.st[t] = upload<R>(m.stretch[t]);
        // CPML
        d.has_pml = m.has_pml;
        d.cpmlE = (R)m.cpmlE;
        d.cpmlH = (R)m.cpmlH;
        d.hsign_fixed = (m.flags & CFD_FIX_CPML_HSIGN) ? 1 : 0;
        const int N[3] = {m.Nx, m.Ny, m.Nz};
        for (int a = 0; a < 3; a++) {
            size_t box[2] = {0, 0};
            for (int e = 0; e < 2; e++) {
                const size_t n = (size_t)m.npml[a][e];
                d.npml[a][e] = (int)n;
                box[e] = a == 0 ? n * d.py * (N[2] + 1) : a == 1 ? (size_t)d.px * n * (N[2] + 1) : (size_t)d.slab * n;
            }
            if (!box[0] && !box[1]) continue;
            // low side then high side in one array of (b, c) pairs
            // (coefficients) or of psi pairs (boxes)
            auto pairs = [](const std::vector<double> &b0, const std::vector<double> &c0, const std::vector<double> &b1,
                            const std::vector<double> &c1) {
                std::vector<double> v;
                for (size_t q = 0; q < b0.size(); q++) v.insert(v.end(), {b0[q], c0[q]});
                for (size_t q = 0; q < b1.size(); q++) v.insert(v.end(), {b1[q], c1[q]});
                return v;
            };
            d.pbc[0][a] = reinterpret_cast<const Pair<R> *>(upload<R>(pairs(m.pbE[a][0], m.pcE[a][0], m.pbE[a][1], m.pcE[a][1])));
            d.pbc[1][a] = reinterpret_cast<const Pair<R> *>(upload<R>(pairs(m.pbH[a][0], m.pcH[a][0], m.pbH[a][1], m.pcH[a][1])));
            d.pofs[0][a] = (int)m.pbE[a][0].size();
            d.pofs[1][a] = (int)m.pbH[a][0].size();
            for (int h = 0; h < 2; h++) {
                d.qofs[h][a] = (long long)box[0];
                d.q[h][a] = alloc<Pair<R>>(box[0] + box[1]);
            }
        }
        // Drude-Lorentz and gain cells: slots numbered in device order (x fastest),
        // so the lanes of a wavefront touch consecutive state words
        // (the map is -1 elsewhere: filled and scattered on the device)
        if (m.has_dl || m.has_gain) {
            int *aux = alloc<int>(ncell);
            HIP_OK(hipMemsetAsync(aux, 0xff, ncell * sizeof(int), stream));
            auto assign = [&](const std::vector<long long> &cells, std::vector<int> &slot_of) {
                std::vector<std::pair<long long, int>> order;
                order.reserve(cells.size());
                for (size_t r = 0; r < cells.size(); r++) {
                    const size_t f = (size_t)cells[r];
                    const int i = (int)(f / ((size_t)(m.Ny + 1) * (m.Nz + 1)));
                    const int j = (int)((f / (m.Nz + 1)) % (m.Ny + 1)), k = (int)(f % (m.Nz + 1));
                    order.emplace_back(dev_index(i, j, k), (int)r);
                }
                std::sort(order.begin(), order.end());
                slot_of.assign(cells.size(), 0);
                std::vector<long long> at(order.size());
                for (size_t s2 = 0; s2 < order.size(); s2++) {
                    at[s2] = order[s2].first;
                    slot_of[order[s2].second] = (int)s2;
                }
                if (at.empty()) return;
                long long *dat = upload<long long>(at);
                CFD_LAUNCH(k_slot_scatter, dim3((unsigned)((at.size() + 255) / 256)), dim3(256), 0, stream, aux, dat,
                           (long long)at.size());
                HIP_OK(hipGetLastError());
            };
            assign(m.dl_cells, dl_slot);
            assign(m.gain_cells, gain_slot);
            d.aux = aux;
        }
        d.dt = (R)m.dt;
        if (m.has_dl) {
            d.nD = m.nD; d.nL = m.nL; d.ndl = (int)m.dl_cells.size();
            d.C1 = (R)m.dlC1; d.C2 = (R)m.dlC2; d.C3 = (R)m.dlC3;
            for (int q = 0; q < m.nD; q++) { d.C4[q] = (R)m.dlC4[q]; d.kap[q] = (R)m.dlKap[q]; d.beta[q] = (R)m.dlBeta[q]; }
            for (int q = 0; q < m.nL; q++) {
                d.C5[q] = (R)m.dlC5[q]; d.C6[q] = (R)m.dlC6[q]; d.alpha[q] = (R)m.dlAlpha[q]; d.xi[q] = (R)m.dlXi[q]; d.gam[q] = (R)m.dlGam[q];
            }
            d.dlst = alloc<R>((size_t)(m.nD + 2 * m.nL + 1) * 3 * std::max<size_t>(d.ndl, 1));
        }
        if (m.has_gain) {
            d.ngain = (int)m.gain_cells.size();
            d.CPa1 = (R)m.gCPa1; d.CPa2 = (R)m.gCPa2; d.g_CPa3 = m.gCPa3;
            d.g_dt = m.dt;
            d.g_hw = kHPlanck / (2.0 * kPi) * m.g_omega;
            d.g_Ntot = m.g_Ntot;
            d.g_h32 = m.dt / (2.0 * m.g_t32);
            d.g_h21 = m.dt / (2.0 * m.g_t21);
            d.g_h10 = m.dt / (2.0 * m.g_t10);
            d.g_dtWp = m.dt * m.g_Wp;
            const size_t ng = std::max<size_t>(d.ngain, 1);
            d.gP = alloc<R>(3 * ng); d.gP1 = alloc<R>(3 * ng); d.gEprev = alloc<R>(3 * ng);
            std::vector<double> N(4 * ng, 0.0);
            for (size_t s = 0; s < (size_t)d.ngain; s++) N[s] = m.g_Ntot;
            d.gN = upload<double>(N);
        }
        // TF/SF
        if (m.has_tfsf) {
            // one thread per corrected value; targets sorted by (component,
            // device address) so a wavefront's read-modify-writes are contiguous
            // (each target keeps its terms in the reference's face order)
            // (the E and H lists on two host threads, each sorted by a 64-bit
            // key: component above the device index, which is below 2^40)
            std::vector<TfTarget> te, th;
            host_parallel([&](int t, int) {
                if (t > 1) return;
                const std::vector<TfsfTarget> &src = t ? m.tfH : m.tfE;
                std::vector<std::pair<unsigned long long, unsigned>> key(src.size());
                for (size_t q = 0; q < src.size(); q++)
                    key[q] = {(unsigned long long)src[q].comp << 60 | (unsigned long long)dev_index(src[q].i, src[q].j, src[q].k),
                              (unsigned)q};
                std::sort(key.begin(), key.end());
                std::vector<TfTarget> &dst = t ? th : te;
                dst.reserve(src.size());
                for (const auto &kq : key) {
                    const TfsfTarget &g = src[kq.second];
                    dst.push_back({dev_index(g.i, g.j, g.k), g.comp, g.first, g.count, 0});
                }
            });
            // terms laid out in sorted-target order too
            std::vector<double> c, dl, pr;
            for (std::vector<TfTarget> *tv : {&te, &th})
                for (TfTarget &t : *tv) {
                    const int first = (int)c.size();
                    for (int q = t.first; q < t.first + t.count; q++) {
                        c.push_back(m.tf_terms[q].coef);
                        dl.push_back(m.tf_terms[q].delay);
                        pr.push_back(m.tf_terms[q].prof);
                    }
                    t.first = first;
                }
            ntfE = (int)te.size(); ntfH = (int)th.size();
            tfE = alloc<TfTarget>(te.size());
            tfH = alloc<TfTarget>(th.size());
            HIP_OK(hipMemcpyAsync(tfE, te.data(), te.size() * sizeof(TfTarget), hipMemcpyHostToDevice, stream));
            HIP_OK(hipMemcpyAsync(tfH, th.data(), th.size() * sizeof(TfTarget), hipMemcpyHostToDevice, stream));
            tcoef = upload<double>(c); tdelay = upload<double>(dl); tprof = upload<double>(pr);
            einc = upload<double>(m.Einc); hinc = upload<double>(m.Hinc);
            tablen = (int)m.Einc.size();
            sig_add(tf_sig, te); sig_add(tf_sig, th); sig_add(tf_sig, c); sig_add(tf_sig, dl); sig_add(tf_sig, pr);
            sig_add(tf_sig, m.Einc); sig_add(tf_sig, m.Hinc);
            if constexpr (std::is_same<R, float>::value) compact_tfsf(te, th, c, dl, pr);
        }
        // sources
        if (!m.src.empty()) {
            const int ns = (int)(m.src.size() / 4);
            std::vector<std::pair<long long, int>> key(ns);  // (cell * 3 + comp, source)
            for (int q = 0; q < ns; q++)
                key[q] = {dev_index(m.src[4 * q], m.src[4 * q + 1], m.src[4 * q + 2]) * 3 + m.src[4 * q + 3], q};
            std::stable_sort(key.begin(), key.end(),
                             [](const std::pair<long long, int> &a, const std::pair<long long, int> &b) { return a.first < b.first; });
            std::vector<SourceGroup> groups;
            std::vector<int> idx;
            for (int r = 0; r < ns; r++) {
                if (r == 0 || key[r].first != key[r - 1].first)
                    groups.push_back({key[r].first / 3, (int)(key[r].first % 3), r, 0, 0});
                groups.back().count++;
                idx.push_back(key[r].second);
            }
            nsgroup = (int)groups.size();
            sgroup = alloc<SourceGroup>(groups.size());
            HIP_OK(hipMemcpyAsync(sgroup, groups.data(), groups.size() * sizeof(SourceGroup), hipMemcpyHostToDevice, stream));
            sidx = upload<int>(idx);
            swave = upload<double>(m.src_wave);
        }
        // ports
        if (!m.freqs.empty()) { twE = upload<double>(m.twE); twH = upload<double>(m.twH); }
        for (const Port &P : m.ports) {
            PortDesc<R> dp{};
            std::vector<long long> cells;
            for (size_t q = 0; q < P.ijk.size(); q += 3) cells.push_back(dev_index(P.ijk[q], P.ijk[q + 1], P.ijk[q + 2]));
            dp.npts = P.npts();
            dp.nblk = (dp.npts + 255) / 256;
            dp.bstart = port_blocks;
            port_blocks += dp.nblk;
            dp.cell = upload<long long>(cells);
            dp.samples = m.record ? alloc<R>((size_t)m.nt * dp.npts * 6) : nullptr;
            dp.batch = m.freqs.empty() ? nullptr : alloc<R>((size_t)kDftBatch * 6 * dp.npts);
            dp.dft = m.freqs.empty() ? nullptr : alloc<double>((size_t)dp.npts * m.freqs.size() * 12);
            dp.ipart = alloc<double>((size_t)m.nt * std::max(dp.nblk, 1));
            dports.push_back(dp);
            gather_max = std::max(gather_max, dp.npts);
            sig_add(port_sig, P.ijk);
        }
        sig_add(port_sig, m.freqs);
        sig_add(port_sig, &m.dft_stride, sizeof m.dft_stride);
        sig_add(port_sig, &m.record, sizeof m.record);
        if (gather_max) gather_buf = alloc<R>((size_t)gather_max * 6);
        setup_passes();
        nbase_dev = alloc<int>(1);
        count_dev = alloc<int>(1);
        if (const char *v = std::getenv("CFD_GRAPHS")) use_graphs = std::string(v) != "0";
        if (!dports.empty()) {
            port_tab = alloc<PortDesc<R>>(dports.size());
            HIP_OK(hipMemcpyAsync(port_tab, dports.data(), dports.size() * sizeof(PortDesc<R>), hipMemcpyHostToDevice, stream));
        }
        setup_post();
        HIP_OK(hipStreamSynchronize(stream));
    }

    TileGrid tiles() const {
        TileGrid g;
        g.gx = (m.Nx + TX - 1) / TX;
        g.gy = (m.Ny + TY - 1) / TY;
        g.gz = m.Nz;
        return g;
    }
    // Tiles of a merged half-step (see MGrid): 32 x 8 tiles in the x-edge
    // columns that hold the x CPML slabs, 64 x 4 tiles between them; one
    // plane's tile origins, the 64 x 4 tiles first (ordered as below), then
    // the edge tiles, x fastest.
    // CFD_TILE_EDGE32=0 gives 64-wide tiles everywhere (A/B runs).
    // y_org: the first tile row (the E pass starts at its first updated row,
    // j = 1 without y periodicity: row 0 would idle a quarter of the first
    // row of tiles; measured 142.5 -> 139.6 us for k_full_e on 256^3)
    // efield, L: the half-step the table is for and its interior box (the
    // flags of each tile, kTile*)
    std::vector<MTileEntry> mgrid_tiles(int y_org, bool efield, const Box &L) const {
        const char *v = std::getenv("CFD_TILE_EDGE32");
        const bool edge32 = !v || std::string(v) != "0";
        const int n0 = edge32 ? m.npml[0][0] : 0, n1 = edge32 ? m.npml[0][1] : 0;
        const int xa = n0 ? (n0 + 1 + 31) / 32 * 32 : 0;            // past the low slab (E cells up to n0)
        const int wide = std::max(0, (n1 ? m.Nx - n1 : m.Nx) - xa);  // up to the high slab
        const int nw = n1 ? wide / 64 : (wide + 63) / 64;
        const int xb = xa + 64 * nw;
        if (m.Nx > 0x7fff || m.Ny > 0xffff) throw Error("grid too large for the merged tile table");
        const int o = efield ? 1 : 0;  // E slabs start at index 1 on the low side
        auto entry = [&](int x0, int y0, bool narrow) {
            const int tw = narrow ? 32 : 64, th = narrow ? 8 : 4;
            unsigned f = 0;
            if ((m.npml[0][0] && x0 < m.npml[0][0] + o) || (m.npml[0][1] && x0 + tw > m.Nx - m.npml[0][1])) f |= kTileSlabX;
            if ((m.npml[1][0] && y0 < m.npml[1][0] + o) || (m.npml[1][1] && y0 + th > m.Ny - m.npml[1][1])) f |= kTileSlabY;
            if (x0 >= L.i0 && x0 + tw <= L.i1 && y0 >= L.j0 && y0 + th <= L.j1) f |= kTileInnerXY;
            return MTileEntry{(unsigned)x0 | (unsigned)y0 << 15 | (narrow ? 1u << 31 : 0u), f};
        };
        // 64 x 4 tiles by groups of eight tile rows, each group's tiles x
        // major: workgroups go to the XCDs in turn, so every tile of a row
        // then runs on the same XCD and finds the neighbour line of the tile
        // beside it (i -+ 1) in that XCD's L2 (256^3 FP32: both merged kernels
        // 1.3 % faster); with a multiple of 8 tiles per plane the tile of the
        // plane below (k -+ 1) runs there too
        std::vector<MTileEntry> t;
        const int nrow = (m.Ny - y_org + 3) / 4;
        for (int g = 0; g < nrow; g += 8)
            for (int b = 0; b < nw; b++)
                for (int r = g; r < g + 8 && r < nrow; r++) t.push_back(entry(xa + 64 * b, y_org + 4 * r, false));
        for (int y0 = y_org; y0 < m.Ny; y0 += 8) {
            for (int x0 = 0; x0 < xa; x0 += 32) t.push_back(entry(x0, y0, true));
            for (int x0 = xb; x0 < m.Nx; x0 += 32) t.push_back(entry(x0, y0, true));
        }
        return t;
    }
    MGrid mgH{}, mgE{};

    // Fused-step update forms (CFD_UPDATE_KERNEL selects one, for A/B checks):
    //   split             lean interior box, CPML shell boxes, special cells
    //   merged            one launch per half-step over all tiles, lean stencil
    //                     on interior tiles, CPML terms on the others, plus the
    //                     special pass for Drude-Lorentz / gain cells (default)
    //   grid              the single full-physics grid kernel per half-step
    enum { kFormMerged, kFormSplit, kFormGrid };
    int form = kFormMerged;
    Box leanH{}, leanE{};
    std::vector<Box> shellH, shellE;
    BoxSet shellH_set{}, shellE_set{};
    int shellH_blocks = 0, shellE_blocks = 0;
    LeanStretch<R> leanH_st{}, leanE_st{};
    bool leanH_cs = false, leanE_cs = false;
    long long *special = nullptr;
    int nspecial = 0;
    int spoles[2] = {-1, -1};  // special pass instantiation (launch_special)
    // the passes after the E update in one launch (k_post_e), when their
    // cells allow it (post_fusable)
    // post_srcs / post_ports: the sources / the port sampling are in it
    bool post_fused = false, post_srcs = false, post_ports = false;
    PostE<R> post{};
    int post_blocks = 0;
    bool nt_fields = false;  // streamed own-cell field traffic (see stream_fields)

    static Box make_box(int i0, int i1, int j0, int j1, int k0, int k1, int tw) {
        Box b{};
        b.i0 = i0; b.i1 = i1; b.j0 = j0; b.j1 = j1; b.k0 = k0; b.k1 = k1;
        b.tw = tw; b.th = 256 / tw;
        b.xbase = (i0 / tw) * tw;
        b.gx = (i1 - b.xbase + tw - 1) / tw;
        b.gy = (j1 - j0 + b.th - 1) / b.th;
        return b;
    }
    static bool empty(const Box &b) { return b.i1 <= b.i0 || b.j1 <= b.j0 || b.k1 <= b.k0; }
    static long long blocks(const Box &b) { return empty(b) ? 0 : (long long)b.gx * b.gy * (b.k1 - b.k0); }

    // cell classes of one half-step: interior box and shell boxes.  lo/hi give the
    // update range, a/b the first and last interior index per axis.
    void classify(bool efield, Box &lean, std::vector<Box> &shell) {
        const int N[3] = {m.Nx, m.Ny, m.Nz};
        int lo[3], a[3], b[3];
        for (int ax = 0; ax < 3; ax++) {
            const int n1 = m.npml[ax][0], n2 = m.npml[ax][1];
            lo[ax] = efield ? (ax == 0 ? d.ie0 : ax == 1 ? d.je0 : 1) : 0;
            a[ax] = n1 ? (efield ? n1 + 1 : n1) : lo[ax];
            b[ax] = N[ax] - n2;
        }
        lean = make_box(a[0], b[0], a[1], b[1], a[2], b[2], TX);
        auto add = [&](int i0, int i1, int j0, int j1, int k0, int k1, int pml) {
            const int tw = (i1 - i0 <= 16) ? 16 : TX;
            Box bx = make_box(i0, i1, j0, j1, k0, k1, tw);
            bx.pml = pml;
            if (!empty(bx)) shell.push_back(bx);
        };
        // a box spanning the whole range of an axis may touch that axis's slabs
        const int mx = m.npml[0][0] || m.npml[0][1] ? 1 : 0, my = m.npml[1][0] || m.npml[1][1] ? 2 : 0;
        add(lo[0], N[0], lo[1], N[1], lo[2], a[2], 4 | mx | my);  // z low
        add(lo[0], N[0], lo[1], N[1], b[2], N[2], 4 | mx | my);   // z high
        add(lo[0], N[0], lo[1], a[1], a[2], b[2], 2 | mx);        // y low
        add(lo[0], N[0], b[1], N[1], a[2], b[2], 2 | mx);         // y high
        add(lo[0], a[0], a[1], b[1], a[2], b[2], 1);              // x low
        add(b[0], N[0], a[1], b[1], a[2], b[2], 1);               // x high
    }

    // Are the stretch tables the lean kernel of one half-step reads constant
    // (bitwise, in R) over the interior box?  Then their values go in c.
    bool lean_constant(const Box &L, bool efield, LeanStretch<R> &c) const {
        if (empty(L)) return false;
        const int lo[3] = {L.i0, L.j0, L.k0}, hi[3] = {L.i1, L.j1, L.k1};
        // axis each table is indexed by: cExy, cEyx by k; cEzx, cExz by j; cEzy, cEyz by i
        static const int axis_of[6] = {2, 2, 1, 1, 0, 0};
        for (int t = 0; t < 6; t++) {
            const int tt = t + (efield ? 0 : 6);
            const std::vector<double> &tab = m.stretch[tt];
            const int ax = axis_of[t];
            if ((int)tab.size() < hi[ax]) return false;
            const R v0 = (R)tab[lo[ax]];
            for (int q = lo[ax]; q < hi[ax]; q++)
                if (!((R)tab[q] == v0) || std::signbit((R)tab[q]) != std::signbit(v0)) return false;
            c.s[tt] = v0;
        }
        return true;
    }

    void setup_passes() {
        classify(false, leanH, shellH);
        classify(true, leanE, shellE);
        auto box_set = [&](std::vector<Box> &v, BoxSet &set, int &nblk) {
            if (v.size() > (size_t)kMaxBoxes) throw Error("internal: too many shell boxes");
            nblk = 0;
            set = BoxSet{};
            for (Box &b : v) { b.bstart = nblk; nblk += (int)blocks(b); }
            for (size_t q = 0; q < v.size(); q++) set.b[q] = v[q];
            set.n = (int)v.size();
        };
        box_set(shellH, shellH_set, shellH_blocks);
        box_set(shellE, shellE_set, shellE_blocks);
        leanH_cs = lean_constant(leanH, false, leanH_st);
        leanE_cs = lean_constant(leanE, true, leanE_st);
        {
            const std::vector<MTileEntry> t = mgrid_tiles(0, false, leanH), te = mgrid_tiles(m.pby ? 0 : 1, true, leanE);
            mgH = MGrid{upload<MTileEntry>(t), (int)t.size(), 0, m.Nz};
            mgE = MGrid{upload<MTileEntry>(te), (int)te.size(), 1, std::max(0, m.Nz - 1)};
        }
        // every Drude-Lorentz / gain cell, in device order; cells outside the
        // interior box carry the CPML flag
        std::vector<long long> sp;
        auto interior = [&](int i, int j, int k) {
            return i >= leanE.i0 && i < leanE.i1 && j >= leanE.j0 && j < leanE.j1 && k >= leanE.k0 && k < leanE.k1;
        };
        for (const std::vector<long long> *cells : {&m.dl_cells, &m.gain_cells})
            for (const long long f : *cells) {
                const long long plane = (long long)(m.Ny + 1) * (m.Nz + 1);
                const int i = (int)(f / plane), j = (int)((f / (m.Nz + 1)) % (m.Ny + 1)), k = (int)(f % (m.Nz + 1));
                sp.push_back(dev_index(i, j, k) | (interior(i, j, k) ? 0LL : 1LL << kSpecialPml));
            }
        const long long mask = (1LL << kSpecialPml) - 1;
        std::sort(sp.begin(), sp.end(), [&](long long a, long long b) { return (a & mask) < (b & mask); });
        nspecial = (int)sp.size();
        {
            int pl[3] = {0, 0, 0};
            merge_poles(pl, d.ndl, d.nD, d.nL);
            spoles[0] = pl[0];
            spoles[1] = pl[1];
        }
        if (nspecial) special = upload<long long>(sp);
        HIP_OK(hipStreamSynchronize(stream));
    }

    // h: the H half-step, else E (and, unless the post-E launch does them,
    // the special cells)
    void launch_update(bool h) {
        if (form == kFormGrid) {
            if (h) launch_h(H_ALL, "update_h");
            else launch_e(E_ALL, "update_e");
            return;
        }
        const Box &L = h ? leanH : leanE;
        if (form == kFormMerged) {
            const MGrid &g = h ? mgH : mgE;
            if (g.per_plane && g.planes) {
                timer.run(stream, h ? "full_h" : "full_e", [&] {
                    launch_full<R, false>(stream, h, h ? leanH_cs : leanE_cs, nt_fields, 1, d, nullptr, g, L,
                                          h ? leanH_st : leanE_st);
                });
                HIP_OK(hipGetLastError());
            }
            if (!h && nspecial && !post_fused) {
                timer.run(stream, "special_e", [&] {
                    launch_special<R, false>(stream, dim3((nspecial + 255) / 256), d, nullptr, special, nspecial, nullptr,
                                             spoles);
                });
                HIP_OK(hipGetLastError());
            }
            return;
        }
        const long long nl = blocks(L);
        if (nl) {
            const int skip5 = m.has_gain ? 1 : 0;
            timer.run(stream, h ? "lean_h" : "lean_e", [&] {
                const dim3 grid((unsigned)nl), blk(TX, TY);
#define CFD_LEAN_H(CS, NT) CFD_LAUNCH((k_lean_h<R, CS, NT>), grid, blk, 0, stream, d, L, leanH_st)
#define CFD_LEAN_E(CS, NT) CFD_LAUNCH((k_lean_e<R, CS, NT>), grid, blk, 0, stream, d, L, leanE_st, skip5)
                if (h) {
                    if (leanH_cs) { if (nt_fields) CFD_LEAN_H(true, true); else CFD_LEAN_H(true, false); }
                    else { if (nt_fields) CFD_LEAN_H(false, true); else CFD_LEAN_H(false, false); }
                } else {
                    if (leanE_cs) { if (nt_fields) CFD_LEAN_E(true, true); else CFD_LEAN_E(true, false); }
                    else { if (nt_fields) CFD_LEAN_E(false, true); else CFD_LEAN_E(false, false); }
                }
#undef CFD_LEAN_H
#undef CFD_LEAN_E
            });
            HIP_OK(hipGetLastError());
        }
        const int nb = h ? shellH_blocks : shellE_blocks;
        if (nb) {
            timer.run(stream, h ? "shell_h" : "shell_e", [&] {
                if (h && nt_fields) CFD_LAUNCH((k_shell_h<R, true>), dim3(nb), dim3(TX, TY), 0, stream, d, shellH_set);
                else if (h) CFD_LAUNCH((k_shell_h<R, false>), dim3(nb), dim3(TX, TY), 0, stream, d, shellH_set);
                else if (nt_fields) CFD_LAUNCH((k_shell_e<R, true>), dim3(nb), dim3(TX, TY), 0, stream, d, shellE_set);
                else CFD_LAUNCH((k_shell_e<R, false>), dim3(nb), dim3(TX, TY), 0, stream, d, shellE_set);
            });
            HIP_OK(hipGetLastError());
        }
        if (!h && nspecial && !post_fused) {
            timer.run(stream, "special_e", [&] {
                launch_special<R, false>(stream, dim3((nspecial + 255) / 256), d, nullptr, special, nspecial, nullptr,
                                         spoles);
            });
            HIP_OK(hipGetLastError());
        }
    }

    void launch_h(unsigned parts, const char *name) {
        const TileGrid g = tiles();
        timer.run(stream, name, [&] { CFD_LAUNCH(k_update_h<R>, dim3(g.gx * g.gy * g.gz), dim3(TX, TY), 0, stream, d, parts, g); });
        HIP_OK(hipGetLastError());
    }
    void launch_e(unsigned parts, const char *name) {
        const TileGrid g = tiles();
        timer.run(stream, name, [&] { CFD_LAUNCH(k_update_e<R>, dim3(g.gx * g.gy * g.gz), dim3(TX, TY), 0, stream, d, parts, g); });
        HIP_OK(hipGetLastError());
    }
    void tfsf(bool h, int n) override {
        if (!m.has_tfsf) return;
        const int nt = h ? ntfH : ntfE;
        if (!nt) return;
        timer.run(stream, h ? "tfsf_h" : "tfsf_e", [&] {
            if constexpr (std::is_same<R, float>::value) {
                if (tf32) {
                    CFD_LAUNCH((k_tfsf32<false>), dim3((nt + 255) / 256), dim3(256), 0, stream, d, nullptr,
                                       h ? tfH32 : tfE32, nt, tterm32, h ? einc32 : hinc32, tablen, nbase_cur, n);
                    return;
                }
            }
            CFD_LAUNCH((k_tfsf<R>), dim3((nt + 255) / 256), dim3(256), 0, stream, d, nullptr, h ? tfH : tfE, nt,
                               tcoef, tdelay, tprof, h ? einc : hinc, tablen, nbase_cur, n);
        });
        HIP_OK(hipGetLastError());
    }
    // TF_SF::{x,y,z}{1,2}Update{E,H}Field: the corrections of one face of
    // the box (face = 2 * axis + high), each value corrected by one thread
    void tfsf_face(bool h, int face, int n) override {
        if (!m.has_tfsf) throw Error("no TF/SF set-up");
        if (face < 0 || face > 5) throw Error("TF/SF face must be 0..5 (x1 x2 y1 y2 z1 z2)");
        if (!fcoef) {
            std::vector<double> c, dl, pr;
            for (const TfsfTerm &t : m.tf_face_terms) { c.push_back(t.coef); dl.push_back(t.delay); pr.push_back(t.prof); }
            for (int f = 0; f < 6; f++)
                for (int hh = 0; hh < 2; hh++) {
                    std::vector<TfTarget> tv;
                    std::set<std::pair<int, long long>> seen;
                    for (const TfsfTarget &t : m.tfFace[f][hh]) {
                        tv.push_back({dev_index(t.i, t.j, t.k), t.comp, t.first, t.count, 0});
                        if (!seen.insert({t.comp, tv.back().cell}).second) throw Error("TF/SF face lists a value twice");
                    }
                    ntfFace[f][hh] = (int)tv.size();
                    tfFaceT[f][hh] = tv.empty() ? nullptr : upload<TfTarget>(tv);
                }
            fdelay = upload<double>(dl); fprof = upload<double>(pr);
            fcoef = upload<double>(c);
        }
        const int nt = ntfFace[face][h];
        if (!nt) return;
        timer.run(stream, h ? "tfsf_face_h" : "tfsf_face_e", [&] {
            CFD_LAUNCH((k_tfsf<R>), dim3((nt + 255) / 256), dim3(256), 0, stream, d, nullptr, tfFaceT[face][h], nt,
                               fcoef, fdelay, fprof, h ? einc : hinc, tablen, nullptr, n);
        });
        HIP_OK(hipGetLastError());
    }
    void sources(int n) override {
        if (!nsgroup) return;
        if (!nbase_cur && (n < 0 || n > m.nt)) throw Error("source time index out of range");
        timer.run(stream, "sources", [&] {
            CFD_LAUNCH(k_sources<R>, dim3((nsgroup + 63) / 64), dim3(64), 0, stream, d, sgroup, nsgroup, sidx,
                               swave, m.nt + 1, nbase_cur, n);
        });
        HIP_OK(hipGetLastError());
    }
    void ports(int n) override {
        if (n < 0 || n >= m.nt) throw Error("port time index out of range");
        if (dports.empty()) return;
        int slot = -1;
        if (!m.freqs.empty() && n % m.dft_stride == 0) {
            if (dft_pending == 0) dft_first = n;
            slot = dft_pending++;
        }
        timer.run(stream, "ports", [&] {
            CFD_LAUNCH((k_port<R>), dim3(port_blocks), dim3(256), 0, stream, d, nullptr, port_tab, (int)dports.size(),
                               nbase_cur, n, slot);
        });
        HIP_OK(hipGetLastError());
        if (dft_pending == kDftBatch) dft_flush();
    }
    void dft_flush() {
        if (!dft_pending) return;
        const int nb = dft_pending;
        dft_pending = 0;
        timer.run(stream, "port_dft", [&] {
            CFD_LAUNCH((k_dft<R>), dim3(4 * port_blocks), dim3(256), 0, stream, port_tab, (int)dports.size(), nb,
                               nbase_cur, dft_first, m.dft_stride, twE, twH, (int)m.freqs.size(), m.nt);
        });
        HIP_OK(hipGetLastError());
    }
    void basic_h() override { launch_h(H_BASIC, "basic_h"); }
    void cpml_h() override { if (m.has_pml) launch_h(H_PML, "cpml_h"); }
    void basic_e() override { launch_e(E_BASIC, "basic_e"); }
    void cpml_e() override { if (m.has_pml) launch_e(E_PML, "cpml_e"); }
    void e_parts(unsigned parts) override { launch_e(parts, "e_module"); }

    // main.cpp:267-343: one fused H pass (basic + CPML), TF/SF-H, one fused E
    // pass (gain polarisation, basic, Drude-Lorentz, gain current and
    // populations, CPML), TF/SF-E, sources, ports
    void one_step(int n) {
        launch_update(true);
        tfsf(true, n);
        launch_update(false);
        if (post_fused) {
            post_e(n);
            return;
        }
        tfsf(false, n);
        sources(n);
        if (!dports.empty()) ports(n);
    }
    // special cells, TF/SF-E, (post_srcs) sources and (post_ports) the port
    // sampling of step n in one launch (k_post_e), with the port bookkeeping
    // of ports(); the passes left out follow as their own launches, in the
    // reference's order
    void post_e(int n) {
        if (post_ports && (n < 0 || n >= m.nt)) throw Error("port time index out of range");
        if (post_srcs && nsgroup && !nbase_cur && (n < 0 || n > m.nt)) throw Error("source time index out of range");
        PostE<R> P = post;
        P.slot = -1;
        if (post_ports && !m.freqs.empty() && n % m.dft_stride == 0) {
            if (dft_pending == 0) dft_first = n;
            P.slot = dft_pending++;
        }
        timer.run(stream, "post_e", [&] {
#define CFD_POST(ND, NL) CFD_LAUNCH((k_post_e<R, ND, NL>), dim3(post_blocks), dim3(256), 0, stream, d, P, nbase_cur, n)
            if (!P.nspecial) CFD_POST(-2, -2);
            else if (spoles[0] == 1 && spoles[1] == 2) CFD_POST(1, 2);
            else if (spoles[0] == 1 && spoles[1] == 0) CFD_POST(1, 0);
            else if (spoles[0] == 0 && spoles[1] == 0) CFD_POST(0, 0);
            else CFD_POST(-1, -1);
#undef CFD_POST
        });
        HIP_OK(hipGetLastError());
        if (!post_srcs) sources(n);
        if (!post_ports) {
            if (!dports.empty()) ports(n);
            return;
        }
        if (dft_pending == kDftBatch) dft_flush();
    }
    // the post-E launch's tables and workgroup ranges, when post_fusable
    // allows it (CFD_FUSE_POST=0: separate launches, for A/B runs)
    void setup_post() {
        post_fused = post_srcs = post_ports = false;
        if (const char *v = std::getenv("CFD_FUSE_POST"))
            if (std::string(v) == "0") return;
        std::vector<long long> te, sp;
        if (m.has_tfsf)
            for (const TfsfTarget &t : m.tfE) te.push_back(dev_index(t.i, t.j, t.k));
        const long long plane = (long long)(m.Ny + 1) * (m.Nz + 1);
        for (const std::vector<long long> *cells : {&m.dl_cells, &m.gain_cells})
            for (const long long f : *cells)
                sp.push_back(dev_index((int)(f / plane), (int)((f / (m.Nz + 1)) % (m.Ny + 1)), (int)(f % (m.Nz + 1))));
        std::vector<std::vector<long long>> pcs;
        for (const Port &P : m.ports) {
            pcs.emplace_back();
            for (size_t q = 0; q < P.ijk.size(); q += 3) pcs.back().push_back(dev_index(P.ijk[q], P.ijk[q + 1], P.ijk[q + 2]));
        }
        const int parts = post_fusable(te, sp, pcs);
        if (!(parts & 1)) return;
        post_srcs = (parts & 2) != 0;
        post_ports = (parts & 4) && !dports.empty();
        PostE<R> P{};
        P.special = special;
        P.nspecial = form == kFormGrid ? 0 : nspecial;
        P.b1 = (P.nspecial + 255) / 256;
        P.tg = tfE;
        if constexpr (std::is_same<R, float>::value) P.tg32 = tf32 ? tfE32 : nullptr;
        P.ntg = m.has_tfsf ? ntfE : 0;
        P.b2 = P.b1 + (P.ntg + 255) / 256;
        P.coef = tcoef; P.delay = tdelay; P.prof = tprof; P.tab = hinc;
        if constexpr (std::is_same<R, float>::value) { P.terms32 = tterm32; P.tab32 = hinc32; }
        P.tablen = tablen;
        P.grp = sgroup;
        P.ngroups = post_srcs ? nsgroup : 0;
        P.b3 = P.b2 + (P.ngroups + 255) / 256;
        P.sidx = sidx;
        P.swave = swave;
        P.sstride = m.nt + 1;
        P.ports = port_tab;
        P.nports = post_ports ? (int)dports.size() : 0;
        post = P;
        post_blocks = P.b3 + (post_ports ? port_blocks : 0);
        // one launch only where it replaces two or more: a pass alone keeps
        // its own kernel (and register budget)
        const int roles = (P.nspecial > 0) + (P.ntg > 0) + (P.ngroups > 0) + (post_ports ? 1 : 0);
        post_fused = roles >= 2;
        if (!post_fused) post_srcs = post_ports = false;
    }
    // Which post-E passes may run side by side (bit mask; cell-level test,
    // conservative)?  Bit 0: the special cells and the TF/SF-E targets are
    // distinct cells (else no fused launch); bit 1: the sources' cells are
    // neither, so they join (else they follow as their own launch); bit 2:
    // the sources joined and the ports sample none of those cells, so the
    // port sampling joins too (else it follows as its own launch).
    // te: TF/SF-E target cells; sp: special cells (flag bits included)
    int post_fusable(std::vector<long long> te, const std::vector<long long> &sp,
                     const std::vector<std::vector<long long>> &port_cells) const {
        const long long mask = (1LL << kSpecialPml) - 1;
        std::vector<long long> a, c, pc;
        if (form != kFormGrid)
            for (long long v : sp) a.push_back(v & mask);
        for (int q = 0; q < (int)(m.src.size() / 4); q++) c.push_back(dev_index(m.src[4 * q], m.src[4 * q + 1], m.src[4 * q + 2]));
        for (const std::vector<long long> &v : port_cells) pc.insert(pc.end(), v.begin(), v.end());
        for (std::vector<long long> *v : {&a, &te, &c, &pc}) {
            std::sort(v->begin(), v->end());
            v->erase(std::unique(v->begin(), v->end()), v->end());
        }
        auto meet = [](const std::vector<long long> &x, const std::vector<long long> &y) {
            size_t i = 0, j = 0;
            while (i < x.size() && j < y.size()) {
                if (x[i] == y[j]) return true;
                if (x[i] < y[j]) i++; else j++;
            }
            return false;
        };
        if (meet(a, te)) return 0;
        if (meet(c, a) || meet(c, te)) return 1;
        return meet(pc, a) || meet(pc, te) || meet(pc, c) ? 3 : 7;
    }

    // Step blocks replayed from a HIP graph: a block of graph_block() steps
    // starting at a multiple of its length is captured once, its
    // step-dependent kernels reading the block start from *nbase_dev and
    // their offset from the captured arguments.  The block length is a
    // multiple of dft_stride x kDftBatch, so the DFT staging pattern (slots,
    // flushes) is the same in every block.  The kernels and arguments are
    // those of the eager path, so replays are bit-identical to it.
    int graph_block() const {
        if (!use_graphs) return 0;
        if (m.freqs.empty()) return 128;
        const int unit = m.dft_stride * kDftBatch;
        return unit >= 128 ? unit : unit * (128 / unit);
    }
    void capture_block(int G) {
        const bool timing = timer.on;
        timer.on = false;
        nbase_cur = nbase_dev;
        hipGraph_t g = nullptr;
        std::lock_guard<std::mutex> lk(capture_mutex());
        HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        try {
            for (int i = 0; i < G; i++) one_step(i);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            nbase_cur = nullptr;
            timer.on = timing;
            throw;
        }
        nbase_cur = nullptr;
        timer.on = timing;
        HIP_OK(hipStreamEndCapture(stream, &g));
        const hipError_t e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_OK(e);
        graph_steps = G;
    }
    void step(int count) override {
        while (count > 0) {
            const int n = done;
            if (n >= m.nt) throw Error("run finished: all " + std::to_string(m.nt) + " steps done");
            const int G = graph_block();
            if (G && count >= G && n % G == 0 && n + G <= m.nt && dft_pending == 0 && !timer.on) {
                if (!gexec) capture_block(G);
                CFD_LAUNCH(k_set_int, dim3(1), dim3(64), 0, stream, nbase_dev, n);
                HIP_OK(hipGetLastError());
                HIP_OK(hipGraphLaunch(gexec, stream));
                done += G;
                count -= G;
                continue;
            }
            one_step(n);
            done++;
            count--;
        }
    }

    void read_field(int comp, double *host) override {
        const size_t ncell = (size_t)d.slab * (m.Nz + 1);
        std::vector<R> tmp(ncell);
        HIP_OK(hipMemcpyAsync(tmp.data(), d.F[comp], ncell * sizeof(R), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for (int i = 0; i <= m.Nx; i++)
            for (int j = 0; j <= m.Ny; j++)
                for (int k = 0; k <= m.Nz; k++) host[m.flat(i, j, k)] = (double)tmp[dev_index(i, j, k)];
    }
    // cells [i0,i1) x [j0,j1) x [k0,k1) of one component, host [i][j][k]
    void read_region(int comp, const int *b, double *host) override {
        const int ni = b[1] - b[0], nj = b[3] - b[2], nk = b[5] - b[4];
        if (ni <= 0 || nj <= 0 || nk <= 0 || b[0] < 0 || b[2] < 0 || b[4] < 0 || b[1] > m.Nx + 1 || b[3] > m.Ny + 1 ||
            b[5] > m.Nz + 1)
            throw Error("region outside the grid");
        std::vector<R> row((size_t)ni * nj * nk);
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                HIP_OK(hipMemcpyAsync(row.data() + ((size_t)k * nj + j) * ni, d.F[comp] + dev_index(b[0], b[2] + j, b[4] + k),
                                      ni * sizeof(R), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < nj; j++)
                for (int k = 0; k < nk; k++) host[((size_t)i * nj + j) * nk + k] = (double)row[((size_t)k * nj + j) * ni + i];
    }
    void write_field(int comp, const double *host) override {
        const size_t ncell = (size_t)d.slab * (m.Nz + 1);
        std::vector<R> tmp(ncell, (R)0);
        for (int i = 0; i <= m.Nx; i++)
            for (int j = 0; j <= m.Ny; j++)
                for (int k = 0; k <= m.Nz; k++) tmp[dev_index(i, j, k)] = (R)host[m.flat(i, j, k)];
        HIP_OK(hipMemcpyAsync(d.F[comp], tmp.data(), ncell * sizeof(R), hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }
    void port_samples(int q, double *host) override {
        const PortDesc<R> &p = dports.at(q);
        if (!p.samples) throw Error("port samples were not recorded (record_ports = 0)");
        const size_t n = (size_t)done * p.npts * 6;
        std::vector<R> tmp(n);
        if (n) HIP_OK(hipMemcpyAsync(tmp.data(), p.samples, n * sizeof(R), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for (size_t t = 0; t < n; t++) host[t] = (double)tmp[t];
    }
    long long nonfinite() override {
        const long long n = d.slab * (m.Nz + 1);
        CFD_LAUNCH(k_set_int, dim3(1), dim3(64), 0, stream, count_dev, 0);
        CFD_LAUNCH(k_nonfinite<R>, dim3(1024), dim3(256), 0, stream, d, n, count_dev);
        HIP_OK(hipGetLastError());
        int h = 0;
        HIP_OK(hipMemcpyAsync(&h, count_dev, sizeof h, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return h;
    }
    void port_gather(int q, double *host) override {
        const PortDesc<R> &p = dports.at(q);
        if (!p.npts) return;
        CFD_LAUNCH(k_gather<R>, dim3((p.npts + 255) / 256), dim3(256), 0, stream, d, p.cell, p.npts, gather_buf);
        HIP_OK(hipGetLastError());
        std::vector<R> tmp((size_t)p.npts * 6);
        HIP_OK(hipMemcpyAsync(tmp.data(), gather_buf, tmp.size() * sizeof(R), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for (size_t t = 0; t < tmp.size(); t++) host[t] = (double)tmp[t];
    }
    void port_gather_via(int q, double *out, double *host) override {
        const PortDesc<R> &p = dports.at(q);
        if (!p.npts) return;
        CFD_LAUNCH((k_gather<R, double>), dim3((p.npts + 255) / 256), dim3(256), 0, stream, d, p.cell, p.npts, out);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(host, out, (size_t)p.npts * 6 * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }
    int port_npts(int q) const override { return dports.at(q).npts; }
    void port_dft(int q, double *host) override {
        const PortDesc<R> &p = dports.at(q);
        if (!p.dft) throw Error("no DFT frequencies configured");
        dft_flush();
        // device [f][c][re,im][point] -> host [point][f][c][re,im]: transposed
        // on the device in chunks of points (32 MiB each), each copied to one
        // of two pinned staging buffers while the host copies the previous
        // chunk out of the other
        const int rows = (int)m.freqs.size() * 12;
        const long long np_ = p.npts;
        if (!np_) return;
        const long long chunk = std::min(np_, std::max(32LL, (long long)(kStageBytes / 8) / rows / 32 * 32));
        const size_t cbytes = (size_t)chunk * rows * sizeof(double);
        std::lock_guard<std::mutex> stage_lk(stage_mutex());
        double **pin = pinned_stage();
        double *buf = nullptr;
        hipEvent_t ev[2] = {nullptr, nullptr};
        {
            std::lock_guard<std::mutex> lk(capture_mutex());
            keep_pool_memory();
            HIP_OK(hipMallocAsync(&buf, 2 * cbytes, stream));
        }
        auto release = [&] {
            (void)hipStreamSynchronize(stream);
            for (hipEvent_t &e : ev)
                if (e) (void)hipEventDestroy(e);
            std::lock_guard<std::mutex> lk(capture_mutex());
            (void)hipFreeAsync(buf, stream);
            (void)hipStreamSynchronize(stream);
        };
        try {
            for (hipEvent_t &e : ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            const long long nch = (np_ + chunk - 1) / chunk;
            auto copy_out = [&](long long c) {  // chunk c: staged -> host
                const long long t0 = c * chunk;
                const int n = (int)std::min(chunk, np_ - t0);
                HIP_OK(hipEventSynchronize(ev[c & 1]));
                // split over host threads: the first touch of fresh result
                // pages (page faults) dominates a single-threaded copy
                const size_t bytes = (size_t)n * rows * sizeof(double);
                char *dst = reinterpret_cast<char *>(host + t0 * rows);
                const char *src = reinterpret_cast<const char *>(pin[c & 1]);
                host_parallel([&](int t, int nth) {
                    const size_t a = bytes * t / nth / 64 * 64, b = t + 1 == nth ? bytes : bytes * (t + 1) / nth / 64 * 64;
                    std::memcpy(dst + a, src + a, b - a);
                });
            };
            for (long long c = 0; c < nch; c++) {
                const long long t0 = c * chunk;
                const int n = (int)std::min(chunk, np_ - t0);
                double *db = buf + (c & 1) * (size_t)chunk * rows;
                CFD_LAUNCH(k_dft_transpose, dim3((unsigned)((n + 31) / 32), (unsigned)((rows + 31) / 32)), dim3(256), 0,
                           stream, p.dft, np_, rows, t0, n, db);
                HIP_OK(hipGetLastError());
                HIP_OK(hipMemcpyAsync(pin[c & 1], db, (size_t)n * rows * sizeof(double), hipMemcpyDeviceToHost, stream));
                HIP_OK(hipEventRecord(ev[c & 1], stream));
                if (c > 0) copy_out(c - 1);
            }
            copy_out(nch - 1);
        } catch (...) {
            release();
            throw;
        }
        release();
    }
    void port_intensity(int q, double *host) override {
        const PortDesc<R> &p = dports.at(q);
        std::vector<double> tmp((size_t)done * p.nblk);
        if (!tmp.empty()) HIP_OK(hipMemcpyAsync(tmp.data(), p.ipart, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for (int n = 0; n < done; n++) {
            double s = 0.0;
            for (int b = 0; b < p.nblk; b++) s = s + tmp[(size_t)n * p.nblk + b];
            host[n] = s;
        }
    }
    void gain_read(double *host) override {
        if (!m.has_gain) throw Error("no gain medium");
        const size_t ng = d.ngain;
        std::vector<double> tmp(4 * std::max<size_t>(ng, 1));
        HIP_OK(hipMemcpyAsync(tmp.data(), d.gN, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for (size_t r = 0; r < ng; r++)
            for (int l = 0; l < 4; l++) host[r * 4 + l] = tmp[l * ng + gain_slot[r]];
    }
    // device value (variable v, component c, slot s) at dlst[(v * 3 + c) * ndl + s]
    int dl_state(double *host) override {
        if (!m.has_dl) throw Error("no Drude-Lorentz medium");
        const size_t nd = d.ndl, per = (size_t)m.nD + 2 * m.nL + 1;
        if (host && nd) {
            std::vector<R> tmp(per * 3 * nd);
            HIP_OK(hipMemcpyAsync(tmp.data(), d.dlst, tmp.size() * sizeof(R), hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            for (size_t r = 0; r < nd; r++)
                for (size_t c = 0; c < 3; c++)
                    for (size_t v = 0; v < per; v++) host[(r * 3 + c) * per + v] = (double)tmp[(v * 3 + c) * nd + dl_slot[r]];
        }
        return (int)nd;
    }
    int gain_state(double *host) override {
        if (!m.has_gain) throw Error("no gain medium");
        const size_t ng = d.ngain;
        if (host && ng) {
            std::vector<R> p(3 * ng), p1(3 * ng);
            HIP_OK(hipMemcpyAsync(p.data(), d.gP, p.size() * sizeof(R), hipMemcpyDeviceToHost, stream));
            HIP_OK(hipMemcpyAsync(p1.data(), d.gP1, p1.size() * sizeof(R), hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            for (size_t r = 0; r < ng; r++)
                for (size_t c = 0; c < 3; c++) {
                    host[r * 6 + c] = (double)p[c * ng + gain_slot[r]];
                    host[r * 6 + 3 + c] = (double)p1[c * ng + gain_slot[r]];
                }
        }
        return (int)ng;
    }
};

// ----------------------------------------------------------------------------
// A batch of simulations of one shape (grid, CPML, TF/SF box and waveform,
// ports) whose media and geometry may differ: every pass of a step is one
// launch over all instances (blockIdx.y = instance), on the batch's stream.
// Each instance is an ordinary Sim (same set-up, same readers) that shares
// that stream; the batch owns its stepping.
class BatchBase {
public:
    virtual ~BatchBase() = default;
    hipStream_t stream = nullptr;
    int done = 0;
    KernelTimer timer;
    virtual int size() const = 0;
    virtual SimBase *member(int i) = 0;
    virtual void step(int count) = 0;
    virtual size_t device_bytes() const = 0;
};

template <class R>
class Batch : public BatchBase {
public:
    std::vector<std::unique_ptr<Sim<R>>> sims;
    std::vector<void *> allocs;
    Dev<R> *devs = nullptr;
    SpecialList *lists = nullptr;
    int max_special = 0;
    int special_poles[2] = {-1, -1};  // one special pass instantiation for every instance
    PortDesc<R> *ports_all = nullptr;  // [instance][port]
    int nports = 0;
    int dft_pending = 0, dft_first = 0;
    size_t own_bytes = 0;
    bool nt_fields = false;  // by the field bytes of all instances together

    Batch(std::vector<Model> &&models, int prec) {
        if (models.empty()) throw Error("empty batch");
        HIP_OK(hipStreamCreate(&stream));
        try {
            for (Model &m : models) {
                sims.emplace_back(new Sim<R>(std::move(m), prec, stream));
                sims.back()->member = true;
            }
            check_shape();
            nt_fields = stream_fields(6 * (size_t)sims[0]->d.slab * (sims[0]->m.Nz + 1) * sizeof(R) * sims.size());
            std::vector<Dev<R>> dv;
            std::vector<SpecialList> sl;
            std::vector<PortDesc<R>> pa;
            int pl[3] = {0, 0, 0};
            for (auto &sp : sims) {
                dv.push_back(sp->d);
                // the batched kernels read this view from device memory: the
                // coefficient tables, not 32 more scalars per workgroup
                dv.back().nsel = 0;
                sl.push_back(SpecialList{sp->special, sp->nspecial, 0});
                max_special = std::max(max_special, sp->nspecial);
                pa.insert(pa.end(), sp->dports.begin(), sp->dports.end());
                merge_poles(pl, sp->d.ndl, sp->d.nD, sp->d.nL);
            }
            special_poles[0] = pl[0];
            special_poles[1] = pl[1];
            nports = (int)sims[0]->dports.size();
            devs = upload(dv);
            lists = upload(sl);
            if (nports) ports_all = upload(pa);
        } catch (...) {
            release();
            throw;
        }
    }
    ~Batch() override { release(); }
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        sims.clear();
        std::lock_guard<std::mutex> lk(capture_mutex());
        for (void *p : allocs) (void)hipFree(p);
        allocs.clear();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
    template <class T>
    T *upload(const std::vector<T> &v) {
        void *p = nullptr;
        std::lock_guard<std::mutex> lk(capture_mutex());
        HIP_OK(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)));
        allocs.push_back(p);
        own_bytes += v.size() * sizeof(T);
        HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return static_cast<T *>(p);
    }
    void check_shape() {
        const Sim<R> &a = *sims[0];
        for (size_t i = 1; i < sims.size(); i++) {
            const Sim<R> &b = *sims[i];
            const std::string who = "batch instance " + std::to_string(i) + ": ";
            if (b.m.Nx != a.m.Nx || b.m.Ny != a.m.Ny || b.m.Nz != a.m.Nz) throw Error(who + "grid size differs from instance 0");
            if (std::memcmp(b.m.npml, a.m.npml, sizeof a.m.npml) || b.m.pbx != a.m.pbx || b.m.pby != a.m.pby)
                throw Error(who + "CPML or periodic axes differ from instance 0");
            if (b.m.nt != a.m.nt || b.m.dt != a.m.dt) throw Error(who + "time step or run length differs from instance 0");
            if (b.m.has_tfsf != a.m.has_tfsf || b.tf_sig != a.tf_sig) throw Error(who + "TF/SF set-up differs from instance 0");
            if (b.dports.size() != a.dports.size() || b.port_sig != a.port_sig) throw Error(who + "ports differ from instance 0");
            if (b.leanH_cs != a.leanH_cs || b.leanE_cs != a.leanE_cs ||
                std::memcmp(&b.leanH_st, &a.leanH_st, sizeof a.leanH_st) || std::memcmp(&b.leanE_st, &a.leanE_st, sizeof a.leanE_st))
                throw Error(who + "interior stretch factors differ from instance 0");
        }
    }
    int size() const override { return (int)sims.size(); }
    SimBase *member(int i) override { return sims.at(i).get(); }
    size_t device_bytes() const override {
        size_t b = own_bytes;
        for (const auto &sp : sims) b += sp->dev_bytes;
        return b;
    }

    void step(int count) override {
        const Sim<R> &a = *sims[0];
        const unsigned ni = (unsigned)sims.size();
        const Dev<R> d0 = a.d;  // by-value argument, unused by the batched kernels
        auto update = [&](bool h) {
            const MGrid &mg = h ? a.mgH : a.mgE;
            if (mg.per_plane && mg.planes)
                launch_full<R, true>(stream, h, h ? a.leanH_cs : a.leanE_cs, nt_fields, ni, d0, devs, mg,
                                     h ? a.leanH : a.leanE, h ? a.leanH_st : a.leanE_st);
        };
        for (int c = 0; c < count; c++) {
            const int n = done;
            if (n >= a.m.nt) throw Error("run finished: all " + std::to_string(a.m.nt) + " steps done");
            timer.run(stream, "full_h", [&] { update(true); });
            HIP_OK(hipGetLastError());
            tfsf(true, n);
            timer.run(stream, "full_e", [&] { update(false); });
            HIP_OK(hipGetLastError());
            if (max_special) {
                timer.run(stream, "special_e", [&] {
                    launch_special<R, true>(stream, dim3((max_special + 255) / 256, ni), d0, devs, nullptr, 0, lists,
                                            special_poles);
                });
                HIP_OK(hipGetLastError());
            }
            tfsf(false, n);
            for (auto &sp : sims) sp->sources(n);  // instances with point sources (rare in a batch)
            ports(n);
            done++;
            for (auto &sp : sims) sp->done++;
        }
        dft_flush();  // members' DFT readers then see every sampled step
    }
    void tfsf(bool h, int n) {
        const Sim<R> &a = *sims[0];
        if (!a.m.has_tfsf) return;
        const int nt = h ? a.ntfH : a.ntfE;
        if (!nt) return;
        timer.run(stream, h ? "tfsf_h" : "tfsf_e", [&] {
            if constexpr (std::is_same<R, float>::value) {
                if (a.tf32) {
                    CFD_LAUNCH((k_tfsf32<true>), dim3((nt + 255) / 256, (unsigned)sims.size()), dim3(256), 0, stream,
                                       a.d, devs, h ? a.tfH32 : a.tfE32, nt, a.tterm32, h ? a.einc32 : a.hinc32, a.tablen,
                                       nullptr, n);
                    return;
                }
            }
            CFD_LAUNCH((k_tfsf<R, true>), dim3((nt + 255) / 256, (unsigned)sims.size()), dim3(256), 0, stream, a.d, devs,
                               h ? a.tfH : a.tfE, nt, a.tcoef, a.tdelay, a.tprof, h ? a.einc : a.hinc, a.tablen, nullptr, n);
        });
        HIP_OK(hipGetLastError());
    }
    void ports(int n) {
        const Sim<R> &a = *sims[0];
        if (!nports) return;
        int slot = -1;
        if (!a.m.freqs.empty() && n % a.m.dft_stride == 0) {
            if (dft_pending == 0) dft_first = n;
            slot = dft_pending++;
        }
        timer.run(stream, "ports", [&] {
            CFD_LAUNCH((k_port<R, true>), dim3(a.port_blocks, (unsigned)sims.size()), dim3(256), 0, stream, a.d, devs,
                               ports_all, nports, nullptr, n, slot);
        });
        HIP_OK(hipGetLastError());
        if (dft_pending == Sim<R>::kDftBatch) dft_flush();
    }
    void dft_flush() {
        if (!dft_pending) return;
        const Sim<R> &a = *sims[0];
        const int nbatch = dft_pending;
        dft_pending = 0;
        timer.run(stream, "port_dft", [&] {
            CFD_LAUNCH((k_dft<R, true>), dim3(4 * a.port_blocks, (unsigned)sims.size()), dim3(256), 0, stream, ports_all,
                               nports, nbatch, nullptr, dft_first, a.m.dft_stride, a.twE, a.twH, (int)a.m.freqs.size(), a.m.nt);
        });
        HIP_OK(hipGetLastError());
    }
};

// ----------------------------------------------------------------------------
// A pack: simulations of any grid size and physics (one precision) stepped
// together, every pass of a step one launch over all members (the packed
// kernels in kernels.hip).  Members keep their own device state and stay
// owned by their handles; the pack holds the tables that describe them to the
// packed kernels, rebuilt when a member joins or leaves, and replays
// sequences of kPackBlock steps from a captured HIP graph while the
// membership is unchanged (a member's time index comes from device memory).
// No reference counterpart: the reference runs one simulation per process
// (main.cpp:115-343).
class PackBase {
public:
    virtual ~PackBase() = default;
    hipStream_t stream = nullptr;
    KernelTimer timer;
    virtual void add(SimBase *s) = 0;
    virtual void remove(SimBase *s) = 0;
    virtual void step(int count) = 0;
    virtual int size() const = 0;
    // non-finite field values of each given member (the serving loop's
    // watchdog), counted on the pack's stream with one synchronisation
    virtual void nonfinite(SimBase *const *sims, int n, int *counts) = 0;
};

// Steps per captured sequence.  Short, so that a membership that changes at
// every quantum (a serving loop) costs one short capture instead of
// thousands of launches: a 256-step quantum is 16 replays.
constexpr int kPackBlock = 16;

template <class R>
class Pack : public PackBase {
public:
    std::vector<Sim<R> *> mem;
    std::vector<void *> allocs;  // tables of the current membership
    PackMember<R> *dmem = nullptr;
    int *nstep = nullptr;
    unsigned long long *workH = nullptr, *workE = nullptr;
    int nworkH = 0, nworkE = 0;
    int *st_tfH = nullptr, *st_tfE = nullptr, *st_sp = nullptr, *st_src = nullptr, *st_port = nullptr;
    int nb_tfH = 0, nb_tfE = 0, nb_sp = 0, nb_src = 0, nb_port = 0;
    bool any_dft = false, nt_fields = false, dirty = true;
    bool use_graphs = true;
    hipGraphExec_t gexec = nullptr;

    Pack() {
        HIP_OK(hipStreamCreate(&stream));
        if (const char *v = std::getenv("CFD_GRAPHS")) use_graphs = std::string(v) != "0";
    }
    ~Pack() override {
        if (stream) (void)hipStreamSynchronize(stream);
        for (Sim<R> *s : mem) s->packed = false;
        drop_tables();
        if (stream) (void)hipStreamDestroy(stream);
    }
    int size() const override { return (int)mem.size(); }

    void drop_tables() {
        std::lock_guard<std::mutex> lk(capture_mutex());
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        for (void *p : allocs) (void)hipFreeAsync(p, stream);
        allocs.clear();
        if (stream) (void)hipStreamSynchronize(stream);
    }
    template <class T>
    T *upload(const std::vector<T> &v) {
        void *p = nullptr;
        const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
        {
            std::lock_guard<std::mutex> lk(capture_mutex());
            HIP_OK(hipMallocAsync(&p, bytes, stream));
        }
        allocs.push_back(p);
        if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream));
        return static_cast<T *>(p);
    }

    void add(SimBase *b) override {
        Sim<R> *s = dynamic_cast<Sim<R> *>(b);
        if (!s) throw Error("pack: the simulation's precision differs from the pack's");
        if (s->member) throw Error("pack: a batch member is stepped by its batch");
        if (s->packed) throw Error("pack: the simulation is already in a pack");
        if (s->done >= s->m.nt) throw Error("pack: the simulation's run has finished");
        if (!s->m.freqs.empty()) {
            // the packed port kernels stage DFT samples in slots counted from
            // step 0 in groups of 16: join on such a boundary, nothing staged
            s->dft_flush();
            if (s->done % (16 * s->m.dft_stride)) throw Error("pack: join at step 0 or a multiple of 16 DFT strides");
        }
        HIP_OK(hipStreamSynchronize(s->stream));  // set-up or earlier steps done before the pack uses it
        HIP_OK(hipStreamSynchronize(stream));
        mem.push_back(s);
        s->packed = true;
        dirty = true;
    }
    void remove(SimBase *b) override {
        auto it = std::find(mem.begin(), mem.end(), b);
        if (it == mem.end()) throw Error("pack: not a member");
        HIP_OK(hipStreamSynchronize(stream));  // its last steps done before its handle is used again
        (*it)->packed = false;
        mem.erase(it);
        dirty = true;
    }

    void rebuild() {
        HIP_OK(hipStreamSynchronize(stream));
        drop_tables();
        std::vector<PackMember<R>> pm;
        std::vector<int> nst;
        std::vector<unsigned long long> wh, we;
        std::vector<int> tfH{0}, tfE{0}, sp{0}, src{0}, port{0};
        size_t field_bytes = 0;
        any_dft = false;
        for (size_t q = 0; q < mem.size(); q++) {
            Sim<R> &S = *mem[q];
            PackMember<R> P{};
            P.d = S.d;
            P.mh = S.mgH; P.me = S.mgE;
            P.Lh = S.leanH; P.Le = S.leanE;
            P.sh = S.leanH_st; P.se = S.leanE_st;
            P.csH = S.leanH_cs; P.csE = S.leanE_cs;
            P.nt = S.m.nt;
            P.ntfE = S.ntfE; P.ntfH = S.ntfH; P.tablen = S.tablen;
            P.tfE = S.tfE; P.tfH = S.tfH;
            P.tcoef = S.tcoef; P.tdelay = S.tdelay; P.tprof = S.tprof; P.einc = S.einc; P.hinc = S.hinc;
            P.tf32 = S.tf32 ? 1 : 0;
            P.tfE32 = S.tfE32; P.tfH32 = S.tfH32; P.tterm32 = S.tterm32; P.einc32 = S.einc32; P.hinc32 = S.hinc32;
            P.special = S.special; P.nspecial = S.nspecial;
            P.nsgroup = S.nsgroup; P.sgroup = S.sgroup; P.sidx = S.sidx; P.swave = S.swave;
            P.ports = S.port_tab; P.nports = (int)S.dports.size();
            P.dft_stride = S.m.dft_stride; P.nf = (int)S.m.freqs.size();
            P.twE = S.twE; P.twH = S.twH;
            pm.push_back(P);
            nst.push_back(S.done);
            any_dft = any_dft || P.nf > 0;
            const unsigned long long sq = (unsigned long long)q << 48;
            for (int k = 0; k < S.mgH.planes; k++)
                for (int t = 0; t < S.mgH.per_plane; t++)
                    wh.push_back(sq | (unsigned long long)(S.mgH.k_lo + k) << kPackPlaneShift | (unsigned long long)t);
            for (int k = 0; k < S.mgE.planes; k++)
                for (int t = 0; t < S.mgE.per_plane; t++)
                    we.push_back(sq | (unsigned long long)(S.mgE.k_lo + k) << kPackPlaneShift | (unsigned long long)t);
            tfH.push_back(tfH.back() + (S.ntfH + 255) / 256);
            tfE.push_back(tfE.back() + (S.ntfE + 255) / 256);
            sp.push_back(sp.back() + (S.nspecial + 255) / 256);
            src.push_back(src.back() + (S.nsgroup + 63) / 64);
            port.push_back(port.back() + S.port_blocks);
            field_bytes += 6 * (size_t)S.d.slab * (S.m.Nz + 1) * sizeof(R);
        }
        if (wh.size() > 0x7fffffffu || we.size() > 0x7fffffffu) throw Error("pack: too many workgroups");
        nt_fields = stream_fields(field_bytes);
        dmem = upload(pm);
        nstep = upload(nst);
        workH = upload(wh); nworkH = (int)wh.size();
        workE = upload(we); nworkE = (int)we.size();
        st_tfH = upload(tfH); nb_tfH = tfH.back();
        st_tfE = upload(tfE); nb_tfE = tfE.back();
        st_sp = upload(sp); nb_sp = sp.back();
        st_src = upload(src); nb_src = src.back();
        st_port = upload(port); nb_port = port.back();
        HIP_OK(hipStreamSynchronize(stream));
        dirty = false;
    }

    void launch_step(int q) {
        const int nm = (int)mem.size();
        const dim3 blk(256);
        if (nworkH) timer.run(stream, "pack_h", [&] {
            if (nt_fields) CFD_LAUNCH((k_pack_h<R, true>), dim3(nworkH), blk, 0, stream, dmem, workH, nstep, q);
            else CFD_LAUNCH((k_pack_h<R, false>), dim3(nworkH), blk, 0, stream, dmem, workH, nstep, q);
        });
        if (nb_tfH) timer.run(stream, "pack_tfsf_h", [&] {
            CFD_LAUNCH(k_pack_tfsf<R>, dim3(nb_tfH), blk, 0, stream, dmem, st_tfH, nm, nstep, q, 1);
        });
        if (nworkE) timer.run(stream, "pack_e", [&] {
            if (nt_fields) CFD_LAUNCH((k_pack_e<R, true>), dim3(nworkE), blk, 0, stream, dmem, workE, nstep, q);
            else CFD_LAUNCH((k_pack_e<R, false>), dim3(nworkE), blk, 0, stream, dmem, workE, nstep, q);
        });
        if (nb_sp) timer.run(stream, "pack_special_e", [&] {
            CFD_LAUNCH(k_pack_special<R>, dim3(nb_sp), blk, 0, stream, dmem, st_sp, nm, nstep, q);
        });
        if (nb_tfE) timer.run(stream, "pack_tfsf_e", [&] {
            CFD_LAUNCH(k_pack_tfsf<R>, dim3(nb_tfE), blk, 0, stream, dmem, st_tfE, nm, nstep, q, 0);
        });
        if (nb_src) timer.run(stream, "pack_sources", [&] {
            CFD_LAUNCH(k_pack_sources<R>, dim3(nb_src), dim3(64), 0, stream, dmem, st_src, nm, nstep, q);
        });
        if (nb_port) {
            timer.run(stream, "pack_ports", [&] {
                CFD_LAUNCH(k_pack_port<R>, dim3(nb_port), blk, 0, stream, dmem, st_port, nm, nstep, q);
            });
            if (any_dft) timer.run(stream, "pack_dft", [&] {
                CFD_LAUNCH(k_pack_dft<R>, dim3(4 * nb_port), blk, 0, stream, dmem, st_port, nm, nstep, q);
            });
        }
        HIP_OK(hipGetLastError());
    }
    void launch_sequence(int count) {
        for (int q = 0; q < count; q++) launch_step(q);
        const int nm = (int)mem.size();
        CFD_LAUNCH(k_pack_advance, dim3((nm + 63) / 64), dim3(64), 0, stream, nstep, nm, count);
        HIP_OK(hipGetLastError());
    }
    void capture() {
        hipGraph_t g = nullptr;
        std::lock_guard<std::mutex> lk(capture_mutex());
        HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        try {
            launch_sequence(kPackBlock);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIP_OK(hipStreamEndCapture(stream, &g));
        const hipError_t e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_OK(e);
    }

    void nonfinite(SimBase *const *sims, int n, int *counts) override {
        std::vector<Sim<R> *> ms;
        for (int q = 0; q < n; q++) {
            Sim<R> *s = dynamic_cast<Sim<R> *>(sims[q]);
            if (!s || std::find(mem.begin(), mem.end(), s) == mem.end()) throw Error("pack: not a member");
            CFD_LAUNCH(k_set_int, dim3(1), dim3(64), 0, stream, s->count_dev, 0);
            CFD_LAUNCH(k_nonfinite<R>, dim3(1024), dim3(256), 0, stream, s->d, s->d.slab * (s->m.Nz + 1),
                               s->count_dev);
            ms.push_back(s);
        }
        HIP_OK(hipGetLastError());
        for (int q = 0; q < n; q++)
            HIP_OK(hipMemcpyAsync(counts + q, ms[q]->count_dev, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    // count steps of every member (asynchronous); a member whose run ends
    // within them idles for the rest
    void step(int count) override {
        if (count <= 0 || mem.empty()) return;
        if (dirty) rebuild();
        const bool graphs = use_graphs && !timer.on;
        int left = count;
        while (left > 0) {
            if (graphs && left >= kPackBlock) {
                if (!gexec) capture();
                HIP_OK(hipGraphLaunch(gexec, stream));
                left -= kPackBlock;
            } else {
                const int c = std::min(left, kPackBlock);
                launch_sequence(c);
                left -= c;
            }
        }
        for (Sim<R> *s : mem) {
            s->done = std::min(s->done + count, s->m.nt);
            s->dft_pending = 0;
            if (s->m.freqs.empty() || s->done == 0) continue;
            // staged DFT samples not yet added in: the sampled steps after the
            // last flush (flushes come at every 16th sampled step and at the
            // run's last sampled step)
            const int st = s->m.dft_stride;
            const int last_sampled = (s->done - 1) / st * st;
            const int slot = (last_sampled / st) % 16;
            const int run_last = (s->m.nt - 1) / st * st;
            if (slot != 15 && last_sampled != run_last) {
                s->dft_pending = slot + 1;
                s->dft_first = last_sampled - slot * st;
            }
        }
    }
};

}  // namespace cfd

// ============================================================================
// Staging memory of cfd_port_gather_static: a file-scope device array, so that
// entry point allocates nothing.  Each GPU the library runs on holds its own
// copy (kStaticGatherDoubles doubles, 8 MiB) until the process ends.  Host
// code never passes the array's name to a kernel or a copy (on the host it is
// a stand-in, not a device address): it uses the address hipGetSymbolAddress
// returns, looked up once per device with that device current.
namespace cfd {
constexpr size_t kStaticGatherDoubles = size_t(1) << 20;
}
static __device__ __align__(256) double scratch[cfd::kStaticGatherDoubles];

namespace cfd {
constexpr int kMaxDevices = 64;
struct StaticScratch {
    std::mutex mu;  // calls on one device share its array: one at a time
    double *ptr[kMaxDevices] = {};
};
inline StaticScratch &static_scratch() {
    static StaticScratch s;
    return s;
}
}  // namespace cfd

// ============================================================================
// C ABI
// ============================================================================
using cfd::SimBase;

struct cfd_sim {
    std::unique_ptr<SimBase> owned;  // null for a batch member
    SimBase *impl = nullptr;
};

struct cfd_batch {
    std::unique_ptr<cfd::BatchBase> impl;
    std::vector<std::unique_ptr<cfd_sim>> members;  // borrowed views of the instances
};

struct cfd_pack {
    std::unique_ptr<cfd::PackBase> impl;
    int precision = 0;
};

template <class F>
static int guard(F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        cfd::g_error = e.what();
    } catch (...) {
        cfd::g_error = "unknown error";
    }
    return -1;
}

#define NEED(s)                                          \
    do {                                                 \
        if (!(s) || !(s)->impl) {                        \
            cfd::g_error = "null simulation handle";     \
            return -1;                                   \
        }                                                \
    } while (0)

// calls that advance or update a simulation: not on a batch member
#define NEED_SOLO(s)                                                                  \
    do {                                                                              \
        NEED(s);                                                                      \
        if ((s)->impl->member) {                                                      \
            cfd::g_error = "batch member: the batch steps its instances (cfd_batch_step)"; \
            return -1;                                                                \
        }                                                                             \
        if ((s)->impl->packed) {                                                      \
            cfd::g_error = "pack member: the pack steps it (cfd_pack_step)";          \
            return -1;                                                                \
        }                                                                             \
    } while (0)

extern "C" {

const char *cfd_last_error(void) { return cfd::g_error.c_str(); }

int cfd_device_info(char *name, int cap, int *n_cu, size_t *hbm_bytes)
{
    return guard([&] {
        int dev = 0;
        HIP_OK(hipGetDevice(&dev));
        hipDeviceProp_t p;
        HIP_OK(hipGetDeviceProperties(&p, dev));
        if (name && cap > 0) snprintf(name, cap, "%s (%s)", p.name, p.gcnArchName);
        if (n_cu) *n_cu = p.multiProcessorCount;
        if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    });
}

int cfd_create(const cfd_config *cfg, cfd_sim **out)
{
    if (!cfg || !out) { cfd::g_error = "null argument"; return -1; }
    *out = nullptr;
    return guard([&] {
        if (cfg->precision != CFD_F32 && cfg->precision != CFD_F64) throw cfd::Error("precision must be 32 or 64");
        cfd::Model m = cfd::build_model(*cfg);
        auto s = std::make_unique<cfd_sim>();
        if (cfg->precision == CFD_F32) s->owned = std::make_unique<cfd::Sim<float>>(std::move(m), 32);
        else s->owned = std::make_unique<cfd::Sim<double>>(std::move(m), 64);
        s->impl = s->owned.get();
        *out = s.release();
    });
}

int cfd_create_from_model(cfd_model *model, int precision, cfd_sim **out)
{
    if (!out) { delete model; cfd::g_error = "null argument"; return -1; }
    *out = nullptr;
    if (!model) { cfd::g_error = "null argument"; return -1; }
    std::unique_ptr<cfd_model> owned(model);
    return guard([&] {
        if (precision != CFD_F32 && precision != CFD_F64) throw cfd::Error("precision must be 32 or 64");
        cfd::Model m = std::move(owned->m);
        owned.reset();
        auto s = std::make_unique<cfd_sim>();
        if (precision == CFD_F32) s->owned = std::make_unique<cfd::Sim<float>>(std::move(m), 32);
        else s->owned = std::make_unique<cfd::Sim<double>>(std::move(m), 64);
        s->impl = s->owned.get();
        *out = s.release();
    });
}

size_t cfd_model_device_bytes(cfd_model *p, int precision)
{
    if (!p || (precision != CFD_F32 && precision != CFD_F64)) return 0;
    const cfd::Model &m = p->m;
    const size_t S = precision == CFD_F32 ? 4 : 8;
    const size_t px = ((size_t)m.Nx + 1 + 31) / 32 * 32, py = (size_t)m.Ny + 1, slab = px * py;
    const size_t ncell = slab * ((size_t)m.Nz + 1);
    size_t b = ncell * (6 * S + 1);                      // fields, material IDs
    b += 4 * (size_t)m.NoMedia * S;                      // aE bE aH bH
    for (int t = 0; t < cfd::kNumStretch; t++) b += m.stretch[t].size() * S;
    const size_t N[3] = {(size_t)m.Nx, (size_t)m.Ny, (size_t)m.Nz};
    for (int a = 0; a < 3; a++)
        for (int e2 = 0; e2 < 2; e2++) {
            const size_t n = (size_t)m.npml[a][e2];
            if (!n) continue;
            b += (m.pbE[a][e2].size() + m.pcE[a][e2].size() + m.pbH[a][e2].size() + m.pcH[a][e2].size()) * S;
            const size_t box = a == 0 ? n * py * (N[2] + 1) : a == 1 ? px * n * (N[2] + 1) : slab * n;
            b += 4 * box * S;                                // psi, E and H, two per slab
        }
    const size_t ndl = m.dl_cells.size(), ng = m.gain_cells.size();
    if (m.has_dl || m.has_gain) b += ncell * sizeof(int) + (ndl + ng) * sizeof(long long);  // slot map, cell list
    if (m.has_dl) b += ((size_t)m.nD + 2 * m.nL + 1) * 3 * std::max<size_t>(ndl, 1) * S;
    if (m.has_gain) b += 9 * std::max<size_t>(ng, 1) * S + 4 * std::max<size_t>(ng, 1) * sizeof(double);
    if (m.has_tfsf) {
        const size_t targets = m.tfE.size() + m.tfH.size(), terms = m.tf_terms.size(), tab = m.Einc.size();
        b += targets * sizeof(cfd::TfTarget) + terms * 3 * sizeof(double) + 2 * tab * sizeof(double);
        if (precision == CFD_F32) b += targets * sizeof(cfd::TfTarget32) + terms * sizeof(cfd::TfTerm32) + 2 * tab * sizeof(float);
    }
    const size_t nsrc = m.src.size() / 4;
    b += nsrc * (sizeof(cfd::SourceGroup) + sizeof(int)) + m.src_wave.size() * sizeof(double);
    b += (m.twE.size() + m.twH.size()) * sizeof(double);
    size_t gather = 0;
    for (const cfd::Port &P : m.ports) {
        const size_t np = (size_t)P.npts(), nblk = (np + 255) / 256, nf = m.freqs.size();
        b += np * sizeof(long long) + (size_t)m.nt * std::max<size_t>(nblk, 1) * sizeof(double);
        if (m.record) b += (size_t)m.nt * np * 6 * S;
        if (nf) b += (size_t)cfd::Sim<float>::kDftBatch * 6 * np * S + np * nf * 12 * sizeof(double);
        gather = std::max(gather, np);
    }
    b += gather * 6 * S;
    // tile tables, pass boxes, counters: small; allowance for them
    return b + b / 64 + ((size_t)1 << 20);
}

int cfd_destroy(cfd_sim *s)
{
    if (!s) return 0;
    if (!s->owned) { cfd::g_error = "a batch member is released with its batch (cfd_batch_destroy)"; return -1; }
    if (s->impl->packed) { cfd::g_error = "a pack member: remove it from its pack first (cfd_pack_remove)"; return -1; }
    delete s;
    return 0;
}

int cfd_info(cfd_sim *s, int *nxyz_nt, double *dd, int *counts)
{
    NEED(s);
    const cfd::Model &m = s->impl->m;
    if (nxyz_nt) { nxyz_nt[0] = m.Nx; nxyz_nt[1] = m.Ny; nxyz_nt[2] = m.Nz; nxyz_nt[3] = m.nt; }
    if (dd) { dd[0] = m.dx; dd[1] = m.dy; dd[2] = m.dz; dd[3] = m.dt; }
    if (counts) {
        counts[0] = (int)m.ports.size();
        counts[1] = (int)m.dl_cells.size();
        counts[2] = (int)m.gain_cells.size();
        counts[3] = (int)(m.tfE.size() + m.tfH.size());
    }
    return 0;
}

size_t cfd_device_bytes(const cfd_sim *s) { return (s && s->impl) ? s->impl->dev_bytes : 0; }

int cfd_step(cfd_sim *s, int count) { NEED_SOLO(s); return guard([&] { s->impl->step(count); }); }
int cfd_steps_done(cfd_sim *s) { NEED(s); return s->impl->done; }
int cfd_check_finite(cfd_sim *s)
{
    NEED(s);
    long long n = 0;
    const int rc = guard([&] { n = s->impl->nonfinite(); });
    return rc ? rc : (int)std::min<long long>(n, 0x7fffffff);
}
int cfd_read_region(cfd_sim *s, int comp, const int *box, double *host)
{
    NEED(s);
    if (comp < 0 || comp > 5 || !box || !host) { cfd::g_error = "bad component or null buffer"; return -1; }
    return guard([&] { s->impl->read_region(comp, box, host); });
}
int cfd_query(cfd_sim *s)
{
    NEED(s);
    const hipError_t e = hipStreamQuery(s->impl->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    cfd::g_error = std::string("HIP: ") + hipGetErrorString(e);
    return -1;
}
int cfd_sync(cfd_sim *s)
{
    NEED(s);
    return guard([&] {
        HIP_OK(hipStreamSynchronize(s->impl->stream));
        s->impl->timer.collect();
    });
}
int cfd_basic_h(cfd_sim *s) { NEED_SOLO(s); return guard([&] { s->impl->basic_h(); }); }
int cfd_basic_e(cfd_sim *s) { NEED_SOLO(s); return guard([&] { s->impl->basic_e(); }); }
int cfd_cpml_h(cfd_sim *s) { NEED_SOLO(s); return guard([&] { s->impl->cpml_h(); }); }
int cfd_cpml_e(cfd_sim *s) { NEED_SOLO(s); return guard([&] { s->impl->cpml_e(); }); }
int cfd_tfsf_h(cfd_sim *s, int n) { NEED_SOLO(s); return guard([&] { s->impl->tfsf(true, n); }); }
int cfd_tfsf_e(cfd_sim *s, int n) { NEED_SOLO(s); return guard([&] { s->impl->tfsf(false, n); }); }
int cfd_tfsf_face_h(cfd_sim *s, int face, int n) { NEED_SOLO(s); return guard([&] { s->impl->tfsf_face(true, face, n); }); }
int cfd_tfsf_face_e(cfd_sim *s, int face, int n) { NEED_SOLO(s); return guard([&] { s->impl->tfsf_face(false, face, n); }); }
int cfd_dl_e(cfd_sim *s) { NEED_SOLO(s); return guard([&] { if (s->impl->m.has_dl) s->impl->e_parts(cfd::E_DL); }); }
int cfd_gain_pol(cfd_sim *s) { NEED_SOLO(s); return guard([&] { if (s->impl->m.has_gain) s->impl->e_parts(cfd::E_GAINPOL); }); }
int cfd_gain_e(cfd_sim *s) { NEED_SOLO(s); return guard([&] { if (s->impl->m.has_gain) s->impl->e_parts(cfd::E_GAINE); }); }
int cfd_gain_pop(cfd_sim *s) { NEED_SOLO(s); return guard([&] { if (s->impl->m.has_gain) s->impl->e_parts(cfd::E_GAINPOP); }); }
int cfd_sources(cfd_sim *s, int n) { NEED_SOLO(s); return guard([&] { s->impl->sources(n); }); }
int cfd_ports_sample(cfd_sim *s, int n) { NEED_SOLO(s); return guard([&] { s->impl->ports(n); }); }
int cfd_advance(cfd_sim *s) { NEED_SOLO(s); s->impl->done++; return 0; }

int cfd_read_field(cfd_sim *s, int comp, double *host)
{
    NEED(s);
    if (comp < 0 || comp > 5 || !host) { cfd::g_error = "bad component or buffer"; return -1; }
    return guard([&] { s->impl->read_field(comp, host); });
}
int cfd_write_field(cfd_sim *s, int comp, const double *host)
{
    NEED(s);
    if (comp < 0 || comp > 5 || !host) { cfd::g_error = "bad component or buffer"; return -1; }
    return guard([&] { s->impl->write_field(comp, host); });
}
int cfd_read_id(cfd_sim *s, int16_t *host)
{
    NEED(s);
    if (!host) { cfd::g_error = "null buffer"; return -1; }
    std::memcpy(host, s->impl->m.id.data(), s->impl->m.id.size() * sizeof(int16_t));
    return 0;
}
int cfd_port_npts(cfd_sim *s, int port)
{
    NEED(s);
    if (port < 0 || port >= (int)s->impl->m.ports.size()) { cfd::g_error = "no such port"; return -1; }
    return s->impl->m.ports[port].npts();
}
int cfd_port_points(cfd_sim *s, int port, int *ijk)
{
    NEED(s);
    if (port < 0 || port >= (int)s->impl->m.ports.size() || !ijk) { cfd::g_error = "no such port"; return -1; }
    const auto &v = s->impl->m.ports[port].ijk;
    std::memcpy(ijk, v.data(), v.size() * sizeof(int));
    return 0;
}
int cfd_port_samples(cfd_sim *s, int port, double *host) { NEED(s); return guard([&] { s->impl->port_samples(port, host); }); }
int cfd_port_gather(cfd_sim *s, int port, double *host)
{
    NEED(s);
    if (port < 0 || port >= (int)s->impl->m.ports.size() || !host) { cfd::g_error = "no such port"; return -1; }
    return guard([&] { s->impl->port_gather(port, host); });
}
int cfd_port_gather_static(cfd_sim *s, int port, double *host)
{
    NEED(s);
    if (port < 0 || port >= (int)s->impl->m.ports.size() || !host) { cfd::g_error = "no such port"; return -1; }
    return guard([&] {
        const size_t need = (size_t)s->impl->port_npts(port) * 6;
        if (need > cfd::kStaticGatherDoubles)
            throw cfd::Error("port has " + std::to_string(need / 6) + " points; the static gather holds at most " +
                             std::to_string(cfd::kStaticGatherDoubles / 6) + " (use cfd_port_gather)");
        int dev = 0;
        HIP_OK(hipGetDevice(&dev));
        if (dev < 0 || dev >= cfd::kMaxDevices) throw cfd::Error("device index out of range for the static gather");
        cfd::StaticScratch &st = cfd::static_scratch();
        std::lock_guard<std::mutex> lk(st.mu);
        if (!st.ptr[dev]) {  // first call on this device: its copy of the array
            void *p = nullptr;
            HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            st.ptr[dev] = static_cast<double *>(p);
        }
        s->impl->port_gather_via(port, st.ptr[dev], host);
    });
}
int cfd_port_dft(cfd_sim *s, int port, double *host) { NEED(s); return guard([&] { s->impl->port_dft(port, host); }); }
int cfd_port_intensity(cfd_sim *s, int port, double *host) { NEED(s); return guard([&] { s->impl->port_intensity(port, host); }); }
int cfd_port_flux(cfd_sim *s, int port, int axis, double *host)
{
    NEED(s);
    if (port < 0 || port >= (int)s->impl->m.ports.size() || axis < 0 || axis > 2 || !host) {
        cfd::g_error = "no such port, bad axis or null buffer";
        return -1;
    }
    return guard([&] {
        const cfd::Model &m = s->impl->m;
        const int npts = m.ports[port].npts(), nf = (int)m.freqs.size();
        if (!nf) throw cfd::Error("the run has no DFT frequencies");
        std::vector<double> dft((size_t)npts * nf * 12);
        s->impl->port_dft(port, dft.data());
        const double d[3] = {m.dx, m.dy, m.dz};
        const double h = m.dft_stride * m.dt;  // sample spacing of the running DFT
        if (cfd_dft_flux(dft.data(), npts, nf, axis, 2.0 * h * h * d[(axis + 1) % 3] * d[(axis + 2) % 3], host) < 0)
            throw cfd::Error("flux: bad arguments");
    });
}
int cfd_gain_read(cfd_sim *s, double *host) { NEED(s); return guard([&] { s->impl->gain_read(host); }); }
int cfd_dl_state(cfd_sim *s, double *host)
{
    NEED(s);
    int n = 0;
    const int rc = guard([&] { n = s->impl->dl_state(host); });
    return rc < 0 ? rc : n;
}
int cfd_gain_state(cfd_sim *s, double *host)
{
    NEED(s);
    int n = 0;
    const int rc = guard([&] { n = s->impl->gain_state(host); });
    return rc < 0 ? rc : n;
}
int cfd_special_cells(cfd_sim *s, int kind, int *ijk)
{
    NEED(s);
    int n = 0;
    const int rc = guard([&] {
        const cfd::Model &m = s->impl->m;
        if (kind != 0 && kind != 1) throw cfd::Error("kind must be 0 (Drude-Lorentz) or 1 (gain)");
        const std::vector<long long> &cells = kind ? m.gain_cells : m.dl_cells;
        n = (int)cells.size();
        if (ijk)
            for (size_t r = 0; r < cells.size(); r++) {
                const long long p = cells[r], kk = p % (m.Nz + 1), jj = (p / (m.Nz + 1)) % (m.Ny + 1);
                ijk[3 * r] = (int)(p / ((long long)(m.Ny + 1) * (m.Nz + 1)));
                ijk[3 * r + 1] = (int)jj;
                ijk[3 * r + 2] = (int)kk;
            }
    });
    return rc < 0 ? rc : n;
}
int cfd_tfsf_wave(cfd_sim *s, double *einc, double *hinc)
{
    NEED(s);
    int n = 0;
    const int rc = guard([&] {
        const cfd::Model &m = s->impl->m;
        if (!m.has_tfsf) throw cfd::Error("no TF/SF set-up");
        n = (int)m.Einc.size();
        if (einc) std::copy(m.Einc.begin(), m.Einc.end(), einc);
        if (hinc) std::copy(m.Hinc.begin(), m.Hinc.end(), hinc);
    });
    return rc < 0 ? rc : n;
}

int cfd_set_timing(cfd_sim *s, int on)
{
    NEED(s);
    return guard([&] {
        HIP_OK(hipStreamSynchronize(s->impl->stream));
        s->impl->timer.collect();
        s->impl->timer.acc.clear();
        s->impl->timer.on = on != 0;
    });
}

int cfd_kernel_times(cfd_sim *s, char (*names)[32], double *ms, int *launches, int cap)
{
    NEED(s);
    int n = 0;
    int rc = guard([&] {
        HIP_OK(hipStreamSynchronize(s->impl->stream));
        s->impl->timer.collect();
        for (auto &kv : s->impl->timer.acc) {
            if (n >= cap) break;
            snprintf(names[n], 32, "%s", kv.first.c_str());
            ms[n] = kv.second.first;
            launches[n] = kv.second.second;
            n++;
        }
    });
    return rc ? rc : n;
}

// ---------------------------------------------------------------------------
// batches
int cfd_batch_create(const cfd_config *cfgs, int n, cfd_batch **out)
{
    if (!cfgs || !out || n <= 0) { cfd::g_error = "null argument or empty batch"; return -1; }
    *out = nullptr;
    return guard([&] {
        const int prec = cfgs[0].precision;
        if (prec != CFD_F32 && prec != CFD_F64) throw cfd::Error("precision must be 32 or 64");
        std::vector<cfd::Model> models;
        for (int i = 0; i < n; i++) {
            if (cfgs[i].precision != prec) throw cfd::Error("batch instances must share one precision");
            models.push_back(cfd::build_model(cfgs[i]));
        }
        auto b = std::make_unique<cfd_batch>();
        if (prec == CFD_F32) b->impl = std::make_unique<cfd::Batch<float>>(std::move(models), 32);
        else b->impl = std::make_unique<cfd::Batch<double>>(std::move(models), 64);
        for (int i = 0; i < n; i++) {
            auto m = std::make_unique<cfd_sim>();
            m->impl = b->impl->member(i);
            b->members.push_back(std::move(m));
        }
        *out = b.release();
    });
}
int cfd_batch_destroy(cfd_batch *b)
{
    if (!b) return 0;
    b->members.clear();
    delete b;
    return 0;
}
#define NEED_BATCH(b)                                   \
    do {                                                \
        if (!(b) || !(b)->impl) {                       \
            cfd::g_error = "null batch handle";         \
            return -1;                                  \
        }                                               \
    } while (0)
int cfd_batch_size(cfd_batch *b) { NEED_BATCH(b); return b->impl->size(); }
int cfd_batch_step(cfd_batch *b, int count) { NEED_BATCH(b); return guard([&] { b->impl->step(count); }); }
int cfd_batch_sync(cfd_batch *b) { NEED_BATCH(b); return guard([&] { HIP_OK(hipStreamSynchronize(b->impl->stream)); }); }
int cfd_batch_member(cfd_batch *b, int i, cfd_sim **out)
{
    NEED_BATCH(b);
    if (!out || i < 0 || i >= (int)b->members.size()) { cfd::g_error = "no such batch member"; return -1; }
    *out = b->members[i].get();
    return 0;
}
int cfd_batch_device_bytes(cfd_batch *b, size_t *bytes)
{
    NEED_BATCH(b);
    if (bytes) *bytes = b->impl->device_bytes();
    return 0;
}
int cfd_batch_set_timing(cfd_batch *b, int on)
{
    NEED_BATCH(b);
    return guard([&] {
        HIP_OK(hipStreamSynchronize(b->impl->stream));
        b->impl->timer.collect();
        b->impl->timer.acc.clear();
        b->impl->timer.on = on != 0;
    });
}
int cfd_batch_kernel_times(cfd_batch *b, char (*names)[32], double *ms, int *launches, int cap)
{
    NEED_BATCH(b);
    int n = 0;
    int rc = guard([&] {
        HIP_OK(hipStreamSynchronize(b->impl->stream));
        b->impl->timer.collect();
        for (auto &kv : b->impl->timer.acc) {
            if (n >= cap) break;
            snprintf(names[n], 32, "%s", kv.first.c_str());
            ms[n] = kv.second.first;
            launches[n] = kv.second.second;
            n++;
        }
    });
    return rc ? rc : n;
}

// ---------------------------------------------------------------------------
// packs
int cfd_pack_create(int precision, cfd_pack **out)
{
    if (!out) { cfd::g_error = "null argument"; return -1; }
    *out = nullptr;
    return guard([&] {
        if (precision != CFD_F32 && precision != CFD_F64) throw cfd::Error("precision must be 32 or 64");
        auto p = std::make_unique<cfd_pack>();
        if (precision == CFD_F32) p->impl = std::make_unique<cfd::Pack<float>>();
        else p->impl = std::make_unique<cfd::Pack<double>>();
        p->precision = precision;
        *out = p.release();
    });
}
int cfd_pack_destroy(cfd_pack *p)
{
    if (!p) return 0;
    delete p;
    return 0;
}
#define NEED_PACK(p)                                    \
    do {                                                \
        if (!(p) || !(p)->impl) {                       \
            cfd::g_error = "null pack handle";          \
            return -1;                                  \
        }                                               \
    } while (0)
int cfd_pack_add(cfd_pack *p, cfd_sim *s)
{
    NEED_PACK(p);
    NEED(s);
    return guard([&] { p->impl->add(s->impl); });
}
int cfd_pack_remove(cfd_pack *p, cfd_sim *s)
{
    NEED_PACK(p);
    NEED(s);
    return guard([&] { p->impl->remove(s->impl); });
}
int cfd_pack_size(cfd_pack *p) { NEED_PACK(p); return p->impl->size(); }
int cfd_pack_check_finite(cfd_pack *p, cfd_sim *const *sims, int n, int *counts)
{
    NEED_PACK(p);
    if (n < 0 || (n && (!sims || !counts))) { cfd::g_error = "null argument"; return -1; }
    return guard([&] {
        std::vector<SimBase *> v;
        for (int q = 0; q < n; q++) {
            if (!sims[q] || !sims[q]->impl) throw cfd::Error("null simulation handle");
            v.push_back(sims[q]->impl);
        }
        p->impl->nonfinite(v.data(), n, counts);
    });
}
int cfd_pack_step(cfd_pack *p, int count) { NEED_PACK(p); return guard([&] { p->impl->step(count); }); }
int cfd_pack_sync(cfd_pack *p) { NEED_PACK(p); return guard([&] { HIP_OK(hipStreamSynchronize(p->impl->stream)); }); }
int cfd_pack_query(cfd_pack *p)
{
    NEED_PACK(p);
    const hipError_t e = hipStreamQuery(p->impl->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    cfd::g_error = std::string("HIP: ") + hipGetErrorString(e);
    return -1;
}
int cfd_pack_set_timing(cfd_pack *p, int on)
{
    NEED_PACK(p);
    return guard([&] {
        HIP_OK(hipStreamSynchronize(p->impl->stream));
        p->impl->timer.collect();
        p->impl->timer.acc.clear();
        p->impl->timer.on = on != 0;
    });
}
int cfd_pack_kernel_times(cfd_pack *p, char (*names)[32], double *ms, int *launches, int cap)
{
    NEED_PACK(p);
    int n = 0;
    int rc = guard([&] {
        HIP_OK(hipStreamSynchronize(p->impl->stream));
        p->impl->timer.collect();
        for (auto &kv : p->impl->timer.acc) {
            if (n >= cap) break;
            snprintf(names[n], 32, "%s", kv.first.c_str());
            ms[n] = kv.second.first;
            launches[n] = kv.second.second;
            n++;
        }
    });
    return rc ? rc : n;
}

}  // extern "C"
