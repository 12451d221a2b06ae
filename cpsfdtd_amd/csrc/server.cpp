// server.cpp -- serving loop over the C ABI: requests (one simulation each)
// are admitted (first come first served, or least work first) while their
// device memory fits a budget, then stepped in quanta, each request on its
// own stream (the streams' small passes overlap on the GPU), or, for
// requests below a size, as members of a pack, where each pass of a step is
// one launch over all of them whatever their sizes (continuous batching
// across shapes; requests join and leave between quanta).  A finished request's
// port results are copied to host memory and its device memory freed at
// once.  submit / poll / result are safe from any thread.
//
// The reference runs one simulation per process (main.cpp:115-343); this has
// no counterpart there.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cpsfdtd.h"

namespace {

using Clock = std::chrono::steady_clock;

// A deep copy of a cfd_config (the caller's strings and arrays may go away).
struct ConfigCopy {
    std::string material, structure, pml, tfsf, dl, gain, source;
    bool has[7] = {};
    std::vector<std::string> ports;
    std::vector<const char *> port_ptrs;
    std::vector<int> kinds, src_ijkc;
    std::vector<double> amp, phase, freqs;
    cfd_config cfg{};

    // c has passed cfd_config_check; port_kinds, src_amp and src_phase are
    // optional (null: plane ports, amplitude 1, phase 0) and stay null here
    explicit ConfigCopy(const cfd_config &c) : cfg(c) {
        const char *in[7] = {c.material, c.structure, c.pml, c.tfsf, c.dl, c.gain, c.source};
        std::string *out[7] = {&material, &structure, &pml, &tfsf, &dl, &gain, &source};
        for (int i = 0; i < 7; i++)
            if (in[i]) { *out[i] = in[i]; has[i] = true; }
        for (int q = 0; q < c.n_ports; q++) {
            ports.push_back(c.port_files[q]);
            if (c.port_kinds) kinds.push_back(c.port_kinds[q]);
        }
        if (c.n_src > 0) {
            src_ijkc.assign(c.src_ijkc, c.src_ijkc + 4 * c.n_src);
            if (c.src_amp) amp.assign(c.src_amp, c.src_amp + c.n_src);
            if (c.src_phase) phase.assign(c.src_phase, c.src_phase + c.n_src);
        }
        if (c.n_freq > 0) freqs.assign(c.freqs, c.freqs + c.n_freq);
    }
    const cfd_config *get() {
        const std::string *s[7] = {&material, &structure, &pml, &tfsf, &dl, &gain, &source};
        const char **f[7] = {&cfg.material, &cfg.structure, &cfg.pml, &cfg.tfsf, &cfg.dl, &cfg.gain, &cfg.source};
        for (int i = 0; i < 7; i++) *f[i] = has[i] ? s[i]->c_str() : nullptr;
        port_ptrs.clear();
        for (const std::string &p : ports) port_ptrs.push_back(p.c_str());
        cfg.port_files = port_ptrs.empty() ? nullptr : port_ptrs.data();
        cfg.port_kinds = kinds.empty() ? nullptr : kinds.data();
        cfg.src_ijkc = src_ijkc.empty() ? nullptr : src_ijkc.data();
        cfg.src_amp = amp.empty() ? nullptr : amp.data();
        cfg.src_phase = phase.empty() ? nullptr : phase.data();
        cfg.freqs = freqs.empty() ? nullptr : freqs.data();
        return &cfg;
    }
};

// A result array filled by a device read-back (no zero fill first: a
// plane port's per-point DFT runs to hundreds of MB).
struct ResultBuf {
    std::unique_ptr<double[]> v;
    size_t n = 0;
    explicit ResultBuf(size_t count = 0) : v(count ? new double[count] : nullptr), n(count) {}
};

struct Request {
    long long id = 0;
    int state = CFD_REQ_QUEUED;
    std::unique_ptr<ConfigCopy> cfg;
    cfd_model *model = nullptr;  // built ahead of admission
    size_t estimate = 0;
    double work = 0;             // cells x steps, once the model is built
    double cells = 0;            // Nx Ny Nz, once the model is built
    cfd_sim *sim = nullptr;
    size_t bytes = 0;
    int nt = 0, done = 0;
    bool inflight = false;      // per-stream mode: a quantum is running
    bool in_pack = false;       // packed mode: a member of a pack
    int pack = 0;               // which pack
    double t_submit = 0, t_start = 0, t_done = 0;
    std::string error;
    std::vector<ResultBuf> intensity, dft;  // per port
};

}  // namespace

// Two threads: the stepping thread only launches quanta and polls for their
// completion; the host thread builds models, creates simulations (within the
// budget), reads back finished requests and releases them, so host work never
// holds up the GPU queue.
struct cfd_server {
    static constexpr int kShortestWindow = 64;  // queued requests a shortest-first choice looks at
    cfd_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv_host, cv_step, cv_done;
    std::deque<long long> queue;          // submitted, not yet running
    std::map<long long, std::unique_ptr<Request>> reqs;
    std::vector<Request *> active;        // stepping
    std::deque<Request *> retiring;       // stepping finished, results not yet on the host
    size_t active_bytes = 0;
    bool budget_blocked = false;          // the queue head waits for memory to be released
    long long next_id = 1;
    bool stop = false;
    Clock::time_point t0 = Clock::now();
    std::thread stepper, host;
    std::atomic<long long> steps_launched{0};
    size_t peak_bytes = 0;
    int peak_active = 0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }


    // host thread: results to host memory, device memory released
    void finish(Request &r, bool ok, std::string err) {
        if (ok && r.sim) {
            int counts[4] = {0, 0, 0, 0};
            cfd_info(r.sim, nullptr, nullptr, counts);
            for (int q = 0; q < counts[0] && ok; q++) {
                const int npts = cfd_port_npts(r.sim, q);
                ResultBuf in((size_t)r.nt), df;
                ok = cfd_port_intensity(r.sim, q, in.v.get()) >= 0;
                const int nf = r.cfg->cfg.n_freq;
                if (ok && nf > 0 && npts > 0) {
                    df = ResultBuf((size_t)npts * nf * 12);
                    ok = cfd_port_dft(r.sim, q, df.v.get()) >= 0;
                }
                r.intensity.push_back(std::move(in));
                r.dft.push_back(std::move(df));
            }
            if (!ok) err = cfd_last_error();
        }
        if (r.sim) cfd_destroy(r.sim);
        r.sim = nullptr;
        if (r.model) cfd_model_free(r.model);
        r.model = nullptr;
        std::lock_guard<std::mutex> lk(mu);
        active_bytes -= r.bytes;
        budget_blocked = false;
        r.state = ok ? CFD_REQ_DONE : CFD_REQ_FAILED;
        r.error = ok ? std::string() : err;
        r.t_done = now();
        r.cfg.reset();
        cv_done.notify_all();
    }

    // host thread: a queued request's model, its device-memory estimate and
    // work; a request whose input fails to build fails alone (false)
    bool ensure_model(Request &r) {
        if (r.model) return true;
        if (cfd_model_build(r.cfg->get(), &r.model) < 0) {
            const std::string err = cfd_model_last_error();
            r.model = nullptr;
            {
                std::lock_guard<std::mutex> lk(mu);
                queue.erase(std::find(queue.begin(), queue.end(), r.id));
            }
            finish(r, false, err);
            return false;
        }
        int n[4] = {0, 0, 0, 0};
        cfd_model_dims(r.model, n, nullptr);
        // device bytes from the built model (ports, DFT, recorded samples,
        // CPML, dispersive state included); the simulation's own figure
        // replaces it once it exists
        r.estimate = cfd_model_device_bytes(r.model, r.cfg->cfg.precision);
        r.cells = (double)n[0] * n[1] * n[2];
        r.work = r.cells * std::max(n[3], 1);
        return true;
    }

    // host thread: start the next request by the policy (the queue head, or
    // the queued request with the least work) if it fits; false when it must
    // wait
    bool admit_head() {
        Request *r = nullptr;
        std::vector<Request *> queued;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (queue.empty() || (int)active.size() >= opts.max_active) return false;
            if (opts.policy == CFD_POLICY_SHORTEST) {
                // the least work among the oldest kShortestWindow requests: only
                // their models are held, whatever the queue's length
                for (long long id : queue) {
                    if ((int)queued.size() == kShortestWindow) break;
                    queued.push_back(reqs[id].get());
                }
            } else
                r = reqs[queue.front()].get();
        }
        if (r) {
            if (!ensure_model(*r)) return true;
        } else {
            for (Request *x : queued)
                if (!ensure_model(*x)) return true;
            std::lock_guard<std::mutex> lk(mu);
            for (Request *x : queued)  // all still queued: only this thread removes requests
                if (!r || x->work < r->work) r = x;
            if (!r) return true;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            if (!active.empty() && active_bytes + r->estimate > opts.hbm_budget) {
                budget_blocked = true;  // wait for a release
                return false;
            }
            queue.erase(std::find(queue.begin(), queue.end(), r->id));
            r->state = CFD_REQ_RUNNING;
            r->t_start = now();
        }
        cfd_model *m = r->model;
        r->model = nullptr;
        cfd_sim *s = nullptr;
        if (cfd_create_from_model(m, r->cfg->cfg.precision, &s) < 0) {
            finish(*r, false, cfd_last_error());
            return true;
        }
        r->sim = s;
        r->bytes = cfd_device_bytes(s);
        int info[4];
        cfd_info(s, info, nullptr, nullptr);
        r->nt = info[3];
        std::lock_guard<std::mutex> lk(mu);
        active.push_back(r);
        active_bytes += r->bytes;
        peak_bytes = std::max(peak_bytes, active_bytes);
        peak_active = std::max(peak_active, (int)active.size());
        cv_step.notify_all();
        return true;
    }

    void host_loop() {
        for (;;) {
            Request *ret = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_host.wait(lk, [&] {
                    return stop || !retiring.empty() ||
                           (!queue.empty() && !budget_blocked && (int)active.size() < opts.max_active);
                });
                if (stop) break;
                if (!retiring.empty()) {
                    ret = retiring.front();
                    retiring.pop_front();
                }
            }
            if (ret) finish(*ret, ret->error.empty(), ret->error);
            else admit_head();
        }
    }

    // Stepping thread.  Requests below opts.pack_cells cells are members of
    // a pack of their precision: each pass of a step is one launch over all
    // of them, whatever their sizes.  Other requests run their quanta on
    // their own streams, beside the packs.  Every idle pack and stream
    // request gets its next quantum (asynchronous); completed quanta are then
    // checked (non-finite watchdog) and finished or failed requests handed to
    // the host thread.
    // A pack quantum has opts.quantum steps; a member whose run ends inside it
    // idles for the rest (equal lengths let a stable membership replay
    // captured sequences).  New members join between quanta.
    // Two packs per precision, on their own streams, so that one pack's
    // short passes (TF/SF, special cells, ports) overlap the other's updates;
    // a new member joins the one with fewer cells.  Pack 2 w + h: precision w
    // (0 FP32, 1 FP64), half h.
    static constexpr int kPacks = 4;
    cfd_pack *packs[kPacks] = {};
    bool pack_busy[kPacks] = {};
    double pack_load[kPacks] = {};           // member cells
    std::vector<Request *> pack_run[kPacks];  // members of the quantum in flight

    bool packed(const Request *r) const { return r->cells < opts.pack_cells; }

    void launch_packs(const std::vector<Request *> &act) {
        for (Request *r : act) {
            if (!packed(r) || r->in_pack) continue;
            const int base = r->cfg->cfg.precision == CFD_F64 ? 2 : 0;
            const int w = pack_load[base] <= pack_load[base + 1] ? base : base + 1;  // the lighter half
            if (pack_busy[w]) continue;  // joins between quanta
            cfd_pack *&p = packs[w];
            if ((!p && cfd_pack_create(r->cfg->cfg.precision, &p) < 0) || cfd_pack_add(p, r->sim) < 0) {
                retire(r, cfd_last_error());
                continue;
            }
            r->in_pack = true;
            r->pack = w;
            pack_load[w] += r->cells;
        }
        for (int w = 0; w < kPacks; w++) {
            if (pack_busy[w] || !packs[w]) continue;
            pack_run[w].clear();
            for (Request *r : act)
                if (r->in_pack && r->pack == w) pack_run[w].push_back(r);
            if (pack_run[w].empty()) continue;
            if (cfd_pack_step(packs[w], opts.quantum) < 0) {
                const std::string e = cfd_last_error();
                for (Request *r : pack_run[w]) {
                    cfd_pack_remove(packs[w], r->sim);
                    r->in_pack = false;
                    pack_load[w] -= r->cells;
                    retire(r, e);
                }
                continue;
            }
            pack_busy[w] = true;
        }
    }
    // a pack quantum completed (or failed with err): watchdog, step counts,
    // finished and failed members leave the pack
    void finish_pack_quantum(int w, std::string err) {
        std::vector<Request *> &mem = pack_run[w];
        std::vector<cfd_sim *> sims;
        for (Request *r : mem) sims.push_back(r->sim);
        std::vector<int> bad(sims.size(), 0);
        // watchdog: a request whose fields went non-finite fails now, alone,
        // instead of running on and returning garbage
        if (err.empty() && cfd_pack_check_finite(packs[w], sims.data(), (int)sims.size(), bad.data()) < 0)
            err = cfd_last_error();
        for (size_t k = 0; k < mem.size(); k++) {
            Request *r = mem[k];
            const int ran = std::min(opts.quantum, r->nt - r->done);
            r->done += ran;
            steps_launched += ran;
            std::string e = err;
            if (e.empty() && bad[k] > 0)
                e = "non-finite field values (" + std::to_string(bad[k]) + ") by step " + std::to_string(r->done);
            if (!e.empty() || r->done >= r->nt) {
                cfd_pack_remove(packs[w], r->sim);
                r->in_pack = false;
                pack_load[w] -= r->cells;
                retire(r, e);
            }
        }
        mem.clear();
        pack_busy[w] = false;
    }

    void step_loop() {
        for (;;) {
            std::vector<Request *> act;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_step.wait(lk, [&] {
                    return stop || !active.empty() || std::any_of(pack_busy, pack_busy + kPacks, [](bool b) { return b; });
                });
                if (stop) break;
                act = active;
            }
            launch_packs(act);
            // every idle stream request gets its next quantum; a request
            // retired here belongs to the host thread from then on (it may be
            // finished and released at once), so the loop below must not
            // touch it
            std::vector<Request *> stepping;
            double cells_in_flight = 0;
            for (Request *r : act)
                if (!packed(r) && r->inflight) cells_in_flight += r->cells;
            for (Request *r : act) {
                if (packed(r)) continue;
                if (r->inflight) {
                    stepping.push_back(r);
                    continue;
                }
                // the step_cells budget: requests in admission order, so the
                // earliest keep stepping and later ones wait for room
                if (opts.step_cells > 0 && cells_in_flight > 0 && cells_in_flight + r->cells > opts.step_cells) continue;
                const int k = std::min(opts.quantum, r->nt - r->done);
                if (k <= 0 || cfd_step(r->sim, k) < 0) {
                    retire(r, k > 0 ? std::string(cfd_last_error()) : std::string());
                    continue;
                }
                r->done += k;
                r->inflight = true;
                cells_in_flight += r->cells;
                steps_launched += k;
                stepping.push_back(r);
            }
            bool any = false;
            for (int w = 0; w < kPacks; w++) {
                if (!pack_busy[w]) continue;
                const int q = cfd_pack_query(packs[w]);
                if (q == 0) continue;
                any = true;
                finish_pack_quantum(w, q < 0 ? std::string(cfd_last_error()) : std::string());
            }
            for (Request *r : stepping) {
                const int q = cfd_query(r->sim);
                if (q == 0) continue;
                any = true;
                r->inflight = false;
                if (q < 0) { retire(r, cfd_last_error()); continue; }
                // watchdog, as for pack members
                const int bad = cfd_check_finite(r->sim);
                if (bad != 0) {
                    retire(r, bad < 0 ? std::string(cfd_last_error())
                                      : "non-finite field values (" + std::to_string(bad) + ") by step " + std::to_string(r->done));
                    continue;
                }
                if (r->done >= r->nt) retire(r, std::string());
            }
            if (!any) std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
        // stopping: the packs let go of their members (shutdown fails them)
        for (int w = 0; w < kPacks; w++) {
            if (packs[w]) {
                cfd_pack_sync(packs[w]);
                cfd_pack_destroy(packs[w]);
            }
            packs[w] = nullptr;
            pack_busy[w] = false;
            pack_load[w] = 0;
            pack_run[w].clear();
        }
    }
    // stepping thread: hand a request over to the host thread
    void retire(Request *r, const std::string &err) {
        std::lock_guard<std::mutex> lk(mu);
        auto it = std::find(active.begin(), active.end(), r);
        if (it != active.end()) active.erase(it);
        r->error = err;
        retiring.push_back(r);
        cv_host.notify_all();
    }
    void shutdown() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_host.notify_all();
        cv_step.notify_all();
        if (stepper.joinable()) stepper.join();
        if (host.joinable()) host.join();
        std::vector<Request *> left(active.begin(), active.end());
        left.insert(left.end(), retiring.begin(), retiring.end());
        active.clear();
        retiring.clear();
        for (Request *r : left) finish(*r, false, "server stopped");
        for (long long id : queue) finish(*reqs[id], false, "server stopped");
        queue.clear();
    }
};

static thread_local std::string g_server_error;

extern "C" {

int cfd_server_create(const cfd_server_opts *o, cfd_server **out)
{
    if (!out) { g_server_error = "null argument"; return -1; }
    auto s = std::make_unique<cfd_server>();
    s->opts.hbm_budget = o && o->hbm_budget ? o->hbm_budget : (size_t)200 << 30;
    // defaults measured on the first 120 requests of the mixed-size FP32
    // stream (MI355X): one stream per request, 8 at once, 256-step quanta:
    // 39.0 Gcell/s, p50 14.1 s, p99 24.5 s.  4 at once: 38.6; 64 at once:
    // the same rate and p50 32 s (the GPU is saturated either way, fewer at
    // once finish early arrivals sooner).  Packing requests below 2e6 cells
    // (two packs per precision): 36.3, below 4e6 cells: 34.2 (the small
    // grids' passes are CPML-bound in the kernels themselves; packing removes
    // launches, which were not the limit).
    s->opts.max_active = o && o->max_active > 0 ? o->max_active : 8;
    s->opts.quantum = o && o->quantum > 0 ? o->quantum : 256;
    s->opts.policy = o && o->policy == CFD_POLICY_SHORTEST ? CFD_POLICY_SHORTEST : CFD_POLICY_FIFO;
    s->opts.pack_cells = o && o->pack_cells > 0 ? o->pack_cells : 0;
    // at most 2e7 cells stepping at once: a grid that size fills the GPU
    // alone, and two of them at once share the caches (four whole 256^3
    // runs submitted together: 51.6 Gcell/s served without the limit, 54.4
    // one at a time, with the next one's set-up and the last one's results
    // overlapping the stepping; the mixed-size stream: 41.6-42.7 with and
    // without it)
    s->opts.step_cells = !o || o->step_cells == 0 ? 2e7 : o->step_cells > 0 ? o->step_cells : 0;
    cfd_server *p = s.get();
    p->stepper = std::thread([p] { p->step_loop(); });
    p->host = std::thread([p] { p->host_loop(); });
    *out = s.release();
    return 0;
}

int cfd_server_destroy(cfd_server *s)
{
    if (!s) return 0;
    s->shutdown();
    delete s;
    return 0;
}

int cfd_server_submit(cfd_server *s, const cfd_config *cfg, long long *id)
{
    if (!s || !cfg || !id) { g_server_error = "null argument"; return -1; }
    if (cfd_config_check(cfg) < 0) {  // a malformed request is refused here, before it is copied
        g_server_error = std::string("request rejected: ") + cfd_model_last_error();
        return -1;
    }
    auto r = std::make_unique<Request>();
    r->cfg = std::make_unique<ConfigCopy>(*cfg);
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->t_submit = s->now();
    *id = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    s->cv_host.notify_all();
    return 0;
}

int cfd_server_poll(cfd_server *s, long long id, double *times)
{
    if (!s) { g_server_error = "null server"; return -1; }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { g_server_error = "no such request"; return -1; }
    const Request &r = *it->second;
    if (times) { times[0] = r.t_submit; times[1] = r.t_start; times[2] = r.t_done; }
    return r.state;
}

int cfd_server_wait(cfd_server *s, long long id, double timeout_s)
{
    if (!s) { g_server_error = "null server"; return -1; }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { g_server_error = "no such request"; return -1; }
    Request *r = it->second.get();
    s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s),
                   [&] { return r->state == CFD_REQ_DONE || r->state == CFD_REQ_FAILED; });
    return r->state;
}

int cfd_server_result(cfd_server *s, long long id, int port, int what, double *host, size_t *count)
{
    if (!s) { g_server_error = "null server"; return -1; }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { g_server_error = "no such request"; return -1; }
    const Request &r = *it->second;
    if (r.state == CFD_REQ_FAILED) { g_server_error = "request failed: " + r.error; return -1; }
    if (r.state != CFD_REQ_DONE) { g_server_error = "request not finished"; return -1; }
    if (port < 0 || port >= (int)r.intensity.size()) { g_server_error = "no such port"; return -1; }
    const ResultBuf &v = what == CFD_RESULT_DFT ? r.dft[port] : r.intensity[port];
    if (count) *count = v.n;
    if (host && v.n) std::memcpy(host, v.v.get(), v.n * sizeof(double));
    return 0;
}

const char *cfd_server_request_error(cfd_server *s, long long id)
{
    if (!s) return "null server";
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return "no such request";
    g_server_error = it->second->error;
    return g_server_error.c_str();
}

int cfd_server_release(cfd_server *s, long long id)
{
    if (!s) { g_server_error = "null server"; return -1; }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { g_server_error = "no such request"; return -1; }
    if (it->second->state != CFD_REQ_DONE && it->second->state != CFD_REQ_FAILED) {
        g_server_error = "request not finished";
        return -1;
    }
    s->reqs.erase(it);
    return 0;
}

int cfd_server_stats(cfd_server *s, long long *steps_launched, size_t *peak_bytes, int *peak_active, int *queued)
{
    if (!s) { g_server_error = "null server"; return -1; }
    std::lock_guard<std::mutex> lk(s->mu);
    if (steps_launched) *steps_launched = s->steps_launched.load();
    if (peak_bytes) *peak_bytes = s->peak_bytes;
    if (peak_active) *peak_active = s->peak_active;
    if (queued) *queued = (int)s->queue.size();
    return 0;
}

const char *cfd_server_last_error(void) { return g_server_error.c_str(); }

}  // extern "C"
