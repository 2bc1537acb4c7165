// Continuous-batching measurement server.
//
// `slots` worker threads each own a library context (its own HIP stream and
// resident meshes/spectra, reused by later requests of the same or smaller N).
// Submitted requests wait in a FIFO; a free worker takes the first request
// whose HBM footprint fits next to the requests already running, so requests
// that arrive while others are on the GPU start as soon as a slot and memory
// free up, and their kernels overlap on the device through the slots' streams.
// A request either brings device particles or asks for a seeded Zel'dovich
// mock generated on the GPU into the slot's own particle buffer.
//
// Every request runs the same kernels with the same launch shapes whatever
// runs beside it, and the deposit and shell sums do not depend on the
// interleaving beyond the order of FP64 additions into the shell sums, so a
// result in the mix equals the result alone to rounding of those sums.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "cm_internal.hpp"
#include "fft_ct.hpp"

namespace cm {

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Device bytes a request needs while it runs: particles (when generated),
// binned records (~2.2 per particle for TSC on 16^3 tiles, 16 B each plus
// headroom), mesh, padded half spectrum, potential and labels, Zel'dovich
// spectra and displacements when generated, and fixed tables.
size_t request_footprint(const cm_request *r) {
    const size_t N = (size_t)r->opts.mas.N;
    const size_t n3 = N * N * N;
    const size_t pitch = (size_t)padded_pitch((int)N);
    size_t n = r->n, b = 0;
    if (!r->pos && r->gen.NL > 0) {
        const size_t NL = (size_t)r->gen.NL, nl3 = NL * NL * NL;
        n = nl3;
        b += 12 * nl3 + 3 * 8 * NL * NL * (NL / 2 + 1) + 3 * 4 * nl3;
    }
    b += 40 * n;                    // records + weights + headroom
    b += 4 * n3 + 8 * N * N * pitch;  // mesh + spectrum
    if (r->opts.do_tweb) b += 4 * n3;
    b += (size_t)64 << 20;          // histograms, tables, work lists
    return b;
}

struct Job {
    cm_request req;
    size_t footprint = 0;
    cm_status status = CM_OK;
    bool done = false;
    std::string err;
    double t_submit = 0, t_start = 0, t_done = 0;
};

}  // namespace cm

struct cm_server {
    int device = 0;
    size_t budget = 0;
    size_t inflight = 0;
    std::vector<cm_ctx *> slots;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, cm::Job> jobs;
    int64_t next = 1;
    bool stop = false;
};

namespace cm {

static cm_status run_job(cm_ctx *ctx, Job &j) {
    const cm_request &r = j.req;
    const void *pos = r.pos;
    size_t n = r.n;
    cm_status s;
    if (!pos && r.gen.NL > 0) {
        const size_t nl3 = (size_t)r.gen.NL * r.gen.NL * r.gen.NL;
        void *buf;
        if ((s = scratch(ctx, "srv.pos", 12 * nl3, &buf)) != CM_OK) return s;
        if ((s = cm_zeldovich(ctx, &r.gen, (float *)buf)) != CM_OK) return s;
        pos = buf;
        n = nl3;
    }
    return measure_run(ctx, &r.opts, pos, r.w, n, r.pk0, r.pk2, r.pk4, r.nmodes, r.labels, r.info);
}

static void worker_main(cm_server *sv, int slot) {
    hipSetDevice(sv->device);
    cm_ctx *ctx = sv->slots[slot];
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        int64_t t = 0;
        sv->cv_work.wait(lk, [&] {
            if (sv->stop) return true;
            for (int64_t id : sv->queue) {
                const Job &j = sv->jobs[id];
                if (sv->inflight == 0 || sv->inflight + j.footprint <= sv->budget) return true;
            }
            return false;
        });
        if (sv->stop) return;
        for (auto it = sv->queue.begin(); it != sv->queue.end(); ++it) {
            const Job &j = sv->jobs[*it];
            if (sv->inflight == 0 || sv->inflight + j.footprint <= sv->budget) {
                t = *it;
                sv->queue.erase(it);
                break;
            }
        }
        Job &job = sv->jobs[t];
        sv->inflight += job.footprint;
        job.t_start = now_ms();
        Job local = job;  // the map may rehash while unlocked
        lk.unlock();
        const cm_status st = run_job(ctx, local);
        const std::string err = st == CM_OK ? std::string() : std::string(cm_last_error(ctx));
        lk.lock();
        Job &done = sv->jobs[t];
        done.status = st;
        done.err = err;
        done.t_done = now_ms();
        done.done = true;
        sv->inflight -= done.footprint;
        sv->cv_done.notify_all();
        sv->cv_work.notify_all();
    }
}

}  // namespace cm

using namespace cm;

extern "C" {

size_t cm_request_footprint(const cm_request *r) { return r ? request_footprint(r) : 0; }

cm_status cm_server_create(int device, int slots, size_t hbm_budget, cm_server **out) {
    if (!out || slots < 1 || slots > 16) return CM_ERR_INVALID;
    *out = nullptr;
    cm_server *sv = new cm_server();
    sv->device = device;
    for (int i = 0; i < slots; ++i) {
        cm_ctx *c = nullptr;
        cm_status s = cm_ctx_create(device, hbm_budget, &c);
        if (s != CM_OK) {
            for (cm_ctx *x : sv->slots) cm_ctx_destroy(x);
            delete sv;
            return s;
        }
        sv->slots.push_back(c);
    }
    sv->budget = sv->slots[0]->budget;
    for (int i = 0; i < slots; ++i) sv->workers.emplace_back(worker_main, sv, i);
    *out = sv;
    return CM_OK;
}

cm_status cm_server_submit(cm_server *sv, const cm_request *r, int64_t *ticket) {
    if (!sv || !r || !ticket) return CM_ERR_INVALID;
    if (!r->pos && r->gen.NL <= 0 && r->n) return CM_ERR_INVALID;
    Job j;
    j.req = *r;
    j.footprint = request_footprint(r);
    if (j.footprint > sv->budget) return CM_ERR_BUDGET;
    j.t_submit = now_ms();
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return CM_ERR_CANCELLED;
    const int64_t t = sv->next++;
    sv->jobs.emplace(t, j);
    sv->queue.push_back(t);
    *ticket = t;
    sv->cv_work.notify_all();
    return CM_OK;
}

cm_status cm_server_wait(cm_server *sv, int64_t ticket, double timeout_s, double *latency_ms, char *err,
                         size_t errlen) {
    if (!sv) return CM_ERR_INVALID;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(ticket);
    if (it == sv->jobs.end()) return CM_ERR_INVALID;
    auto ready = [&] { return sv->jobs[ticket].done; };
    if (timeout_s > 0) {
        if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return CM_ERR_TIMEOUT;
    } else {
        sv->cv_done.wait(lk, ready);
    }
    Job &j = sv->jobs[ticket];
    if (latency_ms) *latency_ms = j.t_done - j.t_submit;
    if (err && errlen) {
        std::strncpy(err, j.err.c_str(), errlen - 1);
        err[errlen - 1] = 0;
    }
    const cm_status st = j.status;
    sv->jobs.erase(ticket);
    return st;
}

cm_status cm_server_destroy(cm_server *sv) {
    if (!sv) return CM_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        for (int64_t t : sv->queue) {
            Job &j = sv->jobs[t];
            j.status = CM_ERR_CANCELLED;
            j.err = "server destroyed";
            j.done = true;
        }
        sv->queue.clear();
        sv->cv_work.notify_all();
        sv->cv_done.notify_all();
    }
    for (auto &w : sv->workers) w.join();
    for (cm_ctx *c : sv->slots) cm_ctx_destroy(c);
    delete sv;
    return CM_OK;
}

cm_status cm_measure_batch(cm_server *sv, const cm_request *reqs, int n, cm_status *statuses, double *latency_ms) {
    if (!sv || (n > 0 && !reqs)) return CM_ERR_INVALID;
    std::vector<int64_t> t((size_t)std::max(n, 0), 0);
    cm_status first = CM_OK;
    for (int i = 0; i < n; ++i) {
        cm_status s = cm_server_submit(sv, &reqs[i], &t[i]);
        if (s != CM_OK) {
            t[i] = 0;
            if (statuses) statuses[i] = s;
            if (first == CM_OK) first = s;
        }
    }
    for (int i = 0; i < n; ++i) {
        if (!t[i]) continue;
        double lat = 0;
        cm_status s = cm_server_wait(sv, t[i], 0.0, &lat, nullptr, 0);
        if (statuses) statuses[i] = s;
        if (latency_ms) latency_ms[i] = lat;
        if (s != CM_OK && first == CM_OK) first = s;
    }
    return first;
}

}  // extern "C"
