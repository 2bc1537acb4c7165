// Continuous-batching measurement server.
//
// `slots` worker threads each own a library context (its own HIP stream and
// resident meshes/spectra, reused by later requests of the same or smaller N).
// Submitted requests wait in a FIFO; a free worker takes the first request
// whose reservation fits in the one HBM budget beside what the other slots
// hold (running jobs' reservations, idle slots' resident buffers, which are
// freed when a queued request needs the room), so requests that arrive while
// others are on the GPU start as soon as a slot and memory free up, and their
// kernels overlap on the device through the slots' streams.
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
    bool alone = false;  // ran out of its reservation once: runs with the whole budget
    std::string err;
    double t_submit = 0, t_start = 0, t_done = 0;
};

// Reservation of a running job: what its slot already holds plus its
// footprint estimate with a quarter and 256 MB of headroom, at most the budget.
// The resident buffers are kept: a request of another shape may grow any of
// them by up to its own footprint, so they do not count towards it.
static size_t reservation(size_t footprint, size_t resident, size_t budget) {
    return std::min(resident + footprint + footprint / 4 + ((size_t)256 << 20), budget);
}

}  // namespace cm

// One HBM budget for all slots.  An idle slot holds its resident work buffers
// (ctx->resident); a running slot holds its reservation, and its context may
// not grow past it (ctx->budget is set to it for the run), so the slots
// together never hold more than the budget.  A queued request that does not
// fit frees the work buffers of idle slots, largest first, until it does.
struct cm_server {
    int device = 0;
    size_t budget = 0;
    std::vector<cm_ctx *> slots;
    std::vector<char> running;
    std::vector<size_t> held;  // running slots: their reservation
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, cm::Job> jobs;
    int64_t next = 1;
    bool stop = false;
    int64_t evictions = 0, reruns = 0;
    size_t evicted_bytes = 0, peak_held = 0;
    std::string last_err;  // message of the most recent failed request
};

namespace cm {

// Bytes the slots other than s hold (under the server lock).
static size_t held_by_others(const cm_server *sv, int s) {
    size_t h = 0;
    for (size_t i = 0; i < sv->slots.size(); ++i) {
        if ((int)i == s) continue;
        h += sv->running[i] ? sv->held[i] : sv->slots[i]->resident;
    }
    return h;
}

static bool any_running(const cm_server *sv) {
    for (char r : sv->running)
        if (r) return true;
    return false;
}

// Whether job j may start on slot s now, and the bytes it would reserve.  A job
// marked alone starts only when no other job runs and may use the whole
// budget.  with_eviction: count idle slots' buffers as free.
static bool fits(const cm_server *sv, int s, const Job &j, bool with_eviction, size_t *resv) {
    if (j.alone) {
        if (any_running(sv)) return false;
        *resv = sv->budget;
        return with_eviction || held_by_others(sv, s) == 0;
    }
    const size_t own = with_eviction ? 0 : sv->slots[s]->resident;
    size_t others = 0;
    for (size_t i = 0; i < sv->slots.size(); ++i) {
        if ((int)i == s) continue;
        if (sv->running[i]) others += sv->held[i];
        else if (!with_eviction) others += sv->slots[i]->resident;
    }
    *resv = reservation(j.footprint, own, sv->budget);
    return others + *resv <= sv->budget;
}

// Frees idle slots' work buffers (largest first, the admitting slot s
// included) until job j fits on s.
static void evict_for(cm_server *sv, int s, const Job &j) {
    size_t resv = 0;
    for (;;) {
        if (fits(sv, s, j, false, &resv)) return;
        int victim = -1;
        for (size_t i = 0; i < sv->slots.size(); ++i) {
            if (sv->running[i] || sv->slots[i]->resident == 0) continue;
            if (victim < 0 || sv->slots[i]->resident > sv->slots[victim]->resident) victim = (int)i;
        }
        if (victim < 0) return;
        const size_t before = sv->slots[victim]->resident;
        sv->evicted_bytes += scratch_release(sv->slots[victim]);
        ++sv->evictions;
        if (sv->slots[victim]->resident == before) return;  // only tables left
    }
}

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
        size_t resv = 0;
        // the first queued job that fits, counting idle slots' buffers as
        // evictable
        sv->cv_work.wait(lk, [&] {
            if (sv->stop) return true;
            for (int64_t id : sv->queue)
                if (fits(sv, slot, sv->jobs.find(id)->second, true, &resv)) return true;
            return false;
        });
        if (sv->stop) return;
        for (auto it = sv->queue.begin(); it != sv->queue.end(); ++it) {
            const Job &j = sv->jobs.find(*it)->second;
            if (fits(sv, slot, j, true, &resv)) {
                t = *it;
                sv->queue.erase(it);
                break;
            }
        }
        Job &job = sv->jobs.find(t)->second;
        evict_for(sv, slot, job);
        fits(sv, slot, job, false, &resv);
        sv->running[slot] = 1;
        sv->held[slot] = resv;
        sv->peak_held = std::max(sv->peak_held, held_by_others(sv, slot) + resv);
        ctx->budget = resv;  // the context may not grow past its reservation
        job.t_start = now_ms();
        Job local = job;  // the map may rehash while unlocked
        lk.unlock();
        const cm_status st = run_job(ctx, local);
        const std::string err = st == CM_OK ? std::string() : std::string(cm_last_error(ctx));
        lk.lock();
        sv->running[slot] = 0;
        sv->held[slot] = 0;
        Job &done = sv->jobs.find(t)->second;
        if (st == CM_ERR_BUDGET && !done.alone) {
            // the footprint estimate was short: run again with the whole budget
            // once nothing else runs
            done.alone = true;
            ++sv->reruns;
            sv->queue.push_front(t);
        } else {
            done.status = st;
            done.err = err;
            if (st != CM_OK) sv->last_err = "request " + std::to_string(t) + ": " + err;
            done.t_done = now_ms();
            done.done = true;
            sv->cv_done.notify_all();
        }
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
    sv->running.assign((size_t)slots, 0);
    sv->held.assign((size_t)slots, 0);
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
    if (sv->jobs.find(ticket) == sv->jobs.end()) return CM_ERR_INVALID;
    // a ticket collected by another waiter meanwhile ends the wait too
    auto ready = [&] {
        auto f = sv->jobs.find(ticket);
        return f == sv->jobs.end() || f->second.done;
    };
    if (timeout_s > 0) {
        if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return CM_ERR_TIMEOUT;
    } else {
        sv->cv_done.wait(lk, ready);
    }
    auto found = sv->jobs.find(ticket);
    if (found == sv->jobs.end()) return CM_ERR_INVALID;
    Job &j = found->second;
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
            Job &j = sv->jobs.find(t)->second;
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

cm_status cm_server_stats(cm_server *sv, int64_t *evictions, size_t *evicted_bytes, size_t *peak_held,
                          size_t *budget, int64_t *reruns) {
    if (!sv) return CM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (reruns) *reruns = sv->reruns;
    if (evictions) *evictions = sv->evictions;
    if (evicted_bytes) *evicted_bytes = sv->evicted_bytes;
    if (peak_held) *peak_held = sv->peak_held;
    if (budget) *budget = sv->budget;
    return CM_OK;
}

cm_status cm_server_last_error(cm_server *sv, char *buf, size_t len) {
    if (!sv || !buf || len == 0) return CM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    std::snprintf(buf, len, "%s", sv->last_err.c_str());
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
