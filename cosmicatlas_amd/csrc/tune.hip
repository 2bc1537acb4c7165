// Launch-shape autotuner and the process-wide table of tuned shapes.
//
// The fused path has a few launch-shape choices with no single best value
// across mesh sizes: the x-pass tile (kz columns, elements per thread), the
// T-web tile width and x segmentation, and the number of resident deposit
// workgroups.  cm_autotune times the real measurement under candidate values,
// one knob at a time, and records the winner per (device, N); cm_measure and
// cm_measure_fkp of every context lay those shapes over the context's own for
// the duration of the call.
#include <map>
#include <mutex>
#include <utility>

#include "cm_internal.hpp"

namespace cm {

static std::mutex g_tune_mu;
static std::map<std::pair<int, int>, cm_tune> g_tunes;  // (device, N) -> shapes

static bool lookup_tune(int device, int N, cm_tune *t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tunes.find({device, N});
    if (it == g_tunes.end()) return false;
    *t = it->second;
    return true;
}

static void store_tune(int device, int N, const cm_tune *t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (t)
        g_tunes[{device, N}] = *t;
    else
        g_tunes.erase({device, N});
}

static void lay_over(cm_ctx *ctx, const cm_tune &t) {
    if (t.xnb) ctx->tune.xnb = t.xnb;
    if (t.xe) ctx->tune.xe = t.xe;
    if (t.tz) ctx->tune.tz = t.tz;
    if (t.tseg) ctx->tune.tseg = t.tseg;
    if (t.dep_wg) ctx->tune.dep_wg = t.dep_wg;
}

TuneScope::TuneScope(cm_ctx *c, int N) : ctx(c), saved(c->tune) {
    cm_tune t;
    if (lookup_tune(c->device, N, &t)) lay_over(c, t);
}

}  // namespace cm

using namespace cm;

extern "C" {

cm_status cm_set_tune(cm_ctx *ctx, int N, const cm_tune *t) {
    if (!ctx) return CM_ERR_INVALID;
    store_tune(ctx->device, N, t);
    return CM_OK;
}

cm_status cm_get_tune(cm_ctx *ctx, int N, cm_tune *t) {
    if (!ctx || !t) return CM_ERR_INVALID;
    if (!lookup_tune(ctx->device, N, t)) *t = cm_tune{0, 0, 0, 0, 0};
    return CM_OK;
}

cm_status cm_autotune(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, const void *w, size_t n,
                      int8_t *labels, cm_tune *best_out, double *ms_before, double *ms_after) {
    if (!ctx) return CM_ERR_INVALID;
    if (!o) return fail(ctx, CM_ERR_INVALID, "cm_autotune: null options");
    const int N = o->mas.N;
    cm_measure_opts oo = *o;
    oo.fused = 1;
    const int ns = o->nshells > 0 ? o->nshells : 1;
    std::vector<double> p0((size_t)ns), p2((size_t)ns), p4((size_t)ns);
    std::vector<int64_t> nm((size_t)ns);
    const bool graphs = ctx->graphs, prof = ctx->profiling;
    cm_tune prev;
    const bool had = lookup_tune(ctx->device, N, &prev);
    ctx->graphs = false;
    // kernel time per mock under shapes t: best of `reps` runs
    auto timed = [&](const cm_tune &t, int reps, double *ms) -> cm_status {
        store_tune(ctx->device, N, &t);
        double bestv = 1e300;
        for (int r = 0; r < reps; ++r) {
            ctx->profiling = true;
            cm_profile_reset(ctx);
            const cm_status s = measure_run(ctx, &oo, pos, w, n, p0.data(), p2.data(), p4.data(), nm.data(), labels,
                                            nullptr);
            const double v = profiled_total_ms(ctx);
            ctx->profiling = false;
            if (s != CM_OK) return s;
            bestv = std::min(bestv, v);
        }
        *ms = bestv;
        return CM_OK;
    };
    auto restore = [&](cm_status s) {
        ctx->graphs = graphs;
        ctx->profiling = prof;
        cm_profile_reset(ctx);
        if (s != CM_OK) store_tune(ctx->device, N, had ? &prev : nullptr);
        return s;
    };
    cm_tune best{0, 0, 0, 0, 0};
    double best_ms = 0.0, t0 = 0.0;
    cm_status s;
    if ((s = timed(best, 1, &t0)) != CM_OK) return restore(s);  // warm-up: buffers, tables
    if ((s = timed(best, 2, &best_ms)) != CM_OK) return restore(s);
    t0 = best_ms;
    // candidate values per knob (0 = built-in)
    const std::vector<int> xnb_c = {16, 8}, xe_c = {16, 8}, dep_c = {4, 8};
    std::vector<int> tz_c, tseg_c = {1, 2, 4, 8};
    for (int tz : {64, 32, 16})
        if (N % tz == 0) tz_c.push_back(tz);
    auto try_knob = [&](int cm_tune::*field, const std::vector<int> &vals) -> cm_status {
        for (int v : vals) {
            if (best.*field == v) continue;
            cm_tune t = best;
            t.*field = v;
            double ms = 0.0;
            const cm_status st = timed(t, 2, &ms);
            if (st == CM_ERR_UNSUPPORTED || st == CM_ERR_INVALID) {  // shape not available for this N
                ctx->err.clear();
                continue;
            }
            if (st != CM_OK) return st;
            if (ms < 0.99 * best_ms) {
                best = t;
                best_ms = ms;
            }
        }
        return CM_OK;
    };
    if ((s = try_knob(&cm_tune::xnb, xnb_c)) != CM_OK) return restore(s);
    if ((s = try_knob(&cm_tune::xe, xe_c)) != CM_OK) return restore(s);
    if (o->do_tweb) {
        if ((s = try_knob(&cm_tune::tz, tz_c)) != CM_OK) return restore(s);
        std::vector<int> segs;
        for (int g : tseg_c)
            if (g <= N / 4) segs.push_back(g);
        if ((s = try_knob(&cm_tune::tseg, segs)) != CM_OK) return restore(s);
    }
    if ((s = try_knob(&cm_tune::dep_wg, dep_c)) != CM_OK) return restore(s);
    store_tune(ctx->device, N, &best);
    if (best_out) *best_out = best;
    if (ms_before) *ms_before = t0;
    if (ms_after) *ms_after = best_ms;
    ctx->graphs = graphs;
    ctx->profiling = prof;
    cm_profile_reset(ctx);
    return CM_OK;
}

}  // extern "C"
