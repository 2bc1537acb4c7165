// Context, device memory, error reporting, per-kernel event timing.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "cm_internal.hpp"

namespace cm {

cm_status fail(cm_ctx *ctx, cm_status s, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return s;
}

cm_status hip_fail(cm_ctx *ctx, hipError_t e, const char *what) {
    cm_status s = (e == hipErrorOutOfMemory) ? CM_ERR_NOMEM : CM_ERR_DEVICE;
    return fail(ctx, s, "%s: %s", what, hipGetErrorString(e));
}

cm_status check_launch(cm_ctx *ctx, const char *name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, CM_ERR_DEVICE, "launch %s: %s", name, hipGetErrorString(e));
    return CM_OK;
}

static hipEvent_t get_event(cm_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_begin(cm_ctx *ctx, const char *name, hipEvent_t *start) {
    *start = get_event(ctx);
    if (*start) hipEventRecord(*start, ctx->stream);
}

void prof_end(cm_ctx *ctx, const char *name, hipEvent_t start) {
    if (!start) return;
    hipEvent_t stop = get_event(ctx);
    if (!stop) return;
    hipEventRecord(stop, ctx->stream);
    ctx->pending.push_back(PendingTiming{name, start, stop});
}

static void resolve_pending(cm_ctx *ctx) {
    for (auto &p : ctx->pending) {
        hipEventSynchronize(p.stop);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            auto &kt = ctx->ktimes[p.name];
            kt.sum_ms += ms;
            kt.count += 1;
        }
        ctx->event_pool.push_back(p.start);
        ctx->event_pool.push_back(p.stop);
    }
    ctx->pending.clear();
}

double profiled_total_ms(cm_ctx *ctx) {
    resolve_pending(ctx);
    double t = 0.0;
    for (auto &kv : ctx->ktimes) t += kv.second.sum_ms;
    return t;
}

cm_status scratch(cm_ctx *ctx, const char *name, size_t bytes, void **out) {
    auto &b = ctx->scratch[name];
    if (b.bytes >= bytes && b.ptr) {
        *out = b.ptr;
        return CM_OK;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        return fail(ctx, CM_ERR_DEVICE, "scratch '%s' would be allocated during graph capture", name);
    if (b.ptr) {
        ++ctx->scratch_gen;  // captured graphs that hold the old address are stale
        hipStreamSynchronize(ctx->stream);
        hipFree(b.ptr);
        ctx->resident -= b.bytes;
        b.ptr = nullptr;
        b.bytes = 0;
    }
    if (ctx->budget && ctx->resident + bytes > ctx->budget)
        return fail(ctx, CM_ERR_BUDGET, "scratch '%s' of %zu bytes exceeds the HBM budget (%zu of %zu held)",
                    name, bytes, ctx->resident, ctx->budget);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc(scratch)");
    b.ptr = p;
    b.bytes = bytes;
    ctx->resident += bytes;
    *out = p;
    return CM_OK;
}

size_t scratch_release(cm_ctx *ctx) {
    size_t freed = 0;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (auto &g : ctx->graph_cache)
        if (g.exec) hipGraphExecDestroy(g.exec);
    ctx->graph_cache.clear();
    for (auto &kv : ctx->scratch) {
        if (kv.second.ptr) {
            hipFree(kv.second.ptr);
            freed += kv.second.bytes;
        }
    }
    ctx->scratch.clear();
    ctx->resident -= freed;
    ++ctx->scratch_gen;
    return freed;
}

const void *find_table(cm_ctx *ctx, const std::string &key) {
    auto it = ctx->ctables.find(key);
    return it == ctx->ctables.end() ? nullptr : it->second.ptr;
}

hipError_t host_upload(cm_ctx *ctx, void *dst, const void *src, size_t bytes) {
    hipError_t e;
    if (!ctx->upload_stream && (e = hipStreamCreateWithFlags(&ctx->upload_stream, hipStreamNonBlocking)) != hipSuccess)
        return e;
    if ((e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->upload_stream)) != hipSuccess) return e;
    return hipStreamSynchronize(ctx->upload_stream);
}

const void *upload_table(cm_ctx *ctx, const std::string &key, const void *host, size_t bytes, cm_status *st) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *st = hip_fail(ctx, e, "hipMalloc(table)");
        return nullptr;
    }
    if (bytes && (e = host_upload(ctx, p, host, bytes)) != hipSuccess) {
        hipFree(p);
        *st = hip_fail(ctx, e, "table upload");
        return nullptr;
    }
    ctx->ctables[key] = Buffer{p, bytes};
    ctx->resident += bytes;
    *st = CM_OK;
    return p;
}

std::string table_key(const char *fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return std::string(buf);
}

cm_status pinned_host(cm_ctx *ctx, size_t bytes, void **out) {
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) {
            CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            hipHostFree(ctx->pinned);
            ctx->pinned = nullptr;
            ctx->pinned_bytes = 0;
        }
        const size_t want = bytes < 4096 ? 4096 : bytes;
        CM_HIP(ctx, hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
        ctx->pinned_bytes = want;
    }
    *out = ctx->pinned;
    return CM_OK;
}

const void *twiddles(cm_ctx *ctx, int n, cm_dtype dt, cm_status *st) {
    uint64_t key = (uint64_t(n) << 8) | uint64_t(dt);
    auto it = ctx->tables.find(key);
    if (it != ctx->tables.end()) {
        *st = CM_OK;
        return it->second;
    }
    // exp(-2 pi i t/n) computed in FP64 and rounded once
    std::vector<double> d(2 * (size_t)n);
    for (int t = 0; t < n; ++t) {
        // reduce the angle to the first octant-exact form via sin/cos of 2*pi*t/n
        double a = 2.0 * M_PI * (double)t / (double)n;
        d[2 * t] = std::cos(a);
        d[2 * t + 1] = -std::sin(a);
    }
    size_t bytes = (dt == CM_F32 ? 8 : 16) * (size_t)n;
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        *st = fail(ctx, CM_ERR_NOMEM, "twiddle table n=%d", n);
        return nullptr;
    }
    hipError_t e;
    if (dt == CM_F32) {
        std::vector<float> f(d.begin(), d.end());
        e = host_upload(ctx, p, f.data(), bytes);
    } else {
        e = host_upload(ctx, p, d.data(), bytes);
    }
    if (e != hipSuccess) {
        hipFree(p);
        *st = hip_fail(ctx, e, "twiddle upload");
        return nullptr;
    }
    ctx->tables[key] = p;
    ctx->resident += bytes;
    *st = CM_OK;
    return p;
}

}  // namespace cm

using namespace cm;

extern "C" {

const char *cm_status_string(cm_status s) {
    switch (s) {
    case CM_OK: return "ok";
    case CM_ERR_INVALID: return "invalid argument";
    case CM_ERR_UNSUPPORTED: return "unsupported";
    case CM_ERR_NOMEM: return "out of device memory";
    case CM_ERR_DEVICE: return "device error";
    case CM_ERR_BUDGET: return "over HBM budget";
    case CM_ERR_TIMEOUT: return "timeout";
    case CM_ERR_CANCELLED: return "cancelled";
    }
    return "unknown";
}

cm_status cm_ctx_create(int device, size_t hbm_budget, cm_ctx **out) {
    if (!out) return CM_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CM_ERR_DEVICE;
    if (device < 0 || device >= ndev) return CM_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return CM_ERR_DEVICE;
    cm_ctx *c = new cm_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->cu_count = prop.multiProcessorCount;
        c->lds_per_block = prop.sharedMemPerBlock;
        size_t total = prop.totalGlobalMem;
        c->budget = hbm_budget ? hbm_budget : (size_t)(0.9 * (double)total);
    } else {
        c->budget = hbm_budget;
    }
    if (const char *env = getenv("CM_TUNE")) {
        // comma-separated key=value pairs; unknown keys are ignored
        std::string t(env);
        size_t pos = 0;
        while (pos < t.size()) {
            size_t end = t.find(',', pos);
            if (end == std::string::npos) end = t.size();
            std::string kv = t.substr(pos, end - pos);
            size_t eq = kv.find('=');
            if (eq != std::string::npos) {
                std::string k = kv.substr(0, eq);
                int v = atoi(kv.c_str() + eq + 1);
                if (k == "xnb") c->tune.xnb = v;
                else if (k == "xe") c->tune.xe = v;
                else if (k == "xtw") c->tune.xtw = v;
                else if (k == "colnb") c->tune.colnb = v;
                else if (k == "zrows") c->tune.zrows = v;
                else if (k == "ty") c->tune.ty = v;
                else if (k == "tz") c->tune.tz = v;
                else if (k == "tseg") c->tune.tseg = v;
                else if (k == "mas_tile") c->tune.mas_tile = v;
                else if (k == "graph") c->graphs = v != 0;
                else if (k == "dep_wg") c->tune.dep_wg = v;
                else if (k == "mas_hist") c->tune.mas_hist = v;
            }
            pos = end + 1;
        }
    }
    if (hipStreamCreate(&c->own_stream) != hipSuccess) {
        delete c;
        return CM_ERR_DEVICE;
    }
    c->stream = c->own_stream;
    *out = c;
    return CM_OK;
}

void cm_ctx_destroy(cm_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->pending) {
        hipEventDestroy(p.start);
        hipEventDestroy(p.stop);
    }
    for (auto e : ctx->event_pool) hipEventDestroy(e);
    for (auto &kv : ctx->scratch)
        if (kv.second.ptr) hipFree(kv.second.ptr);
    for (auto &g : ctx->graph_cache)
        if (g.exec) hipGraphExecDestroy(g.exec);
    for (auto &kv : ctx->tables) hipFree(kv.second);
    for (auto &kv : ctx->ctables)
        if (kv.second.ptr) hipFree(kv.second.ptr);
    if (ctx->pinned) hipHostFree(ctx->pinned);
    if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
    if (ctx->upload_stream) hipStreamDestroy(ctx->upload_stream);
    delete ctx;
}

const char *cm_last_error(const cm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

cm_status cm_set_stream(cm_ctx *ctx, void *stream) {
    if (!ctx) return CM_ERR_INVALID;
    ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
    return CM_OK;
}

void *cm_get_stream(cm_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

cm_status cm_set_graphs(cm_ctx *ctx, int on) {
    if (!ctx) return CM_ERR_INVALID;
    ctx->graphs = on != 0;
    return CM_OK;
}

cm_status cm_graph_counts(cm_ctx *ctx, int64_t *captured, int64_t *launched) {
    if (!ctx) return CM_ERR_INVALID;
    if (captured) *captured = ctx->graph_captures;
    if (launched) *launched = ctx->graph_launches;
    return CM_OK;
}

cm_status cm_synchronize(cm_ctx *ctx) {
    if (!ctx) return CM_ERR_INVALID;
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CM_OK;
}

cm_status cm_dev_alloc(cm_ctx *ctx, size_t bytes, void **ptr) {
    if (!ctx || !ptr) return CM_ERR_INVALID;
    CM_HIP(ctx, hipMalloc(ptr, bytes ? bytes : 1));
    return CM_OK;
}

cm_status cm_dev_free(cm_ctx *ctx, void *ptr) {
    if (!ctx) return CM_ERR_INVALID;
    if (ptr) CM_HIP(ctx, hipFree(ptr));
    return CM_OK;
}

cm_status cm_memset(cm_ctx *ctx, void *dptr, int value, size_t bytes) {
    if (!ctx || (!dptr && bytes)) return CM_ERR_INVALID;
    CM_HIP(ctx, hipMemsetAsync(dptr, value, bytes, ctx->stream));
    return CM_OK;
}

cm_status cm_copy_h2d(cm_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || ((!dst || !src) && bytes)) return CM_ERR_INVALID;
    CM_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CM_OK;
}

cm_status cm_copy_d2h(cm_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || ((!dst || !src) && bytes)) return CM_ERR_INVALID;
    CM_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CM_OK;
}

cm_status cm_profile_enable(cm_ctx *ctx, int on) {
    if (!ctx) return CM_ERR_INVALID;
    ctx->profiling = on != 0;
    return CM_OK;
}

cm_status cm_kernel_time(cm_ctx *ctx, const char *kernel, double *mean_ms, int64_t *count) {
    if (!ctx || !kernel) return CM_ERR_INVALID;
    resolve_pending(ctx);
    auto it = ctx->ktimes.find(kernel);
    if (it == ctx->ktimes.end() || it->second.count == 0) {
        if (mean_ms) *mean_ms = 0.0;
        if (count) *count = 0;
        return CM_OK;
    }
    if (mean_ms) *mean_ms = it->second.sum_ms / (double)it->second.count;
    if (count) *count = it->second.count;
    return CM_OK;
}

cm_status cm_profile_reset(cm_ctx *ctx) {
    if (!ctx) return CM_ERR_INVALID;
    resolve_pending(ctx);
    ctx->ktimes.clear();
    return CM_OK;
}

}  // extern "C"
