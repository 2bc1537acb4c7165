// Whole-mock measurement: particles -> mesh -> r2c -> P(k) shells (+ T-web).
//
// plain (fused=0): the step-by-step GPU path -- global-atomic deposit, FP64
//   mesh sum read back to the host, r2c with the mean removed, shell kernel,
//   then the T-web as separate Poisson / gradient / Hessian+eigen kernels.
// fused (fused=1): tiled LDS deposit; z pass carrying the FP64 mesh sum; y
//   pass; one x pass doing forward FFT + shells + T-web multiplier + inverse
//   FFT; inverse y and c2r z passes; the marching FD-Hessian/label kernel.  The
//   object count stays on the device until the sums are read at the end.
#include <cmath>
#include <cstring>
#include <vector>

#include "cm_internal.hpp"
#include "tweb_common.hpp"

namespace cm {

cm_status poisson_impl(cm_ctx *ctx, const cm_tweb_opts *o, const float *delta, float *phi, double subtract);
cm_status tweb_fused_packed(cm_ctx *ctx, std::vector<TwMock> &mocks, const std::vector<double> &L);

__global__ void k_zero_u32(unsigned *a, int n) {
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0u;
}
cm_status tweb_from_phi_impl(cm_ctx *ctx, const cm_tweb_opts *o, const float *phi, int8_t *labels, float *lam,
                             cm_tweb_stats *st, bool plain);

// scale the half spectrum by -W(k)/k^2 * N^3/N_gal (delta = rho/mean - 1 with
// mean = N_gal/N^3), reading N_gal on the device.
__global__ void __launch_bounds__(256) k_kmul_dev(float2 *dk, int N, double kf, double R2h,
                                                  const double *__restrict__ nobj) {
    const int M1 = N / 2 + 1;
    const size_t total = (size_t)N * N * M1;
    const double n3 = (double)N * N * N;
    const double scale = nobj[0] > 0.0 ? n3 / nobj[0] : 0.0;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(g % M1);
        const size_t r = g / M1;
        const int iy = (int)(r % N), ix = (int)(r / N);
        const double f = smooth_poisson_factor(ix, iy, iz, N, kf, R2h, scale);
        float2 v = dk[g];
        v.x = (float)((double)v.x * f);
        v.y = (float)((double)v.y * f);
        dk[g] = v;
    }
}

static unsigned grid_for(size_t n, int block, size_t cap) {
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// Finalisation shared by both paths: normal_power = N_gal^2/V, SN = V/N_gal
// (raw_sampling, FftwFunctions.cpp:159-174), so SN*normal_power = N_gal.
static void finalize(const cm_measure_opts *o, double nobj_h, const std::vector<double> &hs,
                     const std::vector<unsigned long long> &hc, double *pk0, double *pk2, double *pk4,
                     int64_t *nmodes, cm_measure_info *info) {
    const double L = o->mas.L, V = L * L * L;
    const double normal = nobj_h * nobj_h / V;
    const double SN = (o->sn_correction && nobj_h > 0) ? V / nobj_h : 0.0;
    if (info) {
        info->n_obj = nobj_h;
        info->shot_noise = SN;
        info->normal_power = normal;
    }
    pk_finalize(o->nshells, normal > 0 ? normal : 1.0, SN * normal, hs.data(), hc.data(), pk0, pk2, pk4, nmodes);
}

// Read back the fused results (one copy) and finalise; *redo is set when the
// tiled deposit's record estimate was too small (results discarded, the caller
// repeats with exact sizing).
TolK tweb_tol(int N, double L, double margin_tol, bool fp64, const unsigned *norm);

// The absolute tolerance a fused T-web applied, from the ||phi||_rms it read.
static double tweb_tol_applied(int N, double L, double margin_tol, unsigned norm_bits) {
    const TolK t = tweb_tol(N, L, margin_tol, false, nullptr);
    if (margin_tol >= 0.0) return t.tol;
    float pn;
    std::memcpy(&pn, &norm_bits, sizeof pn);
    return t.per_norm * (double)pn;
}

static cm_status finish_fused(cm_ctx *ctx, const cm_measure_opts *o, const FusedOut &dev, double *pk0, double *pk2,
                              double *pk4, int64_t *nmodes, cm_measure_info *info, bool *redo) {
    const int ns = o->nshells;
    FusedOut h;
    cm_status s;
    if ((s = fused_readback(ctx, dev, ns, &h)) != CM_OK) return s;
    *redo = *h.overflow != 0;
    if (*redo) return CM_OK;
    if (info) {
        for (int i = 0; i < 4; ++i) info->tweb.counts[i] = o->do_tweb ? (int64_t)h.stats[i] : 0;
        info->tweb.n_marginal = o->do_tweb ? (int64_t)h.stats[4] : 0;
        info->tweb.margin_tol = o->do_tweb ? tweb_tol_applied(o->mas.N, o->mas.L, o->margin_tol, *h.phirms) : 0.0;
    }
    const std::vector<double> hs(h.sums, h.sums + 4 * (size_t)ns);
    const std::vector<unsigned long long> hc(h.cnts, h.cnts + (size_t)ns);
    finalize(o, *h.nobj, hs, hc, pk0, pk2, pk4, nmodes, info);
    return CM_OK;
}

// Enqueue one fused measurement: deposit(s) into the mesh (and the interlaced
// mesh), then the fused tail.  exact = false sizes the record buffer from an
// estimate (no host synchronisation; overflow is flagged in the result block).
static cm_status enqueue_fused(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, const void *w, size_t n,
                               void *mesh, void *meshB, int8_t *labels, bool exact, FusedOut *dev) {
    const int N = o->mas.N;
    const double n3 = (double)N * N * N;
    cm_mas_desc md = o->mas;
    md.algo = CM_MAS_AUTO;
    FusedOut fo;
    cm_status s;
    if ((s = fused_out(ctx, o->nshells, &fo)) != CM_OK) return s;
    unsigned *ovf = nullptr;
    // with interlacing, meshB (half a cell on every axis) is deposited from the
    // same binning records
    if ((s = mesh_assign_impl(ctx, &md, pos, w, n, mesh, CM_F32, 0, exact, &ovf, true, fo.overflow, meshB)) !=
        CM_OK)
        return s;
    return fused_measure_mesh(ctx, o, mesh, meshB, (double)n / n3, labels, dev);
}

// hipGraph replay of repeated fused measurements.  A call is keyed by its
// options, pointers, particle count, stream and tuning; the first call of a key
// runs eagerly (allocating every work buffer), the second captures the
// sequence into a graph, later ones launch the graph.  The graph holds work
// buffer addresses, so it is valid only while no buffer of the context has been
// reallocated (ctx->scratch_gen).  Nothing in the sequence uses
// hipMemsetAsync: buffers are cleared by fill kernels.
static std::string graph_key(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, const void *w, size_t n,
                             const int8_t *labels) {
    const cm_mas_desc &m = o->mas;
    return table_key("%d|%d|%d|%d|%.17g|%.17g,%.17g,%.17g|%.17g|%d|%d|%.17g|%.17g|%d|%d|%d|%d|%.17g|%.17g|%.17g|%d|"
                     "%p|%p|%zu|%p|%p|%d,%d,%d,%d,%d|%d|%.17g",
                     m.N, (int)m.scheme, (int)m.layout, (int)m.pos_dtype, m.L, m.xmin[0], m.xmin[1], m.xmin[2],
                     m.interlace_shift, o->nshells, o->mas_correction, o->kmin, o->dk, o->sn_correction,
                     o->float_shell_rule, o->do_tweb, o->interlace, o->R_smooth, o->lambda_th, o->margin_tol,
                     (int)m.algo, pos, w, n, (const void *)labels, (void *)ctx->stream, ctx->tune.xnb,
                     ctx->tune.xe, ctx->tune.tz, ctx->tune.tseg, ctx->tune.dep_wg, o->log_binning, o->dlog);
}

// *first is set on the first sighting of a key: the caller's eager run then
// decides whether the sequence may be captured at all (a sequence that
// synchronises with the host inside, like the x-slab group binning of large
// meshes, is never captured).
static cm_status measure_graph(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, const void *w, size_t n,
                               void *mesh, void *meshB, int8_t *labels, double *pk0, double *pk2, double *pk4,
                               int64_t *nmodes, cm_measure_info *info, bool *done, bool *first) {
    *done = false;
    *first = false;
    const std::string key = graph_key(ctx, o, pos, w, n, labels);
    GraphEntry *e = nullptr;
    for (auto &g : ctx->graph_cache)
        if (g.key == key) e = &g;
    if (!e) {  // first sighting: the eager run warms the work buffers
        if (ctx->graph_cache.size() >= 8) {
            if (ctx->graph_cache.front().exec) hipGraphExecDestroy(ctx->graph_cache.front().exec);
            ctx->graph_cache.erase(ctx->graph_cache.begin());
        }
        ctx->graph_cache.push_back(GraphEntry{key});
        *first = true;
        return CM_OK;
    }
    if (e->bad) return CM_OK;
    if (!e->exec || e->gen != ctx->scratch_gen) {
        if (e->exec) hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
        const uint64_t gen = ctx->scratch_gen;
        if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            e->bad = true;
            return CM_OK;
        }
        FusedOut dev;
        const cm_status s = enqueue_fused(ctx, o, pos, w, n, mesh, meshB, labels, false, &dev);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        for (int k = 0; k < 4 && hipGetLastError() != hipSuccess; ++k) {
        }
        if (s != CM_OK || ec != hipSuccess || !g || gen != ctx->scratch_gen ||
            hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            if (g) hipGraphDestroy(g);
            e->exec = nullptr;
            e->bad = true;  // this key runs eagerly from now on
            ctx->err.clear();
            return CM_OK;
        }
        hipGraphDestroy(g);
        e->gen = gen;
        ++ctx->graph_captures;
    }
    CM_HIP(ctx, hipGraphLaunch(e->exec, ctx->stream));
    ++ctx->graph_launches;
    FusedOut dev;
    cm_status s;
    if ((s = fused_out(ctx, o->nshells, &dev)) != CM_OK) return s;
    bool redo = false;
    if ((s = finish_fused(ctx, o, dev, pk0, pk2, pk4, nmodes, info, &redo)) != CM_OK) return s;
    if (redo) {
        e->bad = true;  // the record estimate is too small for this input: eager exact path
        return CM_OK;
    }
    *done = true;
    return CM_OK;
}

static cm_status measure_impl(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, const void *w, size_t n,
                              double *pk0, double *pk2, double *pk4, int64_t *nmodes, int8_t *labels,
                              cm_measure_info *info) {
    const int N = o->mas.N;
    const size_t n3 = (size_t)N * N * N;
    const size_t nh = (size_t)N * N * (N / 2 + 1);
    cm_status s;
    void *mesh, *dk, *nobj;
    if ((s = scratch(ctx, "m.mesh", sizeof(float) * n3, &mesh)) != CM_OK) return s;
    if ((s = scratch(ctx, "m.dk", sizeof(float2) * nh, &dk)) != CM_OK) return s;
    if ((s = scratch(ctx, "m.nobj", sizeof(double), &nobj)) != CM_OK) return s;

    cm_mas_desc md = o->mas;
    md.algo = o->fused ? CM_MAS_AUTO : CM_MAS_PLAIN;
    if (o->fused) {
        void *meshB = nullptr;
        if (o->interlace && (s = scratch(ctx, "m.meshB", sizeof(float) * n3, &meshB)) != CM_OK) return s;
        bool first = false;
        if (ctx->graphs && !ctx->profiling) {
            bool done = false;
            if ((s = measure_graph(ctx, o, pos, w, n, mesh, meshB, labels, pk0, pk2, pk4, nmodes, info, &done,
                                   &first)) != CM_OK)
                return s;
            if (done) return CM_OK;
        }
        // first attempt without the record-count read-back; repeat exactly on
        // overflow (flag in the result block, read back with the shells)
        for (int attempt = 0; attempt < 2; ++attempt) {
            FusedOut dev;
            const uint64_t syncs0 = ctx->host_syncs;
            if ((s = enqueue_fused(ctx, o, pos, w, n, mesh, meshB, labels, attempt > 0, &dev)) != CM_OK) return s;
            if (first && ctx->host_syncs != syncs0) ctx->graph_cache.back().bad = true;
            bool redo = false;
            if ((s = finish_fused(ctx, o, dev, pk0, pk2, pk4, nmodes, info, &redo)) != CM_OK) return s;
            if (!redo) return CM_OK;
        }
        return fail(ctx, CM_ERR_DEVICE, "cm_measure: record buffer overflow after exact sizing");
    }
    if (o->interlace) return fail(ctx, CM_ERR_UNSUPPORTED, "cm_measure: interlacing needs the fused path");
    if ((s = mesh_assign_impl(ctx, &md, pos, w, n, mesh, CM_F32, 0)) != CM_OK) return s;
    // N_gal = FP64 sum of the mesh (get_nobjects without its float sum)
    double nobj_h = 0.0;
    if ((s = mesh_sum_impl(ctx, mesh, CM_F32, n3, (double *)nobj, &nobj_h)) != CM_OK) return s;
    const double mean_est = nobj_h / (double)n3;  // get_fluctuation (FftwFunctions.cpp:293-294)
    if ((s = fft_r2c_impl(ctx, N, CM_F32, mesh, dk, mean_est)) != CM_OK) return s;

    cm_pk_opts po;
    po.N = N;
    po.L = o->mas.L;
    po.nshells = o->nshells;
    po.kmin = o->kmin;
    po.dk = o->dk;
    po.scheme = o->mas.scheme;
    po.mas_correction = o->mas_correction;
    po.shot_noise = 0.0;  // applied at finalisation from the split sums
    po.normal_power = 1.0;
    po.float_shell_rule = o->float_shell_rule;
    po.log_binning = o->log_binning;
    po.dlog = o->dlog;
    po.window_set = 0;
    double *sums;
    unsigned long long *cnts;
    if ((s = pk_accumulate(ctx, &po, CM_F32, dk, &sums, &cnts)) != CM_OK) return s;
    const int ns = o->nshells;
    std::vector<double> hs(4 * (size_t)ns);
    std::vector<unsigned long long> hc((size_t)ns);
    CM_HIP(ctx, hipMemcpyAsync(hs.data(), sums, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(ctx, hipMemcpyAsync(hc.data(), cnts, sizeof(unsigned long long) * hc.size(), hipMemcpyDeviceToHost,
                               ctx->stream));

    if (o->do_tweb) {
        if (!labels) return fail(ctx, CM_ERR_INVALID, "cm_measure: do_tweb needs a labels buffer");
        void *phi;
        if ((s = scratch(ctx, "m.phi", sizeof(float) * n3, &phi)) != CM_OK) return s;
        const double kf = 2.0 * M_PI / o->mas.L;
        const double R2h = o->R_smooth > 0 ? 0.5 * o->R_smooth * o->R_smooth : 0.0;
        if ((s = launch(ctx, "tweb_kmul", k_kmul_dev, dim3(grid_for(nh, 256, 16384)), dim3(256), 0, (float2 *)dk, N,
                        kf, R2h, (const double *)nobj)) != CM_OK)
            return s;
        if ((s = fft_c2r_impl(ctx, N, CM_F32, dk, phi, 1.0 / (double)n3)) != CM_OK) return s;
        cm_tweb_opts to;
        to.N = N;
        to.L = o->mas.L;
        to.R_smooth = o->R_smooth;
        to.lambda_th = o->lambda_th;
        to.margin_tol = o->margin_tol;
        to.delta_scale = 1.0;
        to.hessian = CM_HESSIAN_FD;
        cm_tweb_stats ts;
        if ((s = tweb_from_phi_impl(ctx, &to, (const float *)phi, labels, nullptr, info ? &ts : nullptr,
                                    !o->fused)) != CM_OK)
            return s;
        if (info) info->tweb = ts;
    }
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));

    finalize(o, nobj_h, hs, hc, pk0, pk2, pk4, nmodes, info);
    return CM_OK;
}

}  // namespace cm

using namespace cm;

namespace cm {
cm_status measure_run(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, const void *w, size_t n, double *pk0,
                      double *pk2, double *pk4, int64_t *nmodes, int8_t *labels, cm_measure_info *info) {
    if (!ctx) return CM_ERR_INVALID;
    if (!o || (!pos && n)) return fail(ctx, CM_ERR_INVALID, "cm_measure: null pointer");
    if (!fft_size_ok(o->mas.N)) return fail(ctx, CM_ERR_INVALID, "cm_measure: N=%d not supported", o->mas.N);
    if (o->mas.pos_dtype != CM_F32) return fail(ctx, CM_ERR_UNSUPPORTED, "cm_measure: float32 positions only");
    cm_status sb = check_binning(ctx, o->nshells, o->kmin, o->dk, o->log_binning, o->dlog, o->float_shell_rule,
                                 "cm_measure");
    if (sb != CM_OK) return sb;
    if (o->do_tweb && !labels) return fail(ctx, CM_ERR_INVALID, "cm_measure: do_tweb needs a labels buffer");
    TuneScope ts(ctx, o->mas.N);
    return measure_impl(ctx, o, pos, w, n, pk0, pk2, pk4, nmodes, labels, info);
}
// ------------------------------------------------------------ packed mocks
// Argument checks of one mock (cm_measure's).
static cm_status check_mock(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, size_t n, const int8_t *labels) {
    if (!o || (!pos && n)) return fail(ctx, CM_ERR_INVALID, "cm_measure: null pointer");
    if (!fft_size_ok(o->mas.N)) return fail(ctx, CM_ERR_INVALID, "cm_measure: N=%d not supported", o->mas.N);
    if (o->mas.pos_dtype != CM_F32) return fail(ctx, CM_ERR_UNSUPPORTED, "cm_measure: float32 positions only");
    cm_status sb = check_binning(ctx, o->nshells, o->kmin, o->dk, o->log_binning, o->dlog, o->float_shell_rule,
                                 "cm_measure");
    if (sb != CM_OK) return sb;
    if (o->do_tweb && !labels) return fail(ctx, CM_ERR_INVALID, "cm_measure: do_tweb needs a labels buffer");
    return CM_OK;
}

// The packed path takes fused, non-interlaced mocks whose mesh bins in one
// range (16^3 tiles, N <= 512) and, with the T-web, N a multiple of 64.
static bool packable(cm_ctx *ctx, const cm_measure_opts *o) {
    return o->fused && !o->interlace && mas_packable(ctx, &o->mas) && (!o->do_tweb || o->mas.N % 64 == 0);
}

// HBM a mock holds from its binning to its results: mesh, potential, records
// (ghost copies of 16^3 tiles with headroom), work list and result block.
static size_t packed_bytes(const cm_packed_mock &m) {
    const cm_measure_opts &o = m.opts;
    const size_t n3 = (size_t)o.mas.N * o.mas.N * o.mas.N;
    const double R = o.mas.scheme == CM_NGP ? 0.0 : (o.mas.scheme == CM_PCS ? 2.0 : 1.0);
    const double g = 1.0 + 2.0 * R / 16.0;
    const size_t recs = (size_t)(1.25 * g * g * g * (double)m.n) + 65536;
    return 4 * n3 * (o.do_tweb ? 2 : 1) + (12 + (m.w ? 4 : 0)) * recs + 16 * (n3 / 4096 + recs / 4096 + 2) +
           8 * fused_out_words(o.nshells) + 4 * (n3 / 4096 + 1);
}

// One group of packable mocks: bin each, deposit all (one launch per scheme
// and weight kind), spectra and shells per mock, one T-web launch, one
// read-back of every result block.
static cm_status packed_group(cm_ctx *ctx, const cm_packed_mock *m, const std::vector<int> &idx, cm_status *st) {
    const int g = (int)idx.size();
    cm_status s;
    std::vector<size_t> off((size_t)g + 1, 0);
    for (int k = 0; k < g; ++k) off[(size_t)k + 1] = off[(size_t)k] + fused_out_words(m[idx[(size_t)k]].opts.nshells);
    void *outs;
    if ((s = scratch(ctx, "pk.out", 8 * off[(size_t)g], &outs)) != CM_OK) return s;
    std::vector<FusedOut> fo((size_t)g);
    std::vector<TileBins> bins((size_t)g);
    std::vector<float *> mesh((size_t)g), phi((size_t)g, nullptr);
    char tag[32];
    for (int k = 0; k < g; ++k) {
        const cm_packed_mock &q = m[idx[(size_t)k]];
        const size_t n3 = (size_t)q.opts.mas.N * q.opts.mas.N * q.opts.mas.N;
        fused_out_at((unsigned long long *)outs + off[(size_t)k], q.opts.nshells, &fo[(size_t)k]);
        void *p;
        std::snprintf(tag, sizeof tag, "pk.mesh.%d", k);
        if ((s = scratch(ctx, tag, 4 * n3, &p)) != CM_OK) return s;
        mesh[(size_t)k] = (float *)p;
        if (q.opts.do_tweb) {
            std::snprintf(tag, sizeof tag, "pk.phi.%d", k);
            if ((s = scratch(ctx, tag, 4 * n3, &p)) != CM_OK) return s;
            phi[(size_t)k] = (float *)p;
        }
    }
    // binning of every mock (exact record sizing)
    for (int k = 0; k < g; ++k) {
        const cm_packed_mock &q = m[idx[(size_t)k]];
        std::snprintf(tag, sizeof tag, "pk.%d", k);
        if ((s = launch(ctx, "fill", k_zero_u32, dim3(1), dim3(64), 0, fo[(size_t)k].overflow, 1)) != CM_OK)
            return s;
        if ((s = mas_bin_mock(ctx, &q.opts.mas, (const float *)q.pos, (const float *)q.w, q.n, tag,
                              fo[(size_t)k].overflow, &bins[(size_t)k])) != CM_OK)
            return s;
    }
    // one deposit launch per scheme and weight kind over the union work list
    for (int sch = 0; sch < 4; ++sch)
        for (int wt = 0; wt < 2; ++wt) {
            std::vector<TileBins> bs;
            std::vector<float *> ms;
            for (int k = 0; k < g; ++k)
                if (bins[(size_t)k].scheme == sch && (int)bins[(size_t)k].weighted == wt) {
                    bs.push_back(bins[(size_t)k]);
                    ms.push_back(mesh[(size_t)k]);
                }
            if (!bs.empty() && (s = mas_deposit_packed(ctx, bs, ms)) != CM_OK) return s;
        }
    // spectra, shells and potentials per mock; the T-web of all in one launch
    std::vector<TwMock> tw;
    std::vector<double> twL;
    std::vector<TwMock> twk((size_t)g);
    for (int k = 0; k < g; ++k) {
        const cm_packed_mock &q = m[idx[(size_t)k]];
        const int N = q.opts.mas.N;
        const double n3 = (double)N * N * N;
        TuneScope ts(ctx, N);
        FusedExt ext;
        ext.out_given = true;
        ext.phi = phi[(size_t)k];
        ext.tweb = q.opts.do_tweb ? &twk[(size_t)k] : nullptr;
        if ((s = fused_measure_mesh(ctx, &q.opts, mesh[(size_t)k], nullptr, (double)q.n / n3, q.labels,
                                    &fo[(size_t)k], false, &ext)) != CM_OK)
            return s;
        if (q.opts.do_tweb) {
            tw.push_back(twk[(size_t)k]);
            twL.push_back(q.opts.mas.L);
        }
    }
    if (!tw.empty() && (s = tweb_fused_packed(ctx, tw, twL)) != CM_OK) return s;
    // one read-back of all result blocks
    void *h;
    if ((s = pinned_host(ctx, 8 * off[(size_t)g], &h)) != CM_OK) return s;
    CM_HIP(ctx, hipMemcpyAsync(h, outs, 8 * off[(size_t)g], hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < g; ++k) {
        const int i = idx[(size_t)k];
        const cm_packed_mock &q = m[i];
        const cm_measure_opts *o = &q.opts;
        FusedOut hb;
        fused_out_at((unsigned long long *)h + off[(size_t)k], o->nshells, &hb);
        if (*hb.overflow) {  // not expected with exact sizing: measure it alone
            st[i] = fail(ctx, CM_ERR_DEVICE, "cm_measure_packed: record buffer overflow in mock %d", i);
            continue;
        }
        if (q.info) {
            for (int c = 0; c < 4; ++c) q.info->tweb.counts[c] = o->do_tweb ? (int64_t)hb.stats[c] : 0;
            q.info->tweb.n_marginal = o->do_tweb ? (int64_t)hb.stats[4] : 0;
            q.info->tweb.margin_tol =
                o->do_tweb ? tweb_tol_applied(o->mas.N, o->mas.L, o->margin_tol, *hb.phirms) : 0.0;
        }
        const std::vector<double> hs(hb.sums, hb.sums + 4 * (size_t)o->nshells);
        const std::vector<unsigned long long> hc(hb.cnts, hb.cnts + (size_t)o->nshells);
        finalize(o, *hb.nobj, hs, hc, q.pk0, q.pk2, q.pk4, q.nmodes, q.info);
        st[i] = CM_OK;
    }
    return CM_OK;
}

cm_status measure_packed(cm_ctx *ctx, const cm_packed_mock *m, int nm, cm_status *statuses) {
    if (!ctx) return CM_ERR_INVALID;
    if (nm < 0 || (nm > 0 && !m)) return fail(ctx, CM_ERR_INVALID, "cm_measure_packed: bad mock list");
    std::vector<cm_status> st((size_t)std::max(nm, 0), CM_OK);
    std::vector<int> pk;
    for (int i = 0; i < nm; ++i) {
        const cm_packed_mock &q = m[i];
        if ((st[(size_t)i] = check_mock(ctx, &q.opts, q.pos, q.n, q.labels)) != CM_OK) continue;
        if (packable(ctx, &q.opts)) {
            pk.push_back(i);
        } else {
            st[(size_t)i] = measure_run(ctx, &q.opts, q.pos, q.w, q.n, q.pk0, q.pk2, q.pk4, q.nmodes, q.labels,
                                        q.info);
        }
    }
    // groups that hold at most half of the context's free budget at once
    const size_t avail = ctx->budget > ctx->resident ? ctx->budget - ctx->resident : 0;
    size_t k0 = 0;
    while (k0 < pk.size()) {
        size_t k1 = k0, held = 0;
        while (k1 < pk.size() && (k1 == k0 || held + packed_bytes(m[pk[k1]]) <= avail / 2))
            held += packed_bytes(m[pk[k1++]]);
        const std::vector<int> grp(pk.begin() + (long)k0, pk.begin() + (long)k1);
        if (packed_group(ctx, m, grp, st.data()) != CM_OK) {
            for (int i : grp)  // the group failed as a whole (e.g. its buffers): each mock alone
                st[(size_t)i] = measure_run(ctx, &m[i].opts, m[i].pos, m[i].w, m[i].n, m[i].pk0, m[i].pk2,
                                            m[i].pk4, m[i].nmodes, m[i].labels, m[i].info);
        }
        k0 = k1;
    }
    cm_status first = CM_OK;
    for (int i = 0; i < nm; ++i) {
        if (statuses) statuses[i] = st[(size_t)i];
        if (first == CM_OK && st[(size_t)i] != CM_OK) first = st[(size_t)i];
    }
    return first;
}

}  // namespace cm

extern "C" cm_status cm_measure_packed(cm_ctx *ctx, const cm_packed_mock *mocks, int nm, cm_status *statuses) {
    return cm::measure_packed(ctx, mocks, nm, statuses);
}

extern "C" cm_status cm_measure(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, const void *w, size_t n,
                                double *pk0, double *pk2, double *pk4, int64_t *nmodes, int8_t *labels,
                                cm_measure_info *info) {
    return cm::measure_run(ctx, o, pos, w, n, pk0, pk2, pk4, nmodes, labels, info);
}
