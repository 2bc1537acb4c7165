// Internal declarations of libcosmo_mesh (gfx950).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/cosmo_mesh.h"
#include "memspace.hpp"

namespace cm {

// Work that rides at the end of a spectrum pass in one workgroup, in place of
// a launch of its own: the FP64 total of part[0, np) into out[0] and cleared
// accumulators zs[0, 4 ns), zc[0, ns) and zt[0, 5) (nullable) for the passes
// after it.  part null: nothing.
struct SumTail {
    const double *part;
    int np;
    double *out;
    double *zs;
    unsigned long long *zc;
    int ns;
    unsigned long long *zt;
};

// SumTail for the calling workgroup (every thread calls it).
__device__ inline void sum_tail(const SumTail &t) {
    __shared__ double red[16];
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < 4 * t.ns; i += nt) t.zs[i] = 0.0;
    for (int i = tid; i < t.ns; i += nt) t.zc[i] = 0ull;
    if (t.zt && tid < 5) t.zt[tid] = 0ull;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = tid;
    for (; i + 3 * nt < t.np; i += 4 * nt) {
        s0 += t.part[i];
        s1 += t.part[i + nt];
        s2 += t.part[i + 2 * nt];
        s3 += t.part[i + 3 * nt];
    }
    for (; i < t.np; i += nt) s0 += t.part[i];
    double sv = (s0 + s1) + (s2 + s3);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sv += __shfl_xor(sv, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = sv;
    __syncthreads();
    if (tid == 0) {
        double tot = 0.0;
        for (int w = 0; w < (nt + 63) / 64; ++w) tot += red[w];
        t.out[0] = tot;
    }
}

// Per-kernel event timing (cm_profile_enable / cm_kernel_time).
struct KernelTimes {
    double sum_ms = 0.0;
    int64_t count = 0;
};
struct PendingTiming {
    std::string name;
    hipEvent_t start, stop;
};

struct Buffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// A captured fused measurement (see measure.hip).
struct GraphEntry {
    std::string key;
    hipGraphExec_t exec = nullptr;
    uint64_t gen = 0;
    bool bad = false;
};

}  // namespace cm

struct cm_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t upload_stream = nullptr;         // non-blocking: host-to-device copies of tables (host_upload)
    hipStream_t stream = nullptr;
    size_t budget = 0;
    size_t resident = 0;  // bytes held in `scratch` and tables
    std::string err;
    bool profiling = false;
    std::map<std::string, cm::KernelTimes> ktimes;
    std::vector<cm::PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, cm::Buffer> scratch;   // named resident work buffers
    std::map<uint64_t, void *> tables;           // twiddle / sinc tables
    std::map<std::string, cm::Buffer> ctables;   // host-computed tables by key (MAS windows, Gaussian, phases)
    void *pinned = nullptr;                      // host staging of result read-backs (hipHostMalloc)
    size_t pinned_bytes = 0;
    uint64_t scratch_gen = 0;                    // bumped whenever a work buffer is (re)allocated
    bool graphs = true;                          // hipGraph replay of repeated fused measurements
    std::vector<cm::GraphEntry> graph_cache;
    int64_t graph_captures = 0, graph_launches = 0;
    uint64_t host_syncs = 0;                     // host synchronisations inside an enqueued sequence
    int cu_count = 256;
    size_t lds_per_block = 65536;
    // launch-shape tuning (0 = built-in choice); set from CM_TUNE at creation,
    // e.g. CM_TUNE="xnb=8,zrows=16,ty=16,tz=64,tseg=4,colnb=16"
    struct Tune {
        int xnb = 0;    // kz columns per workgroup in the fused x pass
        int xe = 0;     // elements per thread in the fused x pass (8: twice the threads)
        int xtw = 0;    // fused x pass twiddles: 0 in LDS when the occupancy allows it, 2 never
        int colnb = 0;  // kz columns per workgroup in the y/x column passes
        int zrows = 0;  // rows per workgroup in the z passes
        int ty = 0, tz = 0, tseg = 0;  // T-web tile and x segments
        int mas_tile = 0;              // cells per tile edge in the deposit
        int dep_wg = 0;                // deposit workgroups per CU
        int mas_hist = 0;              // tile counters per binning workgroup (0: its LDS capacity)
        int mas_groups = 0;            // at least this many x-slab groups in the tiled deposit (0: as needed)
    } tune;
};

namespace cm {

cm_status fail(cm_ctx *ctx, cm_status s, const char *fmt, ...);
cm_status hip_fail(cm_ctx *ctx, hipError_t e, const char *what);

#define CM_HIP(ctx, call)                                              \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return cm::hip_fail((ctx), e_, #call);   \
    } while (0)

// Named persistent device buffer of at least `bytes` (grown on demand).
cm_status scratch(cm_ctx *ctx, const char *name, size_t bytes, void **out);
// Frees every work buffer of the context (after its stream drains) and drops
// its captured graphs; returns the bytes freed.  Tables stay.
size_t scratch_release(cm_ctx *ctx);

// Device copy of a table computed on the host, uploaded on the first request of
// `key` and kept for the context's lifetime (no host work or copy afterwards).
const void *find_table(cm_ctx *ctx, const std::string &key);
const void *upload_table(cm_ctx *ctx, const std::string &key, const void *host, size_t bytes, cm_status *st);
// Synchronous host-to-device copy on the context's non-blocking upload stream:
// unlike hipMemcpy on the legacy stream it takes no implicit dependency on
// other blocking streams, so it is legal while another thread's stream is
// being captured into a graph (server slots).
hipError_t host_upload(cm_ctx *ctx, void *dst, const void *src, size_t bytes);
template <typename T, typename F>
const T *cached_table(cm_ctx *ctx, const std::string &key, F make, cm_status *st) {
    if (const void *p = find_table(ctx, key)) {
        *st = CM_OK;
        return (const T *)p;
    }
    std::vector<T> h;
    make(h);
    return (const T *)upload_table(ctx, key, h.data(), sizeof(T) * h.size(), st);
}
std::string table_key(const char *fmt, ...);

// Pinned host buffer of at least `bytes` owned by the context (result read-backs).
cm_status pinned_host(cm_ctx *ctx, size_t bytes, void **out);

// Events around a launch when profiling is on.
void prof_begin(cm_ctx *ctx, const char *name, hipEvent_t *start);
void prof_end(cm_ctx *ctx, const char *name, hipEvent_t start);
cm_status check_launch(cm_ctx *ctx, const char *name);

template <typename K, typename... Args>
inline cm_status launch(cm_ctx *ctx, const char *name, K kernel, dim3 grid, dim3 block,
                        size_t lds, Args... args) {
    hipEvent_t ev = nullptr;
    if (ctx->profiling) prof_begin(ctx, name, &ev);
    hipLaunchKernelGGL(kernel, grid, block, lds, ctx->stream, args...);
    if (ctx->profiling) prof_end(ctx, name, ev);
    return check_launch(ctx, name);
}

// Twiddle table exp(-2 pi i t / n), t in [0, n), as T2 (float2/double2), cached.
const void *twiddles(cm_ctx *ctx, int n, cm_dtype dt, cm_status *st);

// FFT internals shared with the fused measurement path.
struct FftPlan {
    int n = 0;
    int nst = 0;
    int r[8] = {0};
};
FftPlan make_plan(int n);
bool fft_size_ok(int N);

cm_status fft_r2c_impl(cm_ctx *ctx, int N, cm_dtype dt, const void *in, void *out,
                       double subtract);
cm_status fft_c2r_impl(cm_ctx *ctx, int N, cm_dtype dt, void *in, void *out, double scale);
// c2r of a spectrum that is zero wherever a wave-vector component exceeds K
// (|kx|, |ky|, kz <= K < N/2 hold every non-zero mode): the x pass transforms
// only the lines with |ky| <= K and kz <= K, the y pass only kz <= K and reads
// only |ky| <= K, the z pass reads kz <= K only.  Modes with |ky| > K or
// kz > K of the input are never read; those with kz > K are not written.
cm_status fft_c2r_band(cm_ctx *ctx, int N, cm_dtype dt, void *in, void *out, double scale, int K);
// Strided column passes over the half spectrum, axis 1 (y) or 0 (x).  pitch:
// kz row pitch of the spectrum in complex values (0 = N/2+1, the FFTW layout;
// float32 only may use a padded pitch).
// phase (float32 only, nullable): per-frequency factors float2[N] multiplied
// into the output of the pass (kz index for the z pass, ky for the y pass).
// tail (float32 only, nullable): a reduction and clears that workgroup (0,0)
// of the pass runs after its transform (SumTail).  stream_loads (float32, y
// axis, N >= 512): read the input with the nontemporal hint (the fused
// measurement's y passes, where it speeds up the passes after them).
// ncols (float32, > 0): only the first ncols kz columns; band (float32, >= 0):
// the input is zero at |ky| > band, so the x pass transforms only the lines
// with |ky| <= band and the y pass reads only |ky| <= band.
struct SumTail;
cm_status fft_cols(cm_ctx *ctx, int N, cm_dtype dt, void *data, int axis, bool forward, int pitch = 0,
                   const void *phase = nullptr, const SumTail *tail = nullptr, bool stream_loads = false,
                   int ncols = 0, int band = -1);
cm_status fft_rows_r2c(cm_ctx *ctx, int N, cm_dtype dt, const void *in, void *out,
                       double subtract, double *part, int pitch = 0, const void *phase = nullptr);
int fft_rows_blocks(cm_ctx *ctx, int N, cm_dtype dt);
// sq (float32 only, nullable): per-workgroup sums of squares of the output,
// fft_rows_blocks() FP64 words.  nk (float32, > 0): only the modes kz < nk of
// each row are read, the others taken as zero.
cm_status fft_rows_c2r(cm_ctx *ctx, int N, cm_dtype dt, const void *in, void *out,
                       double scale, int pitch = 0, double *sq = nullptr, int nk = 0);

// exact = false (tiled path only, needs overflow_out): the record buffer is
// sized from an estimate instead of a read-back of the count (no host
// synchronisation); *overflow_out then points to a device flag that is
// non-zero when the estimate was too small and the mesh must be redone.
// overflow_at (nullable): device word used as the overflow flag instead of the
// library's own.
cm_status mesh_assign_impl(cm_ctx *ctx, const cm_mas_desc *d, const void *pos, const void *w,
                           size_t n, void *mesh, cm_dtype mdt, int accumulate, bool exact = true,
                           unsigned **overflow_out = nullptr, bool reset_overflow = true,
                           unsigned *overflow_at = nullptr, void *meshB = nullptr);
cm_status mesh_sum_impl(cm_ctx *ctx, const void *mesh, cm_dtype dt, size_t n, double *out_dev,
                        double *host_out);

// Packed deposits (several meshes, one launch per MAS scheme over the union of
// their tile work lists).  One mock of such a launch: its records, mesh and
// tile geometry.
struct DepMock {
    const uint3 *sorted;
    const float *wsorted;
    float *mesh;
    int N, ntx, tx0, pad;
};
// Tile binning of one mock (records grouped by 16^3 tile, exact sizing, one
// host synchronisation) into buffers named <tag>.*, kept for a later packed
// deposit; the mock's mesh must bin in one range (N <= 512).
struct TileBins {
    int N = 0, T = 0, ntx = 0, ntiles = 0, scheme = 0;
    bool weighted = false;
    void *sorted = nullptr, *wsorted = nullptr, *items = nullptr, *nitems = nullptr, *offsets = nullptr;
    size_t cap = 0, max_items = 0;
};
bool mas_packable(cm_ctx *ctx, const cm_mas_desc *d);
cm_status mas_bin_mock(cm_ctx *ctx, const cm_mas_desc *d, const float *pos, const float *w, size_t n,
                       const char *tag, unsigned *overflow, TileBins *b);
// Zeroes the cells of b's split tiles in mesh (before a deposit that does not
// accumulate), then deposits every mock of `bins` (all of one scheme and weight
// kind) into its mesh with one launch over the union of their work lists.
cm_status mas_deposit_packed(cm_ctx *ctx, const std::vector<TileBins> &bins, const std::vector<float *> &meshes);

// Device results of one fused measurement in one block, so that a single copy
// reads them back: shell sums[4 ns], mode counts[ns], object count, T-web
// counters[5] (4 classes, marginal cells), deposit overflow flag, ||phi||_rms
// (float bits) of the T-web potential (one 8-byte word each for the last four
// groups).
struct FusedOut {
    double *sums = nullptr;
    unsigned long long *cnts = nullptr;
    double *nobj = nullptr;
    unsigned long long *stats = nullptr;
    unsigned *overflow = nullptr;
    unsigned *phirms = nullptr;
    void *base = nullptr;
    size_t bytes = 0;
};
cm_status fused_out(cm_ctx *ctx, int ns, FusedOut *f);
// The fused tail of a measurement: mesh (float32, device) -> z/y/x passes with
// the shells (+ potential and T-web labels when o->do_tweb).  meshB (nullable):
// the interlaced mesh.  Results land in the FusedOut block of o->nshells.
// window_set: the shells visit the window mode set of the random branch
// (stored modes outside the first octant; the zero mode is the caller's).
// ext (nullable, packed measurements): results go to the block *out already
// set up by the caller (out_given), the potential to ext->phi, and the T-web
// launch is left to the caller (tweb_deferred: ext->tweb receives its job).
struct TwMock;
struct FusedExt {
    bool out_given = false;
    float *phi = nullptr;
    TwMock *tweb = nullptr;
};
cm_status fused_measure_mesh(cm_ctx *ctx, const cm_measure_opts *o, void *mesh, void *meshB, double mean_est,
                             int8_t *labels, FusedOut *out, bool window_set = false, const FusedExt *ext = nullptr);
// FusedOut views of a result block at `base` for ns shells (5 ns + 8 words).
size_t fused_out_words(int ns);
void fused_out_at(void *base, int ns, FusedOut *f);
// Copy a FusedOut block to the pinned host buffer and wait; returns its host copy.
cm_status fused_readback(cm_ctx *ctx, const FusedOut &f, int ns, FusedOut *host);

// Sum of the profiled kernel times recorded since the last reset (ms).
double profiled_total_ms(cm_ctx *ctx);

// The process-wide shapes for (device, N) laid over the context's launch
// shapes for the lifetime of the scope.
struct TuneScope {
    cm_ctx *ctx;
    decltype(cm_ctx::tune) saved;
    TuneScope(cm_ctx *c, int N);
    ~TuneScope() { ctx->tune = saved; }
};

// Whole-mock measurement with argument checks (cm_measure's body).
cm_status measure_run(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, const void *w, size_t n, double *pk0,
                      double *pk2, double *pk4, int64_t *nmodes, int8_t *labels, cm_measure_info *info);

// P(k) helpers (pk.hip)
void pk_tables(int N, int scheme, int mas_correction, std::vector<double> &tab);
// pk_tables on the device, cached per (N, scheme, correction)
const double *pk_tables_dev(cm_ctx *ctx, int N, int scheme, int mas_correction, cm_status *st);
cm_status pk_accumulate(cm_ctx *ctx, const cm_pk_opts *o, cm_dtype dt, const void *dk, double **sums,
                        unsigned long long **cnts);
// Argument check of a binning (linear: dk > 0; log: kmin > 0, dlog > 0).
cm_status check_binning(cm_ctx *ctx, int nshells, double kmin, double dk, int log_binning, double dlog,
                        int float_shell_rule, const char *who);
void pk_finalize(int ns, double normal_power, double sn_norm, const double *sum, const unsigned long long *cnt,
                 double *pk0, double *pk2, double *pk4, int64_t *nmodes);

}  // namespace cm
