// Fused x pass of the whole-mock measurement (gfx950).
//
// After the z and y passes, one workgroup owns nb adjacent kz columns of one ky
// row and all N kx values.  In one visit of that tile it
//   1. runs the forward x FFT (global -> registers/LDS),
//   2. bins every mode of the tile into the P(k) shells (the same enumeration,
//      shell rule, MAS deconvolution and shot-noise weight as k_pk_shells): one
//      thread per kx row walks the tile's kz values in order, so |k| grows and
//      a run of equal shells is summed in registers and added once,
//   3. multiplies by exp(-k^2R^2/2)(-1/k^2) N^3/N_gal for the T-web potential,
//   4. runs the inverse x FFT (LDS -> global).
// The spectrum is read and written once instead of four times (x pass, shell
// pass, multiply pass, inverse x pass).  Workgroups are persistent: each loops
// over tiles and keeps its shell sums in LDS until the end, so the global FP64
// atomics are one set per workgroup.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cm_internal.hpp"
#include "fft_ct.hpp"
#include "pk_common.hpp"
#include "tweb_common.hpp"

namespace cm {

// The two FFT phases of a tile live in separate non-inlined functions so that
// each gets its own register allocation (inlined together with the shell code
// they exceed the VGPR budget and spill).
template <int N, int NB, int E, bool TWL>
__device__ __attribute__((noinline)) void x_fft_fwd_to_lds(const float2 *__restrict__ data, size_t base,
                                                            size_t xstride, int valid, float2 *lds,
                                                            const float2 *__restrict__ tw) {
    auto srcG = [&](int b, int e) -> float2 {
        if (b >= valid) return float2{0.f, 0.f};
        return load_pass<N>(&data[base + (size_t)e * xstride + b]);
    };
    auto dstL = [&](int b, int e, float2 v) { lds[e * (NB + 1) + b] = v; };
    ct_fft<float, N, NB, NB + 1, true, true, E, TWL>(lds, tw, srcG, dstL);
}

template <int N, int NB, int E, bool TWL>
__device__ __attribute__((noinline)) void x_fft_inv_from_lds(float2 *__restrict__ data, size_t base, size_t xstride,
                                                              int valid, float2 *lds,
                                                              const float2 *__restrict__ tw) {
    auto srcL = [&](int b, int e) { return lds[e * (NB + 1) + b]; };
    auto dstG = [&](int b, int e, float2 v) {
        if (b < valid) store_stream(&data[base + (size_t)e * xstride + b], v);
    };
    ct_fft<float, N, NB, NB + 1, true, false, E, TWL>(lds, tw, srcL, dstG);
}

// Shell accumulator of one thread: the modes a thread visits along kz have
// non-decreasing |k|, so equal shells form one run that is added to the
// workgroup's LDS sums when the shell changes.
struct ShellAcc {
    int sh;
    float a0, a1, a2, a3;
    unsigned c;
    __device__ __forceinline__ void flush(double *s_acc, unsigned *s_cnt, int ns) {
        if (sh >= 0 && c) {
            atomicAdd(&s_acc[sh], (double)a0);
            atomicAdd(&s_acc[ns + sh], (double)a1);
            atomicAdd(&s_acc[2 * ns + sh], (double)a2);
            atomicAdd(&s_acc[3 * ns + sh], (double)a3);
            atomicAdd(&s_cnt[sh], c);
        }
        a0 = a1 = a2 = a3 = 0.f;
        c = 0;
    }
};

// TWL: the twiddle table is copied into LDS once per workgroup (a stage's
// twiddle loads then wait on LDS, not on the L2), when that leaves the number
// of resident workgroups unchanged.
template <int N, int NB, bool IL, int E, bool TWL>
__global__ void __launch_bounds__((ColShape<N, NB, E>::NT)) k_x_fused(float2 *__restrict__ data,
                                                 const float2 *__restrict__ dataB, int pitch,
                                                 const float2 *__restrict__ tw, PkK pk,
                                                 const double *__restrict__ tabs, float wc, float wa,
                                                 double *__restrict__ gsum,
                                                 unsigned long long *__restrict__ gcnt, int do_tweb, float inv_kf2,
                                                 const float *__restrict__ gauss, const double *__restrict__ nobj,
                                                 const float4 *__restrict__ ztab) {
    // ztab[iz] = {1/W^2, C1/W^2, A/W^2, Gaussian} of the kz index: uniform
    // across the workgroup inside the kz walk, so it is read with scalar loads
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int M1 = N / 2 + 1;
    const int ns = pk.nshells;
    float2 *lds = reinterpret_cast<float2 *>(smem_raw);
    float2 *s_tw = lds + (size_t)(NB + 1) * N;                                   // N (TWL)
    double *s_acc = reinterpret_cast<double *>(s_tw + (TWL ? N : 0));           // 4*ns
    unsigned *s_cnt = reinterpret_cast<unsigned *>(s_acc + 4 * ns);             // ns
    float *s_ic = reinterpret_cast<float *>(s_cnt + ns);                        // M1
    float *s_sq = s_ic + M1;                                                    // M1: C1 (aliasing sum)
    float *s_al = s_sq + M1;                                                    // M1: alternating sum (IL)
    const auto g_gauss = as_global(gauss);  // Gaussian factor per index (two reads per row)
    constexpr int NT = ColShape<N, NB, E>::NT;
    const int ncols = N / 2 + 1;
    const int tid = threadIdx.x;
    for (int i = tid; i < M1; i += NT) {
        s_ic[i] = (float)tabs[i];
        s_sq[i] = (float)tabs[M1 + i];
        if constexpr (IL) s_al[i] = (float)tabs[2 * M1 + i];
    }
    if constexpr (TWL)
        for (int i = tid; i < N; i += NT) s_tw[i] = tw[i];
    const float2 *twp = TWL ? s_tw : tw;
    for (int i = tid; i < 4 * ns; i += NT) s_acc[i] = 0.0;
    for (int i = tid; i < ns; i += NT) s_cnt[i] = 0u;
    const double n3 = (double)N * N * N;
    // -scale/k^2 = -(N^3/N_gal) / (kf^2 q)
    const float fscale = (do_tweb && nobj[0] > 0.0) ? (float)(-(n3 / nobj[0]) * (double)inv_kf2) : 0.0f;
    const size_t xstride = (size_t)N * pitch;
    const int nkzb = (ncols + NB - 1) / NB;
    const int ntiles = nkzb * N;
    const int h = N / 2;
    __syncthreads();
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int kzb = tile % nkzb, iy = tile / nkzb;
        const int kz0 = kzb * NB;
        const size_t base = (size_t)iy * pitch + kz0;
        const int valid = min(NB, ncols - kz0);
        x_fft_fwd_to_lds<N, NB, E, TWL>(data, base, xstride, valid, lds, twp);
        __syncthreads();
        if constexpr (IL) {
            // interlaced meshes: average with the second mesh's finished
            // spectrum (phase exp(i k.(d/2)(1,1,1)) already applied)
            for (int t = tid; t < NB * N; t += NT) {
                const int b = t % NB, e = t / NB;
                if (b < valid) {
                    float2 &a = lds[e * (NB + 1) + b];
                    const float2 v = dataB[base + (size_t)e * xstride + b];
                    a = float2{0.5f * (a.x + v.x), 0.5f * (a.y + v.y)};
                }
            }
            __syncthreads();
        }
        // shells (+ multiply): one thread per kx row, serial along the tile's kz
        // (|k| grows with kz, so shells are visited in order)
        const int ky = iy <= h ? iy : iy - N;
        const bool yl = iy < h;
        const int fy = yl ? iy : N - iy;
        const float gy = g_gauss[iy];
        for (int ix = tid; ix < N; ix += NT) {
            const bool xl = ix < h;
            const int fx = xl ? ix : N - ix;
            // mode_counted: Nyquist x/y planes never; the kz = 0 plane only for
            // (i, j) with xl && yl && (i|j) != 0, or !xl && yl && j > 0
            const bool excl = ix == h || iy == h || (pk.window && xl && yl);
            const bool zero_ok = !excl && yl && (xl ? (ix | iy) != 0 : iy > 0);
            const int qxy = fx * fx + fy * fy;
            // shot-noise weight per mode: (wc prod C1 + wa prod A) / W^2
            // (wa = 0 for one mesh; wc = wa = 1/2 for interlaced meshes)
            const float icxy = s_ic[fx] * s_ic[fy];
            const float sqxy = wc * s_sq[fx] * s_sq[fy] * icxy,
                        alxy = IL ? wa * s_al[fx] * s_al[fy] * icxy : 0.0f;
            const int kx = xl ? ix : ix - N;
            const int q2xy = kx * kx + ky * ky;
            const float gxy = fscale * g_gauss[ix] * gy;
            float2 *row = lds + ix * (NB + 1);
            ShellAcc acc{-1, 0.f, 0.f, 0.f, 0.f, 0u};
            // integer shell rule: raw index r with (r ndel + Dmin)^2 <= q < hiq;
            // log rule: qedge[r] <= q < hiq = qedge[r+1]
            int r = -1, hiq = 0;
            const bool tab = pk.qedge != nullptr;
            if (tab && !excl) {
                const int q0 = qxy + kz0 * kz0;
                r = q0 < pk.qedge[0] ? -1 : log_shell(pk.qedge, ns, q0);
                if (r < 0 && q0 >= pk.qedge[0]) r = ns;  // beyond the last shell
                hiq = r + 1 <= ns ? pk.qedge[r + 1] : 0x7fffffff;
            } else if (pk.int_rule && !excl) {
                const int q0 = qxy + kz0 * kz0;
                r = (int)floorf((__builtin_sqrtf((float)q0) - (float)pk.Dmin_i) * pk.rndel_f);
                r = max(r, -1);
#pragma unroll
                for (int it = 0; it < 2; ++it) {
                    const int lo = r * pk.ndel_i + pk.Dmin_i, hi = (r + 1) * pk.ndel_i + pk.Dmin_i;
                    if (r >= 0 && q0 < lo * lo) --r;
                    else if (q0 >= hi * hi) ++r;
                }
                if (q0 < pk.Dmin_i * pk.Dmin_i) r = -1;
                const int hi = (r + 1) * pk.ndel_i + pk.Dmin_i;
                hiq = hi * hi;
            }
#pragma unroll 4
            for (int b = 0; b < NB; ++b) {
                if (b >= valid) break;
                const int iz = kz0 + b;
                float2 v = row[b];
                const int q = qxy + iz * iz;
                int sh = -1;
                if (!excl && (iz > 0 || zero_ok)) {
                    if (tab) {
                        while (q >= hiq && r < ns) {
                            ++r;
                            hiq = r + 1 <= ns ? pk.qedge[r + 1] : 0x7fffffff;
                        }
                        sh = (r >= 0 && r < ns) ? r : -1;
                    } else if (pk.int_rule) {
                        while (q >= hiq && r < ns) {
                            ++r;
                            const int hi = (r + 1) * pk.ndel_i + pk.Dmin_i;
                            hiq = hi * hi;
                        }
                        sh = (r >= 0 && r < ns) ? r : -1;
                    } else {
                        sh = shell_of(pk, fx, fy, iz);
                    }
                }
                if (sh != acc.sh) {
                    acc.flush(s_acc, s_cnt, ns);
                    acc.sh = sh;
                }
                const int q2 = q2xy + iz * iz;
                const float rq = q2 > 0 ? __builtin_amdgcn_rcpf((float)q2) : 0.0f;
                const float4 zt = ztab[iz];
                if (sh >= 0) {
                    const float pc = (v.x * v.x + v.y * v.y) * icxy * zt.x;
                    const float mu2 = (float)(iz * iz) * rq;
                    acc.a0 += pc;
                    acc.a1 += sqxy * zt.y + alxy * zt.z;
                    acc.a2 += (1.5f * mu2 - 0.5f) * pc;
                    acc.a3 += ((4.375f * mu2 - 3.75f) * mu2 + 0.375f) * pc;
                    acc.c += 1u;
                }
                if (do_tweb) {
                    const float f = gxy * zt.w * rq;
                    v.x *= f;
                    v.y *= f;
                    row[b] = v;
                }
            }
            acc.flush(s_acc, s_cnt, ns);
        }
        __syncthreads();
        if (do_tweb) x_fft_inv_from_lds<N, NB, E, TWL>(data, base, xstride, valid, lds, twp);
        __syncthreads();
    }
    for (int i = tid; i < ns; i += NT) {
        if (s_cnt[i]) {
            unsafeAtomicAdd(&gsum[i], s_acc[i]);
            unsafeAtomicAdd(&gsum[ns + i], s_acc[ns + i]);
            unsafeAtomicAdd(&gsum[2 * ns + i], s_acc[2 * ns + i]);
            unsafeAtomicAdd(&gsum[3 * ns + i], s_acc[3 * ns + i]);
            atomicAdd(&gcnt[i], (unsigned long long)s_cnt[i]);
        }
    }
}

PkK make_pkk(cm_ctx *ctx, const cm_pk_opts *o, cm_status *st);
cm_status tweb_fused_launch(cm_ctx *ctx, int N, double L, const float *phi, double lth, TolK tol, int8_t *labels,
                            float *lam, unsigned long long *stats, bool stats_zeroed);
TolK tweb_tol(int N, double L, double margin_tol, bool fp64, const unsigned *norm);
cm_status phi_rms_final(cm_ctx *ctx, const double *part, int np, size_t n, unsigned *out);

template <int N, int NB, int E>
static cm_status x_fused_run(cm_ctx *ctx, float2 *dk, const float2 *dkB, int pitch,
                             const float2 *tw, const PkK &pk, const double *dtab, double *dsum,
                             unsigned long long *dcnt, int do_tweb, float inv_kf2, const float *dgauss,
                             const double *dnobj, int ns, const float4 *ztab) {
    constexpr int M1 = N / 2 + 1, NT = ColShape<N, NB, E>::NT;
    size_t lds = sizeof(float2) * (size_t)(NB + 1) * N + sizeof(double) * 4 * (size_t)ns +
                 sizeof(unsigned) * (size_t)ns + sizeof(float) * ((dkB ? 3 : 2) * (size_t)M1);
    if (lds > ctx->lds_per_block)
        return fail(ctx, CM_ERR_UNSUPPORTED, "fused x pass needs %zu bytes of LDS (N=%d, %d shells)", lds, N, ns);
    auto kern = dkB ? k_x_fused<N, NB, true, E, false> : k_x_fused<N, NB, false, E, false>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    // twiddles in LDS when the workgroups per CU stay the same
    {
        const size_t lds_tw = lds + sizeof(float2) * (size_t)N;
        auto kern_tw = dkB ? k_x_fused<N, NB, true, E, true> : k_x_fused<N, NB, false, E, true>;
        int per_cu_tw = 0;
        if (ctx->tune.xtw != 2 && lds_tw <= ctx->lds_per_block &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_tw, kern_tw, NT, lds_tw) == hipSuccess &&
            per_cu_tw >= per_cu) {
            kern = kern_tw;
            lds = lds_tw;
        }
    }
    const int ntiles = ((M1 + NB - 1) / NB) * N;
    const unsigned grid = (unsigned)std::min(ntiles, std::max(1, ctx->cu_count) * per_cu);
    const float wc = dkB ? 0.5f : 1.0f, wa = dkB ? 0.5f : 0.0f;
    return launch(ctx, "fused_x", kern, dim3(grid), dim3(NT), lds, dk, dkB, pitch, tw, pk, dtab, wc, wa, dsum, dcnt,
                  do_tweb, inv_kf2, dgauss, dnobj, ztab);
}

template <int N>
static cm_status x_fused_launch(cm_ctx *ctx, int nb, int xe, float2 *dk, const float2 *dkB, int pitch,
                                const float2 *tw, const PkK &pk, const double *dtab, double *dsum,
                                unsigned long long *dcnt, int do_tweb, float inv_kf2, const float *dgauss,
                                const double *dnobj, int ns, const float4 *ztab) {
    constexpr int E0 = ct_E(N);
    // E = 8 (twice the threads) where the plan allows it: radix 8/4/2 only
    if constexpr (E0 == 16) {
        if (xe == 8) {
            if constexpr (16 * N / 8 <= 1024) {
                if (nb == 16)
                    return x_fused_run<N, 16, 8>(ctx, dk, dkB, pitch, tw, pk, dtab, dsum, dcnt, do_tweb, inv_kf2,
                                                 dgauss, dnobj, ns, ztab);
            }
            if constexpr (8 * N / 8 <= 1024)
                return x_fused_run<N, 8, 8>(ctx, dk, dkB, pitch, tw, pk, dtab, dsum, dcnt, do_tweb, inv_kf2, dgauss,
                                            dnobj, ns, ztab);
        }
    }
    if constexpr (ColShape<N, 8, E0>::NT * 2 <= 512) {
        if (nb == 16)
            return x_fused_run<N, 16, E0>(ctx, dk, dkB, pitch, tw, pk, dtab, dsum, dcnt, do_tweb, inv_kf2, dgauss,
                                          dnobj, ns, ztab);
    }
    return x_fused_run<N, 8, E0>(ctx, dk, dkB, pitch, tw, pk, dtab, dsum, dcnt, do_tweb, inv_kf2, dgauss, dnobj,
                                 ns, ztab);
}

size_t fused_out_words(int ns) { return 5 * (size_t)ns + 8; }

void fused_out_at(void *b, int ns, FusedOut *f) {
    auto *w = (unsigned long long *)b;
    f->base = b;
    f->bytes = 8 * fused_out_words(ns);
    f->sums = (double *)w;
    f->cnts = w + 4 * (size_t)ns;
    f->nobj = (double *)(w + 5 * (size_t)ns);
    f->stats = w + 5 * (size_t)ns + 1;
    f->overflow = (unsigned *)(w + 5 * (size_t)ns + 6);
    f->phirms = (unsigned *)(w + 5 * (size_t)ns + 7);
}

cm_status fused_out(cm_ctx *ctx, int ns, FusedOut *f) {
    const size_t words = fused_out_words(ns);
    void *b;
    cm_status s;
    if ((s = scratch(ctx, "f.out", 8 * words, &b)) != CM_OK) return s;
    fused_out_at(b, ns, f);
    return CM_OK;
}

cm_status fused_readback(cm_ctx *ctx, const FusedOut &f, int ns, FusedOut *host) {
    void *h;
    cm_status s;
    if ((s = pinned_host(ctx, f.bytes, &h)) != CM_OK) return s;
    CM_HIP(ctx, hipMemcpyAsync(h, f.base, f.bytes, hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    auto *w = (unsigned long long *)h;
    host->base = h;
    host->bytes = f.bytes;
    host->sums = (double *)w;
    host->cnts = w + 4 * (size_t)ns;
    host->nobj = (double *)(w + 5 * (size_t)ns);
    host->stats = w + 5 * (size_t)ns + 1;
    host->overflow = (unsigned *)(w + 5 * (size_t)ns + 6);
    host->phirms = (unsigned *)(w + 5 * (size_t)ns + 7);
    return CM_OK;
}

// Fused measurement: mesh (float32, device) -> shells (+ labels), results in
// the FusedOut block.  Host tables (MAS windows, interlacing phases, Gaussian
// factors) are cached on the device, so a repeated call issues kernels only.
cm_status fused_measure_mesh(cm_ctx *ctx, const cm_measure_opts *o, void *mesh, void *meshB, double mean_est,
                             int8_t *labels, FusedOut *out, bool window_set, const FusedExt *ext) {
    const int N = o->mas.N;
    const size_t n3 = (size_t)N * N * N;
    const int pitch = padded_pitch(N);  // internal spectrum: kz rows padded to 128-byte multiples
    const size_t nh = (size_t)N * N * pitch;
    const int ns = o->nshells;
    cm_status s;
    void *dk, *part;
    const int nparts = fft_rows_blocks(ctx, N, CM_F32);
    if ((s = scratch(ctx, "m.dk", sizeof(float2) * nh, &dk)) != CM_OK) return s;
    if ((s = scratch(ctx, "f.part", sizeof(double) * (size_t)nparts, &part)) != CM_OK) return s;
    if (!(ext && ext->out_given) && (s = fused_out(ctx, ns, out)) != CM_OK) return s;
    double *dsum = out->sums, *dnobj = out->nobj;
    unsigned long long *dcnt = out->cnts, *dstats = out->stats;
    const double *dtab = pk_tables_dev(ctx, N, (int)o->mas.scheme, o->mas_correction, &s);
    if (!dtab) return s;

    // z pass with the mesh sum; the y pass's first workgroup takes the FP64
    // total and clears the shell sums and T-web counters of the passes after it
    if ((s = fft_rows_r2c(ctx, N, CM_F32, mesh, dk, mean_est, (double *)part, pitch)) != CM_OK) return s;
    const SumTail tail{(const double *)part, nparts, (double *)dnobj, dsum, dcnt, ns, dstats};
    if ((s = fft_cols(ctx, N, CM_F32, dk, 1, true, pitch, nullptr, &tail, true)) != CM_OK) return s;
    // interlaced mesh (particles shifted by d/2 on every axis): its full
    // spectrum, multiplied by exp(+i k.(d/2)(1,1,1)) = prod_a exp(i pi k_a/N)
    // one axis per pass, is averaged with the first in the fused x pass
    void *dkB = nullptr;
    if (meshB) {
        if ((s = scratch(ctx, "m.dkB", sizeof(float2) * nh, &dkB)) != CM_OK) return s;
        const float2 *phase = cached_table<float2>(ctx, table_key("phase:%d", N), [&](std::vector<float2> &ph) {
            ph.resize((size_t)N);
            for (int j = 0; j < N; ++j) {
                const double a = M_PI * (double)(j <= N / 2 ? j : j - N) / (double)N;
                ph[j] = float2{(float)std::cos(a), (float)std::sin(a)};
            }
        }, &s);
        if (!phase) return s;
        if ((s = fft_rows_r2c(ctx, N, CM_F32, meshB, dkB, mean_est, nullptr, pitch, phase)) != CM_OK) return s;
        if ((s = fft_cols(ctx, N, CM_F32, dkB, 1, true, pitch, phase)) != CM_OK) return s;
        if ((s = fft_cols(ctx, N, CM_F32, dkB, 0, true, pitch, phase)) != CM_OK) return s;
    }
    // x pass + shells + multiply + inverse x (sums cleared by fused_sum)
    cm_pk_opts po;
    po.N = N;
    po.L = o->mas.L;
    po.nshells = ns;
    po.kmin = o->kmin;
    po.dk = o->dk;
    po.scheme = o->mas.scheme;
    po.mas_correction = o->mas_correction;
    po.shot_noise = 0.0;
    po.normal_power = 1.0;
    po.float_shell_rule = o->float_shell_rule;
    po.log_binning = o->log_binning;
    po.dlog = o->dlog;
    po.window_set = 0;
    PkK pk = make_pkk(ctx, &po, &s);
    if (s != CM_OK) return s;
    pk.window = window_set ? 1 : 0;
    cm_status st;
    const void *tw = twiddles(ctx, N, CM_F32, &st);
    if (!tw) return st;
    const double kf = 2.0 * M_PI / o->mas.L;
    const double R2h = o->R_smooth > 0 ? 0.5 * o->R_smooth * o->R_smooth : 0.0;
    const int do_tweb = o->do_tweb ? 1 : 0;
    // per-axis Gaussian exp(-k_i^2 R^2/2), k_i = +i kf (i <= N/2) else -(N-i) kf;
    // the kernel multiplies the three factors (exp is separable)
    const float *dgauss = cached_table<float>(ctx, table_key("gauss:%d:%.17g:%.17g", N, kf, R2h),
                                              [&](std::vector<float> &g) {
                                                  g.resize((size_t)N);
                                                  for (int i = 0; i < N; ++i) {
                                                      const double k = kf * (i <= N / 2 ? (double)i : -(double)(N - i));
                                                      g[i] = (float)std::exp(-k * k * R2h);
                                                  }
                                              },
                                              &s);
    if (!dgauss) return s;
    const float inv_kf2 = (float)(1.0 / (kf * kf));
    // per-kz factors {1/W^2, C1/W^2, A/W^2, Gaussian} (kz = 0..N/2)
    const float4 *ztab = cached_table<float4>(
        ctx, table_key("ztab:%d:%d:%d:%.17g:%.17g", N, (int)o->mas.scheme, o->mas_correction, kf, R2h),
        [&](std::vector<float4> &z) {
            std::vector<double> t;
            pk_tables(N, (int)o->mas.scheme, o->mas_correction, t);
            const int M1 = N / 2 + 1;
            z.resize((size_t)M1);
            for (int i = 0; i < M1; ++i) {
                const double k = kf * (double)i;
                z[i] = float4{(float)t[i], (float)(t[M1 + i] * t[i]), (float)(t[2 * M1 + i] * t[i]),
                              (float)std::exp(-k * k * R2h)};
            }
        },
        &s);
    if (!ztab) return s;
    const int nb = ctx->tune.xnb == 8 ? 8 : 16;  // kz columns per tile (16 when the workgroup fits)
    switch (N) {
#define CM_CASE(n)                                                                                                     \
    case n:                                                                                                            \
        s = x_fused_launch<n>(ctx, nb, ctx->tune.xe, (float2 *)dk, (const float2 *)dkB, pitch,                     \
                              (const float2 *)tw, pk, (const double *)dtab,                                            \
                              (double *)dsum, (unsigned long long *)dcnt, do_tweb, inv_kf2, (const float *)dgauss,     \
                              (const double *)dnobj, ns, ztab);                                                        \
        break;
        CM_FFT_SIZES(CM_CASE)
#undef CM_CASE
    default: return fail(ctx, CM_ERR_UNSUPPORTED, "fused x pass: N=%d", N);
    }
    if (s != CM_OK) return s;
    if (do_tweb) {
        void *phi = ext ? ext->phi : nullptr;
        if (!phi && (s = scratch(ctx, "m.phi", sizeof(float) * n3, &phi)) != CM_OK) return s;
        if ((s = fft_cols(ctx, N, CM_F32, dk, 1, false, pitch, nullptr, nullptr, true)) != CM_OK) return s;
        // ||phi||_rms (for a tolerance derived from it): sums of squares ride in
        // the c2r pass (the z pass's partial sums are consumed by now)
        const bool rms = o->margin_tol < 0.0;
        if ((s = fft_rows_c2r(ctx, N, CM_F32, dk, phi, 1.0 / (double)n3, pitch, rms ? (double *)part : nullptr)) !=
            CM_OK)
            return s;
        if (rms && (s = phi_rms_final(ctx, (const double *)part, nparts, n3, out->phirms)) != CM_OK) return s;
        const TolK tk = tweb_tol(N, o->mas.L, o->margin_tol, false, rms ? out->phirms : nullptr);
        if (ext && ext->tweb) {  // the caller launches it with other mocks
            TwMock &t = *ext->tweb;
            t.phi = (const float *)phi;
            t.labels = labels;
            t.stats = (unsigned long long *)dstats;
            t.tk = tk;
            t.lth = o->lambda_th;
            t.N = N;
            return CM_OK;
        }
        if ((s = tweb_fused_launch(ctx, N, o->mas.L, (const float *)phi, o->lambda_th, tk, labels, nullptr,
                                   (unsigned long long *)dstats, true)) != CM_OK)
            return s;
    }
    return CM_OK;
}

}  // namespace cm
