// P(k) shell estimator over an r2c half spectrum (gfx950).
//
// One thread per stored mode (coalesced 8/16-byte loads).  Each mode is mapped
// to the reference's octant enumeration (which stored modes are counted) and
// its |k| shell (exact integer rule or the reference's float32 rule).  Four FP64
// sums per shell are kept: |d|^2 W^-2, Sq W^-2 (the aliased shot-noise weight),
// and the ell=2,4 Legendre-weighted |d|^2 W^-2.  The shot noise is subtracted
// at finalisation, sum(|d|^2 - SN*n*Sq)W^-2 = S0 - SN*n*S1, which is the
// reference's per-mode subtraction (FftwFunctions.cpp:1176) summed.  Modes of
// one wavefront that share
// (row, shell) are combined with a segmented scan over the wavefront (shells
// are monotone along kz, so runs are contiguous); the last lane of each run adds
// the run total into FP64 LDS accumulators, and each workgroup adds its
// accumulators into the FP64 global sums once.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "cm_internal.hpp"
#include "pk_common.hpp"

namespace cm {

template <typename T>
__global__ void __launch_bounds__(256) k_pk_shells(const v2pk<T> *__restrict__ dk, PkK p,
                                                   const double *__restrict__ tabs,
                                                   double *__restrict__ gsum,
                                                   unsigned long long *__restrict__ gcnt) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int M1 = p.N / 2 + 1;
    double *s_ic = reinterpret_cast<double *>(smem_raw);
    double *s_sq = s_ic + M1;
    double *s_acc = s_sq + M1;                                        // 4*nshells
    unsigned *s_cnt = reinterpret_cast<unsigned *>(s_acc + 4 * p.nshells);  // nshells
    for (int i = threadIdx.x; i < M1; i += blockDim.x) {
        s_ic[i] = tabs[i];
        s_sq[i] = tabs[M1 + i];
    }
    for (int i = threadIdx.x; i < 4 * p.nshells; i += blockDim.x) s_acc[i] = 0.0;
    for (int i = threadIdx.x; i < p.nshells; i += blockDim.x) s_cnt[i] = 0u;
    __syncthreads();

    const size_t total = (size_t)p.N * p.N * M1;
    const int lane = threadIdx.x & 63;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    // all lanes of a wavefront iterate together (uniform trip count)
    for (size_t base = (size_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < total; base += step) {
        const size_t g = base + lane;
        int row = -1, sh = -1;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        unsigned c = 0;
        if (g < total) {
            const int iz = (int)(g % M1);
            const size_t r = g / M1;
            const int iy = (int)(r % p.N), ix = (int)(r / p.N);
            int fx, fy;
            bool counted = mode_counted(p.N, ix, iy, iz, fx, fy);
            const bool dc = (ix | iy | iz) == 0;
            if (p.window) counted = dc || (counted && !(ix < p.N / 2 && iy < p.N / 2));
            if (counted) {
                // the zero mode of a window takes shell 0 under log binning
                sh = (dc && p.qedge) ? 0 : shell_of(p, fx, fy, iz);
                if (sh >= 0) {
                    row = (int)r;
                    const v2pk<T> v = dk[g];
                    const double pw = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
                    const double ic = s_ic[fx] * s_ic[fy] * s_ic[iz];
                    const int q = fx * fx + fy * fy + iz * iz;
                    const double mu2 = q > 0 ? (double)iz * (double)iz / (double)q : 0.0;
                    const double pc = pw * ic;
                    a0 = pc;
                    a1 = s_sq[fx] * s_sq[fy] * s_sq[iz] * ic;
                    a2 = 0.5 * (3.0 * mu2 - 1.0) * pc;
                    a3 = (35.0 * mu2 * mu2 - 30.0 * mu2 + 3.0) * 0.125 * pc;
                    c = 1;
                }
            }
        }
        // segmented inclusive scan over equal (row, shell) runs
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int urow = __shfl_up(row, off, 64);
            const int ush = __shfl_up(sh, off, 64);
            const double u0 = __shfl_up(a0, off, 64);
            const double u1 = __shfl_up(a1, off, 64);
            const double u2 = __shfl_up(a2, off, 64);
            const double u3 = __shfl_up(a3, off, 64);
            const unsigned uc = __shfl_up(c, off, 64);
            if (lane >= off && urow == row && ush == sh) {
                a0 += u0;
                a1 += u1;
                a2 += u2;
                a3 += u3;
                c += uc;
            }
        }
        const int nrow = __shfl_down(row, 1, 64);
        const int nsh = __shfl_down(sh, 1, 64);
        const bool tail = (lane == 63) || nrow != row || nsh != sh;
        if (tail && sh >= 0) {
            atomicAdd(&s_acc[sh], a0);
            atomicAdd(&s_acc[p.nshells + sh], a1);
            atomicAdd(&s_acc[2 * p.nshells + sh], a2);
            atomicAdd(&s_acc[3 * p.nshells + sh], a3);
            atomicAdd(&s_cnt[sh], c);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < p.nshells; i += blockDim.x) {
        if (s_cnt[i]) {
            unsafeAtomicAdd(&gsum[i], s_acc[i]);
            unsafeAtomicAdd(&gsum[p.nshells + i], s_acc[p.nshells + i]);
            unsafeAtomicAdd(&gsum[2 * p.nshells + i], s_acc[2 * p.nshells + i]);
            unsafeAtomicAdd(&gsum[3 * p.nshells + i], s_acc[3 * p.nshells + i]);
            atomicAdd(&gcnt[i], (unsigned long long)s_cnt[i]);
        }
    }
}

__global__ void k_pk_zero(double *s, unsigned long long *c, int ns) {
    for (int i = threadIdx.x; i < 4 * ns; i += blockDim.x) s[i] = 0.0;
    for (int i = threadIdx.x; i < ns; i += blockDim.x) c[i] = 0ull;
}

// Per-axis tables: inverse squared MAS window sinc(pi i/N)^-2(p+1) and the
// aliased shot-noise factor Sq(p, sin(pi i/N)) (Miscelanious.cpp:2532-2577).
// Alternating aliasing sum A(x) = sum_n (-1)^n sinc^{2p}(x + n pi), p = scheme + 1,
// x = pi i / N: the aliased shot noise of the interlaced mesh enters as
// (prod C1 + prod A)/2.  p = 1 sums to cos x; for p >= 2 the terms fall as
// n^-2p and 400 of them leave an error below 1e-13.
static double alt_alias(int scheme, double x) {
    const int p = scheme + 1;
    if (p == 1) return std::cos(x);
    double s = 0.0;
    for (int n = -200; n <= 200; ++n) {
        const double y = x + M_PI * n;
        const double sc = std::fabs(y) < 1e-300 ? 1.0 : std::sin(y) / y;
        const double t = std::pow(sc, 2 * p);
        s += (n & 1) ? -t : t;
    }
    return s;
}

void pk_tables(int N, int scheme, int mas_correction, std::vector<double> &tab) {
    const int M1 = N / 2 + 1;
    tab.assign(3 * (size_t)M1, 0.0);
    for (int i = 0; i < M1; ++i) {
        double a = M_PI * (double)i / (double)N;
        double sinc = i == 0 ? 1.0 : std::sin(a) / a;
        double ic = 1.0;
        if (mas_correction) {
            double cp = std::pow(sinc, scheme + 1);
            ic = 1.0 / (cp * cp);
        }
        double s = std::sin(a), s2 = s * s;
        double sq = 1.0;
        switch (scheme) {
        case 0: sq = 1.0; break;
        case 1: sq = 1.0 - (2.0 / 3.0) * s2; break;
        case 2: sq = 1.0 - s2 + (2.0 / 15.0) * s2 * s2; break;
        case 3: sq = 1.0 - (4.0 / 3.0) * s2 + (2.0 / 5.0) * s2 * s2 - (4.0 / 315.0) * s2 * s2 * s2; break;
        }
        tab[i] = ic;
        tab[M1 + i] = sq;
        tab[2 * M1 + i] = alt_alias(scheme, a);
    }
}

const double *pk_tables_dev(cm_ctx *ctx, int N, int scheme, int mas_correction, cm_status *st) {
    return cached_table<double>(ctx, table_key("pk:%d:%d:%d", N, scheme, mas_correction),
                                [&](std::vector<double> &t) { pk_tables(N, scheme, mas_correction, t); }, st);
}

static bool near_int(double v, long &out) {
    double r = std::floor(v + 0.5);
    if (std::fabs(v - r) < 1e-9) {
        out = (long)r;
        return true;
    }
    return false;
}

// Integer thresholds of the logarithmic shells (FftwFunctions.cpp:1103-1108,
// Params.cpp:4106): the mode q = |k/k_f|^2 lies in shell
// s = floor(log10(sqrt(q)/Dmin)/dlog) + 1 exactly when
// (Dmin 10^((s-1) dlog))^2 <= q < (Dmin 10^(s dlog))^2, so shell s starts at
// Q[s] = the smallest integer at or above x_s = (Dmin 10^((s-1) dlog))^2.
// x_s is evaluated in long double; a value within 1e-9 relative of an integer
// (e.g. (N/2)^(2m/n) = 4^m) is that integer, so a mode on an edge goes to the
// shell above as the exact rule says.  Thresholds are capped at qmax + 1.
void log_shell_edges(double Dmin, double dlog, int ns, long qmax, int *Q) {
    for (int s = 0; s <= ns; ++s) {
        const long double e = (long double)Dmin * powl(10.0L, (long double)(s - 1) * (long double)dlog);
        const long double x = e * e;
        long double r = nearbyintl(x);
        long double v = fabsl(x - r) <= 1e-9L * (x > 1.0L ? x : 1.0L) ? r : ceill(x);
        if (v > (long double)(qmax + 1)) v = (long double)(qmax + 1);
        Q[(size_t)s] = (int)v;
    }
}

PkK make_pkk(cm_ctx *ctx, const cm_pk_opts *o, cm_status *st) {
    *st = CM_OK;
    PkK p;
    p.qedge = nullptr;
    p.window = o->window_set ? 1 : 0;
    p.N = o->N;
    p.nshells = o->nshells;
    const double kf = 2.0 * M_PI / o->L;
    p.Dmin = o->kmin / kf;
    p.ndel = o->dk / kf;
    long di = 0, ni = 0;
    // the integer rule squares shell edges in 32-bit arithmetic: keep them below 40000
    p.int_rule = (near_int(p.Dmin, di) && near_int(p.ndel, ni) && ni > 0 && di >= 0 &&
                  di + (long)(o->nshells + 1) * ni < 40000)
                     ? 1
                     : 0;
    p.Dmin_i = (int)di;
    p.ndel_i = (int)ni;
    p.float_rule = o->float_shell_rule ? 1 : 0;
    p.dkx_f = (float)kf;
    p.kmin_f = (float)o->kmin;
    p.rdk_f = (float)(1.0 / (double)(float)o->dk);
    p.rndel_f = p.int_rule ? (float)(1.0 / (double)p.ndel_i) : 0.0f;
    if (o->log_binning) {
        const long h = o->N / 2, qmax = 3 * h * h;
        p.int_rule = 0;
        p.float_rule = 0;
        p.qedge = cached_table<int>(ctx, table_key("logq:%d:%d:%.17g:%.17g", o->N, o->nshells, p.Dmin, o->dlog),
                                    [&](std::vector<int> &q) {
                                        q.resize((size_t)o->nshells + 1);
                                        log_shell_edges(p.Dmin, o->dlog, o->nshells, qmax, q.data());
                                    },
                                    st);
    }
    return p;
}

cm_status check_binning(cm_ctx *ctx, int nshells, double kmin, double dk, int log_binning, double dlog,
                        int float_shell_rule, const char *who) {
    if (nshells < 1) return fail(ctx, CM_ERR_INVALID, "%s: nshells must be >= 1", who);
    if (log_binning) {
        if (!(kmin > 0) || !(dlog > 0) || !std::isfinite(dlog))
            return fail(ctx, CM_ERR_INVALID, "%s: log binning needs kmin > 0 and dlog > 0", who);
        if (float_shell_rule)
            return fail(ctx, CM_ERR_UNSUPPORTED, "%s: the float32 shell rule applies to linear binning", who);
        return CM_OK;
    }
    if (!(dk > 0)) return fail(ctx, CM_ERR_INVALID, "%s: binning (dk must be > 0)", who);
    return CM_OK;
}

cm_status pk_accumulate(cm_ctx *ctx, const cm_pk_opts *o, cm_dtype dt, const void *dk, double **sums,
                        unsigned long long **cnts) {
    const int ns = o->nshells;
    const int M1 = o->N / 2 + 1;
    void *dsum, *dcnt;
    cm_status s;
    const double *dtab = pk_tables_dev(ctx, o->N, (int)o->scheme, o->mas_correction, &s);
    if (!dtab) return s;
    if ((s = scratch(ctx, "pk.sum", sizeof(double) * 4 * (size_t)ns, &dsum)) != CM_OK) return s;
    if ((s = scratch(ctx, "pk.cnt", sizeof(unsigned long long) * (size_t)ns, &dcnt)) != CM_OK) return s;
    if ((s = launch(ctx, "pk_zero", k_pk_zero, dim3(1), dim3(256), 0, (double *)dsum, (unsigned long long *)dcnt,
                    ns)) != CM_OK)
        return s;
    PkK p = make_pkk(ctx, o, &s);
    if (s != CM_OK) return s;
    const size_t total = (size_t)o->N * o->N * M1;
    size_t lds = sizeof(double) * (2 * (size_t)M1 + 4 * (size_t)ns) + sizeof(unsigned) * (size_t)ns;
    if (lds > ctx->lds_per_block)
        return fail(ctx, CM_ERR_UNSUPPORTED, "pk: %d shells need %zu bytes of LDS", ns, lds);
    size_t blocks = (total + 255) / 256;
    unsigned g = (unsigned)std::min<size_t>(blocks, (size_t)ctx->cu_count * 8);
    if (dt == CM_F32)
        s = launch(ctx, "pk_shells", k_pk_shells<float>, dim3(g), dim3(256), lds, (const v2pk<float> *)dk, p,
                   dtab, (double *)dsum, (unsigned long long *)dcnt);
    else
        s = launch(ctx, "pk_shells", k_pk_shells<double>, dim3(g), dim3(256), lds, (const v2pk<double> *)dk, p,
                   dtab, (double *)dsum, (unsigned long long *)dcnt);
    if (s != CM_OK) return s;
    *sums = (double *)dsum;
    *cnts = (unsigned long long *)dcnt;
    return CM_OK;
}

// Normalisation of power_spectrum_fkp (FftwFunctions.cpp:1613-1628) from the
// four shell sums; sn_norm = shot_noise * normal_power (= N_gal in a box).
void pk_finalize(int ns, double normal_power, double sn_norm, const double *sum, const unsigned long long *cnt,
                 double *pk0, double *pk2, double *pk4, int64_t *nmodes) {
    for (int i = 0; i < ns; ++i) {
        double den = (double)cnt[i] * normal_power;
        if (pk0) pk0[i] = cnt[i] ? (sum[i] - sn_norm * sum[ns + i]) / den : 0.0;
        if (pk2) pk2[i] = cnt[i] ? 5.0 * sum[2 * ns + i] / den : 0.0;
        if (pk4) pk4[i] = cnt[i] ? 9.0 * sum[3 * ns + i] / den : 0.0;
        if (nmodes) nmodes[i] = (int64_t)cnt[i];
    }
}

}  // namespace cm

using namespace cm;

extern "C" cm_status cm_pk_shells(cm_ctx *ctx, const cm_pk_opts *o, cm_dtype dtype, const void *dk, double *pk0,
                                  double *pk2, double *pk4, int64_t *nmodes) {
    if (!ctx) return CM_ERR_INVALID;
    if (!o || !dk) return fail(ctx, CM_ERR_INVALID, "cm_pk_shells: null pointer");
    if (o->N < 2 || (o->N & 1)) return fail(ctx, CM_ERR_INVALID, "cm_pk_shells: N=%d must be even", o->N);
    if (!(o->L > 0)) return fail(ctx, CM_ERR_INVALID, "cm_pk_shells: L must be > 0");
    cm_status s = check_binning(ctx, o->nshells, o->kmin, o->dk, o->log_binning, o->dlog, o->float_shell_rule,
                                "cm_pk_shells");
    if (s != CM_OK) return s;
    if ((int)o->scheme < 0 || (int)o->scheme > 3) return fail(ctx, CM_ERR_INVALID, "cm_pk_shells: scheme");
    if (!(o->normal_power > 0) && (pk0 || pk2 || pk4))
        return fail(ctx, CM_ERR_INVALID, "cm_pk_shells: normal_power must be > 0");
    double *sums;
    unsigned long long *cnts;
    s = pk_accumulate(ctx, o, dtype, dk, &sums, &cnts);
    if (s != CM_OK) return s;
    std::vector<double> hs(4 * (size_t)o->nshells);
    std::vector<unsigned long long> hc((size_t)o->nshells);
    CM_HIP(ctx, hipMemcpyAsync(hs.data(), sums, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(ctx, hipMemcpyAsync(hc.data(), cnts, sizeof(unsigned long long) * hc.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    pk_finalize(o->nshells, o->normal_power, o->shot_noise * o->normal_power, hs.data(), hc.data(), pk0, pk2, pk4,
                nmodes);
    return CM_OK;
}

// ------------------------------------------------- shells without allocation
// cm_pk_shells_noalloc keeps everything it needs beyond its inputs and outputs
// in one file-scope device array: the per-axis tables, the shell sums and
// counts, the log-shell thresholds and the finalised results.  The array's
// device address is looked up once per device (hipGetSymbolAddress), and only
// that address is handed to kernels and copies.
constexpr int kNaMaxN = 1024;       // mesh side
constexpr int kNaMaxShells = 1024;  // shells (the shell kernel's LDS holds 4 FP64 sums each)
constexpr int kNaM1 = kNaMaxN / 2 + 1;
// layout in 8-byte words: tables 2*M1 | sums 4*ns | counts ns | pk0,pk2,pk4 3*ns | nmodes ns | thresholds (ns+1) ints
constexpr size_t kNaWords = 2 * (size_t)kNaM1 + 9 * (size_t)kNaMaxShells + kNaMaxShells / 2 + 2;  // 78,880 bytes

// The fixed work area of cm_pk_shells_noalloc: each device holds its own copy
// for the life of the process.  Host code uses only the address
// hipGetSymbolAddress returns for it on the current device.
static __device__ __align__(256) unsigned long long scratch[kNaWords];

namespace cm {

static std::mutex na_mu;
static void *na_ptr[64];  // per device: the array's device address

static cm_status na_scratch(cm_ctx *ctx, void **out) {
    if (ctx->device < 0 || ctx->device >= 64) return fail(ctx, CM_ERR_UNSUPPORTED, "device index %d", ctx->device);
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_ptr[ctx->device]) {
        // with the context's device current; the caller's current device is restored
        int prev = 0;
        CM_HIP(ctx, hipGetDevice(&prev));
        CM_HIP(ctx, hipSetDevice(ctx->device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
        hipSetDevice(prev);
        if (e != hipSuccess || !p) return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
        na_ptr[ctx->device] = p;
    }
    *out = na_ptr[ctx->device];
    return CM_OK;
}

// The tables of pk_tables on the device: 1/W^2 and the aliased shot-noise
// factor C1 per axis index (no interlacing table: the shells do not use it).
__global__ void k_na_tables(int N, int scheme, int mas_correction, double *__restrict__ tab) {
    const int M1 = N / 2 + 1;
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < M1; i += blockDim.x * gridDim.x) {
        const double a = M_PI * (double)i / (double)N;
        const double sinc = i == 0 ? 1.0 : sin(a) / a;
        double ic = 1.0;
        if (mas_correction) {
            const double cp = pow(sinc, scheme + 1);
            ic = 1.0 / (cp * cp);
        }
        const double s = sin(a), s2 = s * s;
        double sq = 1.0;
        if (scheme == 1) sq = 1.0 - (2.0 / 3.0) * s2;
        else if (scheme == 2) sq = 1.0 - s2 + (2.0 / 15.0) * s2 * s2;
        else if (scheme == 3) sq = 1.0 - (4.0 / 3.0) * s2 + (2.0 / 5.0) * s2 * s2 - (4.0 / 315.0) * s2 * s2 * s2;
        tab[i] = ic;
        tab[M1 + i] = sq;
    }
}

// pk_finalize on the device, into the results region.
__global__ void k_na_finalize(int ns, double normal_power, double sn_norm, const double *__restrict__ sum,
                              const unsigned long long *__restrict__ cnt, double *__restrict__ pk,
                              long long *__restrict__ nm) {
#pragma clang fp contract(off)
    for (int i = threadIdx.x; i < ns; i += blockDim.x) {
        const double den = (double)cnt[i] * normal_power;
        pk[i] = cnt[i] ? (sum[i] - sn_norm * sum[ns + i]) / den : 0.0;
        pk[ns + i] = cnt[i] ? 5.0 * sum[2 * ns + i] / den : 0.0;
        pk[2 * ns + i] = cnt[i] ? 9.0 * sum[3 * ns + i] / den : 0.0;
        nm[i] = (long long)cnt[i];
    }
}

}  // namespace cm

extern "C" cm_status cm_pk_shells_noalloc(cm_ctx *ctx, const cm_pk_opts *o, cm_dtype dtype, const void *dk,
                                          double *pk0, double *pk2, double *pk4, int64_t *nmodes) {
    using namespace cm;
    if (!ctx) return CM_ERR_INVALID;
    if (!o || !dk) return fail(ctx, CM_ERR_INVALID, "cm_pk_shells_noalloc: null pointer");
    if (o->N < 2 || (o->N & 1)) return fail(ctx, CM_ERR_INVALID, "cm_pk_shells_noalloc: N=%d must be even", o->N);
    if (o->N > kNaMaxN || o->nshells > kNaMaxShells)
        return fail(ctx, CM_ERR_UNSUPPORTED,
                    "cm_pk_shells_noalloc: N=%d, nshells=%d beyond its fixed scratch (N <= %d, nshells <= %d)", o->N,
                    o->nshells, kNaMaxN, kNaMaxShells);
    if (!(o->L > 0)) return fail(ctx, CM_ERR_INVALID, "cm_pk_shells_noalloc: L must be > 0");
    cm_status s = check_binning(ctx, o->nshells, o->kmin, o->dk, o->log_binning, o->dlog, o->float_shell_rule,
                                "cm_pk_shells_noalloc");
    if (s != CM_OK) return s;
    if ((int)o->scheme < 0 || (int)o->scheme > 3) return fail(ctx, CM_ERR_INVALID, "cm_pk_shells_noalloc: scheme");
    if (dtype != CM_F32 && dtype != CM_F64) return fail(ctx, CM_ERR_INVALID, "cm_pk_shells_noalloc: dtype");
    if (!(o->normal_power > 0) && (pk0 || pk2 || pk4))
        return fail(ctx, CM_ERR_INVALID, "cm_pk_shells_noalloc: normal_power must be > 0");
    void *base;
    if ((s = na_scratch(ctx, &base)) != CM_OK) return s;
    const int N = o->N, M1 = N / 2 + 1, ns = o->nshells;
    auto *w = (unsigned long long *)base;
    double *tab = (double *)w;                                 // 2*M1
    double *sums = (double *)(w + 2 * (size_t)M1);             // 4*ns
    unsigned long long *cnts = w + 2 * (size_t)M1 + 4 * (size_t)ns;
    double *res = (double *)(cnts + ns);                       // 3*ns
    long long *nm = (long long *)(res + 3 * (size_t)ns);       // ns
    int *qedge = (int *)(nm + ns);                             // ns+1
    // shell kernel LDS: the two tables in FP64, 4 FP64 sums and a count per shell
    const size_t lds = sizeof(double) * (2 * (size_t)M1 + 4 * (size_t)ns) + sizeof(unsigned) * (size_t)ns;
    if (lds > ctx->lds_per_block)
        return fail(ctx, CM_ERR_UNSUPPORTED, "cm_pk_shells_noalloc: %d shells need %zu bytes of LDS", ns, lds);
    PkK p;
    const double kf = 2.0 * M_PI / o->L;
    {
        // the linear rules of make_pkk; the log thresholds go into the scratch
        cm_pk_opts lin = *o;
        lin.log_binning = 0;
        p = make_pkk(ctx, &lin, &s);
        if (s != CM_OK) return s;
    }
    if (o->log_binning) {
        int qh[kNaMaxShells + 1];  // host thresholds, copied into the work area
        const long h = N / 2, qmax = 3 * h * h;
        log_shell_edges(o->kmin / kf, o->dlog, ns, qmax, qh);
        CM_HIP(ctx, hipMemcpyAsync(qedge, qh, sizeof(int) * (size_t)(ns + 1), hipMemcpyHostToDevice, ctx->stream));
        CM_HIP(ctx, hipStreamSynchronize(ctx->stream));  // qh lives on this stack frame
        p.int_rule = 0;
        p.float_rule = 0;
        p.qedge = qedge;
    }
    p.window = o->window_set ? 1 : 0;
    // every call shares this device's array: the sums start from zero (fill
    // kernel, no memset node), the tables are rewritten
    if ((s = launch(ctx, "pk_zero", k_pk_zero, dim3(1), dim3(256), 0, sums, cnts, ns)) != CM_OK) return s;
    if ((s = launch(ctx, "pk_na_tables", k_na_tables, dim3((M1 + 255) / 256), dim3(256), 0, N, (int)o->scheme,
                    o->mas_correction, tab)) != CM_OK)
        return s;
    const size_t total = (size_t)N * N * M1;
    const unsigned g = (unsigned)std::min<size_t>((total + 255) / 256, (size_t)ctx->cu_count * 8);
    if (dtype == CM_F32)
        s = launch(ctx, "pk_shells", k_pk_shells<float>, dim3(g), dim3(256), lds, (const v2pk<float> *)dk, p,
                   (const double *)tab, sums, cnts);
    else
        s = launch(ctx, "pk_shells", k_pk_shells<double>, dim3(g), dim3(256), lds, (const v2pk<double> *)dk, p,
                   (const double *)tab, sums, cnts);
    if (s != CM_OK) return s;
    if ((s = launch(ctx, "pk_na_finalize", k_na_finalize, dim3(1), dim3(256), 0, ns, o->normal_power,
                    o->shot_noise * o->normal_power, (const double *)sums, (const unsigned long long *)cnts, res,
                    nm)) != CM_OK)
        return s;
    const size_t b = sizeof(double) * (size_t)ns;
    if (pk0) CM_HIP(ctx, hipMemcpyAsync(pk0, res, b, hipMemcpyDeviceToHost, ctx->stream));
    if (pk2) CM_HIP(ctx, hipMemcpyAsync(pk2, res + ns, b, hipMemcpyDeviceToHost, ctx->stream));
    if (pk4) CM_HIP(ctx, hipMemcpyAsync(pk4, res + 2 * ns, b, hipMemcpyDeviceToHost, ctx->stream));
    if (nmodes) CM_HIP(ctx, hipMemcpyAsync(nmodes, nm, sizeof(int64_t) * (size_t)ns, hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CM_OK;
}
