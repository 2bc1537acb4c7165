// 3-D real<->complex FFT for gfx950: three passes over the FFTW half-spectrum
// layout N x N x (N/2+1) (x slowest, kz fastest).
//   z pass: each row of N reals is packed as N/2 complex values, transformed in
//           LDS and split into N/2+1 outputs (r2c), or the inverse (c2r);
//   y pass / x pass: columns of the half spectrum, nb adjacent kz columns per
//           workgroup so that every global access is a contiguous nb*8-byte run.
#include <cmath>

#include "cm_internal.hpp"
#include "fft_ct.hpp"
#include "fft_engine.hpp"

namespace cm {

FftPlan make_plan(int n) {
    FftPlan p;
    p.n = n;
    int m = n, n3 = 0, k2 = 0;
    while (m % 3 == 0) { m /= 3; ++n3; }
    while (m % 2 == 0) { m /= 2; ++k2; }
    if (m != 1) { p.nst = 0; return p; }
    int n8 = k2 / 3, rem = k2 % 3, n4 = 0, n2 = 0;
    if (rem == 2) n4 = 1;
    if (rem == 1) {
        if (n8 > 0) { --n8; n4 = 2; } else n2 = 1;
    }
    int s = 0;
    for (int i = 0; i < n8; ++i) p.r[s++] = 8;
    for (int i = 0; i < n4; ++i) p.r[s++] = 4;
    for (int i = 0; i < n2; ++i) p.r[s++] = 2;
    for (int i = 0; i < n3; ++i) p.r[s++] = 3;
    p.nst = s;
    return p;
}

bool fft_size_ok(int N) {
    if (N < 8 || N > 1024 || (N % 8)) return false;
    int m = N, n3 = 0;
    while (m % 3 == 0) { m /= 3; ++n3; }
    if (n3 > 1) return false;
    return (m & (m - 1)) == 0;
}

static DevPlan dev_plan(const FftPlan &p) {
    DevPlan d;
    d.n = p.n;
    d.nst = p.nst;
    for (int i = 0; i < 8; ++i) d.r[i] = p.r[i];
    return d;
}

static int plan_E(const FftPlan &p) {
    bool has3 = false, has8 = false;
    for (int i = 0; i < p.nst; ++i) {
        if (p.r[i] == 3) has3 = true;
        if (p.r[i] == 8) has8 = true;
    }
    if (has3) return 24;
    return (has8 && p.n >= 64) ? 16 : 8;
}

// ---------------------------------------------------------------- kernels
// part != nullptr: each workgroup also writes the FP64 sum of its input rows
// (before `sub` is removed) to part[blockIdx.x] -- the mesh sum rides along.
template <typename T, int E>
__global__ void __launch_bounds__(512) k_rows_r2c(const T *__restrict__ in, v2<T> *__restrict__ out,
                                                   DevPlan pl, int N, int nb, T sub,
                                                   const v2<T> *__restrict__ twM,
                                                   const v2<T> *__restrict__ twN,
                                                   double *__restrict__ part) {
    using C = v2<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[8];
    C *lds = reinterpret_cast<C *>(smem_raw);
    const int M = N / 2, stride = row_stride(M);
    const size_t row0 = (size_t)blockIdx.x * nb;
    const C *inrow = reinterpret_cast<const C *>(in + row0 * (size_t)N);
    double acc = 0.0;
    auto src = [&](int b, int m) {
        C z = inrow[(size_t)b * M + m];
        acc += (double)z.x + (double)z.y;
        z.x -= sub;
        z.y -= sub;
        return z;
    };
    auto dst = [&](int b, int k, C v) { lds[b * stride + rpad(k)] = v; };
    fft_run<T, E, false, true>(lds, stride, pl, nb, twM, src, dst);
    if (part) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (part && threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
    C *orow = out + row0 * (size_t)(M + 1);
    const T half = (T)0.5;
    for (int t = threadIdx.x; t < nb * (M + 1); t += blockDim.x) {
        int b = t / (M + 1), k = t - b * (M + 1);
        C zk = lds[b * stride + rpad(k == M ? 0 : k)];
        C zm = cconj(lds[b * stride + rpad(k == 0 ? 0 : M - k)]);
        C e = C{half * (zk.x + zm.x), half * (zk.y + zm.y)};
        C o = C{half * (zk.x - zm.x), half * (zk.y - zm.y)};
        // X_k = E_k + W^k O_k with O_k = (Z_k - conj Z_{M-k}) / (2i)
        C oi = C{o.y, -o.x};  // o / i
        C w = twN[k];
        C x = cadd(e, cmul(w, oi));
        orow[(size_t)b * (M + 1) + k] = x;
    }
}

template <typename T, int E>
__global__ void __launch_bounds__(512) k_rows_c2r(const v2<T> *__restrict__ in, T *__restrict__ out,
                                                   DevPlan pl, int N, int nb, T scale,
                                                   const v2<T> *__restrict__ twM,
                                                   const v2<T> *__restrict__ twN) {
    using C = v2<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *lds = reinterpret_cast<C *>(smem_raw);
    const int M = N / 2, stride = row_stride(M);
    const size_t row0 = (size_t)blockIdx.x * nb;
    const C *irow = in + row0 * (size_t)(M + 1);
    for (int t = threadIdx.x; t < nb * (M + 1); t += blockDim.x) {
        int b = t / (M + 1), k = t - b * (M + 1);
        lds[b * stride + rpad(k)] = irow[(size_t)b * (M + 1) + k];
    }
    __syncthreads();
    const T half = (T)0.5;
    auto src = [&](int b, int m) {
        C xm = lds[b * stride + rpad(m)];
        C xc = cconj(lds[b * stride + rpad(M - m)]);
        C e = C{half * (xm.x + xc.x), half * (xm.y + xc.y)};
        C o = C{half * (xm.x - xc.x), half * (xm.y - xc.y)};
        C w = cconj(twN[m]);
        o = cmul(o, w);
        // Z = E + i O
        return C{e.x - o.y, e.y + o.x};
    };
    C *orow = reinterpret_cast<C *>(out + row0 * (size_t)N);
    const T s2 = (T)2 * scale;
    auto dst = [&](int b, int m, C z) { orow[(size_t)b * M + m] = C{s2 * z.x, s2 * z.y}; };
    fft_run<T, E, false, false>(lds, stride, pl, nb, twM, src, dst);
}

template <typename T, int E, bool FWD>
__global__ void __launch_bounds__(512) k_cols(v2<T> *__restrict__ data, DevPlan pl, int nb, int ncols,
                                               size_t stride, size_t planeStride,
                                               const v2<T> *__restrict__ tw) {
    using C = v2<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *lds = reinterpret_cast<C *>(smem_raw);
    const size_t base = (size_t)blockIdx.y * planeStride + (size_t)blockIdx.x * nb;
    const int valid = min(nb, ncols - (int)blockIdx.x * nb);
    auto src = [&](int b, int e) {
        return b < valid ? data[base + (size_t)e * stride + b] : C{(T)0, (T)0};
    };
    auto dst = [&](int b, int e, C v) {
        if (b < valid) data[base + (size_t)e * stride + b] = v;
    };
    fft_run<T, E, true, FWD>(lds, 0, pl, nb, tw, src, dst);
}

// ---------------------------------------------------- float32, compile-time N
// The same three passes with the plan, block shape and strides fixed at
// compile time (fft_ct.hpp).  `pitch` is the kz row pitch of the half
// spectrum in complex values (N/2+1 for the FFTW layout, padded_pitch(N) for
// the internal one).
template <int N>
__global__ void __launch_bounds__(RowShape<N>::NT) k_rows_r2c_ct(const float *__restrict__ in,
                                                                 float2 *__restrict__ out, int pitch, float sub,
                                                                 const float2 *__restrict__ twM,
                                                                 const float2 *__restrict__ twN,
                                                                 double *__restrict__ part,
                                                                 const float2 *__restrict__ phase) {
    using RS = RowShape<N>;
    constexpr int M = RS::M, NB = RS::NB, STRIDE = RS::STRIDE, NT = RS::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[(NT + 63) / 64];
    float2 *lds = reinterpret_cast<float2 *>(smem_raw);
    const size_t row0 = (size_t)blockIdx.x * NB;
    const float2 *inrow = reinterpret_cast<const float2 *>(in + row0 * (size_t)N);
    float acc = 0.f;  // at most E values per thread; the workgroup total is FP64
    auto src = [&](int b, int m) {
        float2 z = load_pass<N>(&inrow[b * M + m]);
        acc += z.x + z.y;
        z.x -= sub;
        z.y -= sub;
        return z;
    };
    auto dst = [&](int b, int k, float2 v) { lds[b * STRIDE + rpad(k)] = v; };
    ct_fft<float, M, NB, STRIDE, false, true>(lds, twM, src, dst);
    if (part) {
        double a = (double)acc;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (part && threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < (NT + 63) / 64; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
    float2 *orow = out + row0 * (size_t)pitch;
    for (int t = threadIdx.x; t < NB * (M + 1); t += NT) {
        const int b = t / (M + 1), k = t - b * (M + 1);
        const float2 zk = lds[b * STRIDE + rpad(k == M ? 0 : k)];
        const float2 zm = cconj(lds[b * STRIDE + rpad(k == 0 ? 0 : M - k)]);
        const float2 e = float2{0.5f * (zk.x + zm.x), 0.5f * (zk.y + zm.y)};
        const float2 o = float2{0.5f * (zk.x - zm.x), 0.5f * (zk.y - zm.y)};
        // X_k = E_k + W^k O_k with O_k = (Z_k - conj Z_{M-k}) / (2i)
        const float2 oi = float2{o.y, -o.x};
        float2 x = cadd(e, cmul(twN[k], oi));
        if (phase) x = cmul(x, phase[k]);
        orow[(size_t)b * pitch + k] = x;
    }
}

template <int N>
__global__ void __launch_bounds__(RowShape<N>::NT) k_rows_c2r_ct(const float2 *__restrict__ in,
                                                                 float *__restrict__ out, int pitch, float scale,
                                                                 const float2 *__restrict__ twM,
                                                                 const float2 *__restrict__ twN,
                                                                 double *__restrict__ sq, int nk) {
    using RS = RowShape<N>;
    constexpr int M = RS::M, NB = RS::NB, STRIDE = RS::STRIDE, NT = RS::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *lds = reinterpret_cast<float2 *>(smem_raw);
    const size_t row0 = (size_t)blockIdx.x * NB;
    const float2 *irow = in + row0 * (size_t)pitch;
    // stage the rows in LDS: all loads first (one latency), then the stores
    constexpr int NL = (NB * (M + 1) + NT - 1) / NT;
    float2 ld[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int t = threadIdx.x + i * NT;
        if (t < NB * (M + 1)) {
            const int b = t / (M + 1), k = t - b * (M + 1);
            ld[i] = k < nk ? load_pass<N>(&irow[(size_t)b * pitch + k]) : float2{0.f, 0.f};
        }
    }
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int t = threadIdx.x + i * NT;
        if (t < NB * (M + 1)) {
            const int b = t / (M + 1), k = t - b * (M + 1);
            lds[b * STRIDE + rpad(k)] = ld[i];
        }
    }
    __syncthreads();
    auto src = [&](int b, int m) {
        const float2 xm = lds[b * STRIDE + rpad(m)];
        const float2 xc = cconj(lds[b * STRIDE + rpad(M - m)]);
        const float2 e = float2{0.5f * (xm.x + xc.x), 0.5f * (xm.y + xc.y)};
        const float2 o = cmul(float2{0.5f * (xm.x - xc.x), 0.5f * (xm.y - xc.y)}, cconj(twN[m]));
        return float2{e.x - o.y, e.y + o.x};  // Z = E + i O
    };
    float2 *orow = reinterpret_cast<float2 *>(out + row0 * (size_t)N);
    const float s2 = 2.0f * scale;
    double acc = 0.0;  // sum of squares of this thread's outputs
    auto dst = [&](int b, int m, float2 z) {
        const float2 o = float2{s2 * z.x, s2 * z.y};
        store_stream(&orow[b * M + m], o);
        acc += (double)o.x * o.x + (double)o.y * o.y;
    };
    ct_fft<float, M, NB, STRIDE, false, false>(lds, twM, src, dst);
    if (sq) {  // per-workgroup sum of squares (one FP64 word per workgroup)
        __shared__ double red[(NT + 63) / 64];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < (NT + 63) / 64; ++w) t += red[w];
            sq[blockIdx.x] = t;
        }
    }
}

template <int N, int NB, bool FWD, bool EL = false>
__global__ void __launch_bounds__((ColShape<N, NB>::NT)) k_cols_ct(float2 *__restrict__ data, int ncols, size_t stride,
                                                                 size_t planeStride, const float2 *__restrict__ tw,
                                                                 const float2 *__restrict__ phase, SumTail tail,
                                                                 int ntl, int pk, int pskip, int elim) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *lds = reinterpret_cast<float2 *>(smem_raw);
    // planes above pk are offset by pskip (a pass over the two ends of the
    // plane range only)
    const size_t plane = blockIdx.y + ((int)blockIdx.y > pk ? pskip : 0);
    const size_t base = plane * planeStride + (size_t)blockIdx.x * NB;
    const int valid = min(NB, ncols - (int)blockIdx.x * NB);
    // ntl (uniform): nontemporal loads (the fused measurement's y passes of
    // meshes from 512^3 up; in the bispectrum's shell transforms they made
    // the column passes slower, the x passes (columns of planes 1 MB apart)
    // 2.2x)
    auto src = [&](int b, int e) {
        // EL: elements elim < e < N - elim are zero (not read)
        if (b >= valid || (EL && e > elim && e < N - elim)) return float2{0.f, 0.f};
        const float2 *q = &data[base + (size_t)e * stride + b];
        return ntl ? load_stream(q) : *q;
    };
    auto dst = [&](int b, int e, float2 v) {
        if (phase) v = cmul(v, phase[e]);
        if (b < valid) store_stream(&data[base + (size_t)e * stride + b], v);
    };
    ct_fft<float, N, NB, NB, true, FWD>(lds, tw, src, dst);
    if (tail.part && blockIdx.x == 0 && blockIdx.y == 0) {  // uniform per workgroup
        __syncthreads();
        sum_tail(tail);
    }
}

// column block width of the float32 y/x passes
template <int N> constexpr int col_nb() { return N > 768 ? 8 : 16; }

template <int N>
static cm_status rows_r2c_ct(cm_ctx *ctx, const float *in, float2 *out, int pitch, float sub, const float2 *twM,
                             const float2 *twN, double *part, const float2 *phase) {
    using RS = RowShape<N>;
    if (RS::LDS > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "fft z pass: LDS");
    return launch(ctx, "fft_z_r2c", k_rows_r2c_ct<N>, dim3((unsigned)((size_t)N * N / RS::NB)), dim3(RS::NT), RS::LDS,
                  in, out, pitch, sub, twM, twN, part, phase);
}

template <int N>
static cm_status rows_c2r_ct(cm_ctx *ctx, const float2 *in, float *out, int pitch, float scale, const float2 *twM,
                             const float2 *twN, double *sq, int nk) {
    using RS = RowShape<N>;
    if (RS::LDS > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "fft z pass: LDS");
    return launch(ctx, "fft_z_c2r", k_rows_c2r_ct<N>, dim3((unsigned)((size_t)N * N / RS::NB)), dim3(RS::NT), RS::LDS,
                  in, out, pitch, scale, twM, twN, sq, nk > 0 ? nk : N / 2 + 1);
}

template <int N, bool FWD>
static cm_status cols_ct(cm_ctx *ctx, float2 *data, int axis, int pitch, const float2 *tw, const float2 *phase,
                         const SumTail *tail, bool stream_loads, int ncols_used, int band) {
    constexpr int NB = col_nb<N>();
    using CS = ColShape<N, NB>;
    if (CS::LDS > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "fft column pass: LDS");
    const int ncols = ncols_used > 0 ? std::min(ncols_used, N / 2 + 1) : N / 2 + 1;
    unsigned planes = (unsigned)N;
    int pk = N, pskip = 0, elim = N;
    if (band >= 0 && 2 * band + 1 < N) {
        if (axis == 0) {  // y planes 0..band and N-band..N-1
            planes = (unsigned)(2 * band + 1);
            pk = band;
            pskip = N - 2 * band - 1;
        } else {  // y elements band < e < N - band are zero
            elim = band;
        }
    }
    size_t stride, planeStride;
    if (axis == 1) {  // y: fixed x plane, stride over kz rows
        stride = (size_t)pitch;
        planeStride = (size_t)N * pitch;
    } else {  // x: fixed y, stride over planes
        stride = (size_t)N * pitch;
        planeStride = (size_t)pitch;
    }
    dim3 grid((unsigned)((ncols + NB - 1) / NB), planes);
    const char *name = axis == 1 ? (FWD ? "fft_y_fwd" : "fft_y_inv") : (FWD ? "fft_x_fwd" : "fft_x_inv");
    const SumTail t = tail ? *tail : SumTail{nullptr, 0, nullptr, nullptr, nullptr, 0, nullptr};
    return launch(ctx, name, elim < N ? k_cols_ct<N, NB, FWD, true> : k_cols_ct<N, NB, FWD, false>, grid,
                  dim3(CS::NT), CS::LDS, data, ncols, stride, planeStride, tw, phase, t,
                  (stream_loads && N >= 512 && axis == 1) ? 1 : 0, pk, pskip, elim);
}

static cm_status tw_pair(cm_ctx *ctx, int N, const float2 **twM, const float2 **twN) {
    cm_status st;
    *twM = (const float2 *)twiddles(ctx, N / 2, CM_F32, &st);
    if (!*twM) return st;
    *twN = (const float2 *)twiddles(ctx, N, CM_F32, &st);
    if (!*twN) return st;
    return CM_OK;
}

static cm_status rows_r2c_f32(cm_ctx *ctx, int N, const void *in, void *out, int pitch, double subtract,
                              double *part, const float2 *phase) {
    const float2 *twM, *twN;
    cm_status s = tw_pair(ctx, N, &twM, &twN);
    if (s != CM_OK) return s;
    auto *ip = (const float *)in;
    auto *op = (float2 *)out;
    switch (N) {
#define CM_CASE(n) \
    case n: return rows_r2c_ct<n>(ctx, ip, op, pitch, (float)subtract, twM, twN, part, phase);
        CM_FFT_SIZES(CM_CASE)
#undef CM_CASE
    }
    return fail(ctx, CM_ERR_UNSUPPORTED, "fft: N=%d", N);
}

static cm_status rows_c2r_f32(cm_ctx *ctx, int N, const void *in, void *out, int pitch, double scale,
                              double *sq, int nk) {
    const float2 *twM, *twN;
    cm_status s = tw_pair(ctx, N, &twM, &twN);
    if (s != CM_OK) return s;
    auto *ip = (const float2 *)in;
    auto *op = (float *)out;
    switch (N) {
#define CM_CASE(n) \
    case n: return rows_c2r_ct<n>(ctx, ip, op, pitch, (float)scale, twM, twN, sq, nk);
        CM_FFT_SIZES(CM_CASE)
#undef CM_CASE
    }
    return fail(ctx, CM_ERR_UNSUPPORTED, "fft: N=%d", N);
}

static cm_status cols_f32(cm_ctx *ctx, int N, void *data, int axis, bool forward, int pitch, const float2 *phase,
                          const SumTail *tail, bool stream_loads, int ncols, int band) {
    cm_status st;
    const float2 *tw = (const float2 *)twiddles(ctx, N, CM_F32, &st);
    if (!tw) return st;
    auto *dp = (float2 *)data;
    switch (N) {
#define CM_CASE(n)                                                                     \
    case n:                                                                            \
        return forward ? cols_ct<n, true>(ctx, dp, axis, pitch, tw, phase, tail, stream_loads, ncols, band)       \
                       : cols_ct<n, false>(ctx, dp, axis, pitch, tw, phase, tail, stream_loads, ncols, band);
        CM_FFT_SIZES(CM_CASE)
#undef CM_CASE
    }
    return fail(ctx, CM_ERR_UNSUPPORTED, "fft: N=%d", N);
}

static int rows_blocks_f32(int N) {
    switch (N) {
#define CM_CASE(n) \
    case n: return (n * n) / RowShape<n>::NB;
        CM_FFT_SIZES(CM_CASE)
#undef CM_CASE
    }
    return 0;
}

// ---------------------------------------------------------------- host side
struct RowCfg {
    FftPlan plan;
    int E, nb, NT;
    size_t lds;
};

static RowCfg row_cfg(int N, cm_dtype dt, int want_rows = 0) {
    RowCfg c;
    int M = N / 2;
    c.plan = make_plan(M);
    c.E = plan_E(c.plan);
    int target = dt == CM_F32 ? 4096 : 2048;
    if (want_rows > 0) target = want_rows * M;
    int nb = 1;
    const int rows = N * N;
    const int rows_pow2 = rows & -rows;  // nb must divide the row count
    while (nb * 2 * M <= target && nb * 2 <= rows_pow2) nb *= 2;
    while ((nb * M) % c.E) nb *= 2;
    while (nb * M / c.E > 512 && nb > 1) nb /= 2;
    c.nb = nb;
    c.NT = nb * M / c.E;
    c.lds = (size_t)nb * row_stride(M) * (dt == CM_F32 ? 8 : 16);
    return c;
}

struct ColCfg {
    FftPlan plan;
    int E, nb, NT;
    size_t lds;
};

static ColCfg col_cfg(int N, cm_dtype dt, int want_nb = 0) {
    ColCfg c;
    c.plan = make_plan(N);
    c.E = plan_E(c.plan);
    int nb = dt == CM_F32 ? 16 : 8;
    if (N > 512) nb /= 2;
    if (want_nb > 0 && (want_nb & (want_nb - 1)) == 0 && want_nb <= 64) nb = want_nb;
    while (nb * N / c.E > 512 && nb > 1) nb /= 2;
    while ((nb * N) % c.E) nb *= 2;
    c.nb = nb;
    c.NT = nb * N / c.E;
    c.lds = (size_t)nb * N * (dt == CM_F32 ? 8 : 16);
    return c;
}

template <typename T>
static cm_status rows_r2c_t(cm_ctx *ctx, int N, const void *in, void *out, double subtract, double *part) {
    cm_dtype dt = sizeof(T) == 4 ? CM_F32 : CM_F64;
    RowCfg c = row_cfg(N, dt, dt == CM_F32 ? ctx->tune.zrows : 0);
    cm_status st;
    const void *twM = twiddles(ctx, N / 2, dt, &st);
    if (!twM) return st;
    const void *twN = twiddles(ctx, N, dt, &st);
    if (!twN) return st;
    size_t rows = (size_t)N * N;
    if (rows % c.nb || c.lds > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "fft z pass: bad row block");
    dim3 grid((unsigned)(rows / c.nb));
    DevPlan pl = dev_plan(c.plan);
    auto *ip = (const T *)in;
    auto *op = (v2<T> *)out;
    auto *a = (const v2<T> *)twM;
    auto *b = (const v2<T> *)twN;
    switch (c.E) {
    case 8: return launch(ctx, "fft_z_r2c", k_rows_r2c<T, 8>, grid, dim3(c.NT), c.lds, ip, op, pl, N, c.nb, (T)subtract, a, b, part);
    case 16: return launch(ctx, "fft_z_r2c", k_rows_r2c<T, 16>, grid, dim3(c.NT), c.lds, ip, op, pl, N, c.nb, (T)subtract, a, b, part);
    case 24: return launch(ctx, "fft_z_r2c", k_rows_r2c<T, 24>, grid, dim3(c.NT), c.lds, ip, op, pl, N, c.nb, (T)subtract, a, b, part);
    }
    return fail(ctx, CM_ERR_UNSUPPORTED, "fft: no kernel for E=%d", c.E);
}

template <typename T>
static cm_status rows_c2r_t(cm_ctx *ctx, int N, const void *in, void *out, double scale) {
    cm_dtype dt = sizeof(T) == 4 ? CM_F32 : CM_F64;
    RowCfg c = row_cfg(N, dt, dt == CM_F32 ? ctx->tune.zrows : 0);
    cm_status st;
    const void *twM = twiddles(ctx, N / 2, dt, &st);
    if (!twM) return st;
    const void *twN = twiddles(ctx, N, dt, &st);
    if (!twN) return st;
    size_t rows = (size_t)N * N;
    if (rows % c.nb || c.lds > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "fft z pass: bad row block");
    dim3 grid((unsigned)(rows / c.nb));
    DevPlan pl = dev_plan(c.plan);
    auto *ip = (const v2<T> *)in;
    auto *op = (T *)out;
    auto *a = (const v2<T> *)twM;
    auto *b = (const v2<T> *)twN;
    switch (c.E) {
    case 8: return launch(ctx, "fft_z_c2r", k_rows_c2r<T, 8>, grid, dim3(c.NT), c.lds, ip, op, pl, N, c.nb, (T)scale, a, b);
    case 16: return launch(ctx, "fft_z_c2r", k_rows_c2r<T, 16>, grid, dim3(c.NT), c.lds, ip, op, pl, N, c.nb, (T)scale, a, b);
    case 24: return launch(ctx, "fft_z_c2r", k_rows_c2r<T, 24>, grid, dim3(c.NT), c.lds, ip, op, pl, N, c.nb, (T)scale, a, b);
    }
    return fail(ctx, CM_ERR_UNSUPPORTED, "fft: no kernel for E=%d", c.E);
}

template <typename T, bool FWD>
static cm_status cols_t(cm_ctx *ctx, int N, void *data, int axis) {
    cm_dtype dt = sizeof(T) == 4 ? CM_F32 : CM_F64;
    ColCfg c = col_cfg(N, dt, dt == CM_F32 ? ctx->tune.colnb : 0);
    if (c.NT > 512 || c.lds > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "fft column pass: bad block");
    cm_status st;
    const void *tw = twiddles(ctx, N, dt, &st);
    if (!tw) return st;
    const int ncols = N / 2 + 1;
    size_t stride, planeStride;
    if (axis == 1) {  // y: fixed x plane, stride over kz rows
        stride = (size_t)ncols;
        planeStride = (size_t)N * ncols;
    } else {  // x: fixed y, stride over planes
        stride = (size_t)N * ncols;
        planeStride = (size_t)ncols;
    }
    dim3 grid((unsigned)((ncols + c.nb - 1) / c.nb), (unsigned)N);
    DevPlan pl = dev_plan(c.plan);
    auto *dp = (v2<T> *)data;
    auto *twp = (const v2<T> *)tw;
    const char *name = axis == 1 ? (FWD ? "fft_y_fwd" : "fft_y_inv") : (FWD ? "fft_x_fwd" : "fft_x_inv");
    switch (c.E) {
    case 8: return launch(ctx, name, k_cols<T, 8, FWD>, grid, dim3(c.NT), c.lds, dp, pl, c.nb, ncols, stride, planeStride, twp);
    case 16: return launch(ctx, name, k_cols<T, 16, FWD>, grid, dim3(c.NT), c.lds, dp, pl, c.nb, ncols, stride, planeStride, twp);
    case 24: return launch(ctx, name, k_cols<T, 24, FWD>, grid, dim3(c.NT), c.lds, dp, pl, c.nb, ncols, stride, planeStride, twp);
    }
    return fail(ctx, CM_ERR_UNSUPPORTED, "fft: no kernel for E=%d", c.E);
}

cm_status fft_rows_r2c(cm_ctx *ctx, int N, cm_dtype dt, const void *in, void *out, double subtract,
                       double *part, int pitch, const void *phase) {
    if (dt == CM_F32)
        return rows_r2c_f32(ctx, N, in, out, pitch > 0 ? pitch : N / 2 + 1, subtract, part, (const float2 *)phase);
    if (phase) return fail(ctx, CM_ERR_UNSUPPORTED, "fft: phase on an FP64 spectrum");
    if (pitch > 0 && pitch != N / 2 + 1) return fail(ctx, CM_ERR_UNSUPPORTED, "fft: padded FP64 spectrum");
    return rows_r2c_t<double>(ctx, N, in, out, subtract, part);
}

int fft_rows_blocks(cm_ctx *ctx, int N, cm_dtype dt) {
    if (dt == CM_F32) return rows_blocks_f32(N);
    return (N * N) / row_cfg(N, dt, 0).nb;
}

cm_status fft_rows_c2r(cm_ctx *ctx, int N, cm_dtype dt, const void *in, void *out, double scale, int pitch,
                       double *sq, int nk) {
    if (dt == CM_F32) return rows_c2r_f32(ctx, N, in, out, pitch > 0 ? pitch : N / 2 + 1, scale, sq, nk);
    if (sq) return fail(ctx, CM_ERR_UNSUPPORTED, "fft: sum of squares of an FP64 c2r");
    if (pitch > 0 && pitch != N / 2 + 1) return fail(ctx, CM_ERR_UNSUPPORTED, "fft: padded FP64 spectrum");
    return rows_c2r_t<double>(ctx, N, in, out, scale);
}

cm_status fft_cols(cm_ctx *ctx, int N, cm_dtype dt, void *data, int axis, bool forward, int pitch,
                   const void *phase, const SumTail *tail, bool stream_loads, int ncols, int band) {
    if (dt == CM_F32)
        return cols_f32(ctx, N, data, axis, forward, pitch > 0 ? pitch : N / 2 + 1, (const float2 *)phase, tail,
                        stream_loads, ncols, band);
    if (phase || tail) return fail(ctx, CM_ERR_UNSUPPORTED, "fft: phase or tail on an FP64 spectrum");
    if (pitch > 0 && pitch != N / 2 + 1) return fail(ctx, CM_ERR_UNSUPPORTED, "fft: padded FP64 spectrum");
    return forward ? cols_t<double, true>(ctx, N, data, axis) : cols_t<double, false>(ctx, N, data, axis);
}

cm_status fft_r2c_impl(cm_ctx *ctx, int N, cm_dtype dt, const void *in, void *out, double subtract) {
    cm_status s = fft_rows_r2c(ctx, N, dt, in, out, subtract, nullptr);
    if (s != CM_OK) return s;
    if ((s = fft_cols(ctx, N, dt, out, 1, true)) != CM_OK) return s;
    return fft_cols(ctx, N, dt, out, 0, true);
}

cm_status fft_c2r_impl(cm_ctx *ctx, int N, cm_dtype dt, void *in, void *out, double scale) {
    cm_status s = fft_cols(ctx, N, dt, in, 0, false);
    if (s != CM_OK) return s;
    if ((s = fft_cols(ctx, N, dt, in, 1, false)) != CM_OK) return s;
    return fft_rows_c2r(ctx, N, dt, in, out, scale);
}

cm_status fft_c2r_band(cm_ctx *ctx, int N, cm_dtype dt, void *in, void *out, double scale, int K) {
    if (K < 0 || K >= N / 2) return fft_c2r_impl(ctx, N, dt, in, out, scale);
    cm_status s = fft_cols(ctx, N, dt, in, 0, false, 0, nullptr, nullptr, false, K + 1, K);
    if (s != CM_OK) return s;
    if ((s = fft_cols(ctx, N, dt, in, 1, false, 0, nullptr, nullptr, false, K + 1, K)) != CM_OK) return s;
    return fft_rows_c2r(ctx, N, dt, in, out, scale, 0, nullptr, K + 1);
}

}  // namespace cm

using namespace cm;

extern "C" {

int cm_fft_supported(int N) { return fft_size_ok(N) ? 1 : 0; }

cm_status cm_fft_r2c(cm_ctx *ctx, int N, cm_dtype dtype, const void *in, void *out, double subtract) {
    if (!ctx) return CM_ERR_INVALID;
    if (!in || !out) return fail(ctx, CM_ERR_INVALID, "cm_fft_r2c: null pointer");
    if (!fft_size_ok(N)) return fail(ctx, CM_ERR_INVALID, "cm_fft_r2c: N=%d not supported", N);
    if (dtype != CM_F32 && dtype != CM_F64) return fail(ctx, CM_ERR_INVALID, "cm_fft_r2c: dtype");
    return fft_r2c_impl(ctx, N, dtype, in, out, subtract);
}

cm_status cm_fft_c2r(cm_ctx *ctx, int N, cm_dtype dtype, void *in, void *out) {
    if (!ctx) return CM_ERR_INVALID;
    if (!in || !out) return fail(ctx, CM_ERR_INVALID, "cm_fft_c2r: null pointer");
    if (!fft_size_ok(N)) return fail(ctx, CM_ERR_INVALID, "cm_fft_c2r: N=%d not supported", N);
    if (dtype != CM_F32 && dtype != CM_F64) return fail(ctx, CM_ERR_INVALID, "cm_fft_c2r: dtype");
    double n3 = (double)N * N * N;
    return fft_c2r_impl(ctx, N, dtype, in, out, 1.0 / n3);
}

}  // extern "C"
