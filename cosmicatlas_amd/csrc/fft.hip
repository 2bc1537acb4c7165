// 3-D real<->complex FFT for gfx950: three passes over the FFTW half-spectrum
// layout N x N x (N/2+1) (x slowest, kz fastest).
//   z pass: each row of N reals is packed as N/2 complex values, transformed in
//           LDS and split into N/2+1 outputs (r2c), or the inverse (c2r);
//   y pass / x pass: columns of the half spectrum, nb adjacent kz columns per
//           workgroup so that every global access is a contiguous nb*8-byte run.
#include <cmath>

#include "cm_internal.hpp"
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
                       double *part) {
    return dt == CM_F32 ? rows_r2c_t<float>(ctx, N, in, out, subtract, part)
                        : rows_r2c_t<double>(ctx, N, in, out, subtract, part);
}

int fft_rows_blocks(cm_ctx *ctx, int N, cm_dtype dt) {
    return (N * N) / row_cfg(N, dt, dt == CM_F32 ? ctx->tune.zrows : 0).nb;
}

cm_status fft_rows_c2r(cm_ctx *ctx, int N, cm_dtype dt, const void *in, void *out, double scale) {
    return dt == CM_F32 ? rows_c2r_t<float>(ctx, N, in, out, scale)
                        : rows_c2r_t<double>(ctx, N, in, out, scale);
}

cm_status fft_cols(cm_ctx *ctx, int N, cm_dtype dt, void *data, int axis, bool forward) {
    if (dt == CM_F32)
        return forward ? cols_t<float, true>(ctx, N, data, axis) : cols_t<float, false>(ctx, N, data, axis);
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
