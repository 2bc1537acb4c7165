// T-web classification for gfx950 (plain kernels).
//   delta -> r2c -> exp(-k^2R^2/2)(-1/k^2) -> c2r -> phi            (Poisson)
//   phi -> G_a = D_a phi (a = x,y,z)                                 (k_grad3)
//   G -> T_ab = D_b G_a -> FP64 eigenvalues -> label, counts         (k_hess_eig)
// The lower triangle the reference feeds to GSL is A10 = D_x G_y,
// A20 = D_x G_z, A21 = D_y G_z (Miscelanious.cpp:1684-1740).
#include <cmath>

#include "cm_internal.hpp"
#include "tweb_common.hpp"

namespace cm {

template <typename T>
__global__ void __launch_bounds__(256) k_kmul(typename std::conditional<sizeof(T) == 4, float2, double2>::type *dk,
                                              int N, double kf, double R2h, double scale) {
    const int M1 = N / 2 + 1;
    const size_t total = (size_t)N * N * M1;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(g % M1);
        const size_t r = g / M1;
        const int iy = (int)(r % N), ix = (int)(r / N);
        const double f = smooth_poisson_factor(ix, iy, iz, N, kf, R2h, scale);
        auto v = dk[g];
        v.x = (T)((double)v.x * f);
        v.y = (T)((double)v.y * f);
        dk[g] = v;
    }
}

__global__ void __launch_bounds__(256) k_grad3(const float *__restrict__ phi, int N, double ih,
                                               float *__restrict__ gx, float *__restrict__ gy,
                                               float *__restrict__ gz) {
    const size_t n3 = (size_t)N * N * N;
    const size_t NN = (size_t)N;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n3; c += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(c % N), y = (int)((c / N) % N), x = (int)(c / (NN * N));
        auto at = [&](int i, int j, int k) { return phi[((size_t)i * NN + j) * NN + k]; };
        const int xm2 = (x - 2 + N) % N, xm1 = (x - 1 + N) % N, xp1 = (x + 1) % N, xp2 = (x + 2) % N;
        const int ym2 = (y - 2 + N) % N, ym1 = (y - 1 + N) % N, yp1 = (y + 1) % N, yp2 = (y + 2) % N;
        const int zm2 = (z - 2 + N) % N, zm1 = (z - 1 + N) % N, zp1 = (z + 1) % N, zp2 = (z + 2) % N;
        gx[c] = fd4(at(xm2, y, z), at(xm1, y, z), at(xp1, y, z), at(xp2, y, z), ih);
        gy[c] = fd4(at(x, ym2, z), at(x, ym1, z), at(x, yp1, z), at(x, yp2, z), ih);
        gz[c] = fd4(at(x, y, zm2), at(x, y, zm1), at(x, y, zp1), at(x, y, zp2), ih);
    }
}

__global__ void __launch_bounds__(256) k_hess_eig(const float *__restrict__ gx, const float *__restrict__ gy,
                                                  const float *__restrict__ gz, int N, double ih,
                                                  double lth, double tol, int8_t *__restrict__ labels,
                                                  float *__restrict__ lam, unsigned long long *__restrict__ stats) {
    __shared__ unsigned cnt[5];
    if (threadIdx.x < 5) cnt[threadIdx.x] = 0;
    __syncthreads();
    const size_t n3 = (size_t)N * N * N;
    const size_t NN = (size_t)N;
    unsigned loc[5] = {0, 0, 0, 0, 0};
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n3; c += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(c % N), y = (int)((c / N) % N), x = (int)(c / (NN * N));
        const int xm2 = (x - 2 + N) % N, xm1 = (x - 1 + N) % N, xp1 = (x + 1) % N, xp2 = (x + 2) % N;
        const int ym2 = (y - 2 + N) % N, ym1 = (y - 1 + N) % N, yp1 = (y + 1) % N, yp2 = (y + 2) % N;
        const int zm2 = (z - 2 + N) % N, zm1 = (z - 1 + N) % N, zp1 = (z + 1) % N, zp2 = (z + 2) % N;
        auto id = [&](int i, int j, int k) { return ((size_t)i * NN + j) * NN + k; };
        // T_xx = D_x G_x, A10 = D_x G_y, A20 = D_x G_z
        const float txx = fd4(gx[id(xm2, y, z)], gx[id(xm1, y, z)], gx[id(xp1, y, z)], gx[id(xp2, y, z)], ih);
        const float tyx = fd4(gy[id(xm2, y, z)], gy[id(xm1, y, z)], gy[id(xp1, y, z)], gy[id(xp2, y, z)], ih);
        const float tzx = fd4(gz[id(xm2, y, z)], gz[id(xm1, y, z)], gz[id(xp1, y, z)], gz[id(xp2, y, z)], ih);
        // T_yy = D_y G_y, A21 = D_y G_z
        const float tyy = fd4(gy[id(x, ym2, z)], gy[id(x, ym1, z)], gy[id(x, yp1, z)], gy[id(x, yp2, z)], ih);
        const float tzy = fd4(gz[id(x, ym2, z)], gz[id(x, ym1, z)], gz[id(x, yp1, z)], gz[id(x, yp2, z)], ih);
        // T_zz = D_z G_z
        const float tzz = fd4(gz[id(x, y, zm2)], gz[id(x, y, zm1)], gz[id(x, y, zp1)], gz[id(x, y, zp2)], ih);
        double e1, e2, e3;
        eig3(txx, tyy, tzz, tyx, tzx, tzy, e1, e2, e3);
        const int v = (e1 > lth) + (e2 > lth) + (e3 > lth);
        const int lab = 4 - v;
        labels[c] = (int8_t)lab;
        if (lam) {
            lam[c] = (float)e1;
            lam[n3 + c] = (float)e2;
            lam[2 * n3 + c] = (float)e3;
        }
        loc[lab - 1]++;
        const double m = fmin(fabs(e1 - lth), fmin(fabs(e2 - lth), fabs(e3 - lth)));
        if (m < tol) loc[4]++;
    }
    for (int i = 0; i < 5; ++i)
        if (loc[i]) atomicAdd(&cnt[i], loc[i]);
    __syncthreads();
    if (threadIdx.x < 5 && cnt[threadIdx.x]) atomicAdd(&stats[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}


// Fused FD Hessian + labels.  One workgroup owns a TY x TZ column of cells in
// (y,z) and a segment of x planes and marches along x.  LDS holds a ring of
// the last five phi planes (halo 4 in y,z) and the centre plane's G_y, G_z
// (halo 2); each thread keeps, for its positions of the haloed region
// (TY+4) x (TZ+4), the last five G_x, G_y, G_z values in registers.  The next
// phi plane is loaded into registers while the current one is processed.  G and
// T never touch HBM.  Composition order of the reference: G_a = D_a phi, then
// D_x G_x, D_x G_y, D_x G_z, D_y G_y, D_y G_z, D_z G_z
// (Miscelanious.cpp:1684-1700).  Labels from the inertia of T - lambda_th I
// (positive_eigs_f: float with a certified sign, else FP64); closed-form FP64
// eigenvalues only for the lambda output and for cells whose margin bound is
// below tol.
// Rare path of the fused kernel (kept out of line so its FP64 registers do not
// count against the marching loop): closed-form eigenvalues, optional lambda
// output, and the exact margin test.  Returns 1 for a marginal cell.
__device__ __attribute__((noinline)) unsigned eig_margin(float txx, float tyy, float tzz, float a10, float a20,
                                                         float a21, double lth, double tol, bool test,
                                                         float *lam, size_t cell, size_t n3) {
    double e1, e2, e3;
    eig3(txx, tyy, tzz, a10, a20, a21, e1, e2, e3);
    if (lam) {
        lam[cell] = (float)e1;
        lam[n3 + cell] = (float)e2;
        lam[2 * n3 + cell] = (float)e3;
    }
    return (test && fmin(fabs(e1 - lth), fmin(fabs(e2 - lth), fabs(e3 - lth))) < tol) ? 1u : 0u;
}

template <int P>
__global__ void __launch_bounds__(512) k_tweb_fused(const float *__restrict__ phi, int N, int TY, int TZ,
                                                    int segLen, float ih, double lth, double tol,
                                                    int8_t *__restrict__ labels, float *__restrict__ lam,
                                                    unsigned long long *__restrict__ stats) {
    extern __shared__ float sm[];
    const int HY = TY + 8, HZ = TZ + 8, GY = TY + 4, GZ = TZ + 4;
    const int HP = HY * HZ;  // one phi plane
    float *ring = sm;        // 5 planes
    float *sgy = ring + 5 * HP;
    float *sgz = sgy + GY * GZ;
    __shared__ unsigned scnt[5];
    const int tid = threadIdx.x, NT = blockDim.x;
    if (tid < 5) scnt[tid] = 0;
    const int ntz = N / TZ;
    const int ty0 = (blockIdx.x / ntz) * TY, tz0 = (blockIdx.x % ntz) * TZ;
    const int x0 = blockIdx.y * segLen, x1 = min(N, x0 + segLen);
    const int npos = GY * GZ;
    const size_t NN = (size_t)N, n3 = NN * NN * NN;
    const float lthf = (float)lth;
    float gx[P][5], gy[P][5], gz[P][5];
#pragma unroll
    for (int k = 0; k < P; ++k)
#pragma unroll
        for (int i = 0; i < 5; ++i) gx[k][i] = gy[k][i] = gz[k][i] = 0.f;
    unsigned loc[5] = {0, 0, 0, 0, 0};
    constexpr int PF = 4;  // phi values per thread per plane (HP <= PF*NT)
    float pf[PF];
    auto fetch = [&](int s) {
        const int sp = (s % N + N) % N;
        const float *plane = phi + (size_t)sp * NN * NN;
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int q = tid + k * NT;
            if (q < HP) {
                const int yy = q / HZ, zz = q - yy * HZ;
                const int y = (ty0 - 4 + yy + N) % N, z = (tz0 - 4 + zz + N) % N;
                pf[k] = plane[(size_t)y * NN + z];
            }
        }
    };
    fetch(x0 - 4);
    for (int s = x0 - 4; s < x1 + 4; ++s) {
        float *cur = ring + ((s - x0 + 8) % 5) * HP;  // slot of plane s
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int q = tid + k * NT;
            if (q < HP) cur[q] = pf[k];
        }
        if (s + 1 < x1 + 4) fetch(s + 1);  // in flight while plane s is processed
        __syncthreads();
        // planes s-4..s live in the ring; G_x(s-2) from the x-stencil, G_y and
        // G_z(s-2) from plane s-2's y/z neighbours
        const float *pm4 = ring + ((s - x0 + 4) % 5) * HP, *pm3 = ring + ((s - x0 + 5) % 5) * HP;
        const float *pm2 = ring + ((s - x0 + 6) % 5) * HP, *pm1 = ring + ((s - x0 + 7) % 5) * HP;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const int q = tid + k * NT;
            if (q < npos) {
                const int a = q / GZ, b = q - a * GZ;
                const int o = (a + 2) * HZ + (b + 2);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    gx[k][i] = gx[k][i + 1];
                    gy[k][i] = gy[k][i + 1];
                    gz[k][i] = gz[k][i + 1];
                }
                gx[k][4] = fd4f(pm4[o], pm3[o], pm1[o], cur[o], ih);
                gy[k][4] = fd4f(pm2[o - 2 * HZ], pm2[o - HZ], pm2[o + HZ], pm2[o + 2 * HZ], ih);
                gz[k][4] = fd4f(pm2[o - 2], pm2[o - 1], pm2[o + 1], pm2[o + 2], ih);
                sgy[q] = gy[k][2];  // centre plane c = s-4
                sgz[q] = gz[k][2];
            }
        }
        __syncthreads();
        const int c = s - 4;
        if (c < x0) continue;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const int q = tid + k * NT;
            if (q >= npos) continue;
            const int a = q / GZ, b = q - a * GZ;
            if (a < 2 || a >= TY + 2 || b < 2 || b >= TZ + 2) continue;
            const float txx = fd4f(gx[k][0], gx[k][1], gx[k][3], gx[k][4], ih);
            const float a10 = fd4f(gy[k][0], gy[k][1], gy[k][3], gy[k][4], ih);
            const float a20 = fd4f(gz[k][0], gz[k][1], gz[k][3], gz[k][4], ih);
            const float tyy = fd4f(sgy[q - 2 * GZ], sgy[q - GZ], sgy[q + GZ], sgy[q + 2 * GZ], ih);
            const float a21 = fd4f(sgz[q - 2 * GZ], sgz[q - GZ], sgz[q + GZ], sgz[q + 2 * GZ], ih);
            const float tzz = fd4f(sgz[q - 2], sgz[q - 1], sgz[q + 1], sgz[q + 2], ih);
            float det, det_err, fro2;
            const int v = positive_eigs_f(txx - lthf, tyy - lthf, tzz - lthf, a10, a20, a21, det, det_err, fro2);
            const int lab = 4 - v;
            const size_t cell = ((size_t)c * NN + (ty0 + a - 2)) * NN + (tz0 + b - 2);
            labels[cell] = (int8_t)lab;
            loc[lab - 1]++;
            // margin bound min|mu| >= |det B| / ||B||_F^2, taken conservatively
            // (|det| minus its rounding bound); the FP64 eigenvalues settle it
            const bool maybe_marginal = tol > 0.0 && (fabsf(det) - det_err) < (float)tol * fro2 * 1.0001f;
            if (lam || maybe_marginal)
                loc[4] += eig_margin(txx, tyy, tzz, a10, a20, a21, lth, tol, maybe_marginal, lam, cell, n3);
        }
    }
#pragma unroll
    for (int i = 0; i < 5; ++i)
        if (loc[i]) atomicAdd(&scnt[i], loc[i]);
    __syncthreads();
    if (tid < 5 && scnt[tid]) atomicAdd(&stats[tid], (unsigned long long)scnt[tid]);
}

__global__ void k_zero_u64(unsigned long long *a, int n) {
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0ull;
}

static unsigned grid_for(size_t n, int block, size_t cap) {
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}


static int pick_tile(int N, const int *cands, int nc) {
    for (int i = 0; i < nc; ++i)
        if (cands[i] <= N && N % cands[i] == 0) return cands[i];
    return 0;
}

// Fused T-web from phi: labels (and optional lambdas) plus the five counters.
cm_status tweb_fused_launch(cm_ctx *ctx, int N, double L, const float *phi, double lth, double tol, int8_t *labels,
                            float *lam, unsigned long long *stats) {
    static const int cz[] = {64, 32, 16, 8}, cy[] = {16, 8};
    int TZ = pick_tile(N, cz, 4), TY = pick_tile(N, cy, 2);
    if (ctx->tune.tz > 0 && N % ctx->tune.tz == 0 && ctx->tune.tz <= 64) TZ = ctx->tune.tz;
    if (ctx->tune.ty > 0 && N % ctx->tune.ty == 0 && ctx->tune.ty <= 32) TY = ctx->tune.ty;
    if (!TZ || !TY) return fail(ctx, CM_ERR_UNSUPPORTED, "tweb: N=%d has no tile decomposition", N);
    const int npos = (TY + 4) * (TZ + 4);
    const int NT = 512;
    const int P = (npos + NT - 1) / NT;
    int nseg = N >= 256 ? 4 : (N >= 64 ? 2 : 1);
    if (ctx->tune.tseg > 0 && ctx->tune.tseg <= N / 4) nseg = ctx->tune.tseg;
    const int segLen = (N + nseg - 1) / nseg;
    dim3 grid((unsigned)((N / TY) * (N / TZ)), (unsigned)nseg);
    const size_t lds = sizeof(float) * (5 * (size_t)(TY + 8) * (TZ + 8) + 2 * (size_t)npos);
    if ((size_t)(TY + 8) * (TZ + 8) > 4 * 512 || P > 3 || lds > ctx->lds_per_block)
        return fail(ctx, CM_ERR_UNSUPPORTED, "tweb: tile %dx%d too large", TY, TZ);
    const double ih = 1.0 / (12.0 * (L / (double)N));
    cm_status s;
    if ((s = launch(ctx, "tweb_zero", k_zero_u64, dim3(1), dim3(64), 0, stats, 5)) != CM_OK) return s;
    const float ihf = (float)ih;
    switch (P) {
    case 1: return launch(ctx, "tweb_fused", k_tweb_fused<1>, grid, dim3(NT), lds, phi, N, TY, TZ, segLen, ihf, lth, tol, labels, lam, stats);
    case 2: return launch(ctx, "tweb_fused", k_tweb_fused<2>, grid, dim3(NT), lds, phi, N, TY, TZ, segLen, ihf, lth, tol, labels, lam, stats);
    case 3: return launch(ctx, "tweb_fused", k_tweb_fused<3>, grid, dim3(NT), lds, phi, N, TY, TZ, segLen, ihf, lth, tol, labels, lam, stats);
    }
    return fail(ctx, CM_ERR_UNSUPPORTED, "tweb: tile needs %d positions per thread", P);
}

// phi = c2r(W(k)(-1/k^2) r2c(delta)*scale), float32.
cm_status poisson_impl(cm_ctx *ctx, const cm_tweb_opts *o, const float *delta, float *phi, double subtract) {
    const int N = o->N;
    const size_t nh = (size_t)N * N * (N / 2 + 1);
    void *dk;
    cm_status s;
    if ((s = scratch(ctx, "tweb.dk", sizeof(float2) * nh, &dk)) != CM_OK) return s;
    if ((s = fft_r2c_impl(ctx, N, CM_F32, delta, dk, subtract)) != CM_OK) return s;
    const double kf = 2.0 * M_PI / o->L;
    const double R2h = o->R_smooth > 0 ? 0.5 * o->R_smooth * o->R_smooth : 0.0;
    const double scale = o->delta_scale == 0.0 ? 1.0 : o->delta_scale;
    if ((s = launch(ctx, "tweb_kmul", k_kmul<float>, dim3(grid_for(nh, 256, 16384)), dim3(256), 0, (float2 *)dk, N,
                    kf, R2h, scale)) != CM_OK)
        return s;
    const double n3 = (double)N * N * N;
    return fft_c2r_impl(ctx, N, CM_F32, dk, phi, 1.0 / n3);
}

cm_status tweb_from_phi_impl(cm_ctx *ctx, const cm_tweb_opts *o, const float *phi, int8_t *labels, float *lam,
                             cm_tweb_stats *st, bool plain) {
    const int N = o->N;
    const size_t n3 = (size_t)N * N * N;
    void *stats;
    cm_status s;
    if ((s = scratch(ctx, "tweb.stats", sizeof(unsigned long long) * 5, &stats)) != CM_OK) return s;
    const double ih = 1.0 / (12.0 * (o->L / (double)N));
    if (plain) {
        void *g;
        if ((s = scratch(ctx, "tweb.grad", sizeof(float) * 3 * n3, &g)) != CM_OK) return s;
        float *gx = (float *)g, *gy = gx + n3, *gz = gy + n3;
        unsigned gr = grid_for(n3, 256, 65536);
        if ((s = launch(ctx, "tweb_grad", k_grad3, dim3(gr), dim3(256), 0, phi, N, ih, gx, gy, gz)) != CM_OK) return s;
        if ((s = launch(ctx, "tweb_zero", k_zero_u64, dim3(1), dim3(64), 0, (unsigned long long *)stats, 5)) != CM_OK)
            return s;
        if ((s = launch(ctx, "tweb_hess_eig", k_hess_eig, dim3(gr), dim3(256), 0, (const float *)gx,
                        (const float *)gy, (const float *)gz, N, ih, o->lambda_th, o->margin_tol, labels, lam,
                        (unsigned long long *)stats)) != CM_OK)
            return s;
    } else {
        if ((s = tweb_fused_launch(ctx, N, o->L, phi, o->lambda_th, o->margin_tol, labels, lam,
                                   (unsigned long long *)stats)) != CM_OK)
            return s;
    }
    if (st) {
        unsigned long long h[5];
        CM_HIP(ctx, hipMemcpyAsync(h, stats, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
        CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < 4; ++i) st->counts[i] = (int64_t)h[i];
        st->n_marginal = (int64_t)h[4];
    }
    return CM_OK;
}

static cm_status check_opts(cm_ctx *ctx, const cm_tweb_opts *o) {
    if (!o) return fail(ctx, CM_ERR_INVALID, "tweb: null options");
    if (!fft_size_ok(o->N)) return fail(ctx, CM_ERR_INVALID, "tweb: N=%d not supported by the FFT", o->N);
    if (!(o->L > 0)) return fail(ctx, CM_ERR_INVALID, "tweb: L must be > 0");
    return CM_OK;
}

}  // namespace cm

using namespace cm;

extern "C" {

cm_status cm_poisson(cm_ctx *ctx, const cm_tweb_opts *o, const float *delta, float *phi) {
    if (!ctx) return CM_ERR_INVALID;
    cm_status s = check_opts(ctx, o);
    if (s != CM_OK) return s;
    if (!delta || !phi) return fail(ctx, CM_ERR_INVALID, "cm_poisson: null pointer");
    return poisson_impl(ctx, o, delta, phi, 0.0);
}

cm_status cm_tweb_from_phi(cm_ctx *ctx, const cm_tweb_opts *o, const float *phi, int8_t *labels, float *lambdas,
                           cm_tweb_stats *stats) {
    if (!ctx) return CM_ERR_INVALID;
    cm_status s = check_opts(ctx, o);
    if (s != CM_OK) return s;
    if (!phi || !labels) return fail(ctx, CM_ERR_INVALID, "cm_tweb_from_phi: null pointer");
    return tweb_from_phi_impl(ctx, o, phi, labels, lambdas, stats, false);
}

cm_status cm_tweb(cm_ctx *ctx, const cm_tweb_opts *o, const float *delta, int8_t *labels, float *lambdas,
                  cm_tweb_stats *stats) {
    if (!ctx) return CM_ERR_INVALID;
    cm_status s = check_opts(ctx, o);
    if (s != CM_OK) return s;
    if (!delta || !labels) return fail(ctx, CM_ERR_INVALID, "cm_tweb: null pointer");
    void *phi;
    const size_t n3 = (size_t)o->N * o->N * o->N;
    if ((s = scratch(ctx, "tweb.phi", sizeof(float) * n3, &phi)) != CM_OK) return s;
    if ((s = poisson_impl(ctx, o, delta, (float *)phi, 0.0)) != CM_OK) return s;
    return tweb_from_phi_impl(ctx, o, (const float *)phi, labels, lambdas, stats, false);
}

}  // extern "C"
