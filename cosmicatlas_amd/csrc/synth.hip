// Synthetic mocks on the GPU (gfx950): seeded Gaussian random field with fixed
// amplitude, Zel'dovich displacement, and uniform randoms.
//
// Field: |delta_k| = sqrt(A k^ns NL^6 / L^3) on the r2c half spectrum with a
// uniform phase per mode (the create_GARFIELD_FIXED_AMP semantics of
// Source/NumericalMethods.cpp:2147-2189).  The phase comes from the
// counter-based generator Philox4x32-10 (Salmon et al., SC11) keyed by the seed
// and counted by the mode index, so every mode's draw is independent of the
// launch shape (the reference's shared-generator OpenMP loop is a race).  The
// kz = 0 and kz = NL/2 planes are Hermitian: of a mode and its partner (-i,-j)
// the smaller linear index draws, the other takes the conjugate;
// self-conjugate modes keep amp cos(theta).
// Displacement: Psi_a(k) = i k_a delta(k)/k^2 (Nyquist component of k_a set to
// zero), three c2r transforms, x = (i + 1/2) d + Psi(q) wrapped into [0, L)
// (LPT::get_displacement, Source/LPT.cpp:285, at first order).
#include <cmath>

#include "cm_internal.hpp"

namespace cm {

struct Philox4 {
    unsigned x[4];
};

// Philox4x32 with 10 rounds: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
// increments 0x9E3779B9 / 0xBB67AE85 (the published constants of the method).
__device__ __forceinline__ Philox4 philox4x32_10(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0,
                                                 unsigned k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
        const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
        const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
        const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
        const unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return Philox4{{c0, c1, c2, c3}};
}

// Half spectra of the three displacement components.
__global__ void __launch_bounds__(256) k_zel_spectra(int NL, double L, double A, double ns, unsigned long long seed,
                                                     float2 *__restrict__ sx, float2 *__restrict__ sy,
                                                     float2 *__restrict__ sz) {
    const int M1 = NL / 2 + 1, h = NL / 2;
    const size_t total = (size_t)NL * NL * M1;
    const double kf = 2.0 * M_PI / L;
    const double norm = (double)NL * NL * NL / sqrt(L * L * L);  // sqrt(NL^6 / V)
    const unsigned k0 = (unsigned)seed, k1 = (unsigned)(seed >> 32);
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(g % M1);
        const size_t r = g / M1;
        const int iy = (int)(r % NL), ix = (int)(r / NL);
        const double kx = kf * (ix <= h ? ix : ix - NL), ky = kf * (iy <= h ? iy : iy - NL), kz = kf * iz;
        const double k2 = kx * kx + ky * ky + kz * kz;
        float2 ox = {0.f, 0.f}, oy = {0.f, 0.f}, oz = {0.f, 0.f};
        if (k2 > 0.0) {
            const bool plane = iz == 0 || iz == h;
            const size_t partner = ((size_t)((NL - ix) % NL) * NL + (NL - iy) % NL) * M1 + iz;
            const size_t canon = plane && partner < g ? partner : g;
            const Philox4 w = philox4x32_10((unsigned)canon, (unsigned)(canon >> 32), 0u, 0u, k0, k1);
            const double theta = ((double)w.x[0] + 0.5) * (2.0 * M_PI / 4294967296.0);
            const double amp = sqrt(A * pow(sqrt(k2), ns)) * norm;
            double s, c;
            sincos(theta, &s, &c);
            double re = amp * c, im = amp * s;
            if (plane && partner < g) im = -im;
            if (plane && partner == g) im = 0.0;
            // Psi_a = i k_a delta / k^2, Nyquist component of k_a zeroed
            const double ikx = ix == h ? 0.0 : kx / k2, iky = iy == h ? 0.0 : ky / k2, ikz = iz == h ? 0.0 : kz / k2;
            ox = float2{(float)(-im * ikx), (float)(re * ikx)};
            oy = float2{(float)(-im * iky), (float)(re * iky)};
            oz = float2{(float)(-im * ikz), (float)(re * ikz)};
        }
        sx[g] = ox;
        sy[g] = oy;
        sz[g] = oz;
    }
}

__global__ void __launch_bounds__(256) k_zel_displace(int NL, double L, const float *__restrict__ px,
                                                      const float *__restrict__ py, const float *__restrict__ pz,
                                                      float *__restrict__ pos) {
    const size_t n = (size_t)NL * NL * NL;
    const double d = L / NL;
    const float Lf = (float)L;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(p % NL);
        const size_t r = p / NL;
        const int j = (int)(r % NL), i = (int)(r / NL);
        const double q[3] = {(i + 0.5) * d, (j + 0.5) * d, (k + 0.5) * d};
        const float ps[3] = {px[p], py[p], pz[p]};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double x = fmod(q[a] + (double)ps[a], L);
            if (x < 0.0) x += L;
            float xf = (float)x;
            if (xf >= Lf) xf = 0.0f;
            pos[3 * p + a] = xf;
        }
    }
}

__global__ void __launch_bounds__(256) k_uniform(size_t n, double L, unsigned long long seed, float *__restrict__ pos) {
    const float Lf = (float)L;
    const unsigned k0 = (unsigned)seed, k1 = (unsigned)(seed >> 32);
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const Philox4 w = philox4x32_10((unsigned)p, (unsigned)(p >> 32), 1u, 0u, k0, k1);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float x = (float)(((double)w.x[a] + 0.5) * (1.0 / 4294967296.0) * L);
            if (x >= Lf) x = 0.0f;
            pos[3 * p + a] = x;
        }
    }
}

static unsigned grid_of(size_t n) {
    size_t g = (n + 255) / 256;
    return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

}  // namespace cm

using namespace cm;

extern "C" {

cm_status cm_zeldovich(cm_ctx *ctx, const cm_zeldovich_opts *o, float *pos) {
    if (!ctx) return CM_ERR_INVALID;
    if (!o || !pos) return fail(ctx, CM_ERR_INVALID, "cm_zeldovich: null pointer");
    const int NL = o->NL;
    if (!fft_size_ok(NL)) return fail(ctx, CM_ERR_INVALID, "cm_zeldovich: NL=%d not supported", NL);
    if (!(o->L > 0) || !(o->A >= 0)) return fail(ctx, CM_ERR_INVALID, "cm_zeldovich: L, A");
    const size_t nh = (size_t)NL * NL * (NL / 2 + 1), n3 = (size_t)NL * NL * NL;
    void *spec, *psi;
    cm_status s;
    if ((s = scratch(ctx, "zel.spec", sizeof(float2) * 3 * nh, &spec)) != CM_OK) return s;
    if ((s = scratch(ctx, "zel.psi", sizeof(float) * 3 * n3, &psi)) != CM_OK) return s;
    auto *sp = (float2 *)spec;
    auto *ps = (float *)psi;
    if ((s = launch(ctx, "zel_spectra", k_zel_spectra, dim3(grid_of(nh)), dim3(256), 0, NL, o->L, o->A, o->ns,
                    (unsigned long long)o->seed, sp, sp + nh, sp + 2 * nh)) != CM_OK)
        return s;
    const double scale = 1.0 / (double)n3;  // irfftn convention
    for (int a = 0; a < 3; ++a)
        if ((s = fft_c2r_impl(ctx, NL, CM_F32, sp + a * nh, ps + a * n3, scale)) != CM_OK) return s;
    return launch(ctx, "zel_displace", k_zel_displace, dim3(grid_of(n3)), dim3(256), 0, NL, o->L,
                  (const float *)ps, (const float *)(ps + n3), (const float *)(ps + 2 * n3), pos);
}

cm_status cm_uniform(cm_ctx *ctx, size_t n, double L, uint64_t seed, float *pos) {
    if (!ctx) return CM_ERR_INVALID;
    if (!pos && n) return fail(ctx, CM_ERR_INVALID, "cm_uniform: null pointer");
    if (!(L > 0)) return fail(ctx, CM_ERR_INVALID, "cm_uniform: L");
    if (!n) return CM_OK;
    return launch(ctx, "uniform", k_uniform, dim3(grid_of(n)), dim3(256), 0, n, L, (unsigned long long)seed, pos);
}

}  // extern "C"
