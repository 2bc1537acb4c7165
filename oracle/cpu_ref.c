/*
 * cpu_ref.c -- OpenMP restatement of the reference's CPU path, used only as the
 * CPU baseline that bench.py times (test infrastructure, never the product).
 *
 * It keeps the reference's parallel structure and precision (real_prec=float):
 *   MAS     getDensity_TSC: thread-private N^3 grids + omp critical reduction
 *           (Source/Miscelanious.cpp:693-815)
 *   FFT     3-D r2c/c2r in float, lines distributed with OpenMP (stands in for
 *           FFTW with FFTW_ESTIMATE plans, Source/NumericalMethods.cpp:1718-1856;
 *           FFTW is not installed here)
 *   shells  power_spectrum_fkp with omp atomic float accumulation
 *           (Source/FftwFunctions.cpp:958-1657)
 *   T-web   PoissonSolver + 12 gradfindif passes + per-cell eigen solve with a
 *           per-cell workspace allocation like gsl_eigen_symmv_alloc/free
 *           (Source/Miscelanious.cpp:1189-1274,1668-1769; Cwclass.cpp:874-894)
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { float re, im; } fc;

/* in-place radix-2 FFT of length n (power of two), sign -1/+1, float */
static void fft1(fc *a, int n, int sign, const fc *tw /* n/2 entries exp(-2pi i k/n) */)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { fc t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int step = n / len;
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                fc w = tw[k * step];
                if (sign > 0) w.im = -w.im;
                fc u = a[i + k], v = a[i + k + len / 2];
                fc t = {v.re * w.re - v.im * w.im, v.re * w.im + v.im * w.re};
                a[i + k].re = u.re + t.re; a[i + k].im = u.im + t.im;
                a[i + k + len / 2].re = u.re - t.re; a[i + k + len / 2].im = u.im - t.im;
            }
    }
}

static fc *twiddles(int n)
{
    fc *t = (fc *)malloc(sizeof(fc) * (size_t)(n / 2 + 1));
    for (int k = 0; k <= n / 2; ++k) {
        double a = -2.0 * M_PI * k / n;
        t[k].re = (float)cos(a);
        t[k].im = (float)sin(a);
    }
    return t;
}

/* r2c: along z each real row is transformed as a complex row (imag 0), then y, x */
static void r2c3(int N, const float *in, fc *out)
{
    const size_t NN = (size_t)N, nh = NN / 2 + 1;
    fc *tw = twiddles(N);
#pragma omp parallel
    {
        fc *line = (fc *)malloc(sizeof(fc) * NN);
#pragma omp for collapse(2) schedule(static)
        for (size_t i = 0; i < NN; ++i)
            for (size_t j = 0; j < NN; ++j) {
                const float *r = in + (i * NN + j) * NN;
                for (size_t k = 0; k < NN; ++k) { line[k].re = r[k]; line[k].im = 0.f; }
                fft1(line, N, -1, tw);
                memcpy(out + (i * NN + j) * nh, line, sizeof(fc) * nh);
            }
#pragma omp for collapse(2) schedule(static)
        for (size_t i = 0; i < NN; ++i)
            for (size_t k = 0; k < nh; ++k) {
                for (size_t j = 0; j < NN; ++j) line[j] = out[(i * NN + j) * nh + k];
                fft1(line, N, -1, tw);
                for (size_t j = 0; j < NN; ++j) out[(i * NN + j) * nh + k] = line[j];
            }
#pragma omp for collapse(2) schedule(static)
        for (size_t j = 0; j < NN; ++j)
            for (size_t k = 0; k < nh; ++k) {
                for (size_t i = 0; i < NN; ++i) line[i] = out[(i * NN + j) * nh + k];
                fft1(line, N, -1, tw);
                for (size_t i = 0; i < NN; ++i) out[(i * NN + j) * nh + k] = line[i];
            }
        free(line);
    }
    free(tw);
}

/* c2r / N^3 of a Hermitian half spectrum (destroys `in`) */
static void c2r3(int N, fc *in, float *out)
{
    const size_t NN = (size_t)N, nh = NN / 2 + 1;
    fc *tw = twiddles(N);
    const float inv = 1.0f / ((float)N * N * N);
#pragma omp parallel
    {
        fc *line = (fc *)malloc(sizeof(fc) * NN);
#pragma omp for collapse(2) schedule(static)
        for (size_t j = 0; j < NN; ++j)
            for (size_t k = 0; k < nh; ++k) {
                for (size_t i = 0; i < NN; ++i) line[i] = in[(i * NN + j) * nh + k];
                fft1(line, N, +1, tw);
                for (size_t i = 0; i < NN; ++i) in[(i * NN + j) * nh + k] = line[i];
            }
#pragma omp for collapse(2) schedule(static)
        for (size_t i = 0; i < NN; ++i)
            for (size_t k = 0; k < nh; ++k) {
                for (size_t j = 0; j < NN; ++j) line[j] = in[(i * NN + j) * nh + k];
                fft1(line, N, +1, tw);
                for (size_t j = 0; j < NN; ++j) in[(i * NN + j) * nh + k] = line[j];
            }
#pragma omp for collapse(2) schedule(static)
        for (size_t i = 0; i < NN; ++i)
            for (size_t j = 0; j < NN; ++j) {
                const fc *h = in + (i * NN + j) * nh;
                for (size_t k = 0; k < nh; ++k) line[k] = h[k];
                for (size_t k = nh; k < NN; ++k) { line[k] = h[NN - k]; line[k].im = -line[k].im; }
                fft1(line, N, +1, tw);
                float *r = out + (i * NN + j) * NN;
                for (size_t k = 0; k < NN; ++k) r[k] = line[k].re * inv;
            }
        free(line);
    }
    free(tw);
}

static double now(void) { return omp_get_wtime(); }

/* getDensity_TSC structure: private grids + critical sum */
static void mas_tsc(int N, float L, const float *pos, size_t n, float *delta)
{
    const size_t NN = (size_t)N, n3 = NN * NN * NN;
    const float d = L / N;
    memset(delta, 0, sizeof(float) * n3);
#pragma omp parallel
    {
        float *priv = (float *)calloc(n3, sizeof(float));
#pragma omp for schedule(static)
        for (size_t p = 0; p < n; ++p) {
            float x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
            int i = (int)floorf(x / d), j = (int)floorf(y / d), k = (int)floorf(z / d);
            float dx = x / d - (i + 0.5f), dy = y / d - (j + 0.5f), dz = z / d - (k + 0.5f);
            float hx[3] = {0.5f * (0.5f - dx) * (0.5f - dx), 0.75f - dx * dx, 0.5f * (0.5f + dx) * (0.5f + dx)};
            float hy[3] = {0.5f * (0.5f - dy) * (0.5f - dy), 0.75f - dy * dy, 0.5f * (0.5f + dy) * (0.5f + dy)};
            float hz[3] = {0.5f * (0.5f - dz) * (0.5f - dz), 0.75f - dz * dz, 0.5f * (0.5f + dz) * (0.5f + dz)};
            for (int a = 0; a < 3; ++a) {
                size_t ii = (size_t)((i + a - 1 + N) % N);
                for (int b = 0; b < 3; ++b) {
                    size_t jj = (size_t)((j + b - 1 + N) % N);
                    for (int c = 0; c < 3; ++c) {
                        size_t kk = (size_t)((k + c - 1 + N) % N);
                        priv[(ii * NN + jj) * NN + kk] += hx[a] * hy[b] * hz[c];
                    }
                }
            }
        }
#pragma omp critical
        for (size_t q = 0; q < n3; ++q) delta[q] += priv[q];
        free(priv);
    }
}

static void shells(int N, float L, const fc *dk, int ns, float kmin, float dkb, float SN, float normal,
                   float *pk, int *modes)
{
    const float kf = (float)(2.0 * M_PI / L);
    const float rdk = 1.0f / dkb;
    memset(pk, 0, sizeof(float) * (size_t)ns);
    memset(modes, 0, sizeof(int) * (size_t)ns);
    const size_t NN = (size_t)N, nh = NN / 2 + 1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N / 2; ++i)
        for (int j = 0; j < N / 2; ++j)
            for (int k = 0; k <= N / 2; ++k) {
                float kv = sqrtf((float)(i * i) * kf * kf + (float)(j * j) * kf * kf + (float)(k * k) * kf * kf);
                int s = (int)floorf((kv - kmin) * rdk);
                if (s < 0 || s >= ns) continue;
                float sx = i ? sinf((float)M_PI * i / N) / ((float)M_PI * i / N) : 1.f;
                float sy = j ? sinf((float)M_PI * j / N) / ((float)M_PI * j / N) : 1.f;
                float sz = k ? sinf((float)M_PI * k / N) / ((float)M_PI * k / N) : 1.f;
                float corr = powf(sx * sy * sz, 3.f);
                float ic2 = 1.f / (corr * corr);
                float a = sinf((float)M_PI * i / N), b = sinf((float)M_PI * j / N), c = sinf((float)M_PI * k / N);
                float sq = (1 - a * a + 2.f / 15 * a * a * a * a) * (1 - b * b + 2.f / 15 * b * b * b * b) *
                           (1 - c * c + 2.f / 15 * c * c * c * c);
                size_t lps[4];
                int nl = 0;
                size_t lp = ((size_t)i * NN + j) * nh + k;
                if (lp) lps[nl++] = lp;
                if (j > 0 && k > 0) lps[nl++] = ((size_t)i * NN + (NN - j)) * nh + k;
                if (i > 0 && (j > 0 || k > 0)) lps[nl++] = ((NN - i) * NN + j) * nh + k;
                if (i > 0 && j > 0 && k > 0) lps[nl++] = ((NN - i) * NN + (NN - j)) * nh + k;
                for (int q = 0; q < nl; ++q) {
                    fc v = dk[lps[q]];
                    float P = (v.re * v.re + v.im * v.im - SN * sq * normal) * ic2;
#pragma omp atomic
                    pk[s] += P;
#pragma omp atomic
                    modes[s]++;
                }
            }
    for (int s = 0; s < ns; ++s) pk[s] = modes[s] ? pk[s] / (modes[s] * normal) : 0.f;
}

static void grad(int N, float L, const float *in, float *out, int dim)
{
    const size_t NN = (size_t)N;
    const float ih = 1.0f / (12.0f * (L / N));
#pragma omp parallel for collapse(2) schedule(static)
    for (int x = 0; x < N; ++x)
        for (int y = 0; y < N; ++y)
            for (int z = 0; z < N; ++z) {
                int c[3] = {x, y, z}, a = dim;
                int r[3], l[3], rr[3], ll[3];
                memcpy(r, c, sizeof c); memcpy(l, c, sizeof c); memcpy(rr, c, sizeof c); memcpy(ll, c, sizeof c);
                r[a] = (c[a] + 1) % N; rr[a] = (c[a] + 2) % N; l[a] = (c[a] - 1 + N) % N; ll[a] = (c[a] - 2 + N) % N;
#define ID(v) (((size_t)(v)[0] * NN + (v)[1]) * NN + (v)[2])
                out[ID(c)] = (float)((-(double)in[ID(rr)] + 8.0 * in[ID(r)] - 8.0 * in[ID(l)] + in[ID(ll)]) * ih);
#undef ID
            }
}

static void tweb(int N, float L, const float *delta, float R, float lth, int8_t *labels, double *t_fft)
{
    const size_t NN = (size_t)N, n3 = NN * NN * NN, nh = NN * NN * (NN / 2 + 1);
    fc *dk = (fc *)malloc(sizeof(fc) * nh);
    float *phi = (float *)malloc(sizeof(float) * n3);
    double t0 = now();
    r2c3(N, delta, dk);
    const double kf = 2.0 * M_PI / L;
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < nh; ++q) {
        int k = (int)(q % (NN / 2 + 1)), j = (int)((q / (NN / 2 + 1)) % NN), i = (int)(q / ((NN / 2 + 1) * NN));
        double kx = kf * (i <= N / 2 ? i : i - N), ky = kf * (j <= N / 2 ? j : j - N), kz = kf * k;
        double k2 = kx * kx + ky * ky + kz * kz;
        float f = k2 > 0 ? (float)(-exp(-0.5 * k2 * R * R) / k2) : 0.f;
        dk[q].re *= f;
        dk[q].im *= f;
    }
    c2r3(N, dk, phi);
    *t_fft = now() - t0;
    free(dk);
    float *g = (float *)malloc(sizeof(float) * n3);
    float *T[9];
    for (int a = 0; a < 9; ++a) T[a] = (float *)malloc(sizeof(float) * n3);
    for (int a = 0; a < 3; ++a) {
        grad(N, L, phi, g, a);
        for (int b = 0; b < 3; ++b) grad(N, L, g, T[a * 3 + b], b);
    }
    int vff[4] = {0, 0, 0, 0};
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n3; ++c) {
        double *m = (double *)malloc(9 * sizeof(double)); /* per-cell workspace, like gsl_*_alloc */
        m[0] = T[0][c]; m[1] = T[3][c]; m[2] = T[6][c];
        m[3] = T[1][c]; m[4] = T[4][c]; m[5] = T[7][c];
        m[6] = T[2][c]; m[7] = T[5][c]; m[8] = T[8][c];
        double ev[3];
        or_eigen3(m, ev);
        free(m);
        int v = (ev[0] > lth) + (ev[1] > lth) + (ev[2] > lth);
        labels[c] = (int8_t)(4 - v);
#pragma omp atomic
        vff[3 - v]++;
    }
    for (int a = 0; a < 9; ++a) free(T[a]);
    free(g);
    free(phi);
}

/* Times each stage once (after one untimed warm-up of the deposit) at N^3. */
void cr_time_pipeline(int N, float L, const float *pos, size_t n, int ns, float ndel_kf, float R, float lth,
                      int threads, double out[5])
{
    if (threads > 0) omp_set_num_threads(threads);
    const size_t NN = (size_t)N, n3 = NN * NN * NN, nh = NN * NN * (NN / 2 + 1);
    float *mesh = (float *)malloc(sizeof(float) * n3);
    fc *dk = (fc *)malloc(sizeof(fc) * nh);
    float *pk = (float *)malloc(sizeof(float) * (size_t)ns);
    int *modes = (int *)malloc(sizeof(int) * (size_t)ns);
    int8_t *labels = (int8_t *)malloc(n3);
    double t0 = now();
    mas_tsc(N, L, pos, n, mesh);
    double t1 = now();
    double nobj = 0;
    for (size_t q = 0; q < n3; ++q) nobj += mesh[q];
    const float mean = (float)(nobj / (double)n3);
    for (size_t q = 0; q < n3; ++q) mesh[q] -= mean;
    r2c3(N, mesh, dk);
    double t2 = now();
    const float V = L * L * L;
    shells(N, L, dk, ns, 0.f, ndel_kf, V / (float)nobj, (float)(nobj * nobj / V), pk, modes);
    double t3 = now();
    for (size_t q = 0; q < n3; ++q) mesh[q] = mesh[q] / mean;
    double tf = 0;
    tweb(N, L, mesh, R, lth, labels, &tf);
    double t4 = now();
    out[0] = t1 - t0;
    out[1] = t2 - t1;
    out[2] = t3 - t2;
    out[3] = t4 - t3;
    out[4] = tf;
    free(mesh); free(dk); free(pk); free(modes); free(labels);
}
