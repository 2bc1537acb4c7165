/*
 * oracle.c -- FP64 CPU restatement of the COSMICATLAS mesh-measurement path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.  It is
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg and
 * by nothing in the product (cosmicatlas_amd/).
 *
 * Built with -ffp-contract=off so the float32 emulations (periodic wrap of
 * float32 positions, the reference's float32 shell rule) round exactly like the
 * reference's scalar code.
 *
 * Pins (tests/test_oracle.py): numpy.fft for the FFT, brute-force DFT, Parseval
 * and Hermitian symmetry; closed-form MAS stencils and mass conservation; exact
 * per-shell mode counts; closed-form single-mode P(k); numpy.linalg.eigvalsh and
 * plane-wave tidal tensors for the T-web.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* MAS kernels: Source/Miscelanious.cpp:1277-1321                            */
/* ------------------------------------------------------------------------- */
double or_mas_kernel(int scheme, double x)
{
    double y = fabs(x);
    switch (scheme) {
    case OR_NGP: return y < 0.5 ? 1.0 : 0.0;
    case OR_CIC: return y < 1.0 ? 1.0 - y : 0.0;
    case OR_TSC:
        if (y < 0.5) return 0.75 - y * y;
        if (y < 1.5) { double r = 1.5 - y; return 0.5 * r * r; }
        return 0.0;
    case OR_PCS:
        if (y < 1.0) return (4.0 - 6.0 * y * y + 3.0 * y * y * y) / 6.0;
        if (y < 2.0) { double r = 2.0 - y; return r * r * r / 6.0; }
        return 0.0;
    }
    return 0.0;
}

/* Periodic wrap with the reference's while-loops (Miscelanious.cpp:2608-2613),
 * performed in the position dtype. */
static double wrap_f64(double x, double L)
{
    while (x < 0) x += L;
    while (x >= L) x -= L;
    return x;
}
static float wrap_f32(float x, float L)
{
    while (x < 0) x += L;
    while (x >= L) x -= L;
    return x;
}

/* get_bin (Miscelanious.cpp:1447-1457): floor in double, ==N -> N-1. */
static long get_cell(double x, double d, int N)
{
    long c = (long)floor(x / d);
    if (c >= N) c = N - 1;
    if (c < 0) c = 0;
    return c;
}

size_t or_mas(int N, double L, const double xmin[3], const void *pos, int pos_f32,
              int layout_soa, size_t n, const void *w, int scheme,
              double interlace_shift, double *mesh)
{
    const double d = L / (double)N;
    const size_t NN = (size_t)N;
    size_t deposited = 0;
    for (size_t p = 0; p < n; ++p) {
        double xyz[3];
        double wt = 1.0;
        for (int a = 0; a < 3; ++a) {
            size_t off = layout_soa ? (size_t)a * n + p : 3 * p + (size_t)a;
            if (pos_f32) {
                float v = ((const float *)pos)[off];
                /* x -= (Xoffset - L/2) evaluated in double, stored to float */
                float s = (float)((double)v - xmin[a]);
                xyz[a] = (double)wrap_f32(s, (float)L);
            } else {
                double v = ((const double *)pos)[off];
                xyz[a] = wrap_f64(v - xmin[a], L);
            }
            if (interlace_shift != 0.0) {
                double t = xyz[a] + interlace_shift * d;
                if (t >= L) t -= L;
                xyz[a] = t;
            }
        }
        if (w) wt = pos_f32 ? (double)((const float *)w)[p] : ((const double *)w)[p];

        long c[3];
        for (int a = 0; a < 3; ++a) c[a] = get_cell(xyz[a], d, N);

        if (scheme == OR_NGP) {
            size_t idx = (size_t)c[2] + NN * ((size_t)c[1] + NN * (size_t)c[0]);
            mesh[idx] += wt;
            ++deposited;
            continue;
        }
        /* CIC/TSC: nodes c-1, c, c+1 at (c+1/2)d (Miscelanious.cpp:2656-2691,
         * 2720-2756); PCS: nodes c-2..c+2 (Miscelanious.cpp:2771-2868). */
        int lo = (scheme == OR_PCS) ? -2 : -1, hi = (scheme == OR_PCS) ? 2 : 1;
        double W[3][5];
        long I[3][5];
        for (int a = 0; a < 3; ++a) {
            for (int m = lo; m <= hi; ++m) {
                double node = d * ((double)(c[a] + m) + 0.5);
                W[a][m - lo] = or_mas_kernel(scheme, (node - xyz[a]) / d);
                long ii = c[a] + m;
                if (ii < 0) ii += N;
                if (ii >= N) ii -= N;
                I[a][m - lo] = ii;
            }
        }
        int nm = hi - lo + 1;
        for (int a = 0; a < nm; ++a)
            for (int b = 0; b < nm; ++b)
                for (int e = 0; e < nm; ++e) {
                    size_t idx = (size_t)I[2][e] + NN * ((size_t)I[1][b] + NN * (size_t)I[0][a]);
                    mesh[idx] += wt * W[0][a] * W[1][b] * W[2][e];
                }
        ++deposited;
    }
    return deposited;
}

/* ------------------------------------------------------------------------- */
/* FFT: own mixed-radix recursive DIT FFT in FP64                             */
/* ------------------------------------------------------------------------- */
typedef struct { double re, im; } cplx;

static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* tw[j] = exp(sign*2*pi*i*j/ntop); sub-length n uses tw[(e*ntop/n) % ntop]. */
static void fft_rec(const cplx *in, size_t istride, cplx *out, int n, int ntop,
                    const cplx *tw, cplx *tmp)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n);
    int m = n / p;
    for (int r = 0; r < p; ++r)
        fft_rec(in + (size_t)r * istride, istride * (size_t)p, out + (size_t)r * m, m, ntop, tw, tmp);
    int step = ntop / n;
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) tmp[r] = out[(size_t)r * m + k];
        for (int q = 0; q < p; ++q) {
            double sr = 0, si = 0;
            int kk = k + q * m;
            for (int r = 0; r < p; ++r) {
                long e = ((long)r * kk) % n;
                cplx w = tw[(size_t)e * step];
                sr += tmp[r].re * w.re - tmp[r].im * w.im;
                si += tmp[r].re * w.im + tmp[r].im * w.re;
            }
            out[(size_t)q * m + k].re = sr;
            out[(size_t)q * m + k].im = si;
        }
    }
}

static cplx *make_twiddles(int n, int sign)
{
    cplx *tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < n; ++j) {
        /* exact angle reduction keeps the table accurate for every j */
        double a = 2.0 * M_PI * (double)j / (double)n;
        tw[j].re = cos(a);
        tw[j].im = sign * sin(a);
    }
    return tw;
}

void or_fft_1d(int n, const double *in, double *out, int sign)
{
    cplx *tw = make_twiddles(n, sign);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    fft_rec((const cplx *)in, 1, (cplx *)out, n, n, tw, tmp);
    free(tmp);
    free(tw);
}

/* In-place 1-D transforms along one axis of a full N^3 complex array. */
static void fft_axis(cplx *a, int N, int axis, int sign)
{
    cplx *tw = make_twiddles(N, sign);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    cplx *res = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    size_t NN = (size_t)N;
    size_t stride = axis == 0 ? NN * NN : (axis == 1 ? NN : 1);
    for (size_t u = 0; u < NN; ++u)
        for (size_t v = 0; v < NN; ++v) {
            size_t base;
            if (axis == 0) base = u * NN + v;          /* (., u, v) */
            else if (axis == 1) base = u * NN * NN + v; /* (u, ., v) */
            else base = (u * NN + v) * NN;              /* (u, v, .) */
            for (size_t t = 0; t < NN; ++t) line[t] = a[base + t * stride];
            fft_rec(line, 1, res, N, N, tw, tmp);
            for (size_t t = 0; t < NN; ++t) a[base + t * stride] = res[t];
        }
    free(res);
    free(line);
    free(tmp);
    free(tw);
}

void or_fft_r2c(int N, const double *in, double *out)
{
    size_t NN = (size_t)N, n3 = NN * NN * NN, nh = NN / 2 + 1;
    cplx *a = (cplx *)malloc(sizeof(cplx) * n3);
    for (size_t i = 0; i < n3; ++i) { a[i].re = in[i]; a[i].im = 0; }
    for (int ax = 0; ax < 3; ++ax) fft_axis(a, N, ax, -1);
    cplx *o = (cplx *)out;
    for (size_t i = 0; i < NN; ++i)
        for (size_t j = 0; j < NN; ++j)
            for (size_t k = 0; k < nh; ++k) o[(i * NN + j) * nh + k] = a[(i * NN + j) * NN + k];
    free(a);
}

void or_fft_c2r(int N, const double *in, double *out)
{
    size_t NN = (size_t)N, n3 = NN * NN * NN, nh = NN / 2 + 1;
    const cplx *h = (const cplx *)in;
    cplx *a = (cplx *)malloc(sizeof(cplx) * n3);
    for (size_t i = 0; i < NN; ++i)
        for (size_t j = 0; j < NN; ++j)
            for (size_t k = 0; k < NN; ++k) {
                cplx v;
                if (k < nh) v = h[(i * NN + j) * nh + k];
                else {
                    size_t ii = (NN - i) % NN, jj = (NN - j) % NN, kk = NN - k;
                    v = h[(ii * NN + jj) * nh + kk];
                    v.im = -v.im;
                }
                a[(i * NN + j) * NN + k] = v;
            }
    for (int ax = 0; ax < 3; ++ax) fft_axis(a, N, ax, +1);
    double inv = 1.0 / (double)n3;
    for (size_t i = 0; i < n3; ++i) out[i] = a[i].re * inv;
    free(a);
}

/* ------------------------------------------------------------------------- */
/* P(k): Source/FftwFunctions.cpp:958-1657                                    */
/* ------------------------------------------------------------------------- */

/* Sq(exp, s) of Miscelanious.cpp:2532-2554 (exp = scheme index). */
static double sq_corr(int e, double x)
{
    switch (e) {
    case 0: return 1.0;
    case 1: return 1.0 - (2.0 / 3.0) * x * x;
    case 2: return 1.0 - x * x + (2.0 / 15.0) * x * x * x * x;
    case 3: return 1.0 - (4.0 / 3.0) * x * x + (2.0 / 5.0) * x * x * x * x - (4.0 / 315.0) * x * x * x * x * x * x;
    }
    return 1.0;
}

static double sinc_pi(int i, int N)
{
    if (i == 0) return 1.0;
    double a = (double)i * M_PI / (double)N;
    return sin(a) / a;
}

static int near_int(double v, long *out)
{
    double r = floor(v + 0.5);
    if (fabs(v - r) < 1e-9) { *out = (long)r; return 1; }
    return 0;
}

int or_shell_index(const or_pk_opts *o, int i, int j, int k)
{
    const double kf = 2.0 * M_PI / o->L;
    if (o->log_binning) {
        /* FftwFunctions.cpp:1103-1108: kv = |k|^2/kmin^2, shell
         * floor(log10(sqrt(kv))/Deltal) + 1 for kv != 0, evaluated in FP64;
         * a quotient within 1e-10 of an integer is that integer (a mode on a
         * shell edge goes to the shell above, as the exact rule says). */
        long q = (long)i * i + (long)j * j + (long)k * k;
        if (q == 0) return -1;
        double t = log10(sqrt((double)q) / (o->kmin / kf)) / o->dlog;
        double r = floor(t + 0.5);
        if (fabs(t - r) < 1e-10) t = r;
        long s = (long)floor(t) + 1;
        if (s < 0 || s >= o->nshells) return -1;
        return (int)s;
    }
    if (o->float_shell_rule) {
        /* The reference's float32 arithmetic (real_prec=float):
         * Params members rounded to float (Params.h:745,760), terms
         * i*i*dk*dk in float (FftwFunctions.cpp:1068,1079,1089), sqrt in
         * float, floor(float((kv-kmin)*rDeltaK)) (FftwFunctions.cpp:1097). */
        float dkx = (float)kf;
        float kmin_f = (float)o->kmin;
        float rdk = (float)(1.0 / (double)(float)o->dk);
        float ti = (float)(i * i) * dkx; ti = ti * dkx;
        float tj = (float)(j * j) * dkx; tj = tj * dkx;
        float tk = (float)(k * k) * dkx; tk = tk * dkx;
        float s = ti + tj; s = s + tk;
        float kv = sqrtf(s);
        float t = (kv - kmin_f) * rdk;
        int sh = (int)floorf(t);
        if (sh < 0 || sh >= o->nshells) return -1;
        return sh;
    }
    /* Exact rule: floor((sqrt(q) - Dmin)/ndel), q = i^2+j^2+k^2, Dmin = kmin/kf,
     * ndel = dk/kf; boundaries checked in integer arithmetic when Dmin and ndel
     * are integers (the reference's DOUBLE_PREC behaviour). */
    long q = (long)i * i + (long)j * j + (long)k * k;
    double Dmin = o->kmin / kf, ndel = o->dk / kf;
    double t = (sqrt((double)q) - Dmin) / ndel;
    long s = (long)floor(t);
    long Di, ni;
    if (near_int(Dmin, &Di) && near_int(ndel, &ni) && ni > 0) {
        for (int it = 0; it < 3; ++it) {
            long lo = s * ni + Di, hi = (s + 1) * ni + Di;
            if (lo > 0 && q < lo * lo) { --s; continue; }
            if (hi >= 0 && q >= hi * hi) { ++s; continue; }
            break;
        }
    }
    if (t < 0 && s >= 0) s = -1;
    if (s < 0 || s >= o->nshells) return -1;
    return (int)s;
}

void or_pk_shells(const double *dk, const or_pk_opts *o, double *pk0, double *pk2,
                  double *pk4, int64_t *nmodes)
{
    const int N = o->N;
    const size_t NN = (size_t)N, nh = NN / 2 + 1;
    const double kf = 2.0 * M_PI / o->L;
    const int ns = o->nshells;
    double *s0 = (double *)calloc((size_t)ns, sizeof(double));
    double *s2 = (double *)calloc((size_t)ns, sizeof(double));
    double *s4 = (double *)calloc((size_t)ns, sizeof(double));
    int64_t *cnt = (int64_t *)calloc((size_t)ns, sizeof(int64_t));
    const cplx *d = (const cplx *)dk;
    for (int i = 0; i < N / 2; ++i)
        for (int j = 0; j < N / 2; ++j)
            for (int k = 0; k <= N / 2; ++k) {
                int sh = or_shell_index(o, i, j, k);
                double corr = 1.0;
                if (o->mas_correction)
                    corr = pow(sinc_pi(i, N) * sinc_pi(j, N) * sinc_pi(k, N), o->scheme + 1);
                double icorr2 = 1.0 / (corr * corr);
                double sx = sin(i * M_PI / N), sy = sin(j * M_PI / N), sz = sin(k * M_PI / N);
                double sn = o->shot_noise * sq_corr(o->scheme, sx) * sq_corr(o->scheme, sy) * sq_corr(o->scheme, sz);
                double kv = kf * sqrt((double)i * i + (double)j * j + (double)k * k);
                double mu = kv > 0 ? (double)k * kf / kv : 0.0;
                double leg2 = 0.5 * (3.0 * mu * mu - 1.0);
                double leg4 = (35.0 * mu * mu * mu * mu - 30.0 * mu * mu + 3.0) / 8.0;
                /* the four octant copies (FftwFunctions.cpp:1180,1273,1373,1474) */
                size_t lps[4];
                int nl = 0;
                size_t lp = ((size_t)i * NN + (size_t)j) * nh + (size_t)k;
                if (lp != 0) lps[nl++] = lp;
                if (j > 0 && k > 0) lps[nl++] = ((size_t)i * NN + (NN - (size_t)j)) * nh + (size_t)k;
                if (i > 0 && (j > 0 || k > 0)) lps[nl++] = ((NN - (size_t)i) * NN + (size_t)j) * nh + (size_t)k;
                if (i > 0 && j > 0 && k > 0) lps[nl++] = ((NN - (size_t)i) * NN + (NN - (size_t)j)) * nh + (size_t)k;
                if (sh < 0) continue;
                for (int c = 0; c < nl; ++c) {
                    double p2 = d[lps[c]].re * d[lps[c]].re + d[lps[c]].im * d[lps[c]].im;
                    double Pk = (p2 - sn * o->normal_power) * icorr2;
                    double Pk2 = p2 * icorr2;
                    s0[sh] += Pk;
                    s2[sh] += leg2 * Pk2;
                    s4[sh] += leg4 * Pk2;
                    cnt[sh] += 1;
                }
            }
    /* normalisation (FftwFunctions.cpp:1613-1628) */
    for (int s = 0; s < ns; ++s) {
        double den = (double)cnt[s] * o->normal_power;
        if (pk0) pk0[s] = cnt[s] == 0 ? 0.0 : s0[s] / den;
        if (pk2) pk2[s] = cnt[s] == 0 ? 0.0 : 5.0 * s2[s] / den;
        if (pk4) pk4[s] = cnt[s] == 0 ? 0.0 : 9.0 * s4[s] / den;
        if (nmodes) nmodes[s] = cnt[s];
    }
    free(s0); free(s2); free(s4); free(cnt);
}

/* Window shells of power_spectrum_fkp's random branch (FftwFunctions.cpp:
 * 1248-1262, 1354-1362, 1454-1462, 1557-1568, normalised at :1635-1638): the
 * first-octant loop adds only the zero mode to powerk_r (the else of
 * if(lp != 0)); the (i, N-j, k), (N-i, j, k) and (N-i, N-j, k) copies add their
 * modes.  |d|^2 over the squared MAS window, no shot noise; the zero mode
 * lands in the shell of k = 0 (shell 0 when kmin = 0 under linear binning,
 * always under log binning, :1110-1113).  Divided by n_modes * normal_power. */
void or_window_shells(const double *dk, const or_pk_opts *o, double *pw, int64_t *nmodes)
{
    const int N = o->N, ns = o->nshells;
    const size_t NN = (size_t)N, nh = NN / 2 + 1;
    double *s0 = (double *)calloc((size_t)ns, sizeof(double));
    int64_t *cnt = (int64_t *)calloc((size_t)ns, sizeof(int64_t));
    const cplx *d = (const cplx *)dk;
    for (int i = 0; i < N / 2; ++i)
        for (int j = 0; j < N / 2; ++j)
            for (int k = 0; k <= N / 2; ++k) {
                int sh = or_shell_index(o, i, j, k);
                if ((i | j | k) == 0 && o->log_binning) sh = 0;
                if (sh < 0) continue;
                double corr = 1.0;
                if (o->mas_correction)
                    corr = pow(sinc_pi(i, N) * sinc_pi(j, N) * sinc_pi(k, N), o->scheme + 1);
                const double icorr2 = 1.0 / (corr * corr);
                size_t lps[4];
                int nl = 0;
                if ((i | j | k) == 0) lps[nl++] = 0;
                if (j > 0 && k > 0) lps[nl++] = ((size_t)i * NN + (NN - (size_t)j)) * nh + (size_t)k;
                if (i > 0 && (j > 0 || k > 0)) lps[nl++] = ((NN - (size_t)i) * NN + (size_t)j) * nh + (size_t)k;
                if (i > 0 && j > 0 && k > 0) lps[nl++] = ((NN - (size_t)i) * NN + (NN - (size_t)j)) * nh + (size_t)k;
                for (int c = 0; c < nl; ++c) {
                    s0[sh] += (d[lps[c]].re * d[lps[c]].re + d[lps[c]].im * d[lps[c]].im) * icorr2;
                    cnt[sh] += 1;
                }
            }
    for (int s = 0; s < ns; ++s) {
        if (pw) pw[s] = cnt[s] == 0 ? 0.0 : s0[s] / ((double)cnt[s] * o->normal_power);
        if (nmodes) nmodes[s] = cnt[s];
    }
    free(s0); free(cnt);
}

/* ------------------------------------------------------------------------- */
/* T-web                                                                      */
/* ------------------------------------------------------------------------- */
static double kcoord(int i, int N, double kf)
{
    /* PoissonSolver's coords (Miscelanious.cpp:1236-1241): i<=N/2 -> +i */
    return kf * (i <= N / 2 ? (double)i : -(double)(N - i));
}

void or_smooth_poisson(int N, double L, double R, double *dk)
{
    const size_t NN = (size_t)N, nh = NN / 2 + 1;
    const double kf = 2.0 * M_PI / L;
    cplx *d = (cplx *)dk;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < (int)nh; ++k) {
                double kx = kcoord(i, N, kf), ky = kcoord(j, N, kf), kz = kcoord(k, N, kf);
                double k2 = kx * kx + ky * ky + kz * kz;
                double f = 0.0;
                if (k2 > 0) {
                    f = -1.0 / k2;
                    if (R > 0) f *= exp(-0.5 * k2 * R * R);
                }
                size_t idx = ((size_t)i * NN + (size_t)j) * nh + (size_t)k;
                d[idx].re *= f;
                d[idx].im *= f;
            }
}

void or_gradfindif(int N, double L, const double *in, double *out, int dim)
{
    const size_t NN = (size_t)N;
    const double ideltaH = 1.0 / (12.0 * (L / (double)N));
    for (int x = 0; x < N; ++x)
        for (int y = 0; y < N; ++y)
            for (int z = 0; z < N; ++z) {
                int c[3] = {x, y, z};
                int a = dim - 1;
                int cr[3], cl[3], crr[3], cll[3];
                memcpy(cr, c, sizeof c); memcpy(cl, c, sizeof c);
                memcpy(crr, c, sizeof c); memcpy(cll, c, sizeof c);
                cr[a] = (c[a] + 1) % N;
                crr[a] = (c[a] + 2) % N;
                cl[a] = (c[a] - 1 + N) % N;
                cll[a] = (c[a] - 2 + N) % N;
#define IDX(v) ((size_t)(v)[2] + NN * ((size_t)(v)[1] + NN * (size_t)(v)[0]))
                out[IDX(c)] = (-in[IDX(crr)] + 8.0 * in[IDX(cr)] - 8.0 * in[IDX(cl)] + in[IDX(cll)]) * ideltaH;
#undef IDX
            }
}

void or_eigen3(const double m[9], double ev[3])
{
    /* symmetric matrix from the lower triangle A(i,j)=m[i+3j], i>=j */
    double a[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j <= i; ++j) { a[i][j] = m[i + 3 * j]; a[j][i] = a[i][j]; }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        double scale = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2] + off;
        if (off <= 1e-34 * scale || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < 3; ++r) {
                    double arp = a[r][p], arq = a[r][q];
                    a[r][p] = c * arp - s * arq;
                    a[r][q] = s * arp + c * arq;
                }
                for (int r = 0; r < 3; ++r) {
                    double apr = a[p][r], aqr = a[q][r];
                    a[p][r] = c * apr - s * aqr;
                    a[q][r] = s * apr + c * aqr;
                }
            }
    }
    double e0 = a[0][0], e1 = a[1][1], e2 = a[2][2], t;
    if (e0 < e1) { t = e0; e0 = e1; e1 = t; }
    if (e1 < e2) { t = e1; e1 = e2; e2 = t; }
    if (e0 < e1) { t = e0; e0 = e1; e1 = t; }
    ev[0] = e0; ev[1] = e1; ev[2] = e2;
}

void or_tweb_from_phi(int N, double L, const double *phi, double lambda_th,
                      int8_t *labels, double *lambdas, int64_t counts[4])
{
    const size_t n3 = (size_t)N * N * N;
    double *g = (double *)malloc(sizeof(double) * n3);
    double *T[9];
    for (int q = 0; q < 9; ++q) T[q] = (double *)malloc(sizeof(double) * n3);
    /* EigenValuesTweb FD branch (Miscelanious.cpp:1684-1700): T[a*3+b] =
     * gradfindif(gradfindif(phi, a+1), b+1); in_matrix ordering
     * (Miscelanious.cpp:1732-1740) m = {Txx, Tyx, Tzx, Txy, Tyy, Tzy, Txz, Tyz, Tzz}
     * where Tyx = D_x(D_y phi). */
    for (int a = 0; a < 3; ++a) {
        or_gradfindif(N, L, phi, g, a + 1);
        for (int b = 0; b < 3; ++b) or_gradfindif(N, L, g, T[a * 3 + b], b + 1);
    }
    if (counts) memset(counts, 0, 4 * sizeof(int64_t));
    for (size_t c = 0; c < n3; ++c) {
        /* m[i+3j] = A(i,j); A(1,0)=Tyx=T[1*3+0], A(2,0)=Tzx=T[2*3+0], A(2,1)=Tzy=T[2*3+1] */
        double m[9];
        m[0] = T[0][c];          /* A00 = Txx */
        m[1] = T[1 * 3 + 0][c];  /* A10 = D_x(D_y) */
        m[2] = T[2 * 3 + 0][c];  /* A20 = D_x(D_z) */
        m[3] = T[0 * 3 + 1][c];  /* A01 = D_y(D_x) */
        m[4] = T[1 * 3 + 1][c];
        m[5] = T[2 * 3 + 1][c];  /* A21 = D_y(D_z) */
        m[6] = T[0 * 3 + 2][c];
        m[7] = T[1 * 3 + 2][c];
        m[8] = T[2 * 3 + 2][c];
        double ev[3];
        or_eigen3(m, ev);
        int v = 0;
        for (int q = 0; q < 3; ++q) if (ev[q] > lambda_th) ++v;
        int lab = 4 - v;
        if (labels) labels[c] = (int8_t)lab;
        if (lambdas) { lambdas[c] = ev[0]; lambdas[n3 + c] = ev[1]; lambdas[2 * n3 + c] = ev[2]; }
        if (counts) counts[lab - 1]++;
    }
    for (int q = 0; q < 9; ++q) free(T[q]);
    free(g);
}

void or_tweb(int N, double L, const double *delta, double R, double lambda_th,
             int8_t *labels, double *lambdas, double *phi, int64_t counts[4])
{
    const size_t NN = (size_t)N, n3 = NN * NN * NN, nh = NN * NN * (NN / 2 + 1);
    double *dk = (double *)malloc(sizeof(double) * 2 * nh);
    double *ph = phi ? phi : (double *)malloc(sizeof(double) * n3);
    or_fft_r2c(N, delta, dk);
    or_smooth_poisson(N, L, R, dk);
    or_fft_c2r(N, dk, ph);
    or_tweb_from_phi(N, L, ph, lambda_th, labels, lambdas, counts);
    if (!phi) free(ph);
    free(dk);
}

void or_tweb_spectral(int N, double L, const double *delta, double R, double lambda_th,
                      int8_t *labels, double *lambdas, int64_t counts[4])
{
    /* EigenValuesTweb _USE_GFFT_EIGENV_ branch (Miscelanious.cpp:1703-1718):
     * T_ab = c2r(k_a k_b / k^2 delta_k) per component (calc_LapPhiv,
     * NumericalMethods.cpp:2605-2679: coords +i k_f for i <= N/2, zero at
     * k = 0), the Gaussian of the pre-smoothing applied to delta_k; the matrix
     * is symmetric, so the six components ab = xx, yy, zz, yx, zx, zy fill it. */
    const size_t NN = (size_t)N, n3 = NN * NN * NN, nh = NN / 2 + 1, nhh = NN * NN * nh;
    const double kf = 2.0 * M_PI / L;
    static const int ab[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 0}, {2, 0}, {2, 1}};
    cplx *dk = (cplx *)malloc(sizeof(cplx) * nhh), *tmp = (cplx *)malloc(sizeof(cplx) * nhh);
    double *T[6];
    or_fft_r2c(N, delta, (double *)dk);
    for (int c = 0; c < 6; ++c) {
        T[c] = (double *)malloc(sizeof(double) * n3);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < (int)nh; ++k) {
                    const double kv[3] = {kcoord(i, N, kf), kcoord(j, N, kf), kcoord(k, N, kf)};
                    const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                    double f = 0.0;
                    if (k2 > 0) {
                        f = kv[ab[c][0]] * kv[ab[c][1]] / k2;
                        if (R > 0) f *= exp(-0.5 * k2 * R * R);
                    }
                    const size_t idx = ((size_t)i * NN + (size_t)j) * nh + (size_t)k;
                    tmp[idx].re = dk[idx].re * f;
                    tmp[idx].im = dk[idx].im * f;
                }
        or_fft_c2r(N, (const double *)tmp, T[c]);
    }
    if (counts) memset(counts, 0, 4 * sizeof(int64_t));
    for (size_t c = 0; c < n3; ++c) {
        double m[9];
        m[0] = T[0][c]; m[4] = T[1][c]; m[8] = T[2][c];
        m[1] = m[3] = T[3][c];
        m[2] = m[6] = T[4][c];
        m[5] = m[7] = T[5][c];
        double ev[3];
        or_eigen3(m, ev);
        int v = 0;
        for (int q = 0; q < 3; ++q) if (ev[q] > lambda_th) ++v;
        const int lab = 4 - v;
        if (labels) labels[c] = (int8_t)lab;
        if (lambdas) { lambdas[c] = ev[0]; lambdas[n3 + c] = ev[1]; lambdas[2 * n3 + c] = ev[2]; }
        if (counts) counts[lab - 1]++;
    }
    for (int c = 0; c < 6; ++c) free(T[c]);
    free(tmp);
    free(dk);
}

/* bin of |q|^2 = q2 (units of k_f), -1 outside [Dmin, Dmin + nb ndel) */
static int bkd_bin(long q2, double Dmin, double ndel, int nb, int int_rule, long Di, long ni)
{
    int s;
    if (int_rule) {
        if (q2 < Di * Di) return -1;
        s = (int)floor((sqrt((double)q2) - (double)Di) / (double)ni);
        while (s > 0 && q2 < (Di + s * ni) * (Di + s * ni)) --s;
        while (q2 >= (Di + (s + 1) * ni) * (Di + (s + 1) * ni)) ++s;
    } else {
        double t = (sqrt((double)q2) - Dmin) / ndel;
        if (t < 0) return -1;
        s = (int)floor(t);
    }
    return s < nb ? s : -1;
}

/* the half-spectrum value of q (remap, Source/Miscelanious.cpp:43-62): stored
 * at (q mod N) when q_z >= 0, else the conjugate of -q's */
static void bkd_value(int N, const double *dk, int qx, int qy, int qz, double *re, double *im)
{
    const int M1 = N / 2 + 1;
    double sg = 1.0;
    if (qz < 0) { qx = -qx; qy = -qy; qz = -qz; sg = -1.0; }
    const int ix = ((qx % N) + N) % N, iy = ((qy % N) + N) % N;
    const size_t g = ((size_t)ix * N + iy) * M1 + qz;
    *re = dk[2 * g];
    *im = sg * dk[2 * g + 1];
}

void or_bk_direct(int N, const double *dk, double Dmin, double ndel, int nb, int scheme, int mas_corr,
                  double *bsum, double *snsum, int64_t *cnt, double *psum, int64_t *pcnt)
{
    const int h = N / 2, M = N + 1;
    long Di = 0, ni = 0;
    const int int_rule = near_int(Dmin, &Di) && near_int(ndel, &ni) && ni > 0;
    const size_t G = (size_t)M * M * M;
    double *vr = (double *)malloc(sizeof(double) * G), *vi = (double *)malloc(sizeof(double) * G);
    int *bin = (int *)malloc(sizeof(int) * G);
    for (int a = -h; a <= h; ++a)
        for (int b = -h; b <= h; ++b)
            for (int c = -h; c <= h; ++c) {
                const size_t t = ((size_t)(a + h) * M + (b + h)) * M + (c + h);
                double re, im;
                bkd_value(N, dk, a, b, c, &re, &im);
                double w = 1.0;
                if (mas_corr) w = pow(sinc_pi(a, N) * sinc_pi(b, N) * sinc_pi(c, N), scheme + 1);
                vr[t] = re / w;
                vi[t] = im / w;
                const long q2 = (long)a * a + (long)b * b + (long)c * c;
                bin[t] = q2 == 0 ? -1 : bkd_bin(q2, Dmin, ndel, nb, int_rule, Di, ni);
            }
    for (int a = -h; a <= h; ++a)
        for (int b = -h; b <= h; ++b)
            for (int c = -h; c <= h; ++c) {
                const size_t t1 = ((size_t)(a + h) * M + (b + h)) * M + (c + h);
                const int s1 = bin[t1];
                if (s1 < 0) continue;
                const double p1 = vr[t1] * vr[t1] + vi[t1] * vi[t1];
                psum[s1] += p1;
                pcnt[s1] += 1;
                for (int x = -h; x <= h; ++x)
                    for (int y = -h; y <= h; ++y)
                        for (int z = -h; z <= h; ++z) {
                            const int u = -a - x, v = -b - y, w = -c - z;
                            if (abs(u) > h || abs(v) > h || abs(w) > h) continue;
                            const size_t t2 = ((size_t)(x + h) * M + (y + h)) * M + (z + h);
                            const size_t t3 = ((size_t)(u + h) * M + (v + h)) * M + (w + h);
                            int s2 = bin[t2], s3 = bin[t3];
                            if (s2 < 0 || s3 < 0) continue;
                            /* Re(d1 d2 d3) */
                            const double r12 = vr[t1] * vr[t2] - vi[t1] * vi[t2];
                            const double i12 = vr[t1] * vi[t2] + vi[t1] * vr[t2];
                            const double re = r12 * vr[t3] - i12 * vi[t3];
                            const double sn = p1 + vr[t2] * vr[t2] + vi[t2] * vi[t2] + vr[t3] * vr[t3] + vi[t3] * vi[t3];
                            int lo = s1, mid = s2, hi = s3, tmp;
                            if (lo > mid) { tmp = lo; lo = mid; mid = tmp; }
                            if (mid > hi) { tmp = mid; mid = hi; hi = tmp; }
                            if (lo > mid) { tmp = lo; lo = mid; mid = tmp; }
                            const size_t o = ((size_t)lo * nb + mid) * nb + hi;
                            bsum[o] += re;
                            snsum[o] += sn;
                            cnt[o] += 1;
                        }
            }
    free(vr);
    free(vi);
    free(bin);
}
