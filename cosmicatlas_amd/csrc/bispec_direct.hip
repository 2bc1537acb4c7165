// Direct (brute-force) bispectrum of a periodic box (gfx950).
//
// The estimator of FftwFunctions::bispectrum_fkp (Source/FftwFunctions.cpp:
// 3105-3297) and its normalisation in get_bispectrum_fkp (:3318-3350): every
// ordered pair (k1, k2) of the grid q in [-N/2, N/2]^3 (N+1 points per axis, so
// the Nyquist planes appear as +N/2 and -N/2) whose k3 = -k1-k2 lies in the same
// grid, no vector zero, summed in the bin triple sorted ascending:
//   bsum += Re(d1 d2 d3),  snsum += |d1|^2 + |d2|^2 + |d3|^2,  ntri += 1.
//
// Layout: one table over the (N+1)^3 grid, indexed by
// t = ((qx+h) M + (qy+h)) M + (qz+h) (M = N+1, h = N/2), holding the
// MAS-corrected mode d(q) in double (the half-spectrum lookup of remap,
// Source/Miscelanious.cpp:43-62: q_z < 0 reads the conjugate of -q) and its bin
// (int8, -1 for q = 0, modes below the first bin and at or past the last).
//
// Pair kernel: one workgroup per k1 (bin s1 fixed).  For a fixed k1 the valid
// k2 form a box (k3 inside the grid) cut by two spheres (|k2| and |k3| below
// the last bin edge); each thread walks whole z lines of that region, so
// successive k2 (and the mirrored k3) are adjacent table entries, and keeps a
// running sum while the pair (s2, s3) stays the same, flushing it into an
// LDS tile [s2][s3] when it changes.  The tile is then added into the sorted
// triple's global sums.  Cost: sum over kept k1 of the kept k2, about
// (4/3 pi R^3)^2 pairs for a last bin edge of R k_f.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cm_internal.hpp"

namespace cm {

constexpr int kBkdMaxBins = 64;

// bin of |q|^2 = q2 (units of k_f): exact integer comparisons when the bin
// edges Di + s ni are integers, else FP64 floor((|q| - Dmin)/ndel)
__device__ __forceinline__ int bkd_bin(long q2, int int_rule, long Di, long ni, double Dmin, double ndel, int nb) {
    int s;
    if (int_rule) {
        if (q2 < Di * Di) return -1;
        s = (int)floor((sqrt((double)q2) - (double)Di) / (double)ni);
        if (s < 0) s = 0;
        while (s > 0 && q2 < (Di + s * ni) * (Di + s * ni)) --s;
        while (q2 >= (Di + (s + 1) * ni) * (Di + (s + 1) * ni)) ++s;
    } else {
        const double t = (sqrt((double)q2) - Dmin) / ndel;
        if (t < 0) return -1;
        s = (int)floor(t);
    }
    return s < nb ? s : -1;
}

__device__ __forceinline__ double bkd_sinc(int q, int N) {
    if (q == 0) return 1.0;
    const double a = (double)q * M_PI / (double)N;
    return sin(a) / a;
}

// Table of d(q)/W(q) and bins over the (N+1)^3 grid, plus the per-bin power
// sums over k1 (psum, pcnt).
__global__ void __launch_bounds__(256) k_bkd_table(const float2 *__restrict__ dk, int N, int nb, int int_rule, long Di,
                                                   long ni, double Dmin, double ndel, int expo, int mas,
                                                   double2 *__restrict__ tv, int8_t *__restrict__ tb,
                                                   double *__restrict__ psum, unsigned long long *__restrict__ pcnt) {
    __shared__ double s_p[kBkdMaxBins];
    __shared__ unsigned s_n[kBkdMaxBins];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        s_p[i] = 0.0;
        s_n[i] = 0;
    }
    __syncthreads();
    const int M = N + 1, h = N / 2, M1 = N / 2 + 1;
    const size_t G = (size_t)M * M * M;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < G) {
        const int c = (int)(t % M) - h;
        const int b = (int)((t / M) % M) - h;
        const int a = (int)(t / ((size_t)M * M)) - h;
        int qx = a, qy = b, qz = c;
        double sg = 1.0;
        if (qz < 0) {
            qx = -qx;
            qy = -qy;
            qz = -qz;
            sg = -1.0;
        }
        const int ix = qx < 0 ? qx + N : qx;
        const int iy = qy < 0 ? qy + N : qy;
        const float2 v = dk[((size_t)ix * N + iy) * M1 + qz];
        double w = 1.0;
        if (mas) w = pow(bkd_sinc(a, N) * bkd_sinc(b, N) * bkd_sinc(c, N), (double)expo);
        const double re = (double)v.x / w, im = sg * (double)v.y / w;
        tv[t] = double2{re, im};
        const long q2 = (long)a * a + (long)b * b + (long)c * c;
        const int s = q2 == 0 ? -1 : bkd_bin(q2, int_rule, Di, ni, Dmin, ndel, nb);
        tb[t] = (int8_t)s;
        if (s >= 0) {
            atomicAdd(&s_p[s], re * re + im * im);
            atomicAdd(&s_n[s], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if (s_n[i]) {
            unsafeAtomicAdd(&psum[i], s_p[i]);
            atomicAdd(&pcnt[i], (unsigned long long)s_n[i]);
        }
}

__device__ __forceinline__ int isqrt_floor(long r) {
    long s = (long)sqrt((double)r);
    while (s * s > r) --s;
    while ((s + 1) * (s + 1) <= r) ++s;
    return (int)s;
}

// One workgroup per k1: table index blockIdx.y M^2 + blockIdx.x.  LDS tile: re[nb^2],
// sn[nb^2] (double), n[nb^2] (u32).  r2max: the largest |q|^2 of any kept
// mode (prunes k2 and k3 to spheres; the bins decide).
__global__ void __launch_bounds__(256) k_bkd_pairs(const double2 *__restrict__ tv, const int8_t *__restrict__ tb,
                                                   int N, int nb, long r2max, double *__restrict__ bsum,
                                                   double *__restrict__ snsum, unsigned long long *__restrict__ cnt) {
    const size_t t1 = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
    const int s1 = tb[t1];
    if (s1 < 0) return;  // uniform over the workgroup, before any barrier
    extern __shared__ double s_tile[];
    const int nb2 = nb * nb;
    double *s_re = s_tile;
    double *s_sn = s_tile + nb2;
    unsigned *s_n = (unsigned *)(s_tile + 2 * nb2);
    for (int i = threadIdx.x; i < nb2; i += blockDim.x) {
        s_re[i] = 0.0;
        s_sn[i] = 0.0;
        s_n[i] = 0;
    }
    __syncthreads();
    const int M = N + 1, h = N / 2;
    const int c = (int)(t1 % M) - h;
    const int b = (int)((t1 / M) % M) - h;
    const int a = (int)(t1 / ((size_t)M * M)) - h;
    const double2 v1 = tv[t1];
    const double p1 = v1.x * v1.x + v1.y * v1.y;
    // k3 = -k1-k2 inside the grid: per axis max(-h, -h - q1) <= q2 <= min(h, h - q1)
    const int xlo = max(-h, -h - a), xhi = min(h, h - a);
    const int ylo = max(-h, -h - b), yhi = min(h, h - b);
    const int zlo0 = max(-h, -h - c), zhi0 = min(h, h - c);
    const int nly = yhi - ylo + 1;
    const int nlines = (xhi - xlo + 1) * nly;
    for (int line = threadIdx.x; line < nlines; line += blockDim.x) {
        const int x = xlo + line / nly, y = ylo + line % nly;
        const long r2 = r2max - (long)x * x - (long)y * y;
        const int u = -a - x, v = -b - y;
        const long r3 = r2max - (long)u * u - (long)v * v;
        if (r2 < 0 || r3 < 0) continue;
        const int l2 = isqrt_floor(r2), l3 = isqrt_floor(r3);
        const int zlo = max(zlo0, max(-l2, -c - l3));
        const int zhi = min(zhi0, min(l2, -c + l3));
        if (zlo > zhi) continue;
        const size_t base2 = ((size_t)(x + h) * M + (y + h)) * M + h;  // + z
        const size_t base3 = ((size_t)(u + h) * M + (v + h)) * M + h;  // + (-c - z)
        int key = -1;
        double re = 0.0, sn = 0.0;
        unsigned n = 0;
        for (int z = zlo; z <= zhi; ++z) {
            const size_t t2 = base2 + z, t3 = base3 - c - z;
            const int s2 = tb[t2], s3 = tb[t3];
            if ((s2 | s3) < 0) continue;
            const int k = s2 * nb + s3;
            if (k != key) {
                if (n) {
                    atomicAdd(&s_re[key], re);
                    atomicAdd(&s_sn[key], sn);
                    atomicAdd(&s_n[key], n);
                }
                key = k;
                re = sn = 0.0;
                n = 0;
            }
            const double2 v2 = tv[t2], v3 = tv[t3];
            const double r12 = v1.x * v2.x - v1.y * v2.y;
            const double i12 = v1.x * v2.y + v1.y * v2.x;
            re += r12 * v3.x - i12 * v3.y;
            sn += p1 + (v2.x * v2.x + v2.y * v2.y) + (v3.x * v3.x + v3.y * v3.y);
            ++n;
        }
        if (n) {
            atomicAdd(&s_re[key], re);
            atomicAdd(&s_sn[key], sn);
            atomicAdd(&s_n[key], n);
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nb2; e += blockDim.x) {
        const unsigned n = s_n[e];
        if (!n) continue;
        int lo = s1, mid = e / nb, hi = e % nb;
        if (lo > mid) { const int tt = lo; lo = mid; mid = tt; }
        if (mid > hi) { const int tt = mid; mid = hi; hi = tt; }
        if (lo > mid) { const int tt = lo; lo = mid; mid = tt; }
        const size_t o = ((size_t)lo * nb + mid) * nb + hi;
        unsafeAtomicAdd(&bsum[o], s_re[e]);
        unsafeAtomicAdd(&snsum[o], s_sn[e]);
        atomicAdd(&cnt[o], (unsigned long long)n);
    }
}

}  // namespace cm

using namespace cm;

extern "C" cm_status cm_bispectrum_direct(cm_ctx *ctx, const cm_bkd_opts *o, const void *dk, double *bk, double *sn,
                                          int64_t *ntri, double *pk, int64_t *nk, double *bsum_out,
                                          double *snsum_out) {
    if (!ctx) return CM_ERR_INVALID;
    if (!o || !dk || !bk || !sn || !ntri || !pk || !nk)
        return fail(ctx, CM_ERR_INVALID, "cm_bispectrum_direct: null pointer");
    const int N = o->N, nb = o->nbins;
    if (N < 2 || N > 1024 || (N & 1)) return fail(ctx, CM_ERR_INVALID, "cm_bispectrum_direct: N=%d (even, 2..1024)", N);
    if (nb < 1 || nb > kBkdMaxBins || !(o->dk > 0) || !(o->kmin >= 0) || !(o->n_obj > 0) || !(o->L > 0))
        return fail(ctx, CM_ERR_INVALID, "cm_bispectrum_direct: nbins (1..%d), kmin, dk, L or n_obj", kBkdMaxBins);
    if ((int)o->scheme < 0 || (int)o->scheme > 3) return fail(ctx, CM_ERR_INVALID, "cm_bispectrum_direct: scheme");
    const double L = o->L, V = L * L * L, kf = 2.0 * M_PI / L;
    const double Dmin = o->kmin / kf, ndel = o->dk / kf;
    const long Di = std::lround(Dmin), ni = std::lround(ndel);
    const int int_rule = (std::fabs(Dmin - (double)Di) < 1e-9 && std::fabs(ndel - (double)ni) < 1e-9 && ni > 0) ? 1 : 0;
    // largest |q|^2 that can fall in a kept bin (a conservative bound for the FP rule)
    const long qmax2 = 3L * (N / 2) * (N / 2);
    long r2max;
    if (int_rule) {
        const long hi = Di + (long)nb * ni;
        r2max = std::min(qmax2, hi * hi - 1);
    } else {
        const double hi = Dmin + nb * ndel;
        r2max = std::min(qmax2, (long)std::ceil(hi * hi) + 1);
    }
    const int M = N + 1;
    const size_t G = (size_t)M * M * M, nb3 = (size_t)nb * nb * nb;
    cm_status s;
    void *tv, *tb, *sums, *cnts;
    if ((s = scratch(ctx, "bkd.tv", sizeof(double2) * G, &tv)) != CM_OK) return s;
    if ((s = scratch(ctx, "bkd.tb", G, &tb)) != CM_OK) return s;
    if ((s = scratch(ctx, "bkd.sums", sizeof(double) * (2 * nb3 + nb), &sums)) != CM_OK) return s;
    if ((s = scratch(ctx, "bkd.cnts", sizeof(unsigned long long) * (nb3 + nb), &cnts)) != CM_OK) return s;
    auto *dsum = (double *)sums;  // bsum[nb3], snsum[nb3], psum[nb]
    auto *dcnt = (unsigned long long *)cnts;  // ntri[nb3], pcnt[nb]
    CM_HIP(ctx, hipMemsetAsync(dsum, 0, sizeof(double) * (2 * nb3 + nb), ctx->stream));
    CM_HIP(ctx, hipMemsetAsync(dcnt, 0, sizeof(unsigned long long) * (nb3 + nb), ctx->stream));
    const int expo = (int)o->scheme + 1;
    if ((s = launch(ctx, "bkd_table", k_bkd_table, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, (const float2 *)dk,
                    N, nb, int_rule, Di, ni, Dmin, ndel, expo, o->mas_correction ? 1 : 0, (double2 *)tv, (int8_t *)tb,
                    dsum + 2 * nb3, dcnt + nb3)) != CM_OK)
        return s;
    const size_t lds = (2 * sizeof(double) + sizeof(unsigned)) * (size_t)nb * nb;
    if (lds > ctx->lds_per_block)
        return fail(ctx, CM_ERR_UNSUPPORTED, "cm_bispectrum_direct: %d bins need %zu B LDS", nb, lds);
    if ((s = launch(ctx, "bkd_pairs", k_bkd_pairs, dim3((unsigned)M * M, (unsigned)M), dim3(256), lds, (const double2 *)tv,
                    (const int8_t *)tb, N, nb, r2max, dsum, dsum + nb3, dcnt)) != CM_OK)
        return s;
    std::vector<double> hs(2 * nb3 + nb);
    std::vector<unsigned long long> hc(nb3 + nb);
    CM_HIP(ctx, hipMemcpyAsync(hs.data(), dsum, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(ctx, hipMemcpyAsync(hc.data(), dcnt, sizeof(unsigned long long) * hc.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double n = o->n_obj, normal_p = n * n / V, normal_b = n * n * n / (V * V), sn_p = V / n;
    const bool snc = o->sn_correction != 0;
    for (int i = 0; i < nb; ++i) {
        const unsigned long long m = hc[nb3 + i];
        nk[i] = (int64_t)m;
        pk[i] = m ? hs[2 * nb3 + i] / (double)m / normal_p - (snc ? sn_p : 0.0) : 0.0;
    }
    for (size_t e = 0; e < nb3; ++e) {
        const unsigned long long m = hc[e];
        ntri[e] = (int64_t)m;
        if (bsum_out) bsum_out[e] = hs[e];
        if (snsum_out) snsum_out[e] = hs[nb3 + e];
        if (!m) {
            bk[e] = sn[e] = 0.0;
            continue;
        }
        // <P1+P2+P3>/nbar + 1/nbar^2 with P_i = |d_i|^2/normal_p - V/n
        const double snb = snc ? (hs[nb3 + e] / (double)m / normal_p - 3.0 * sn_p) * sn_p + sn_p * sn_p : 0.0;
        sn[e] = snb;
        bk[e] = hs[e] / ((double)m * normal_b) - snb;
    }
    return CM_OK;
}
