// Direct (brute-force) bispectrum of a periodic box (gfx950).
//
// The estimator of FftwFunctions::bispectrum_fkp (Source/FftwFunctions.cpp:
// 3105-3297) and its normalisation in get_bispectrum_fkp (:3318-3350): every
// ordered pair (k1, k2) of the grid q in [-N/2, N/2]^3 (N+1 points per axis, so
// the Nyquist planes appear as +N/2 and -N/2) whose k3 = -k1-k2 lies in the same
// grid, no vector zero, summed in the bin triple sorted ascending:
//   bsum += Re(d1 d2 d3),  snsum += |d1|^2 + |d2|^2 + |d3|^2,  ntri += 1
// (snsum as three times the sum of |d1|^2, an identity of the pair set).
//
// Layout: one table over the (N+1)^3 grid, indexed by
// t = ((qx+h) M + (qz+h)) M + (qy+h) (M = N+1, h = N/2; y fastest), holding the
// MAS-corrected mode d(q) in double (the half-spectrum lookup of remap,
// Source/Miscelanious.cpp:43-62: q_z < 0 reads the conjugate of -q) and its bin
// (int8, -1 for q = 0, modes below the first bin and at or past the last).
//
// Pair kernel: one workgroup per chunk of k1 of one bin s1.  For a fixed k1 the valid
// k2 form a box (k3 inside the grid) cut by two spheres (|k2| and |k3| below
// the last bin edge); the 64 lanes of a wave take 64 adjacent y lines of one x
// and walk them in z, so each load of the wave reads adjacent table entries
// (for k2 and the mirrored k3), and each lane keeps a running sum while its
// pair (s2, s3) stays the same, flushing it into an LDS tile [s2][s3] when it
// changes.  The tile is then added into the sorted triple's global sums.
// Cost: sum over kept k1 of the kept k2, about (4/3 pi R^3)^2 pairs for a last
// bin edge of R k_f.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cm_internal.hpp"

namespace cm {

constexpr int kBkdMaxBins = 64;
constexpr int kBkdZ = 4;  // z steps whose table loads are in flight together
// table padding (entries) before and after the grid: a round of kBkdZ steps
// may read up to (kBkdZ-1) lines past either end of a z line
constexpr unsigned kBkdPad = 4 * 1025;

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
        // table layout ((qx+h) M + (qz+h)) M + (qy+h): y fastest
        const int b = (int)(t % M) - h;
        const int c = (int)((t / M) % M) - h;
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
        tv[t + kBkdPad] = double2{re, im};
        const long q2 = (long)a * a + (long)b * b + (long)c * c;
        const int s = q2 == 0 ? -1 : bkd_bin(q2, int_rule, Di, ni, Dmin, ndel, nb);
        tb[t + kBkdPad] = (int8_t)s;
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

// Table indices of the kept k1 grouped by bin: list[cursor[s]++] = t (cursor
// starts at the bin's offset).  Positions are reserved per workgroup and bin
// (one global atomic per nonempty bin), then ranked inside the workgroup.
__global__ void __launch_bounds__(256) k_bkd_list(const int8_t *__restrict__ tb, unsigned G, int nb,
                                                  unsigned *__restrict__ cursor, unsigned *__restrict__ list) {
    __shared__ unsigned s_cnt[kBkdMaxBins], s_base[kBkdMaxBins];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s_cnt[i] = 0;
    __syncthreads();
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = t < G ? (int)tb[t + kBkdPad] : -1;
    unsigned rank = 0;
    if (s >= 0) rank = atomicAdd(&s_cnt[s], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if (s_cnt[i]) s_base[i] = atomicAdd(&cursor[i], s_cnt[i]);
    __syncthreads();
    if (s >= 0) list[s_base[s] + rank] = t;
}

// One workgroup per chunk of k1 of one bin: chunk[blockIdx.x] = {s1, first,
// count} into list.  LDS tiles: re[nb^2], sum of p1 n[nb^2] (double), n[nb^2]
// and the current k1's n[nb^2] (u32, folded after each k1); the chunk's tiles are
// added once into the global sums.
// r2max: the largest |q|^2 of any kept mode (prunes k2 and k3 to spheres; the
// bins decide).  Work items are (x, 64 consecutive y) of k2, one per wave; lane
// y walks its z line, so every table load of a wave reads 64 adjacent entries
// (the table keeps y fastest), for k2 and for the mirrored k3 alike.
//
// Only sum p(k1) is accumulated: the ordered pairs of one sorted bin triple
// are closed under (k1, k2, k3) -> (k2, k3, k1), so sum p1 = sum p2 = sum p3
// over them and the caller's |d1|^2+|d2|^2+|d3|^2 sum is three times it.
__global__ void __launch_bounds__(256) k_bkd_pairs(const double2 *__restrict__ tv, const int8_t *__restrict__ tb,
                                                   const unsigned *__restrict__ list, const int4 *__restrict__ chunks,
                                                   int N, int nb, long r2max, double *__restrict__ bsum,
                                                   double *__restrict__ psum3, unsigned long long *__restrict__ cnt) {
    const int M = N + 1, h = N / 2;
    const int4 ch = chunks[blockIdx.x];
    const int s1 = ch.x;
    extern __shared__ double s_tile[];
    const int nb2 = nb * nb;
    double *s_re = s_tile;
    double *s_sn = s_tile + nb2;
    unsigned *s_n = (unsigned *)(s_tile + 2 * nb2);  // pairs of the chunk
    unsigned *s_nk = s_n + nb2;                       // pairs of the current k1
    for (int i = threadIdx.x; i < nb2; i += blockDim.x) {
        s_re[i] = 0.0;
        s_sn[i] = 0.0;
        s_n[i] = 0;
        s_nk[i] = 0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const unsigned uM = (unsigned)M;
    const int rmax = isqrt_floor(r2max);
    int item = 0;  // work items continue across the chunk's k1, balancing the waves
    for (int q = 0; q < ch.z; ++q) {
        // table index ((qx+h) M + (qz+h)) M + (qy+h), stored kBkdPad entries in
        const unsigned t1 = list[ch.y + q];
        const int a = (int)(t1 / (uM * uM)) - h;
        const int c = (int)((t1 / uM) % uM) - h;
        const int b = (int)(t1 % uM) - h;
        const double2 v1 = tv[t1 + kBkdPad];
        const double p1 = v1.x * v1.x + v1.y * v1.y;
        // k3 = -k1-k2 inside the grid: per axis max(-h, -h - q1) <= q2 <= min(h, h - q1)
        // ... and |x| <= rmax, |u| = |a + x| <= rmax (the spheres)
        const int xlo = max(max(-h, -h - a), max(-rmax, -a - rmax)), xhi = min(min(h, h - a), min(rmax, -a + rmax));
        const int ylo = max(-h, -h - b), yhi = min(h, h - b);
        const int zlo0 = max(-h, -h - c), zhi0 = min(h, h - c);
        for (int x = xlo; x <= xhi; ++x) {
            const int u = -a - x;
            const long rx2 = r2max - (long)x * x, rx3 = r2max - (long)u * u;
            if (rx2 < 0 || rx3 < 0) continue;
            const int ly2 = isqrt_floor(rx2), ly3 = isqrt_floor(rx3);
            const int y0 = max(ylo, max(-ly2, -b - ly3)), y1 = min(yhi, min(ly2, -b + ly3));
            const int zx0 = max(zlo0, max(-ly2, -c - ly3)), zx1 = min(zhi0, min(ly2, -c + ly3));
            if (y0 > y1 || zx0 > zx1) continue;
            for (int yc = y0; yc <= y1; yc += 64, ++item) {
                if (item % nw != wave) continue;
                const int y = yc + lane, v = -b - y;
                int zl = 0, zh = -1;
                if (y <= y1) {
                    const long r2 = rx2 - (long)y * y, r3 = rx3 - (long)v * v;
                    if (r2 >= 0 && r3 >= 0) {
                        const int l2 = isqrt_floor(r2), l3 = isqrt_floor(r3);
                        zl = max(zx0, max(-l2, -c - l3));
                        zh = min(zx1, min(l2, -c + l3));
                    }
                }
                if (zl > zh) continue;
                // padded indices of k2 = (x, y, z) and k3 = (u, v, -c-z); z + i
                // past zh reads the padding or the next line and is skipped
                unsigned t2 = ((unsigned)(x + h) * uM + (unsigned)(zl + h)) * uM + (unsigned)(y + h) + kBkdPad;
                unsigned t3 = ((unsigned)(u + h) * uM + (unsigned)(-c - zl + h)) * uM + (unsigned)(v + h) + kBkdPad;
                int key = -1;
                double re = 0.0;
                unsigned n = 0;
                for (int z = zl; z <= zh; z += kBkdZ, t2 += kBkdZ * uM, t3 -= kBkdZ * uM) {
                    int bb2[kBkdZ], bb3[kBkdZ];
                    double2 w2[kBkdZ], w3[kBkdZ];
#pragma unroll
                    for (int i = 0; i < kBkdZ; ++i) {
                        bb2[i] = tb[t2 + i * uM];
                        bb3[i] = tb[t3 - i * uM];
                        w2[i] = tv[t2 + i * uM];
                        w3[i] = tv[t3 - i * uM];
                    }
                    // products for every step (keeps all loads ahead of the first use)
                    double val[kBkdZ];
#pragma unroll
                    for (int i = 0; i < kBkdZ; ++i) {
                        const double r12 = v1.x * w2[i].x - v1.y * w2[i].y;
                        const double i12 = v1.x * w2[i].y + v1.y * w2[i].x;
                        val[i] = r12 * w3[i].x - i12 * w3[i].y;
                    }
#pragma unroll
                    for (int i = 0; i < kBkdZ; ++i) {
                        // a skipped step keeps the key and adds nothing (no branch
                        // around the products, so their loads stay hoisted)
                        const bool ok = z + i <= zh && (bb2[i] | bb3[i]) >= 0;
                        const int k = ok ? bb2[i] * nb + bb3[i] : key;
                        if (k != key) {
                            if (n) {
                                atomicAdd(&s_re[key], re);
                                atomicAdd(&s_nk[key], n);
                            }
                            key = k;
                            re = 0.0;
                            n = 0;
                        }
                        re += ok ? val[i] : 0.0;
                        n += ok ? 1u : 0u;
                    }
                }
                if (n) {
                    atomicAdd(&s_re[key], re);
                    atomicAdd(&s_nk[key], n);
                }
            }
        }
        // fold this k1's pair counts: p1 n into the p(k1) sums, n into the totals
        __syncthreads();
        for (int i = threadIdx.x; i < nb2; i += blockDim.x) {
            const unsigned nk = s_nk[i];
            if (nk) {
                s_sn[i] += p1 * (double)nk;
                s_n[i] += nk;
                s_nk[i] = 0;
            }
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nb2; e += blockDim.x) {
        const unsigned n = s_n[e];
        if (!n) continue;
        int lo = s1, mid = e / nb, hi = e % nb;
        if (lo > mid) { const int tt = lo; lo = mid; mid = tt; }
        if (mid > hi) { const int tt = mid; mid = hi; hi = tt; }
        if (lo > mid) { const int tt = lo; lo = mid; mid = tt; }
        const size_t o = ((size_t)lo * nb + mid) * nb + hi;
        unsafeAtomicAdd(&bsum[o], s_re[e]);
        unsafeAtomicAdd(&psum3[o], s_sn[e]);
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
    static_assert(kBkdPad >= (kBkdZ - 1) * 1025u, "padding below the largest read-ahead");
    if ((s = scratch(ctx, "bkd.tv", sizeof(double2) * (G + 2 * kBkdPad), &tv)) != CM_OK) return s;
    if ((s = scratch(ctx, "bkd.tb", G + 2 * kBkdPad, &tb)) != CM_OK) return s;
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
    const size_t lds = (2 * sizeof(double) + 2 * sizeof(unsigned)) * (size_t)nb * nb;
    if (lds > ctx->lds_per_block)
        return fail(ctx, CM_ERR_UNSUPPORTED, "cm_bispectrum_direct: %d bins need %zu B LDS", nb, lds);
    // k1 grouped by bin, in chunks of K sharing one LDS tile (one global flush per
    // chunk instead of per k1); K keeps about 32 chunks per CU
    std::vector<unsigned long long> pc(nb);
    CM_HIP(ctx, hipMemcpyAsync(pc.data(), dcnt + nb3, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost,
                               ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    unsigned long long nk1 = 0;
    for (int i = 0; i < nb; ++i) nk1 += pc[i];
    if (nk1 > 0xffffffffull) return fail(ctx, CM_ERR_UNSUPPORTED, "cm_bispectrum_direct: too many modes");
    const unsigned long long target = (unsigned long long)std::max(1, ctx->cu_count) * 32;
    const int K = (int)std::max<unsigned long long>(1, std::min<unsigned long long>(64, nk1 / target));
    std::vector<unsigned> off(nb);
    std::vector<int4> chunks;
    unsigned acc = 0;
    for (int i = 0; i < nb; ++i) {
        off[i] = acc;
        for (unsigned long long f = 0; f < pc[i]; f += (unsigned long long)K)
            chunks.push_back(int4{i, (int)(acc + f), (int)std::min<unsigned long long>(K, pc[i] - f), 0});
        acc += (unsigned)pc[i];
    }
    if (chunks.size() > (1u << 24)) return fail(ctx, CM_ERR_UNSUPPORTED, "cm_bispectrum_direct: too many modes");
    if (!chunks.empty()) {
        void *dlist, *dcur, *dch;
        if ((s = scratch(ctx, "bkd.list", sizeof(unsigned) * (size_t)nk1, &dlist)) != CM_OK) return s;
        if ((s = scratch(ctx, "bkd.cursor", sizeof(unsigned) * (size_t)nb, &dcur)) != CM_OK) return s;
        if ((s = scratch(ctx, "bkd.chunks", sizeof(int4) * chunks.size(), &dch)) != CM_OK) return s;
        CM_HIP(ctx, hipMemcpyAsync(dcur, off.data(), sizeof(unsigned) * nb, hipMemcpyHostToDevice, ctx->stream));
        CM_HIP(ctx, hipMemcpyAsync(dch, chunks.data(), sizeof(int4) * chunks.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
        if ((s = launch(ctx, "bkd_list", k_bkd_list, dim3((unsigned)((G + 255) / 256)), dim3(256), 0,
                        (const int8_t *)tb, (unsigned)G, nb, (unsigned *)dcur, (unsigned *)dlist)) != CM_OK)
            return s;
        if ((s = launch(ctx, "bkd_pairs", k_bkd_pairs, dim3((unsigned)chunks.size()), dim3(256), lds,
                        (const double2 *)tv, (const int8_t *)tb, (const unsigned *)dlist, (const int4 *)dch, N, nb,
                        r2max, dsum, dsum + nb3, dcnt)) != CM_OK)
            return s;
    }
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
        const double snsum = 3.0 * hs[nb3 + e];  // see k_bkd_pairs
        if (snsum_out) snsum_out[e] = snsum;
        if (!m) {
            bk[e] = sn[e] = 0.0;
            continue;
        }
        // <P1+P2+P3>/nbar + 1/nbar^2 with P_i = |d_i|^2/normal_p - V/n
        const double snb = snc ? (snsum / (double)m / normal_p - 3.0 * sn_p) * sn_p + sn_p * sn_p : 0.0;
        sn[e] = snb;
        bk[e] = hs[e] / ((double)m * normal_b) - snb;
    }
    return CM_OK;
}
