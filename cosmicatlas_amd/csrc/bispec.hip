// Bispectrum of a periodic box by shell inverse transforms (gfx950).
//
// The estimator of FftwFunctions::get_bispectrum_fkp_fast
// (Source/FftwFunctions.cpp:3358-3718, 4878): for every k shell s three fields
// are transformed back to the mesh,
//   D_s(x) = c2r(delta_k / W_k  on the shell),  U_s(x) = c2r(1 on the shell),
//   P_s(x) = c2r(P_k on the shell)    (P_k: the per-mode power estimate),
// and for every triangle bin (i >= j >= k, the reference's loop bounds)
//   B_ijk = sum_x D_i D_j D_k / (normal_b sum_x U_i U_j U_k),
//   T_ijk = N^6 sum_x U_i U_j U_k   (closed triangles k1 + k2 + k3 = 0),
// because sum_x of a product of c2r fields keeps only closed triangles.  The
// shot noise subtracted is the Poisson one, <P_i + P_j + P_k>_tri / nbar +
// 1/nbar^2, with the triangle average taken the same way.  normal_b = n^3/V^2
// (raw_sampling, :159-174).
//
// One kernel holds, for a chunk of 128 cells, the three fields of every shell
// in LDS and accumulates all triangle bins at once, so each field is read once.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cm_internal.hpp"

namespace cm {

// Shell fields of shell s on the FFTW half spectrum: exact integer shell rule
// when kmin and dk are integer multiples of k_f (shell of q = |n|^2 with
// (s ndel + Dmin)^2 <= q < ((s+1) ndel + Dmin)^2), else the FP64 rule.
__global__ void __launch_bounds__(256) k_bk_masks(const float2 *__restrict__ dk, int N, int s, int int_rule,
                                                  int Dmin_i, int ndel_i, double Dmin, double ndel,
                                                  const double *__restrict__ tabs, double sn_norm, double inv_normal,
                                                  float2 *__restrict__ fd, float2 *__restrict__ fu,
                                                  float2 *__restrict__ fp) {
    const int M1 = N / 2 + 1, h = N / 2;
    const size_t total = (size_t)N * N * M1;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(g % M1);
        const size_t r = g / M1;
        const int iy = (int)(r % N), ix = (int)(r / N);
        const int fx = ix <= h ? ix : N - ix, fy = iy <= h ? iy : N - iy;
        const int q = fx * fx + fy * fy + iz * iz;
        bool in = false;
        if (q > 0) {
            if (int_rule) {
                const int lo = s * ndel_i + Dmin_i, hi = lo + ndel_i;
                in = q >= lo * lo && q < hi * hi;
            } else {
                const double t = (sqrt((double)q) - Dmin) / ndel;
                in = t >= (double)s && t < (double)(s + 1);
            }
        }
        float2 d = {0.f, 0.f}, u = {0.f, 0.f}, p = {0.f, 0.f};
        if (in) {
            const int M1t = M1;
            const double ic = tabs[fx] * tabs[fy] * tabs[iz];  // 1/W^2
            const double sq = tabs[M1t + fx] * tabs[M1t + fy] * tabs[M1t + iz];
            const double iw = sqrt(ic);
            const float2 v = dk[g];
            d = float2{(float)(v.x * iw), (float)(v.y * iw)};
            u = float2{1.f, 0.f};
            const double pw = ((double)v.x * v.x + (double)v.y * v.y - sn_norm * sq) * ic * inv_normal;
            p = float2{(float)pw, 0.f};
        }
        fd[g] = d;
        fu[g] = u;
        fp[g] = p;
    }
}

constexpr int kBkChunk = 128;
// LDS row pitch of one field: one word of padding, so that lanes reading the
// same cell of different shells hit different banks
constexpr int kBkPitch = kBkChunk + 1;

// Triangle sums over cells.  fields: [3][ns][n3] (D, U, P); tri: (i, j, k) per
// bin; part[blockIdx][bin][3] = partial sums of DDD, UUU and the cyclic PUU.
__global__ void __launch_bounds__(256) k_bk_sums(const float *__restrict__ fields, int ns, size_t n3,
                                                 const int4 *__restrict__ tri, int nbins,
                                                 double *__restrict__ part) {
    extern __shared__ float sf[];  // [3][ns][kBkPitch]
    const int tid = threadIdx.x;
    const int NT = blockDim.x;
    // each thread owns bins tid, tid + NT, ... (at most 8 per thread here)
    double acc[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a) acc[a][0] = acc[a][1] = acc[a][2] = 0.0;
    const size_t nchunks = (n3 + kBkChunk - 1) / kBkChunk;
    for (size_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const size_t x0 = c * kBkChunk;
        const int len = (int)std::min<size_t>(kBkChunk, n3 - x0);
        __syncthreads();
        for (int t = tid; t < 3 * ns * kBkChunk; t += NT) {
            const int f = t / kBkChunk, xi = t % kBkChunk;
            sf[f * kBkPitch + xi] = xi < len ? fields[(size_t)f * n3 + x0 + xi] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const int b = tid + a * NT;
            if (b >= nbins) break;
            const int4 t = tri[b];
            const float *Di = sf + t.x * kBkPitch, *Dj = sf + t.y * kBkPitch, *Dk = sf + t.z * kBkPitch;
            const float *Ui = sf + (ns + t.x) * kBkPitch, *Uj = sf + (ns + t.y) * kBkPitch,
                        *Uk = sf + (ns + t.z) * kBkPitch;
            const float *Pi = sf + (2 * ns + t.x) * kBkPitch, *Pj = sf + (2 * ns + t.y) * kBkPitch,
                        *Pk = sf + (2 * ns + t.z) * kBkPitch;
            float s0 = 0.f, s1 = 0.f, s2 = 0.f;
            double d0 = 0.0, d1 = 0.0, d2 = 0.0;
            for (int xi = 0; xi < kBkChunk; ++xi) {
                const float ui = Ui[xi], uj = Uj[xi], uk = Uk[xi];
                s0 += Di[xi] * Dj[xi] * Dk[xi];
                s1 += ui * uj * uk;
                s2 += Pi[xi] * uj * uk + Pj[xi] * ui * uk + Pk[xi] * ui * uj;
                if ((xi & 15) == 15) {  // FP64 every 16 terms
                    d0 += s0;
                    d1 += s1;
                    d2 += s2;
                    s0 = s1 = s2 = 0.f;
                }
            }
            acc[a][0] += d0;
            acc[a][1] += d1;
            acc[a][2] += d2;
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int b = tid + a * NT;
        if (b >= nbins) break;
        double *o = part + ((size_t)blockIdx.x * nbins + b) * 3;
        o[0] = acc[a][0];
        o[1] = acc[a][1];
        o[2] = acc[a][2];
    }
}

__global__ void __launch_bounds__(256) k_bk_reduce(const double *__restrict__ part, int nblk, int nbins,
                                                   double *__restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= 3 * nbins) return;
    double s = 0.0;
    for (int g = 0; g < nblk; ++g) s += part[(size_t)g * 3 * nbins + b];
    out[b] = s;
}

// Triangle bins of the reference's loop (Source/FftwFunctions.cpp:3529-3544).
static void triangle_bins(int ns, int smin, std::vector<int4> &tri) {
    tri.clear();
    for (int i = smin; i < ns; ++i) {
        const int nstep2 = i / 2;
        const int jstart = nstep2 == 0 ? i : nstep2;
        for (int j = jstart; j <= i; ++j) {
            const int nstep3 = i - j - 1;
            int kstart = nstep3 > j ? j : nstep3;
            if (nstep3 <= 0) kstart = smin;
            for (int k = kstart; k <= j; ++k) tri.push_back(int4{i, j, k, 0});
        }
    }
}

}  // namespace cm

using namespace cm;

extern "C" cm_status cm_bispectrum(cm_ctx *ctx, const cm_bk_opts *o, const void *dk, int max_bins, int32_t *ijk,
                                   double *bk, double *sn, double *ntri, int *nbins) {
    if (!ctx) return CM_ERR_INVALID;
    if (!o || !dk || !nbins) return fail(ctx, CM_ERR_INVALID, "cm_bispectrum: null pointer");
    const int N = o->N, ns = o->nshells;
    if (!fft_size_ok(N)) return fail(ctx, CM_ERR_INVALID, "cm_bispectrum: N=%d not supported", N);
    if (ns < 1 || ns > 64 || o->shell_min < 0 || o->shell_min >= ns || !(o->dk > 0) || !(o->n_obj > 0))
        return fail(ctx, CM_ERR_INVALID, "cm_bispectrum: shells (1..64), shell_min, dk or n_obj");
    std::vector<int4> tri;
    triangle_bins(ns, o->shell_min, tri);
    const int nb = (int)tri.size();
    *nbins = nb;
    if (nb > 8 * 256) return fail(ctx, CM_ERR_UNSUPPORTED, "cm_bispectrum: %d triangle bins (at most 2048)", nb);
    if (max_bins < nb || !ijk || !bk || !sn || !ntri)
        return fail(ctx, CM_ERR_INVALID, "cm_bispectrum: output arrays hold %d bins, %d needed", max_bins, nb);
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * (N / 2 + 1);
    const size_t lds = sizeof(float) * 3 * (size_t)ns * kBkPitch;
    if (lds > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "cm_bispectrum: %d shells need %zu B LDS", ns, lds);
    cm_status s;
    void *fields, *specs, *dtri, *part, *dsum;
    if ((s = scratch(ctx, "bk.fields", sizeof(float) * 3 * (size_t)ns * n3, &fields)) != CM_OK) return s;
    if ((s = scratch(ctx, "bk.specs", sizeof(float2) * 3 * nh, &specs)) != CM_OK) return s;
    const double *dtab = pk_tables_dev(ctx, N, (int)o->scheme, o->mas_correction, &s);
    if (!dtab) return s;
    const double L = o->L, V = L * L * L, kf = 2.0 * M_PI / L;
    const double n = o->n_obj, normal = n * n / V;
    const double Dmin = o->kmin / kf, ndel = o->dk / kf;
    // closed triangles on the mesh are sums k1 + k2 + k3 = 0 mod N: below N/3
    // per component no wrapped triangle exists
    if (Dmin + ns * ndel >= N / 3.0)
        return fail(ctx, CM_ERR_INVALID, "cm_bispectrum: shells must end below N/3 = %.2f k_f (here %.2f k_f)", N / 3.0,
                    Dmin + ns * ndel);
    const long Di = std::lround(Dmin), ndi = std::lround(ndel);
    const int int_rule = (std::fabs(Dmin - Di) < 1e-9 && std::fabs(ndel - ndi) < 1e-9 && ndi > 0 && Di >= 0 &&
                          Di + (long)(ns + 1) * ndi < 40000)
                             ? 1
                             : 0;
    const double sn_norm = o->sn_correction ? n : 0.0;  // Poisson |delta|^2 of n objects
    auto *sp = (float2 *)specs;
    auto *fl = (float *)fields;
    size_t g = (nh + 255) / 256;
    if (g > 16384) g = 16384;
    for (int sh = 0; sh < ns; ++sh) {
        if ((s = launch(ctx, "bk_masks", k_bk_masks, dim3((unsigned)g), dim3(256), 0, (const float2 *)dk, N, sh,
                        int_rule, (int)Di, (int)ndi, Dmin, ndel, dtab, sn_norm, 1.0 / normal, sp,
                        sp + nh, sp + 2 * nh)) != CM_OK)
            return s;
        for (int f = 0; f < 3; ++f)
            if ((s = fft_c2r_impl(ctx, N, CM_F32, sp + f * nh, fl + ((size_t)f * ns + sh) * n3, 1.0 / (double)n3)) !=
                CM_OK)
                return s;
    }
    if ((s = scratch(ctx, "bk.tri", sizeof(int4) * (size_t)nb, &dtri)) != CM_OK) return s;
    CM_HIP(ctx, hipMemcpyAsync(dtri, tri.data(), sizeof(int4) * (size_t)nb, hipMemcpyHostToDevice, ctx->stream));
    const int nthreads = 256;
    const size_t nchunks = (n3 + kBkChunk - 1) / kBkChunk;
    const int nblk = (int)std::min<size_t>(nchunks, (size_t)std::max(1, ctx->cu_count) * 4);
    if ((s = scratch(ctx, "bk.part", sizeof(double) * 3 * (size_t)nblk * nb, &part)) != CM_OK) return s;
    if ((s = scratch(ctx, "bk.sum", sizeof(double) * 3 * (size_t)nb, &dsum)) != CM_OK) return s;
    if ((s = launch(ctx, "bk_sums", k_bk_sums, dim3(nblk), dim3(nthreads), lds, (const float *)fl, ns, n3,
                    (const int4 *)dtri, nb, (double *)part)) != CM_OK)
        return s;
    if ((s = launch(ctx, "bk_reduce", k_bk_reduce, dim3((3 * nb + 255) / 256), dim3(256), 0, (const double *)part, nblk,
                    nb, (double *)dsum)) != CM_OK)
        return s;
    std::vector<double> h(3 * (size_t)nb);
    CM_HIP(ctx, hipMemcpyAsync(h.data(), dsum, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double normal_b = n * n * n / (V * V), nbar = n / V;
    for (int b = 0; b < nb; ++b) {
        ijk[3 * b] = tri[b].x;
        ijk[3 * b + 1] = tri[b].y;
        ijk[3 * b + 2] = tri[b].z;
        const double ddd = h[3 * b], uuu = h[3 * b + 1], puu = h[3 * b + 2];
        const double T = uuu * (double)n3 * (double)n3;  // sum_x U U U = T / N^6
        ntri[b] = std::round(T);
        if (T < 0.5) {
            bk[b] = sn[b] = 0.0;
            continue;
        }
        const double snb = o->sn_correction ? (puu / uuu) / nbar + 1.0 / (nbar * nbar) : 0.0;
        sn[b] = snb;
        bk[b] = ddd / (uuu * normal_b) - snb;
    }
    return CM_OK;
}
