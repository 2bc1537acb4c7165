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
// One kernel holds, for a chunk of cells, the three fields of every shell in
// LDS and accumulates all triangle bins at once, so each field is read once.
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
                                                  float2 *__restrict__ fp, int nz, int ny) {
    // writes the kz columns 0..nz-1 of the rows (kx, ky) with ky in
    // 0..(ny-1)/2 and N-(ny-1)/2..N-1 (ny odd, or ny = N: every row)
    const int M1 = N / 2 + 1, h = N / 2, yh = (ny - 1) / 2;
    const size_t total = (size_t)N * ny * nz;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(t % nz);
        const size_t rb = t / nz;
        const int yb = (int)(rb % ny), ix = (int)(rb / ny);
        const int iy = ny == N || yb <= yh ? yb : yb + N - ny;
        const size_t g = ((size_t)ix * N + iy) * M1 + iz;
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

// Triangle sums over cells, tiled over the bins.  A tile holds the bins
// (i, j, k) of one shell i, two consecutive shells j and four consecutive
// shells k (slots not in the bin list are computed and never read).  Per cell
// it forms the pair products DD = D_i D_j, UU = U_i U_j and Q = P_i U_j + U_i P_j
// once per j and adds, for every k,
//   DDD += DD D_k,   UUU += UU U_k,   PUU += Q U_k + UU P_k
// (PUU is the cyclic sum P_i U_j U_k + U_i P_j U_k + U_i U_j P_k).  A workgroup
// holds a chunk of C cells of all 3 ns fields in LDS (row f: C floats); each
// wave takes every nw-th tile, each lane two adjacent cells per step, so the
// 24 sums of a tile sit in registers for the whole chunk and are reduced
// across the wave once per chunk (two lane-half swaps, then row sums) into
// FP64 tile sums in LDS.  part[blockIdx][tile][slot][3] receives them.
constexpr int kBkJW = 2, kBkKW = 4, kBkSlots = kBkJW * kBkKW, kBkAcc = 3 * kBkSlots;

struct alignas(32) BkTile {
    int i, j[kBkJW], k[kBkKW], pad;  // shell indices (clamped to ns - 1)
};

__device__ __forceinline__ float bk_dpp_add(float v, int ctrl) {
    // v + (v of the lane that ctrl names); quad_perm, row_half_mirror and
    // row_mirror patterns stay inside a row of 16 lanes
    switch (ctrl) {
        case 0xB1: return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, true));
        case 0x4E: return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, true));
        case 0x141: return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, true));
        default: return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, true));
    }
}

template <int C>
__global__ void __launch_bounds__(1024) k_bk_sums(const float *__restrict__ fields, int ns, size_t n3,
                                                  const BkTile *__restrict__ tiles, int ntiles,
                                                  double *__restrict__ part) {
    static_assert(C % 128 == 0, "two cells per lane per step");
    extern __shared__ __attribute__((aligned(16))) float sf[];  // [3 ns][C], then double [ntiles][kBkAcc]
    const int nf = 3 * ns, nth = blockDim.x;
    double *accd = (double *)(sf + (size_t)nf * C);
    for (int q = threadIdx.x; q < ntiles * kBkAcc; q += nth) accd[q] = 0.0;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = nth >> 6;
    const size_t nchunks = (n3 + C - 1) / C;
    for (size_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const size_t x0 = c * C;
        __syncthreads();
        if ((n3 & 3) == 0 && x0 + C <= n3) {
            for (int t = threadIdx.x; t < nf * (C / 4); t += nth) {
                const int f = t / (C / 4), q = t % (C / 4);
                *(float4 *)(sf + f * C + 4 * q) = *(const float4 *)(fields + (size_t)f * n3 + x0 + 4 * q);
            }
        } else {
            for (int t = threadIdx.x; t < nf * C; t += nth) {
                const int f = t / C, xi = t % C;
                sf[f * C + xi] = x0 + xi < n3 ? fields[(size_t)f * n3 + x0 + xi] : 0.f;
            }
        }
        __syncthreads();
        for (int t = wave; t < ntiles; t += nw) {
            const BkTile tl = tiles[t];
            const float2 *base = (const float2 *)sf + lane;
            const float2 *Di = base + tl.i * (C / 2), *Ui = base + (ns + tl.i) * (C / 2),
                         *Pi = base + (2 * ns + tl.i) * (C / 2);
            const float2 *Dj[kBkJW], *Uj[kBkJW], *Pj[kBkJW], *Dk[kBkKW], *Uk[kBkKW], *Pk[kBkKW];
#pragma unroll
            for (int a = 0; a < kBkJW; ++a) {
                Dj[a] = base + tl.j[a] * (C / 2);
                Uj[a] = base + (ns + tl.j[a]) * (C / 2);
                Pj[a] = base + (2 * ns + tl.j[a]) * (C / 2);
            }
#pragma unroll
            for (int a = 0; a < kBkKW; ++a) {
                Dk[a] = base + tl.k[a] * (C / 2);
                Uk[a] = base + (ns + tl.k[a]) * (C / 2);
                Pk[a] = base + (2 * ns + tl.k[a]) * (C / 2);
            }
            float acc[kBkAcc];
#pragma unroll
            for (int a = 0; a < kBkAcc; ++a) acc[a] = 0.f;
#pragma unroll
            for (int st = 0; st < C / 128; ++st) {
                const int o = st * 64;
                const float2 di = Di[o], ui = Ui[o], pi = Pi[o];
                float2 dk[kBkKW], uk[kBkKW], pk[kBkKW];
#pragma unroll
                for (int b = 0; b < kBkKW; ++b) {
                    dk[b] = Dk[b][o];
                    uk[b] = Uk[b][o];
                    pk[b] = Pk[b][o];
                }
#pragma unroll
                for (int a = 0; a < kBkJW; ++a) {
                    const float2 dj = Dj[a][o], uj = Uj[a][o], pj = Pj[a][o];
                    const float ddx = di.x * dj.x, uux = ui.x * uj.x, qx = fmaf(pi.x, uj.x, ui.x * pj.x);
                    const float ddy = di.y * dj.y, uuy = ui.y * uj.y, qy = fmaf(pi.y, uj.y, ui.y * pj.y);
#pragma unroll
                    for (int b = 0; b < kBkKW; ++b) {
                        float *r = acc + 3 * (a * kBkKW + b);
                        r[0] = fmaf(ddy, dk[b].y, fmaf(ddx, dk[b].x, r[0]));
                        r[1] = fmaf(uuy, uk[b].y, fmaf(uux, uk[b].x, r[1]));
                        r[2] = fmaf(qy, uk[b].y, fmaf(uuy, pk[b].y, fmaf(qx, uk[b].x, fmaf(uux, pk[b].x, r[2]))));
                    }
                }
            }
            // 24 sums x 64 lanes -> lane 16 r + a holds sum a + 6 r (a < 6):
            // lanes 32-63 of acc[a] trade with lanes 0-31 of acc[a + 12], then
            // odd rows of r1[a] with even rows of r1[a + 6]; rows then sum up
            float r1[kBkAcc / 2], r2[kBkAcc / 4];
#pragma unroll
            for (int a = 0; a < kBkAcc / 2; ++a) {
                const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc[a]),
                                                                __float_as_uint(acc[a + kBkAcc / 2]), false, false);
                r1[a] = __uint_as_float(p[0]) + __uint_as_float(p[1]);
            }
#pragma unroll
            for (int a = 0; a < kBkAcc / 4; ++a) {
                const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(r1[a]),
                                                                __float_as_uint(r1[a + kBkAcc / 4]), false, false);
                float v = __uint_as_float(p[0]) + __uint_as_float(p[1]);
                v = bk_dpp_add(v, 0xB1);
                v = bk_dpp_add(v, 0x4E);
                v = bk_dpp_add(v, 0x141);
                r2[a] = bk_dpp_add(v, 0x140);
            }
            const int a16 = lane & 15;
            float mine = r2[0];
#pragma unroll
            for (int a = 1; a < kBkAcc / 4; ++a) mine = a16 == a ? r2[a] : mine;
            if (a16 < kBkAcc / 4) accd[t * kBkAcc + a16 + (kBkAcc / 4) * (lane >> 4)] += (double)mine;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < ntiles * kBkAcc; q += nth) part[(size_t)blockIdx.x * ntiles * kBkAcc + q] = accd[q];
}

// out[3 b + q] = sum over workgroups of the tile sum at slot pos[b].
__global__ void __launch_bounds__(256) k_bk_reduce(const double *__restrict__ part, int nblk, int rowlen,
                                                   const int *__restrict__ pos, int nbins,
                                                   double *__restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= 3 * nbins) return;
    const int at = 3 * pos[b / 3] + b % 3;
    double s = 0.0;
    for (int g = 0; g < nblk; ++g) s += part[(size_t)g * rowlen + at];
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

// Tiles of kBkJW x kBkKW bins over the bin list (bins of one shell i; j
// windows from the first j of that i, k windows from the first k of the j
// window); pos[b] = tile * kBkSlots + slot of bin b.
static void bin_tiles(const std::vector<int4> &tri, int ns, std::vector<BkTile> &tiles, std::vector<int> &pos) {
    tiles.clear();
    pos.assign(tri.size(), -1);
    size_t b0 = 0;
    while (b0 < tri.size()) {
        const int i = tri[b0].x;
        size_t b1 = b0;
        while (b1 < tri.size() && tri[b1].x == i) ++b1;
        for (int j0 = tri[b0].y; j0 <= i; j0 += kBkJW) {
            int kmin = ns, kmax = -1;
            for (size_t b = b0; b < b1; ++b)
                if (tri[b].y >= j0 && tri[b].y < j0 + kBkJW) {
                    kmin = std::min(kmin, tri[b].z);
                    kmax = std::max(kmax, tri[b].z);
                }
            for (int k0 = kmin; k0 <= kmax; k0 += kBkKW) {
                BkTile t{};
                t.i = i;
                for (int a = 0; a < kBkJW; ++a) t.j[a] = std::min(j0 + a, ns - 1);
                for (int a = 0; a < kBkKW; ++a) t.k[a] = std::min(k0 + a, ns - 1);
                bool any = false;
                for (size_t b = b0; b < b1; ++b) {
                    const int jj = tri[b].y - j0, kk = tri[b].z - k0;
                    if (jj >= 0 && jj < kBkJW && kk >= 0 && kk < kBkKW) {
                        pos[b] = (int)tiles.size() * kBkSlots + jj * kBkKW + kk;
                        any = true;
                    }
                }
                if (any) tiles.push_back(t);
            }
        }
        b0 = b1;
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
    std::vector<BkTile> tiles;
    std::vector<int> pos;
    bin_tiles(tri, ns, tiles, pos);
    const int ntl = (int)tiles.size();
    for (int v : pos)
        if (v < 0) return fail(ctx, CM_ERR_UNSUPPORTED, "cm_bispectrum: triangle bin outside the tiles");
    // cells per chunk: the largest of 512, 256, 128 whose fields and FP64 tile
    // sums fit the LDS of one workgroup
    int C = 0;
    size_t lds = 0;
    for (int c : {512, 256, 128}) {
        lds = sizeof(float) * 3 * (size_t)ns * c + sizeof(double) * kBkAcc * (size_t)ntl;
        if (lds <= ctx->lds_per_block) {
            C = c;
            break;
        }
    }
    if (!C) return fail(ctx, CM_ERR_UNSUPPORTED, "cm_bispectrum: %d shells need %zu B LDS", ns, lds);
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
    // every shell mode has |kx|, |ky|, kz < Dmin + ns ndel <= K: the masks
    // write the modes |ky| <= K, kz <= K (zeros outside the shell), the only
    // ones the band-limited c2r reads
    const int K = std::min(N / 2, (int)std::ceil(Dmin + ns * ndel));
    const int nz = std::min(K + 1, N / 2 + 1), ny = 2 * K + 1 < N ? 2 * K + 1 : N;
    for (int sh = 0; sh < ns; ++sh) {
        if ((s = launch(ctx, "bk_masks", k_bk_masks, dim3((unsigned)g), dim3(256), 0, (const float2 *)dk, N, sh,
                        int_rule, (int)Di, (int)ndi, Dmin, ndel, dtab, sn_norm, 1.0 / normal, sp,
                        sp + nh, sp + 2 * nh, nz, ny)) != CM_OK)
            return s;
        for (int f = 0; f < 3; ++f)
            if ((s = fft_c2r_band(ctx, N, CM_F32, sp + f * nh, fl + ((size_t)f * ns + sh) * n3, 1.0 / (double)n3,
                                  K)) != CM_OK)
                return s;
    }
    void *dpos;
    if ((s = scratch(ctx, "bk.tri", sizeof(BkTile) * (size_t)ntl, &dtri)) != CM_OK) return s;
    if ((s = scratch(ctx, "bk.pos", sizeof(int) * (size_t)nb, &dpos)) != CM_OK) return s;
    CM_HIP(ctx, host_upload(ctx, dtri, tiles.data(), sizeof(BkTile) * (size_t)ntl));
    CM_HIP(ctx, host_upload(ctx, dpos, pos.data(), sizeof(int) * (size_t)nb));
    const size_t nchunks = (n3 + C - 1) / C;
    const int nblk = (int)std::min<size_t>(nchunks, (size_t)std::max(1, ctx->cu_count));
    const int rowlen = kBkAcc * ntl;
    if ((s = scratch(ctx, "bk.part", sizeof(double) * (size_t)nblk * rowlen, &part)) != CM_OK) return s;
    if ((s = scratch(ctx, "bk.sum", sizeof(double) * 3 * (size_t)nb, &dsum)) != CM_OK) return s;
    auto kern = C == 512 ? k_bk_sums<512> : C == 256 ? k_bk_sums<256> : k_bk_sums<128>;
    if ((s = launch(ctx, "bk_sums", kern, dim3(nblk), dim3(1024), lds, (const float *)fl, ns, n3,
                    (const BkTile *)dtri, ntl, (double *)part)) != CM_OK)
        return s;
    if ((s = launch(ctx, "bk_reduce", k_bk_reduce, dim3((3 * nb + 255) / 256), dim3(256), 0, (const double *)part, nblk,
                    rowlen, (const int *)dpos, nb, (double *)dsum)) != CM_OK)
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
