// Particle -> mesh mass assignment (NGP/CIC/TSC/PCS) for gfx950.
//
// plain : one thread per particle, 1/8/27/125 global atomics (the plain GPU
//         path; any dtype).
// tiled : float32 positions and mesh.  Particles are binned to T^3-cell tiles
//         (histogram and slot cursors in LDS per workgroup, so no global
//         atomics; integer LDS atomics are full rate on gfx950);
//         a particle whose stencil reaches a neighbouring tile is also copied to
//         that tile ("ghost" copy).  One workgroup per (tile, chunk) accumulates
//         its particles into an LDS grid (T+2R)^3 with 64-bit fixed-point LDS
//         atomics and writes
//         the T^3 interior with plain coalesced stores, so the mesh needs no
//         zero fill and sees no global atomics (tiles holding more than CH
//         particles are split into chunks that add with global atomics).
#include <algorithm>
#include <cmath>

#include "cm_internal.hpp"
#include "mas_common.hpp"

namespace cm {

// ---------------------------------------------------------------- plain path
template <typename PT, typename MT>
__global__ void __launch_bounds__(256) k_mas_plain(const PT *__restrict__ pos, const PT *__restrict__ w,
                                                   MasK p, MT *__restrict__ mesh,
                                                   unsigned long long *dropped) {
    const size_t N = (size_t)p.N;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n;
         i += (size_t)gridDim.x * blockDim.x) {
        int c[3];
        MT f[3];
        bool ok = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            PT s = (PT)0;
            ok = ok && wrap_coord<PT>(load_coord(pos, i, a, p.soa, p.n), p.xmin[a], p.L, s);
            double x = shifted((double)s, p.ishift, p.L);
            c[a] = cell_of(x, p.d, p.N);
            f[a] = (MT)(x / p.d - (double)c[a] - 0.5);
        }
        if (!ok) {
            atomicAdd(dropped, 1ull);
            continue;
        }
        const MT wt = w ? (MT)w[i] : (MT)1;
        if (p.scheme == 0) {
            unsafeAtomicAdd(&mesh[((size_t)c[0] * N + c[1]) * N + c[2]], wt);
            continue;
        }
        MT wx[5], wy[5], wz[5];
        axis_weights<MT>(p.scheme, f[0], wx);
        axis_weights<MT>(p.scheme, f[1], wy);
        axis_weights<MT>(p.scheme, f[2], wz);
        const int R = mas_radius(p.scheme);
        for (int a = -R; a <= R; ++a) {
            MT va = wt * wx[a + 2];
            if (va == (MT)0) continue;
            size_t ia = (size_t)((c[0] + a + p.N) % p.N);
            for (int b = -R; b <= R; ++b) {
                MT vb = va * wy[b + 2];
                if (vb == (MT)0) continue;
                size_t ib = (size_t)((c[1] + b + p.N) % p.N);
                for (int e = -R; e <= R; ++e) {
                    MT v = vb * wz[e + 2];
                    if (v == (MT)0) continue;
                    size_t ie = (size_t)((c[2] + e + p.N) % p.N);
                    unsafeAtomicAdd(&mesh[(ia * N + ib) * N + ie], v);
                }
            }
        }
    }
}

template <typename T>
__global__ void k_fill(T *__restrict__ a, size_t n, T v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// ---------------------------------------------------------------- tiled path
struct TileK {
    int T;      // tile edge in cells
    int R;      // stencil radius
    int ntx;    // tiles per side
    int CH;     // max particles per workgroup
};

// Tiles touched by a particle in cell c along one axis: home tile and, when the
// stencil crosses the tile edge, the neighbour.  Returns the count (1 or 2).
__device__ __forceinline__ int axis_tiles(int c, const TileK &t, int out[2]) {
    int home = c / t.T;
    int lc = c - home * t.T;
    out[0] = home;
    if (lc < t.R) { out[1] = (home - 1 + t.ntx) % t.ntx; return 2; }
    if (lc >= t.T - t.R) { out[1] = (home + 1) % t.ntx; return 2; }
    return 1;
}

// Cell and in-cell offset f = x/d - c - 1/2 (FP64, as the plain kernel) of a
// particle; s[] receives the offsets rounded to float.
__device__ __forceinline__ bool particle_cells(const float *pos, size_t i, const MasK &p, int c[3], float s[3]) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float v = 0.0f;
        ok = ok && wrap_coord<float>(load_coord(pos, i, a, p.soa, p.n), p.xmin[a], p.L, v);
        const double x = shifted((double)v, p.ishift, p.L);
        c[a] = cell_of(x, p.d, p.N);
        s[a] = (float)(x / p.d - (double)c[a] - 0.5);
    }
    return ok;
}

// Sorted record: packed global cell (10 bits per axis) and the three offsets.
__device__ __forceinline__ float4 make_record(const int c[3], const float f[3]) {
    const unsigned pk = (unsigned)c[0] | ((unsigned)c[1] << 10) | ((unsigned)c[2] << 20);
    return make_float4(__uint_as_float(pk), f[0], f[1], f[2]);
}

__global__ void __launch_bounds__(256) k_tile_count(const float *__restrict__ pos, MasK p, TileK t,
                                                    unsigned *__restrict__ counts,
                                                    unsigned long long *dropped) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n;
         i += (size_t)gridDim.x * blockDim.x) {
        int c[3];
        float s[3];
        if (!particle_cells(pos, i, p, c, s)) {
            atomicAdd(dropped, 1ull);
            continue;
        }
        int tx[2], ty[2], tz[2];
        int nx = axis_tiles(c[0], t, tx), ny = axis_tiles(c[1], t, ty), nz = axis_tiles(c[2], t, tz);
        for (int a = 0; a < nx; ++a)
            for (int b = 0; b < ny; ++b)
                for (int e = 0; e < nz; ++e)
                    atomicAdd(&counts[((size_t)tx[a] * t.ntx + ty[b]) * t.ntx + tz[e]], 1u);
    }
}

// Single-workgroup exclusive scan of the tile counts into offsets (and cursors),
// and of the per-tile chunk counts into the work list.
__global__ void __launch_bounds__(1024) k_tile_scan(const unsigned *__restrict__ counts, int ntiles, int CH,
                                                    unsigned *__restrict__ offsets,
                                                    unsigned *__restrict__ cursor,
                                                    unsigned *__restrict__ item_off,
                                                    unsigned *__restrict__ nitems) {
    __shared__ unsigned long long sA[1024];
    __shared__ unsigned sB[1024];
    const int tid = threadIdx.x, NT = blockDim.x;
    const int per = (ntiles + NT - 1) / NT;
    const int beg = min(ntiles, tid * per), end = min(ntiles, beg + per);
    unsigned long long a = 0;
    unsigned b = 0;
    for (int i = beg; i < end; ++i) {
        unsigned cnt = counts[i];
        a += cnt;
        b += cnt == 0 ? 1u : (cnt + CH - 1) / CH;
    }
    sA[tid] = a;
    sB[tid] = b;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
        unsigned long long va = tid >= off ? sA[tid - off] : 0;
        unsigned vb = tid >= off ? sB[tid - off] : 0;
        __syncthreads();
        sA[tid] += va;
        sB[tid] += vb;
        __syncthreads();
    }
    unsigned long long ra = sA[tid] - a;
    unsigned rb = sB[tid] - b;
    for (int i = beg; i < end; ++i) {
        unsigned cnt = counts[i];
        offsets[i] = (unsigned)ra;
        cursor[i] = (unsigned)ra;
        item_off[i] = rb;
        ra += cnt;
        rb += cnt == 0 ? 1u : (cnt + CH - 1) / CH;
    }
    if (tid == NT - 1) {
        offsets[ntiles] = (unsigned)sA[NT - 1];
        item_off[ntiles] = sB[NT - 1];
        *nitems = sB[NT - 1];
    }
}

__global__ void __launch_bounds__(256) k_tile_items(const unsigned *__restrict__ item_off, int ntiles,
                                                    uint2 *__restrict__ items) {
    int tile = blockIdx.x * blockDim.x + threadIdx.x;
    if (tile >= ntiles) return;
    unsigned b = item_off[tile], e = item_off[tile + 1];
    for (unsigned k = b; k < e; ++k) items[k] = make_uint2((unsigned)tile, k - b);
}

__global__ void __launch_bounds__(256) k_tile_scatter(const float *__restrict__ pos, const float *__restrict__ w,
                                                      MasK p, TileK t, unsigned *__restrict__ cursor,
                                                      float4 *__restrict__ sorted, float *__restrict__ wsorted,
                                                      unsigned cap, unsigned *__restrict__ overflow) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n;
         i += (size_t)gridDim.x * blockDim.x) {
        int c[3];
        float s[3];
        if (!particle_cells(pos, i, p, c, s)) continue;
        const float4 rec = make_record(c, s);
        int tx[2], ty[2], tz[2];
        int nx = axis_tiles(c[0], t, tx), ny = axis_tiles(c[1], t, ty), nz = axis_tiles(c[2], t, tz);
        for (int a = 0; a < nx; ++a)
            for (int b = 0; b < ny; ++b)
                for (int e = 0; e < nz; ++e) {
                    unsigned slot = atomicAdd(&cursor[((size_t)tx[a] * t.ntx + ty[b]) * t.ntx + tz[e]], 1u);
                    if (slot < cap) {
                        sorted[slot] = rec;
                        if (w) wsorted[slot] = w[i];
                    } else {
                        atomicOr(overflow, 1u);
                    }
                }
    }
}


// Count with an LDS histogram per workgroup (no global atomics): workgroup g
// owns particles [g*chunk, (g+1)*chunk) and writes its histogram row
// hist[g*ntiles + tile] with coalesced stores.
__global__ void __launch_bounds__(1024) k_tile_count_lds(const float *__restrict__ pos, MasK p, TileK t, int ntiles,
                                                         size_t chunk, unsigned *__restrict__ hist,
                                                         unsigned long long *dropped) {
    extern __shared__ unsigned lh[];
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x) lh[i] = 0u;
    __syncthreads();
    const size_t i0 = (size_t)blockIdx.x * chunk, i1 = min(p.n, i0 + chunk);
    unsigned ndrop = 0;
    for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        int c[3];
        float s[3];
        if (!particle_cells(pos, i, p, c, s)) {
            ++ndrop;
            continue;
        }
        int tx[2], ty[2], tz[2];
        int nx = axis_tiles(c[0], t, tx), ny = axis_tiles(c[1], t, ty), nz = axis_tiles(c[2], t, tz);
        for (int a = 0; a < nx; ++a)
            for (int b = 0; b < ny; ++b)
                for (int e = 0; e < nz; ++e) atomicAdd(&lh[(tx[a] * t.ntx + ty[b]) * t.ntx + tz[e]], 1u);
    }
    if (ndrop) atomicAdd(dropped, (unsigned long long)ndrop);
    __syncthreads();
    unsigned *row = hist + (size_t)blockIdx.x * ntiles;
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x) row[i] = lh[i];
}

// Per tile: exclusive prefix over the workgroups' histogram rows (in place) and
// the tile total.  One thread per tile; for a fixed row consecutive threads
// touch consecutive words.
__global__ void __launch_bounds__(256) k_tile_prefix(unsigned *__restrict__ hist, int nrows, int ntiles,
                                                     unsigned *__restrict__ totals) {
    const int tile = blockIdx.x * blockDim.x + threadIdx.x;
    if (tile >= ntiles) return;
    unsigned run = 0;
    for (int g = 0; g < nrows; ++g) {
        unsigned *h = hist + (size_t)g * ntiles + tile;
        const unsigned v = *h;
        *h = run;
        run += v;
    }
    totals[tile] = run;
}

// Scatter with LDS cursors: cursor[tile] = offsets[tile] + prefix of earlier
// workgroups; integer LDS atomics hand out the slots.
__global__ void __launch_bounds__(1024) k_tile_scatter_lds(const float *__restrict__ pos, const float *__restrict__ w,
                                                           MasK p, TileK t, int ntiles, size_t chunk,
                                                           const unsigned *__restrict__ pref,
                                                           const unsigned *__restrict__ offsets,
                                                           float4 *__restrict__ sorted, float *__restrict__ wsorted,
                                                           unsigned cap, unsigned *__restrict__ overflow) {
    extern __shared__ unsigned cur[];
    const unsigned *row = pref + (size_t)blockIdx.x * ntiles;
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x) cur[i] = offsets[i] + row[i];
    __syncthreads();
    const size_t i0 = (size_t)blockIdx.x * chunk, i1 = min(p.n, i0 + chunk);
    for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        int c[3];
        float s[3];
        if (!particle_cells(pos, i, p, c, s)) continue;
        const float4 rec = make_record(c, s);
        int tx[2], ty[2], tz[2];
        int nx = axis_tiles(c[0], t, tx), ny = axis_tiles(c[1], t, ty), nz = axis_tiles(c[2], t, tz);
        for (int a = 0; a < nx; ++a)
            for (int b = 0; b < ny; ++b)
                for (int e = 0; e < nz; ++e) {
                    const unsigned slot = atomicAdd(&cur[(tx[a] * t.ntx + ty[b]) * t.ntx + tz[e]], 1u);
                    if (slot < cap) {
                        sorted[slot] = rec;
                        if (w) wsorted[slot] = w[i];
                    } else {
                        atomicOr(overflow, 1u);
                    }
                }
    }
}

// Zero the mesh cells of tiles that will be deposited by several workgroups.
__global__ void __launch_bounds__(256) k_tile_zero_split(const unsigned *__restrict__ offsets, TileK t, int N,
                                                         float *__restrict__ mesh) {
    const int tile = blockIdx.x;
    unsigned cnt = offsets[tile + 1] - offsets[tile];
    if (cnt <= (unsigned)t.CH) return;
    const int tz = tile % t.ntx, ty = (tile / t.ntx) % t.ntx, tx = tile / (t.ntx * t.ntx);
    const int T = t.T;
    for (int q = threadIdx.x; q < T * T * T; q += blockDim.x) {
        int iz = q % T, iy = (q / T) % T, ix = q / (T * T);
        size_t idx = ((size_t)(tx * T + ix) * N + (ty * T + iy)) * N + (tz * T + iz);
        mesh[idx] = 0.0f;
    }
}

// LDS accumulation is in 64-bit fixed point (2^-32 quantum): integer LDS
// atomics run at full rate on gfx950 (float LDS atomic add does not), and the
// integer sum does not depend on the order of the records, so every tile that
// one workgroup deposits is bit-reproducible.
__device__ __forceinline__ long long to_fixed(float v) { return __double2ll_rn((double)v * 4294967296.0); }
__device__ __forceinline__ float from_fixed(long long q) { return (float)((double)q * (1.0 / 4294967296.0)); }

// Stencil of one record into the LDS grid, unrolled for the scheme.  A record
// of the tile sits in a cell c with local coordinate c-o in [-R, T+R) (ghosts
// come from up to R cells outside), so its stencil spans [-2R, T+2R): the LDS
// grid has S = T+4R cells per axis with local cell u at index u+2R, every
// contribution lands in it without a range test, and only [0,T) is written
// out.
template <int SCHEME, bool UNIT>
__device__ __forceinline__ void deposit_record(long long *lgrid, int S, int base, const float f[3], float wt) {
    constexpr int R = SCHEME == 0 ? 0 : (SCHEME == 3 ? 2 : 1);
    constexpr int W = 2 * R + 1;
    float wa[3][5];
#pragma unroll
    for (int a = 0; a < 3; ++a) axis_weights<float>(SCHEME, f[a], wa[a]);
#pragma unroll
    for (int a = 0; a < W; ++a) {
#pragma unroll
        for (int b = 0; b < W; ++b) {
            const float vab = wt * wa[0][a + 2 - R] * wa[1][b + 2 - R];
            unsigned long long *row = (unsigned long long *)&lgrid[base + (a * S + b) * S];
#pragma unroll
            for (int e = 0; e < W; ++e) {
                const float v = vab * wa[2][e + 2 - R];
                // unit weights: |v| <= 1 fits v*2^30 in an int32 (one conversion);
                // general weights go through FP64
                const long long q = UNIT ? (long long)__float2int_rn(v * 1073741824.0f) * 4ll
                                         : __double2ll_rn((double)v * 4294967296.0);
                atomicAdd(&row[e], (unsigned long long)q);
            }
        }
    }
}

template <int SCHEME, bool UNIT>
__global__ void __launch_bounds__(512) k_tile_deposit(const float4 *__restrict__ sorted,
                                                      const float *__restrict__ wsorted,
                                                      const unsigned *__restrict__ offsets,
                                                      const uint2 *__restrict__ items,
                                                      const unsigned *__restrict__ nitems, MasK p, TileK t,
                                                      unsigned cap, float *__restrict__ mesh, int accumulate) {
    extern __shared__ long long lgrid[];
    constexpr int R = SCHEME == 0 ? 0 : (SCHEME == 3 ? 2 : 1);
    const int T = t.T, S = T + 4 * R, S3 = S * S * S;
    const unsigned total = *nitems;
    const int N = p.N;
    for (unsigned item = blockIdx.x; item < total; item += gridDim.x) {
        const uint2 it = items[item];
        const int tile = (int)it.x;
        const int tz = tile % t.ntx, ty = (tile / t.ntx) % t.ntx, tx = tile / (t.ntx * t.ntx);
        const int o[3] = {tx * T, ty * T, tz * T};
        for (int q = threadIdx.x; q < S3; q += blockDim.x) lgrid[q] = 0ll;
        __syncthreads();
        const unsigned tb = offsets[tile], te = offsets[tile + 1];
        const unsigned beg = tb + it.y * (unsigned)t.CH;
        const unsigned end = min(min(te, beg + (unsigned)t.CH), cap);
        for (unsigned i = beg + threadIdx.x; i < end; i += blockDim.x) {
            const float4 r = sorted[i];
            const unsigned pk = __float_as_uint(r.x);
            const int cc[3] = {(int)(pk & 1023u), (int)((pk >> 10) & 1023u), (int)(pk >> 20)};
            const float f[3] = {r.y, r.z, r.w};
            int l[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                int lc = cc[a] - o[a];
                if (lc < -R) lc += N;
                if (lc >= N - R) lc -= N;
                l[a] = lc + R;  // node c-R (local lc-R) sits at LDS index lc-R+2R
            }
            const int base = (l[0] * S + l[1]) * S + l[2];
            deposit_record<SCHEME, UNIT>(lgrid, S, base, f, UNIT ? 1.0f : wsorted[i]);
        }
        __syncthreads();
        const bool split = (te - tb) > (unsigned)t.CH;
        for (int q = threadIdx.x; q < T * T * T; q += blockDim.x) {
            int iz = q % T, iy = (q / T) % T, ix = q / (T * T);
            float v = from_fixed(lgrid[((ix + 2 * R) * S + (iy + 2 * R)) * S + (iz + 2 * R)]);
            size_t idx = ((size_t)(o[0] + ix) * N + (o[1] + iy)) * N + (o[2] + iz);
            if (split) {
                if (v != 0.0f) unsafeAtomicAdd(&mesh[idx], v);
            } else if (accumulate) {
                mesh[idx] += v;
            } else {
                mesh[idx] = v;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- reductions
template <typename T>
__global__ void __launch_bounds__(256) k_sum_partial(const T *__restrict__ a, size_t n, double *__restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += (double)a[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum_final(const double *__restrict__ part, int np, double *__restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

template <typename T>
__global__ void k_affine(T *__restrict__ a, size_t n, T scale, T shift) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = a[i] * scale + shift;
}

// ---------------------------------------------------------------- host side
static unsigned grid_for(size_t n, int block, int cap = 65536) {
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = cap;
    return (unsigned)g;
}

static MasK make_mask(const cm_mas_desc *d, size_t n) {
    MasK p;
    p.N = d->N;
    p.scheme = (int)d->scheme;
    p.soa = d->layout == CM_SOA ? 1 : 0;
    p.pad = 0;
    p.L = d->L;
    p.d = d->L / (double)d->N;
    for (int a = 0; a < 3; ++a) p.xmin[a] = d->xmin[a];
    p.ishift = d->interlace_shift * p.d;
    p.n = n;
    return p;
}

template <typename T>
static cm_status fill(cm_ctx *ctx, T *a, size_t n, T v) {
    return launch(ctx, "fill", k_fill<T>, dim3(grid_for(n, 256, 16384)), dim3(256), 0, a, n, v);
}

static int tile_edge(int N) {
    if (N > 1024) return 0;  // records pack the cell index in 10 bits per axis
    if (N >= 32 && N % 16 == 0) return 16;
    if (N >= 16 && N % 8 == 0) return 8;
    return 0;
}

// Capacity of the sorted-record buffer: when `exact`, the total number of
// records is read back after the scan (one host synchronisation) and the buffer
// sized to it; otherwise a generous estimate is used and the scatter reports
// overflow through `overflow` (checked by the caller).
static cm_status mas_tiled(cm_ctx *ctx, const MasK &p, const float *pos, const float *w, float *mesh,
                           int accumulate, unsigned long long *dropped, bool exact, unsigned *overflow) {
    TileK t;
    t.T = tile_edge(p.N);
    t.R = mas_radius(p.scheme);
    t.ntx = p.N / t.T;
    t.CH = 4096;
    const int ntiles = t.ntx * t.ntx * t.ntx;
    const size_t max_copies = p.n * (t.R > 0 ? 8 : 1);
    if (max_copies > 0xffffffffull) return fail(ctx, CM_ERR_UNSUPPORTED, "mas tiled: too many particles");
    void *counts, *offsets, *cursor, *item_off, *items, *nitems, *sorted;
    cm_status s;
    if ((s = scratch(ctx, "mas.counts", sizeof(unsigned) * ntiles, &counts)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.offsets", sizeof(unsigned) * (ntiles + 1), &offsets)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.cursor", sizeof(unsigned) * ntiles, &cursor)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.item_off", sizeof(unsigned) * (ntiles + 1), &item_off)) != CM_OK) return s;
    const size_t max_items = (size_t)ntiles + max_copies / t.CH + 1;
    if ((s = scratch(ctx, "mas.items", sizeof(uint2) * max_items, &items)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.nitems", sizeof(unsigned), &nitems)) != CM_OK) return s;
    const unsigned g = grid_for(p.n, 256, 8192);
    // LDS-privatised histogram when a workgroup's LDS holds one counter per tile
    const size_t hist_lds = sizeof(unsigned) * (size_t)ntiles;
    const bool lds_hist = hist_lds <= ctx->lds_per_block && ntiles <= 65536;
    const int nrows = ctx->cu_count;
    const size_t chunk = (p.n + nrows - 1) / nrows;
    void *hist = nullptr;
    if (lds_hist) {
        if ((s = scratch(ctx, "mas.hist", sizeof(unsigned) * (size_t)nrows * ntiles, &hist)) != CM_OK) return s;
        if ((s = launch(ctx, "mas_tile_count", k_tile_count_lds, dim3(nrows), dim3(1024), hist_lds, pos, p, t, ntiles,
                        chunk, (unsigned *)hist, dropped)) != CM_OK)
            return s;
        if ((s = launch(ctx, "mas_tile_prefix", k_tile_prefix, dim3((ntiles + 255) / 256), dim3(256), 0,
                        (unsigned *)hist, nrows, ntiles, (unsigned *)counts)) != CM_OK)
            return s;
    } else {
        if ((s = fill(ctx, (unsigned *)counts, (size_t)ntiles, 0u)) != CM_OK) return s;
        if ((s = launch(ctx, "mas_tile_count", k_tile_count, dim3(g), dim3(256), 0, pos, p, t, (unsigned *)counts,
                        dropped)) != CM_OK)
            return s;
    }
    if ((s = launch(ctx, "mas_tile_scan", k_tile_scan, dim3(1), dim3(1024), 0, (const unsigned *)counts, ntiles,
                    t.CH, (unsigned *)offsets, (unsigned *)cursor, (unsigned *)item_off, (unsigned *)nitems)) != CM_OK)
        return s;
    size_t cap;
    if (exact) {
        unsigned total = 0;
        CM_HIP(ctx, hipMemcpyAsync(&total, (unsigned *)offsets + ntiles, sizeof(unsigned), hipMemcpyDeviceToHost,
                                   ctx->stream));
        CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        cap = (size_t)total;
    } else {
        const double g1 = 1.0 + 2.0 * t.R / t.T;
        cap = std::min(max_copies, (size_t)(1.5 * g1 * g1 * g1 * (double)p.n) + 65536);
    }
    if ((s = scratch(ctx, "mas.sorted", sizeof(float4) * (cap ? cap : 1), &sorted)) != CM_OK) return s;
    void *wsorted = nullptr;
    if (w && (s = scratch(ctx, "mas.wsorted", sizeof(float) * (cap ? cap : 1), &wsorted)) != CM_OK) return s;
    if ((s = launch(ctx, "mas_tile_items", k_tile_items, dim3((ntiles + 255) / 256), dim3(256), 0,
                    (const unsigned *)item_off, ntiles, (uint2 *)items)) != CM_OK)
        return s;
    if (lds_hist) {
        if ((s = launch(ctx, "mas_tile_scatter", k_tile_scatter_lds, dim3(nrows), dim3(1024), hist_lds, pos, w, p, t,
                        ntiles, chunk, (const unsigned *)hist, (const unsigned *)offsets, (float4 *)sorted,
                        (float *)wsorted, (unsigned)cap, overflow)) != CM_OK)
            return s;
    } else if ((s = launch(ctx, "mas_tile_scatter", k_tile_scatter, dim3(g), dim3(256), 0, pos, w, p, t,
                           (unsigned *)cursor, (float4 *)sorted, (float *)wsorted, (unsigned)cap,
                           overflow)) != CM_OK) {
        return s;
    }
    if (!accumulate) {
        if ((s = launch(ctx, "mas_tile_zero_split", k_tile_zero_split, dim3(ntiles), dim3(256), 0,
                        (const unsigned *)offsets, t, p.N, mesh)) != CM_OK)
            return s;
    }
    const int S = t.T + 4 * t.R;
    const size_t lds = sizeof(long long) * (size_t)S * S * S;
    if (lds > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "mas: deposit grid of %zu bytes", lds);
    unsigned dg = (unsigned)std::min<size_t>(max_items, (size_t)ctx->cu_count * 8);
    const float *ws = (const float *)wsorted;
    const dim3 gd(dg), bd(512);
#define CM_DEPOSIT(SCH, UNIT)                                                                                   \
    launch(ctx, "mas_tile_deposit", k_tile_deposit<SCH, UNIT>, gd, bd, lds, (const float4 *)sorted, ws,        \
           (const unsigned *)offsets, (const uint2 *)items, (const unsigned *)nitems, p, t, (unsigned)cap, mesh, \
           accumulate)
    switch (p.scheme * 2 + (w ? 0 : 1)) {
    case 0: return CM_DEPOSIT(0, false);
    case 1: return CM_DEPOSIT(0, true);
    case 2: return CM_DEPOSIT(1, false);
    case 3: return CM_DEPOSIT(1, true);
    case 4: return CM_DEPOSIT(2, false);
    case 5: return CM_DEPOSIT(2, true);
    case 6: return CM_DEPOSIT(3, false);
    case 7: return CM_DEPOSIT(3, true);
    }
#undef CM_DEPOSIT
    return fail(ctx, CM_ERR_INVALID, "mas: scheme %d", p.scheme);
}

cm_status mesh_assign_impl(cm_ctx *ctx, const cm_mas_desc *d, const void *pos, const void *w, size_t n,
                           void *mesh, cm_dtype mdt, int accumulate) {
    MasK p = make_mask(d, n);
    const size_t n3 = (size_t)d->N * d->N * d->N;
    void *dropped;
    cm_status s;
    if ((s = scratch(ctx, "mas.dropped", sizeof(unsigned long long), &dropped)) != CM_OK) return s;
    if ((s = fill(ctx, (unsigned long long *)dropped, 1, 0ull)) != CM_OK) return s;
    bool tiled_ok = d->pos_dtype == CM_F32 && mdt == CM_F32 && tile_edge(d->N) > 0;
    bool use_tiled = d->algo == CM_MAS_TILED || (d->algo == CM_MAS_AUTO && tiled_ok);
    if (use_tiled && !tiled_ok)
        return fail(ctx, CM_ERR_UNSUPPORTED, "tiled mass assignment needs float32 positions and mesh and N%%8==0, N>=16");
    if (use_tiled) {
        void *ovf;
        if ((s = scratch(ctx, "mas.overflow", sizeof(unsigned), &ovf)) != CM_OK) return s;
        if ((s = fill(ctx, (unsigned *)ovf, 1, 0u)) != CM_OK) return s;
        return mas_tiled(ctx, p, (const float *)pos, (const float *)w, (float *)mesh, accumulate,
                         (unsigned long long *)dropped, true, (unsigned *)ovf);
    }
    if (!accumulate) {
        s = mdt == CM_F32 ? fill(ctx, (float *)mesh, n3, 0.0f) : fill(ctx, (double *)mesh, n3, 0.0);
        if (s != CM_OK) return s;
    }
    const unsigned g = grid_for(n, 256, 16384);
    auto *dr = (unsigned long long *)dropped;
    if (d->pos_dtype == CM_F32) {
        if (mdt == CM_F32)
            return launch(ctx, "mas_plain", k_mas_plain<float, float>, dim3(g), dim3(256), 0, (const float *)pos,
                          (const float *)w, p, (float *)mesh, dr);
        return launch(ctx, "mas_plain", k_mas_plain<float, double>, dim3(g), dim3(256), 0, (const float *)pos,
                      (const float *)w, p, (double *)mesh, dr);
    }
    if (mdt == CM_F32)
        return launch(ctx, "mas_plain", k_mas_plain<double, float>, dim3(g), dim3(256), 0, (const double *)pos,
                      (const double *)w, p, (float *)mesh, dr);
    return launch(ctx, "mas_plain", k_mas_plain<double, double>, dim3(g), dim3(256), 0, (const double *)pos,
                  (const double *)w, p, (double *)mesh, dr);
}

cm_status mesh_sum_impl(cm_ctx *ctx, const void *mesh, cm_dtype dt, size_t n, double *out_dev, double *host_out) {
    void *part;
    const int NB = 1024;
    cm_status s;
    if ((s = scratch(ctx, "sum.part", sizeof(double) * NB, &part)) != CM_OK) return s;
    void *res = out_dev;
    if (!res && (s = scratch(ctx, "sum.out", sizeof(double), &res)) != CM_OK) return s;
    unsigned g = grid_for(n, 256, NB);
    if (dt == CM_F32)
        s = launch(ctx, "mesh_sum", k_sum_partial<float>, dim3(g), dim3(256), 0, (const float *)mesh, n, (double *)part);
    else
        s = launch(ctx, "mesh_sum", k_sum_partial<double>, dim3(g), dim3(256), 0, (const double *)mesh, n, (double *)part);
    if (s != CM_OK) return s;
    if ((s = launch(ctx, "mesh_sum_final", k_sum_final, dim3(1), dim3(256), 0, (const double *)part, (int)g,
                    (double *)res)) != CM_OK)
        return s;
    if (host_out) {
        CM_HIP(ctx, hipMemcpyAsync(host_out, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return CM_OK;
}

}  // namespace cm

using namespace cm;

extern "C" {

cm_status cm_mesh_assign(cm_ctx *ctx, const cm_mas_desc *desc, const void *pos, const void *w, size_t n,
                         void *mesh, cm_dtype mesh_dtype, int accumulate) {
    if (!ctx) return CM_ERR_INVALID;
    if (!desc || !mesh || (!pos && n)) return fail(ctx, CM_ERR_INVALID, "cm_mesh_assign: null pointer");
    if (desc->N < 2 || desc->N > 4096) return fail(ctx, CM_ERR_INVALID, "cm_mesh_assign: N=%d", desc->N);
    if (!(desc->L > 0) || !std::isfinite(desc->L)) return fail(ctx, CM_ERR_INVALID, "cm_mesh_assign: L must be > 0");
    if ((int)desc->scheme < 0 || (int)desc->scheme > 3) return fail(ctx, CM_ERR_INVALID, "cm_mesh_assign: scheme");
    if (desc->pos_dtype != CM_F32 && desc->pos_dtype != CM_F64) return fail(ctx, CM_ERR_INVALID, "pos dtype");
    if (mesh_dtype != CM_F32 && mesh_dtype != CM_F64) return fail(ctx, CM_ERR_INVALID, "mesh dtype");
    if (desc->scheme == CM_PCS && desc->N < 5) return fail(ctx, CM_ERR_INVALID, "PCS needs N>=5");
    return mesh_assign_impl(ctx, desc, pos, w, n, mesh, mesh_dtype, accumulate);
}

cm_status cm_mesh_sum(cm_ctx *ctx, const void *mesh, cm_dtype dtype, size_t n, double *sum) {
    if (!ctx) return CM_ERR_INVALID;
    if (!mesh || !sum) return fail(ctx, CM_ERR_INVALID, "cm_mesh_sum: null pointer");
    return mesh_sum_impl(ctx, mesh, dtype, n, nullptr, sum);
}

cm_status cm_mesh_affine(cm_ctx *ctx, void *mesh, cm_dtype dtype, size_t n, double scale, double shift) {
    if (!ctx) return CM_ERR_INVALID;
    if (!mesh) return fail(ctx, CM_ERR_INVALID, "cm_mesh_affine: null pointer");
    unsigned g = grid_for(n, 256, 16384);
    if (dtype == CM_F32)
        return launch(ctx, "mesh_affine", k_affine<float>, dim3(g), dim3(256), 0, (float *)mesh, n, (float)scale,
                      (float)shift);
    return launch(ctx, "mesh_affine", k_affine<double>, dim3(g), dim3(256), 0, (double *)mesh, n, scale, shift);
}

}  // extern "C"
