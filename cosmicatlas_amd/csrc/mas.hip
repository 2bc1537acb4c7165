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

// Clears a particle-drop counter and (nullable) an overflow flag in one launch.
__global__ void k_clear_flags(unsigned long long *__restrict__ dropped, unsigned *__restrict__ ovf) {
    if (threadIdx.x == 0) {
        *dropped = 0ull;
        if (ovf) *ovf = 0u;
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
    int Rup;    // ghost cells taken from the tile below: R, or R + 1 when the
                // records also feed the half-cell-shifted (interlaced) mesh
    int ntx;    // tiles per side
    int CH;     // max records per deposit workgroup
    int tx0;    // first tile along x of the binned range
    int nxl;    // tiles along x in the range (ntx: the whole mesh)
    int rl;     // lattice row length of the scatter's blocked walk (0: particle order)
};

constexpr int kTileCH = 4096;

// A particle in cell c touches along one axis its home tile and, when the
// stencil crosses the tile edge, the neighbour.  The stencil spans cells
// c-R..c+R; shifted by half a cell (the interlaced mesh) it spans c-R..c+R+1,
// so records shared by both meshes go to the tile above from Rup = R+1 cells
// below its edge (axis_pair).

// Cell and in-cell offset f = x/d - c - 1/2 (FP64, as the plain kernel) of a
// particle from its loaded coordinates; s[] receives the offsets in float.
// Out-of-box coordinates (periodic images, non-finite values) take the full
// wrap out of line; in-box ones (the common case) only the subtraction.
// Returns NaN for a dropped coordinate.
__device__ __attribute__((noinline)) float wrap_slow(float v, double xmin, double L) {
    float o = 0.0f;
    return wrap_coord<float>(v, xmin, L, o) ? o : __int_as_float(0x7fc00000);
}

// Cell and in-cell offset of a particle from its loaded coordinates: fr[a] =
// the fraction x/d - c in [0, 1) as a 27-bit fixed-point number (floored, so
// it never reaches the next cell).
constexpr int kFracBits = 27;
__device__ __forceinline__ bool cells_from(const float v[3], const MasK &p, int c[3], unsigned fr[3]) {
    bool ok = true;
    const float Lf = (float)p.L;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // wrap_coord's first step; a value already in [0, L) is returned as is
        float u = (float)((double)v[a] - p.xmin[a]);
        if (__builtin_expect(!(u >= 0.0f && u < Lf), 0)) {
            u = wrap_slow(v[a], p.xmin[a], p.L);
            if (u != u) {
                ok = false;
                u = 0.0f;
            }
        }
        const double x = shifted((double)u, p.ishift, p.L);
        double q;
        c[a] = cell_fast(x, p.d, p.rd, p.N, q);
        const unsigned qf = (unsigned)((q - (double)c[a]) * (double)(1u << kFracBits));
        fr[a] = qf < (1u << kFracBits) ? qf : (1u << kFracBits) - 1u;
    }
    return ok;
}

// The same cells and fractions with the common case in straight-line code.
// A coordinate in the box whose q = x rd lies at least 2^-39 from a cell edge
// takes c = floor(q) and the fraction q - c: x rd is within 2^-40 of x/d for
// every N <= 4096 (two roundings of 2^-53 relative, q < 2^12), so floor(q) is
// get_bin's cell, cell_fast's edge test does not fire for it and cells_from
// returns the same c and fraction.  A particle with any other coordinate
// (outside the box, non-finite, near an edge) goes through cells_from.
__device__ __forceinline__ bool particle_cells(const float v[3], const MasK &p, int c[3], unsigned fr[3]) {
    const float Lf = (float)p.L;
    bool slow = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float u = (float)((double)v[a] - p.xmin[a]);
        double x = (double)u;
        if (p.ishift != 0.0) {  // uniform; shifted()
            x += p.ishift;
            if (x >= p.L) x -= p.L;
        }
        const double q = x * p.rd, fq = floor(q), f = q - fq;
        slow = slow | !(u >= 0.0f && u < Lf) | !(f >= 0x1p-39 && f <= 1.0 - 0x1p-39);
        c[a] = (int)fq;
        fr[a] = (unsigned)(f * (double)(1u << kFracBits));
    }
    if (__builtin_expect(slow, 0)) return cells_from(v, p, c, fr);
    return true;
}

// Destination tiles of a cell c along one axis (axis_tiles in closed form, for
// ntx >= 2): home tile h and, where the stencil crosses the tile edge (two),
// the neighbour nb; lh and ln are the cell's coordinates relative to the two
// tiles' origins (the record words' local coordinates, periodic images
// included: lh + T in the tile below, lh - T in the tile above).
template <int T>
__device__ __forceinline__ void axis_pair(int c, int R, int Rup, int ntx, int &h, int &nb, int &lh, int &ln,
                                          bool &two) {
    h = c / T;
    lh = c - h * T;
    const bool lo = lh < R, hi = lh >= T - Rup;
    two = lo | hi;
    nb = lo ? (h == 0 ? ntx - 1 : h - 1) : (h + 1 == ntx ? 0 : h + 1);
    ln = lo ? lh + T : lh - T;
}

// Sorted record: one 32-bit word per axis.  Bits 27..31 hold the cell's
// coordinate in the destination tile plus kLcBias (the tile's own cells are
// 0..T-1, ghosts from neighbours -Rup..T+R-1, R <= 2, Rup <= R+1); bits 0..26
// the fraction x/d - c.  A coordinate below -Rup or at least N-Rup is moved by
// one period.
constexpr int kLcBias = 3;
__device__ __forceinline__ unsigned record_word(int c, int o, unsigned fr, int N, int Rup) {
    int lc = c - o;
    if (lc < -Rup) lc += N;
    if (lc >= N - Rup) lc -= N;
    return ((unsigned)(lc + kLcBias) << kFracBits) | fr;
}

// In-cell offset f = x/d - c - 1/2 from a record word (midpoint of the 2^-27
// bin; the float conversion rounds by at most 2^-26).
__device__ __forceinline__ float record_offset(unsigned w) {
    const int q = (int)(w & ((1u << kFracBits) - 1u)) - (1 << (kFracBits - 1));
    return (float)q * (1.0f / (float)(1u << kFracBits)) + (0.5f / (float)(1u << kFracBits));
}

constexpr int kBatch = 8;
// Particles i = base + u*stride (u < U, i < i1) are loaded before any is used,
// so one thread keeps U loads in flight; an AoS particle is one 12-byte load.
// A lane whose whole batch lies below i1 (all but the chunk's last batch)
// takes the unguarded loads.
template <int U, bool SOA>
__device__ __forceinline__ void load_batch(const float *__restrict__ pos, size_t n, size_t base, size_t stride,
                                           size_t i1, float v[U][3]) {
    auto one = [&](size_t i, float w[3]) {
        if constexpr (SOA) {
            w[0] = pos[i];
            w[1] = pos[n + i];
            w[2] = pos[2 * n + i];
        } else {
            // nontemporal: the positions are read once per pass (0.28 -> 0.266 ms
            // in the record scatter at 512^3, whose record stores keep the L2)
            const auto g = as_global(pos) + 3 * i;
            w[0] = __builtin_nontemporal_load(g);
            w[1] = __builtin_nontemporal_load(g + 1);
            w[2] = __builtin_nontemporal_load(g + 2);
        }
    };
    if (base + (size_t)(U - 1) * stride < i1) {
#pragma unroll
        for (int u = 0; u < U; ++u) one(base + (size_t)u * stride, v[u]);
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = base + (size_t)u * stride;
            if (i < i1) {
                one(i, v[u]);
            } else {
                v[u][0] = v[u][1] = v[u][2] = 0.0f;
            }
        }
    }
}

// Particles idx[u] (u < U, where valid) loaded before any is used.
template <int U, bool SOA>
__device__ __forceinline__ void load_idx(const float *__restrict__ pos, size_t n, const size_t idx[U], float v[U][3]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t i = idx[u];
        if constexpr (SOA) {
            v[u][0] = pos[i];
            v[u][1] = pos[n + i];
            v[u][2] = pos[2 * n + i];
        } else {
            const auto g = as_global(pos) + 3 * i;
            v[u][0] = __builtin_nontemporal_load(g);
            v[u][1] = __builtin_nontemporal_load(g + 1);
            v[u][2] = __builtin_nontemporal_load(g + 2);
        }
    }
}

// Particle chunk of binning workgroup b of n: workgroups are dispatched to the
// eight XCDs in turn (b mod 8), so consecutive chunks go to workgroups of one
// XCD.  Consecutive chunks of an ordered catalogue share tiles, and their
// record runs of a tile are adjacent: the cache lines where one run ends and
// the next begins are then written through one L2 (scatter 0.215 -> 0.210 ms
// at 512^3).  The count pass uses the same map, so its histogram row of
// chunk c is the one the scatter reads.
__device__ __forceinline__ unsigned chunk_of(unsigned b, unsigned n) {
    return n % 8u == 0u ? (b % 8u) * (n / 8u) + b / 8u : b;
}

// Count records per tile.  LDS form: workgroup g owns particles
// [g*chunk, (g+1)*chunk) and writes its histogram row hist[g*ntiles + tile];
// global form (hist_lds == false, too many tiles for LDS): one row, global
// atomics.
template <int T, bool LDSH, bool SOA>
__global__ void __launch_bounds__(1024) k_tile_count(const float *__restrict__ pos, MasK p, TileK t, int ntiles,
                                                     size_t chunk, unsigned *__restrict__ hist,
                                                     unsigned long long *dropped) {
    extern __shared__ unsigned lh[];
    unsigned *h = hist;
    size_t i0 = 0, i1 = p.n, step = (size_t)gridDim.x * blockDim.x;
    if constexpr (LDSH) {
        for (int i = threadIdx.x; i < ntiles; i += blockDim.x) lh[i] = 0u;
        __syncthreads();
        h = lh;
        i0 = (size_t)chunk_of(blockIdx.x, gridDim.x) * chunk;
        i1 = min(p.n, i0 + chunk);
        step = blockDim.x;
    } else {
        i0 = (size_t)blockIdx.x * blockDim.x;
    }
    unsigned ndrop = 0;
    for (size_t base = i0 + threadIdx.x; base < i1; base += kBatch * step) {
        float v[kBatch][3];
        load_batch<kBatch, SOA>(pos, p.n, base, step, i1, v);
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (base + (size_t)u * step >= i1) continue;
            int c[3];
            unsigned f[3];
            if (!particle_cells(v[u], p, c, f)) {
                ++ndrop;
                continue;
            }
            int th[3], tn[3], lh[3], ln[3];
            bool two[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) axis_pair<T>(c[a], t.R, t.Rup, t.ntx, th[a], tn[a], lh[a], ln[a], two[a]);
            // tiles outside the x range belong to another pass
            const int lx0 = th[0] - t.tx0, lx1 = tn[0] - t.tx0;
            const bool x0 = (unsigned)lx0 < (unsigned)t.nxl, x1 = two[0] && (unsigned)lx1 < (unsigned)t.nxl;
            auto tile = [&](int lx, int y, int z) { return (lx * t.ntx + y) * t.ntx + z; };
            if (x0) atomicAdd(&h[tile(lx0, th[1], th[2])], 1u);
            if (two[0] | two[1] | two[2]) {
                if (x0 && two[2]) atomicAdd(&h[tile(lx0, th[1], tn[2])], 1u);
                if (x0 && two[1]) atomicAdd(&h[tile(lx0, tn[1], th[2])], 1u);
                if (x0 && two[1] && two[2]) atomicAdd(&h[tile(lx0, tn[1], tn[2])], 1u);
                if (x1) {
                    atomicAdd(&h[tile(lx1, th[1], th[2])], 1u);
                    if (two[2]) atomicAdd(&h[tile(lx1, th[1], tn[2])], 1u);
                    if (two[1]) atomicAdd(&h[tile(lx1, tn[1], th[2])], 1u);
                    if (two[1] && two[2]) atomicAdd(&h[tile(lx1, tn[1], tn[2])], 1u);
                }
            }
        }
    }
    if (ndrop) atomicAdd(dropped, (unsigned long long)ndrop);
    if constexpr (LDSH) {
        __syncthreads();
        unsigned *row = hist + (size_t)chunk_of(blockIdx.x, gridDim.x) * ntiles;
        for (int i = threadIdx.x; i < ntiles; i += blockDim.x) row[i] = lh[i];
    }
}

// Per tile (one thread each): exclusive prefix over the histogram rows, in
// place, and the tile total; then a workgroup scan of the totals and of the
// deposit chunk counts.  offsets/item_off receive the workgroup-local
// exclusive values, bsum[2*blk..] the workgroup sums.
__global__ void __launch_bounds__(256) k_tile_prefix(unsigned *__restrict__ hist, int nrows, int ntiles,
                                                     unsigned *__restrict__ offsets, unsigned *__restrict__ item_off,
                                                     unsigned *__restrict__ counts,
                                                     unsigned long long *__restrict__ bsum) {
    __shared__ unsigned long long sa[256], sb[256];  // 64-bit: a block of tiles may hold 2^32 records
    const int tid = threadIdx.x, tile = blockIdx.x * 256 + tid;
    unsigned run = 0;
    if (tile < ntiles) {
        int g = 0;
        for (; g + 8 <= nrows; g += 8) {
            unsigned v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = hist[(size_t)(g + k) * ntiles + tile];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                hist[(size_t)(g + k) * ntiles + tile] = run;
                run += v[k];
            }
        }
        for (; g < nrows; ++g) {
            unsigned *hp = hist + (size_t)g * ntiles + tile;
            const unsigned v = *hp;
            *hp = run;
            run += v;
        }
    }
    const unsigned chunks = tile < ntiles ? (run == 0 ? 1u : (run + kTileCH - 1) / kTileCH) : 0u;
    sa[tid] = run;
    sb[tid] = chunks;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const unsigned long long va = tid >= off ? sa[tid - off] : 0ull, vb = tid >= off ? sb[tid - off] : 0ull;
        __syncthreads();
        sa[tid] += va;
        sb[tid] += vb;
        __syncthreads();
    }
    // workgroup-local exclusive offsets; a value of 2^32 or more only occurs
    // with a total beyond the record capacity, which k_tile_items reports
    if (tile < ntiles) {
        offsets[tile] = (unsigned)(sa[tid] - run);
        item_off[tile] = (unsigned)(sb[tid] - chunks);
        counts[tile] = run;
    }
    if (tid == 255) {
        bsum[2 * blockIdx.x] = sa[255];
        bsum[2 * blockIdx.x + 1] = sb[255];
    }
}

// Totals of the workgroup sums (one workgroup) into bsum[2 nblk], [2 nblk + 1],
// in 64-bit so that a total of 2^32 records or more is seen and reported: the
// host reads them when it sizes the record buffer exactly.
__global__ void __launch_bounds__(1024) k_tile_total(unsigned long long *__restrict__ bsum, int nblk) {
    __shared__ unsigned long long sa[16], sb[16];
    unsigned long long ta = 0, tb = 0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        ta += bsum[2 * i];
        tb += bsum[2 * i + 1];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ta += __shfl_xor(ta, o, 64);
        tb += __shfl_xor(tb, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sa[threadIdx.x >> 6] = ta;
        sb[threadIdx.x >> 6] = tb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            a += sa[w];
            b += sb[w];
        }
        bsum[2 * nblk] = a;
        bsum[2 * nblk + 1] = b;
    }
}

// Final offsets, slot cursors and the deposit work list: one item
// {tile, first record, end record, split} per chunk of at most kTileCH records,
// clamped to the record buffer's capacity.  counts[tile] holds the tile total
// on entry and the slot cursor (= first record) on exit.
// bsum holds the per-workgroup sums of k_tile_prefix; each workgroup adds up
// the sums of the workgroups before it (the exclusive scan, at most 4096
// values) and tile 0's workgroup all of them.
__global__ void __launch_bounds__(256) k_tile_items(int ntiles, const unsigned long long *__restrict__ bsum,
                                                    unsigned *__restrict__ offsets, unsigned *__restrict__ item_off,
                                                    unsigned *__restrict__ counts, uint4 *__restrict__ items,
                                                    unsigned *__restrict__ nitems, unsigned cap,
                                                    unsigned max_items, unsigned *__restrict__ overflow) {
    const int tile = blockIdx.x * 256 + threadIdx.x;
    const int nblk = (ntiles + 255) / 256;
    const int blk = blockIdx.x;
    __shared__ unsigned long long sred[4][4];
    unsigned long long pa = 0, pb = 0, sa = 0, sb = 0;
    for (int j = threadIdx.x; j < (blk == 0 ? nblk : blk); j += 256) {
        const unsigned long long a = bsum[2 * j], b = bsum[2 * j + 1];
        sa += a;
        sb += b;
        if (j < blk) {
            pa += a;
            pb += b;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        pa += __shfl_xor(pa, o, 64);
        pb += __shfl_xor(pb, o, 64);
        sa += __shfl_xor(sa, o, 64);
        sb += __shfl_xor(sb, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sred[threadIdx.x >> 6][0] = pa;
        sred[threadIdx.x >> 6][1] = pb;
        sred[threadIdx.x >> 6][2] = sa;
        sred[threadIdx.x >> 6][3] = sb;
    }
    __syncthreads();
    pa = (sred[0][0] + sred[1][0]) + (sred[2][0] + sred[3][0]);
    pb = (sred[0][1] + sred[1][1]) + (sred[2][1] + sred[3][1]);
    if (tile == 0) {
        const unsigned long long tot = (sred[0][2] + sred[1][2]) + (sred[2][2] + sred[3][2]),
                                 nit = (sred[0][3] + sred[1][3]) + (sred[2][3] + sred[3][3]);
        offsets[ntiles] = (unsigned)min(tot, (unsigned long long)cap);
        item_off[ntiles] = (unsigned)min(nit, (unsigned long long)max_items);
        *nitems = (unsigned)min(nit, (unsigned long long)max_items);
        if (tot > cap || nit > max_items) atomicOr(overflow, 1u);
    }
    if (tile >= ntiles) return;
    const unsigned cnt = counts[tile];
    const unsigned long long tb64 = offsets[tile] + pa;
    const unsigned tb = (unsigned)min(tb64, (unsigned long long)cap);
    const unsigned ib = item_off[tile] + (unsigned)pb;
    const unsigned long long te64 = tb64 + cnt;
    offsets[tile] = tb;
    item_off[tile] = ib;
    counts[tile] = tb;  // slot cursor of the global scatter
    const unsigned nch = cnt == 0 ? 1u : (cnt + kTileCH - 1) / kTileCH;
    const unsigned split = cnt > (unsigned)kTileCH ? 1u : 0u;
    for (unsigned k = 0; k < nch; ++k) {
        if ((unsigned long long)ib + k >= max_items) break;  // reported through overflow above
        const unsigned long long b = min(tb64 + (unsigned long long)k * kTileCH, (unsigned long long)cap);
        const unsigned long long e = min(min(te64, tb64 + (unsigned long long)(k + 1) * kTileCH), (unsigned long long)cap);
        items[ib + k] = make_uint4((unsigned)tile, (unsigned)b, (unsigned)e, split);
    }
}

// Scatter records to their tiles.  LDS form: cursor[tile] = offsets[tile] +
// prefix of earlier workgroups, integer LDS atomics hand out the slots; global
// form: global cursors.
template <int T, bool LDSH, bool SOA>
__global__ void __launch_bounds__(1024) k_tile_scatter(const float *__restrict__ pos, const float *__restrict__ w,
                                                       MasK p, TileK t, int ntiles, size_t chunk,
                                                       const unsigned *__restrict__ pref,
                                                       const unsigned *__restrict__ offsets,
                                                       unsigned *__restrict__ gcursor, uint3 *__restrict__ sorted,
                                                       float *__restrict__ wsorted, unsigned cap,
                                                       unsigned *__restrict__ overflow) {
    extern __shared__ unsigned cur[];
    unsigned *cu = gcursor;
    size_t i0 = 0, i1 = p.n, step = (size_t)gridDim.x * blockDim.x;
    if constexpr (LDSH) {
        const unsigned *row = pref + (size_t)chunk_of(blockIdx.x, gridDim.x) * ntiles;
        for (int i = threadIdx.x; i < ntiles; i += blockDim.x) cur[i] = offsets[i] + row[i];
        __syncthreads();
        cu = cur;
        i0 = (size_t)chunk_of(blockIdx.x, gridDim.x) * chunk;
        i1 = min(p.n, i0 + chunk);
        step = blockDim.x;
    } else {
        i0 = (size_t)blockIdx.x * blockDim.x;
    }
    unsigned ovf = 0;
    // one particle: its records to the tiles its stencil touches
    auto scatter_one = [&](const float vv[3], float wu) {
        int c[3];
        unsigned f[3];
        if (!particle_cells(vv, p, c, f)) return;
        int th[3], tn[3], lh[3], ln[3];
        bool two[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) axis_pair<T>(c[a], t.R, t.Rup, t.ntx, th[a], tn[a], lh[a], ln[a], two[a]);
        // record words (record_word) for the home and the neighbour tile
        unsigned wh[3], wn[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            wh[a] = ((unsigned)(lh[a] + kLcBias) << kFracBits) | f[a];
            wn[a] = ((unsigned)(ln[a] + kLcBias) << kFracBits) | f[a];
        }
        const int lx0 = th[0] - t.tx0, lx1 = tn[0] - t.tx0;
        const bool x0 = (unsigned)lx0 < (unsigned)t.nxl, x1 = two[0] && (unsigned)lx1 < (unsigned)t.nxl;
        auto emit = [&](int lx, int y, int z, unsigned wx, unsigned wy, unsigned wz) {
            const unsigned slot = atomicAdd(&cu[(lx * t.ntx + y) * t.ntx + z], 1u);
            if (slot < cap) {
                sorted[slot] = make_uint3(wx, wy, wz);
                if (w) wsorted[slot] = wu;
            } else {
                ovf = 1u;
            }
        };
        if (x0) emit(lx0, th[1], th[2], wh[0], wh[1], wh[2]);
        if (two[0] | two[1] | two[2]) {
            if (x0 && two[2]) emit(lx0, th[1], tn[2], wh[0], wh[1], wn[2]);
            if (x0 && two[1]) emit(lx0, tn[1], th[2], wh[0], wn[1], wh[2]);
            if (x0 && two[1] && two[2]) emit(lx0, tn[1], tn[2], wh[0], wn[1], wn[2]);
            if (x1) {
                emit(lx1, th[1], th[2], wn[0], wh[1], wh[2]);
                if (two[2]) emit(lx1, th[1], tn[2], wn[0], wh[1], wn[2]);
                if (two[1]) emit(lx1, tn[1], th[2], wn[0], wn[1], wh[2]);
                if (two[1] && two[2]) emit(lx1, tn[1], tn[2], wn[0], wn[1], wn[2]);
            }
        }
    };
    if (LDSH && t.rl > 0) {
        // blocked walk of a lattice-ordered chunk (rows of t.rl particles):
        // blocks of 32 rows x 32 columns, one particle per thread, column
        // blocks fastest, so that a tile's records from this chunk are
        // written within one or two blocks (a scattered walk of whole rows
        // keeps so many record runs open that their cache lines are evicted
        // and written back several times before they fill)
        const int RL = t.rl, nbz = RL / 32, r_in = threadIdx.x / 32, c_in = threadIdx.x % 32;
        const size_t nblocks = (i1 - i0) / 1024;
        for (size_t b0 = 0; b0 < nblocks; b0 += kBatch) {
            size_t idx[kBatch];
            float v[kBatch][3];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const size_t b = min(b0 + (size_t)u, nblocks - 1), rb = b / nbz, cb = b - rb * nbz;
                idx[u] = i0 + (rb * 32 + r_in) * (size_t)RL + cb * 32 + c_in;
            }
            load_idx<kBatch, SOA>(pos, p.n, idx, v);
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
                if (b0 + (size_t)u < nblocks) scatter_one(v[u], w ? w[idx[u]] : 0.0f);
        }
    } else {
        for (size_t base = i0 + threadIdx.x; base < i1; base += kBatch * step) {
            float v[kBatch][3], wv[kBatch];
            load_batch<kBatch, SOA>(pos, p.n, base, step, i1, v);
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const size_t i = base + (size_t)u * step;
                wv[u] = (w && i < i1) ? w[i] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                if (base + (size_t)u * step >= i1) continue;
                scatter_one(v[u], wv[u]);
            }
        }
    }
    if (ovf) atomicOr(overflow, 1u);
}

// ------------------------------------------- x-slab groups (level-1 binning)
// When the tiles of the mesh do not fit one workgroup's LDS histogram (N >= 768
// with 16^3 tiles), or for particles in no spatial order, particles are first
// sorted by their tile row (tx, ty): ntx^2 buckets, with a copy for the row of
// the neighbouring x-slab group whose first or last tile plane the particle's
// stencil reaches.  The rows of gs consecutive tile planes form a group; each
// group then goes through the LDS binning and deposit restricted to its tile
// planes, so every record is emitted by exactly one group.  Sorted by row, the
// particles of one binning workgroup's chunk share a few tile rows, so its
// records are written in runs even when the input has no order.  Raw positions
// are copied (AoS float32), so the second level computes the same cells.
template <int T>
__device__ __forceinline__ int particle_rows(const int c[3], int R, int Rup, int ntx, int gs, int G, int out[2]) {
    const int home = c[0] / T, lc = c[0] % T, ty = c[1] / T;
    out[0] = home * ntx + ty;
    if (G > 1 && lc < R && home % gs == 0) {
        out[1] = (home == 0 ? ntx - 1 : home - 1) * ntx + ty;
        return 2;
    }
    if (G > 1 && lc >= T - Rup && (home % gs == gs - 1 || home == ntx - 1)) {
        out[1] = (home + 1 == ntx ? 0 : home + 1) * ntx + ty;
        return 2;
    }
    return 1;
}

// Per workgroup (a chunk of particles): count per tile row into hist[row][b].
template <int T, bool SOA>
__global__ void __launch_bounds__(1024) k_group_count(const float *__restrict__ pos, MasK p, int R, int Rup, int ntx,
                                                      int gs, int G, size_t chunk, unsigned *__restrict__ hist,
                                                      unsigned long long *dropped) {
    extern __shared__ unsigned lh[];  // ntx^2
    const int nb = ntx * ntx;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) lh[i] = 0u;
    __syncthreads();
    const size_t i0 = (size_t)blockIdx.x * chunk, i1 = min(p.n, i0 + chunk);
    unsigned ndrop = 0;
    for (size_t base = i0 + threadIdx.x; base < i1; base += kBatch * (size_t)blockDim.x) {
        float v[kBatch][3];
        load_batch<kBatch, SOA>(pos, p.n, base, blockDim.x, i1, v);
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (base + (size_t)u * blockDim.x >= i1) continue;
            int c[3];
            unsigned f[3];
            if (!particle_cells(v[u], p, c, f)) {
                ++ndrop;
                continue;
            }
            int g[2];
            const int ng = particle_rows<T>(c, R, Rup, ntx, gs, G, g);
            for (int k = 0; k < ng; ++k) atomicAdd(&lh[g[k]], 1u);
        }
    }
    if (ndrop) atomicAdd(dropped, (unsigned long long)ndrop);
    __syncthreads();
    unsigned *row = hist + (size_t)blockIdx.x * nb;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) row[i] = lh[i];
}

// Exclusive prefix over the workgroups per tile row (in place, one thread per
// row) and the row totals.
__global__ void __launch_bounds__(256) k_group_colscan(unsigned *__restrict__ hist, int nrows, int nb,
                                                       unsigned long long *__restrict__ tot) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    unsigned long long run = 0;
    for (int r = 0; r < nrows; ++r) {
        const unsigned v = hist[(size_t)r * nb + b];
        hist[(size_t)r * nb + b] = (unsigned)run;  // < 2^32: checked against the total on the host
        run += v;
    }
    tot[b] = run;
}

// Row starts rstart[0..nb] (64-bit exclusive scan of the totals), one workgroup
// of 1024 threads, nb <= 4096.
__global__ void __launch_bounds__(1024) k_group_scan(const unsigned long long *__restrict__ tot, int nb,
                                                     unsigned long long *__restrict__ rstart) {
    __shared__ unsigned long long sw[1024];
    const int t = threadIdx.x;
    unsigned long long loc[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int b = t * 4 + k;
        loc[k] = b < nb ? tot[b] : 0ull;
        sum += loc[k];
    }
    sw[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const unsigned long long x = t >= off ? sw[t - off] : 0ull;
        __syncthreads();
        sw[t] += x;
        __syncthreads();
    }
    unsigned long long run = sw[t] - sum;  // exclusive
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int b = t * 4 + k;
        if (b < nb) rstart[b] = run;
        run += loc[k];
    }
    if (t == 1023) rstart[nb] = sw[1023];
}

template <int T, bool SOA>
__global__ void __launch_bounds__(1024) k_group_scatter(const float *__restrict__ pos, const float *__restrict__ w,
                                                        MasK p, int R, int Rup, int ntx, int gs, int G, size_t chunk,
                                                        const unsigned *__restrict__ hist,
                                                        const unsigned long long *__restrict__ rstart,
                                                        float3 *__restrict__ out, float *__restrict__ wout) {
    extern __shared__ unsigned long long rcur[];  // ntx^2
    const int nb = ntx * ntx;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) rcur[i] = rstart[i] + hist[(size_t)blockIdx.x * nb + i];
    __syncthreads();
    const size_t i0 = (size_t)blockIdx.x * chunk, i1 = min(p.n, i0 + chunk);
    for (size_t base = i0 + threadIdx.x; base < i1; base += kBatch * (size_t)blockDim.x) {
        float v[kBatch][3];
        load_batch<kBatch, SOA>(pos, p.n, base, blockDim.x, i1, v);
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const size_t i = base + (size_t)u * blockDim.x;
            if (i >= i1) continue;
            int c[3];
            unsigned f[3];
            if (!particle_cells(v[u], p, c, f)) continue;
            int g[2];
            const int ng = particle_rows<T>(c, R, Rup, ntx, gs, G, g);
            const float wi = w ? w[i] : 0.0f;
            for (int k = 0; k < ng; ++k) {
                const unsigned long long slot = atomicAdd(&rcur[g[k]], 1ull);
                out[slot] = make_float3(v[u][0], v[u][1], v[u][2]);
                if (w) wout[slot] = wi;
            }
        }
    }
}

// Zero the mesh cells of tiles that will be deposited by several workgroups.
template <int T>
__global__ void __launch_bounds__(256) k_tile_zero_split(const unsigned *__restrict__ offsets, TileK t, int N,
                                                         float *__restrict__ mesh) {
    const int tile = blockIdx.x;
    const unsigned cnt = offsets[tile + 1] - offsets[tile];
    if (cnt <= (unsigned)kTileCH) return;
    const int tz = tile % t.ntx, ty = (tile / t.ntx) % t.ntx, tx = t.tx0 + tile / (t.ntx * t.ntx);
    for (int q = threadIdx.x; q < T * T * T; q += blockDim.x) {
        const int iz = q % T, iy = (q / T) % T, ix = q / (T * T);
        mesh[((size_t)(tx * T + ix) * N + (ty * T + iy)) * N + (tz * T + iz)] = 0.0f;
    }
}

// LDS accumulation is in fixed point: integer LDS atomics run at full rate on
// gfx950 (float LDS atomic add does not), and an integer sum does not depend
// on the order of the records, so every tile that one workgroup deposits is
// bit-reproducible.
//   unit weights: 32-bit cells with a per-item quantum 2^-k, k the largest
//     integer with cnt * wmax * 2^k <= 2^31 (cnt = records of the item, wmax
//     = the scheme's largest stencil weight), so no cell can overflow (the
//     factor 2 covers rounding); an add is off by at most 3/4 of a quantum
//     plus the float rounding of the product (k = 22 for a TSC chunk of 730
//     records, 20 at the 4096-record maximum).  NGP adds 2^k per record:
//     counts stay exact.  32-bit atomics move half the LDS
//     words of 64-bit ones and halve the grid, so more workgroups fit a CU.
//   general weights: 64-bit cells, quantum 2^-32.
__device__ __forceinline__ float from_fixed(long long q) { return (float)((double)q * (1.0 / 4294967296.0)); }

template <int SCHEME>
__device__ __forceinline__ int unit_quantum_exp(unsigned cnt) {
    constexpr float wmax = SCHEME == 2 ? 0.421875f : (SCHEME == 3 ? 0.2962963f : 1.0f);
    const float lim = 2147483648.0f / ((float)(cnt ? cnt : 1u) * wmax);
    const int k = (int)floorf(log2f(lim));
    return k < 0 ? 0 : (k > 31 ? 31 : k);
}

// Stencil of one record into the LDS grid, unrolled for the scheme.  A record
// of the tile sits in a cell c with local coordinate c-o in [-R, T+R) (ghosts
// come from up to R cells outside), so its stencil spans [-2R, T+2R): the LDS
// grid has S = T+4R cells per axis with local cell u at index u+2R, every
// contribution lands in it without a range test, and only [0,T) is written
// out.  Unit weights: 32-bit cells, quantum 1/qs (qs = 2^k); the quantum is
// folded into the x weights, and a stencil value v (>= 0) becomes the integer
// trunc(fl(v + 1/2)) with one fused multiply-add and one conversion: within
// 3/4 of a quantum of v, plus the float rounding of the product.
template <int SCHEME, bool UNIT, int S>
__device__ __forceinline__ void deposit_record(void *lgrid, int base, const float f[3], float wt, float qs) {
    constexpr int R = SCHEME == 0 ? 0 : (SCHEME == 3 ? 2 : 1);
    constexpr int W = 2 * R + 1;
    float wa[3][5];
    const float sx = UNIT ? qs : wt;
    if constexpr (SCHEME == 1) {
        // CIC: the two nodes with weight along each axis (c-1 and c when
        // f < 0, else c and c+1; the third node of the 3-node stencil has
        // weight zero), the same weights and products as the general loop: 8
        // LDS adds per record instead of 27
        float wx[3][2];
        int off = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const bool neg = f[a] < 0.0f;
            const float c0 = 1.0f - fabsf(f[a]);
            wx[a][0] = neg ? -f[a] : c0;
            wx[a][1] = neg ? c0 : f[a];
            off = off * S + (neg ? 0 : 1);
        }
        wx[0][0] *= sx;
        wx[0][1] *= sx;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const float vab = wx[0][a] * wx[1][b];
                const int rowi = base + off + (a * S + b) * S;
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    if constexpr (UNIT) {
                        const float vq = fmaf(vab, wx[2][e], 0.5f);
                        atomicAdd((unsigned *)lgrid + rowi + e, (unsigned)vq);
                    } else {
                        const float v = vab * wx[2][e];
                        const long long q = __double2ll_rn((double)v * 4294967296.0);
                        atomicAdd((unsigned long long *)lgrid + rowi + e, (unsigned long long)q);
                    }
                }
            }
        }
        return;
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) axis_weights<float>(SCHEME, f[a], wa[a]);
#pragma unroll
        for (int a = 0; a < W; ++a) wa[0][a + 2 - R] *= sx;
    }
    if constexpr (UNIT && W > 1 && S % 2 == 0) {
        // 32-bit cells: two adjacent z nodes whose pair of cells is 8-byte
        // aligned take one 64-bit add (no carry crosses between the two
        // cells: every cell's total stays below 2^32 by the choice of the
        // quantum), so a row of W nodes takes (W+1)/2 adds instead of W.  The
        // row's first node has the parity of base (S is even); the z weights
        // are permuted once so that the pairs come first and the single node
        // last.
        constexpr int NP = W / 2;
        const int p = base & 1;
        float wp[W];
#pragma unroll
        for (int e = 0; e < W - 1; ++e) wp[e] = p ? wa[2][e + 3 - R] : wa[2][e + 2 - R];
        wp[W - 1] = p ? wa[2][2 - R] : wa[2][2 + R];
        unsigned *const g32 = (unsigned *)lgrid;
        const int pbase = base + p, sbase = base + (p ? 0 : W - 1);
#pragma unroll
        for (int a = 0; a < W; ++a) {
#pragma unroll
            for (int b = 0; b < W; ++b) {
                const float vab = wa[0][a + 2 - R] * wa[1][b + 2 - R];
                const int ro = (a * S + b) * S;
#pragma unroll
                for (int k = 0; k < NP; ++k) {
                    const unsigned lo = (unsigned)fmaf(vab, wp[2 * k], 0.5f),
                                   hi = (unsigned)fmaf(vab, wp[2 * k + 1], 0.5f);
                    atomicAdd((unsigned long long *)(g32 + pbase + ro + 2 * k),
                              ((unsigned long long)hi << 32) | (unsigned long long)lo);
                }
                atomicAdd(g32 + sbase + ro, (unsigned)fmaf(vab, wp[W - 1], 0.5f));
            }
        }
        return;
    }
#pragma unroll
    for (int a = 0; a < W; ++a) {
#pragma unroll
        for (int b = 0; b < W; ++b) {
            const float vab = wa[0][a + 2 - R] * wa[1][b + 2 - R];
            const int rowi = base + (a * S + b) * S;
#pragma unroll
            for (int e = 0; e < W; ++e) {
                if constexpr (UNIT) {
                    const float vq = fmaf(vab, wa[2][e + 2 - R], 0.5f);
                    atomicAdd((unsigned *)lgrid + rowi + e, (unsigned)vq);
                } else {
                    const float v = vab * wa[2][e + 2 - R];
                    const long long q = __double2ll_rn((double)v * 4294967296.0);
                    atomicAdd((unsigned long long *)lgrid + rowi + e, (unsigned long long)q);
                }
            }
        }
    }
}

// One workgroup per work item (a tile, or a kTileCH-record chunk of a crowded
// tile), persistent over items.  The records of the next item are loaded into
// registers while the current tile is written out.  Packed deposits of several
// meshes (dm non-null): item i belongs to mock item_mock[i], whose records,
// mesh and geometry come from dm[]; the arguments sorted .. mesh are unused.
template <int SCHEME, bool UNIT, int T>
__global__ void __launch_bounds__(512) k_tile_deposit(const uint3 *__restrict__ sorted,
                                                      const float *__restrict__ wsorted,
                                                      const uint4 *__restrict__ items,
                                                      const unsigned *__restrict__ nitems, int N, int ntx, int tx0,
                                                      float *__restrict__ mesh, int accumulate, int half,
                                                      const DepMock *__restrict__ dm,
                                                      const unsigned *__restrict__ item_mock) {
    extern __shared__ long long lgrid_raw[];
    void *lgrid = lgrid_raw;
    constexpr int R = SCHEME == 0 ? 0 : (SCHEME == 3 ? 2 : 1);
    constexpr int S = T + 4 * R, S3 = S * S * S, NT = 512, RPT = kTileCH / NT;
    constexpr int WB = UNIT ? 4 : 8;  // bytes per LDS cell
    static_assert((S3 * WB) % 16 == 0, "LDS grid is cleared in 16-byte words");
    const int tid = threadIdx.x;
    const unsigned total = *nitems;
    // half a cell up: fraction + 1/2; a fraction past 1 carries into the cell
    // field above it (bits 27..31 stay below 32 - kLcBias)
    const unsigned hadd = half ? 1u << (kFracBits - 1) : 0u;
    // write-out: thread tid owns tile cells q = tid + k NT, k < NW; q and
    // q + NT differ by NT / T^2 planes along x, so the LDS and mesh positions
    // of cell k are the ones of cell 0 plus k times a constant stride
    constexpr int TT = T * T * T, NW = (TT + NT - 1) / NT;
    static_assert(NT % (T * T) == 0 || TT <= NT, "write-out stride");
    constexpr int XS = NT / (T * T);
    const int wz = tid % T, wy = (tid / T) % T, wx = tid / (T * T);
    const int li0 = ((wx + 2 * R) * S + (wy + 2 * R)) * S + (wz + 2 * R);
    unsigned item = blockIdx.x;
    uint4 it = item < total ? items[item] : make_uint4(0u, 0u, 0u, 0u);
    // the mock of an item (uniform per workgroup)
    auto mock_of = [&](unsigned i) -> DepMock {
        if (dm && i < total) return dm[item_mock[i]];
        return DepMock{sorted, wsorted, mesh, N, ntx, tx0, 0};
    };
    DepMock mk = mock_of(item);
    uint3 rec[RPT];
    float wr[RPT];
    auto load_item = [&](const uint4 &x, const DepMock &m) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const unsigned idx = x.y + (unsigned)(tid + r * NT);
            if (idx < x.z) {
                rec[r] = as_global(m.sorted)[idx];
                wr[r] = UNIT ? 1.0f : as_global(m.wsorted)[idx];
            }
        }
    };
    load_item(it, mk);
    for (; item < total; item += gridDim.x) {
        const unsigned nxt = item + gridDim.x;
        const uint4 itn = nxt < total ? items[nxt] : make_uint4(0u, 0u, 0u, 0u);
        const DepMock mkn = mock_of(nxt);
        const int mN = mk.N, mntx = mk.ntx;
        uint4 *lz = reinterpret_cast<uint4 *>(lgrid);
        for (int q = tid; q < S3 * WB / 16; q += NT) lz[q] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
        const int tile = (int)it.x;
        const int tz = tile % mntx, ty = (tile / mntx) % mntx, tx = mk.tx0 + tile / (mntx * mntx);
        const int o[3] = {tx * T, ty * T, tz * T};
        const int qk = UNIT ? unit_quantum_exp<SCHEME>(it.z - it.y) : 0;
        const float qs = ldexpf(1.0f, qk);
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const unsigned idx = it.y + (unsigned)(tid + r * NT);
            if (idx >= it.z) continue;
            unsigned wd[3] = {rec[r].x + hadd, rec[r].y + hadd, rec[r].z + hadd};
            bool inside = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                // a record whose stencil misses the tile's cells (the extra ghosts
                // kept for the other mesh) is skipped
                const int lc = (int)(wd[a] >> kFracBits) - kLcBias;
                inside = inside && lc >= -R && lc < T + R;
            }
            if (!inside) continue;
            const float f[3] = {record_offset(wd[0]), record_offset(wd[1]), record_offset(wd[2])};
            int l[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)  // node c-R (local lc-R) sits at LDS index lc-R+2R = lc+R
                l[a] = (int)(wd[a] >> kFracBits) - kLcBias + R;
            deposit_record<SCHEME, UNIT, S>(lgrid, (l[0] * S + l[1]) * S + l[2], f, wr[r], qs);
        }
        const bool split = it.w != 0u;
        float *mp = mk.mesh + ((size_t)(o[0] + wx) * mN + (o[1] + wy)) * mN + (o[2] + wz);
        const size_t mstride = (size_t)XS * mN * mN;
        it = itn;
        mk = mkn;
        load_item(it, mk);  // in flight during the write-out
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            if (TT % NT != 0 && tid + k * NT >= TT) break;
            const int li = li0 + k * XS * S * S;
            const float v = UNIT ? ldexpf(__uint2float_rn(((const unsigned *)lgrid)[li]), -qk)
                                 : from_fixed(((const long long *)lgrid)[li]);
            float *dst = mp + k * mstride;
            if (split) {
                if (v != 0.0f) global_add(dst, v);
            } else if (accumulate) {
                *as_global(dst) += v;
            } else {
                store_stream(dst, v);
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
    p.rd = 1.0 / p.d;
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

// Binning of one particle range into tiles: records grouped by tile (buffers
// <tag>.sorted/.wsorted), per-tile offsets (<tag>.offsets) and the deposit work
// list (<tag>.items, <tag>.nitems); the histogram and scan buffers are shared.
// interlaced: records also feed the half-cell-shifted mesh (one more ghost
// cell towards the tile above).
template <int T>
static cm_status mas_bin_range(cm_ctx *ctx, const MasK &p, const float *pos, const float *w, bool interlaced,
                               unsigned long long *dropped, bool exact, unsigned *overflow, int tx0, int nxl,
                               const char *tag, TileBins *b, TileK *tk) {
    TileK t;
    t.T = T;
    t.R = mas_radius(p.scheme);
    t.Rup = t.R + (interlaced ? 1 : 0);
    t.ntx = p.N / T;
    t.CH = kTileCH;
    t.tx0 = tx0;
    t.nxl = nxl;
    const int ntiles = nxl * t.ntx * t.ntx;
    // a particle is copied to at most two tiles per axis: below when its
    // stencil reaches R cells under the tile edge, above when it reaches Rup
    // cells over it (Rup = 1 for NGP with the shifted mesh)
    const size_t max_copies = p.n * (t.R > 0 || t.Rup > 0 ? 8 : 1);
    const int nblk = (ntiles + 255) / 256;
    if (nblk > 4096) return fail(ctx, CM_ERR_UNSUPPORTED, "mas tiled: too many tiles");
    // LDS-privatised histogram and cursors when a workgroup's LDS holds one
    // counter per tile, else one global row
    const size_t hist_lds = sizeof(unsigned) * (size_t)ntiles;
    const bool lds_hist = hist_lds <= ctx->lds_per_block;
    const int nrows = lds_hist ? std::max(1, ctx->cu_count) : 1;
    const size_t chunk = (p.n + nrows - 1) / nrows;
    // n = NL^3 particles (a lattice-based catalogue, rows of NL particles): the
    // scatter walks each chunk in blocks of 32 rows x 32 columns when the
    // chunks are whole multiples of 32 rows (any walk of a chunk gives the same
    // records; this one keeps few record runs open at a time for lattice order)
    t.rl = 0;
    if (lds_hist && nrows > 1 && p.n % (size_t)nrows == 0) {
        size_t nl = (size_t)std::llround(std::cbrt((double)p.n));
        if (nl >= 32 && nl * nl * nl == p.n && nl % 32 == 0 && chunk % (32 * nl) == 0) t.rl = (int)nl;
    }
    *tk = t;
    void *hist, *counts, *offsets, *item_off, *bsum, *items, *nitems, *sorted;
    const std::string tg(tag);
    cm_status s;
    if ((s = scratch(ctx, "mas.hist", sizeof(unsigned) * (size_t)nrows * ntiles, &hist)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.counts", sizeof(unsigned) * ntiles, &counts)) != CM_OK) return s;
    if ((s = scratch(ctx, (tg + ".offsets").c_str(), sizeof(unsigned) * (ntiles + 1), &offsets)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.item_off", sizeof(unsigned) * (ntiles + 1), &item_off)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.bsum", sizeof(unsigned long long) * 2 * (size_t)(nblk + 1), &bsum)) != CM_OK) return s;
    const size_t max_items = (size_t)ntiles + max_copies / kTileCH + 1;
    if (max_items > 0xffffffffull) return fail(ctx, CM_ERR_UNSUPPORTED, "mas tiled: too many particles");
    if ((s = scratch(ctx, (tg + ".items").c_str(), sizeof(uint4) * max_items, &items)) != CM_OK) return s;
    if ((s = scratch(ctx, (tg + ".nitems").c_str(), sizeof(unsigned), &nitems)) != CM_OK) return s;
    auto *hp = (unsigned *)hist;
    if (lds_hist) {
        if ((s = launch(ctx, "mas_tile_count", p.soa ? k_tile_count<T, true, true> : k_tile_count<T, true, false>,
                        dim3(nrows), dim3(1024), hist_lds, pos, p, t, ntiles, chunk, hp, dropped)) != CM_OK)
            return s;
    } else {
        if ((s = fill(ctx, hp, (size_t)ntiles, 0u)) != CM_OK) return s;
        const unsigned g = grid_for(p.n, 1024 * kBatch, 2048);
        if ((s = launch(ctx, "mas_tile_count", p.soa ? k_tile_count<T, false, true> : k_tile_count<T, false, false>,
                        dim3(g), dim3(1024), 0, pos, p, t, ntiles, chunk, hp, dropped)) != CM_OK)
            return s;
    }
    if ((s = launch(ctx, "mas_tile_prefix", k_tile_prefix, dim3(nblk), dim3(256), 0, hp, nrows, ntiles,
                    (unsigned *)offsets, (unsigned *)item_off, (unsigned *)counts,
                    (unsigned long long *)bsum)) != CM_OK)
        return s;
    // record indices are 32-bit
    const size_t max_cap = 0xffffffffull;
    size_t cap;
    if (exact) {
        if ((s = launch(ctx, "mas_tile_scan", k_tile_total, dim3(1), dim3(1024), 0, (unsigned long long *)bsum,
                        nblk)) != CM_OK)
            return s;
        unsigned long long total = 0;
        CM_HIP(ctx, hipMemcpyAsync(&total, (unsigned long long *)bsum + 2 * nblk, sizeof(total),
                                   hipMemcpyDeviceToHost, ctx->stream));
        CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ++ctx->host_syncs;
        if (total > max_cap)
            return fail(ctx, CM_ERR_UNSUPPORTED, "mas tiled: %llu records exceed 32-bit indices", total);
        cap = (size_t)total;
    } else {
        // geometric ghost fraction (1 + 2R/T)^3 with 50% headroom; a larger
        // count sets *overflow and the caller repeats with exact = true
        const double g1 = 1.0 + (double)(t.R + t.Rup) / T;
        cap = std::min(std::min(max_copies, max_cap), (size_t)(1.5 * g1 * g1 * g1 * (double)p.n) + 65536);
    }
    if ((s = scratch(ctx, (tg + ".sorted").c_str(), sizeof(uint3) * (cap ? cap : 1), &sorted)) != CM_OK) return s;
    void *wsorted = nullptr;
    if (w && (s = scratch(ctx, (tg + ".wsorted").c_str(), sizeof(float) * (cap ? cap : 1), &wsorted)) != CM_OK)
        return s;
    if ((s = launch(ctx, "mas_tile_items", k_tile_items, dim3(nblk), dim3(256), 0, ntiles,
                    (const unsigned long long *)bsum, (unsigned *)offsets, (unsigned *)item_off, (unsigned *)counts,
                    (uint4 *)items, (unsigned *)nitems, (unsigned)cap, (unsigned)max_items, overflow)) != CM_OK)
        return s;
    if (lds_hist) {
        if ((s = launch(ctx, "mas_tile_scatter",
                        p.soa ? k_tile_scatter<T, true, true> : k_tile_scatter<T, true, false>, dim3(nrows),
                        dim3(1024), hist_lds, pos, w, p, t, ntiles, chunk, (const unsigned *)hp,
                        (const unsigned *)offsets, (unsigned *)counts, (uint3 *)sorted, (float *)wsorted,
                        (unsigned)cap, overflow)) != CM_OK)
            return s;
    } else {
        const unsigned g = grid_for(p.n, 1024 * kBatch, 2048);
        if ((s = launch(ctx, "mas_tile_scatter",
                        p.soa ? k_tile_scatter<T, false, true> : k_tile_scatter<T, false, false>, dim3(g),
                        dim3(1024), 0, pos, w, p, t, ntiles, chunk, (const unsigned *)hp, (const unsigned *)offsets,
                        (unsigned *)counts, (uint3 *)sorted, (float *)wsorted, (unsigned)cap, overflow)) != CM_OK)
            return s;
    }
    b->N = p.N;
    b->T = T;
    b->ntx = t.ntx;
    b->ntiles = ntiles;
    b->scheme = p.scheme;
    b->weighted = w != nullptr;
    b->sorted = sorted;
    b->wsorted = wsorted;
    b->items = items;
    b->nitems = nitems;
    b->offsets = offsets;
    b->cap = cap;
    b->max_items = max_items;
    return CM_OK;
}

// Zero the cells of the tiles of b (range t) that several workgroups deposit.
template <int T>
static cm_status mas_zero_split(cm_ctx *ctx, const TileBins &b, const TileK &t, float *mesh) {
    return launch(ctx, "mas_tile_zero_split", k_tile_zero_split<T>, dim3(b.ntiles), dim3(256), 0,
                  (const unsigned *)b.offsets, t, b.N, mesh);
}

// Deposit grid: persistent workgroups, dep_wg (default 16) per CU, at most one
// per work item.
static unsigned deposit_grid(cm_ctx *ctx, size_t max_items) {
    const int wg_per_cu = ctx->tune.dep_wg > 0 && ctx->tune.dep_wg <= 64 ? ctx->tune.dep_wg : 16;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(max_items, (size_t)std::max(1, ctx->cu_count) * wg_per_cu));
}

// Deposit launch of one scheme and weight kind: a single mock (dm null: the
// buffers of b, geometry of t) or the union work list of packed mocks.
template <int T>
static cm_status deposit_launch(cm_ctx *ctx, int scheme, bool weighted, const TileBins &b, const TileK &t,
                                const uint4 *items, const unsigned *nitems, size_t max_items, float *dst,
                                int accumulate, int half, const DepMock *dm, const unsigned *item_mock) {
    const int S = T + 4 * mas_radius(scheme);
    const size_t lds = (weighted ? sizeof(long long) : sizeof(unsigned)) * (size_t)S * S * S;
    if (lds > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "mas: deposit grid of %zu bytes", lds);
    const dim3 gd(deposit_grid(ctx, max_items)), bd(512);
    const float *ws = (const float *)b.wsorted;
#define CM_DEPOSIT(SCH, UNIT)                                                                                    \
    launch(ctx, "mas_tile_deposit", k_tile_deposit<SCH, UNIT, T>, gd, bd, lds, (const uint3 *)b.sorted, ws,   \
           items, nitems, b.N, b.ntx, t.tx0, dst, accumulate, half, dm, item_mock)
    switch (scheme * 2 + (weighted ? 0 : 1)) {
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
    return fail(ctx, CM_ERR_INVALID, "mas: scheme %d", scheme);
}

// meshB (nullable): the half-cell-shifted mesh of the same particles, deposited
// from the same records (binned with one more ghost cell towards the tile above).
template <int T>
static cm_status mas_tiled_range(cm_ctx *ctx, const MasK &p, const float *pos, const float *w, float *mesh,
                                 float *meshB, int accumulate, unsigned long long *dropped, bool exact,
                                 unsigned *overflow, int tx0, int nxl) {
    TileBins b;
    TileK t;
    cm_status s;
    if ((s = mas_bin_range<T>(ctx, p, pos, w, meshB != nullptr, dropped, exact, overflow, tx0, nxl, "mas", &b,
                              &t)) != CM_OK)
        return s;
    if (!accumulate) {
        for (float *m : {mesh, meshB})
            if (m && (s = mas_zero_split<T>(ctx, b, t, m)) != CM_OK) return s;
    }
    for (int half = 0; half < (meshB ? 2 : 1); ++half) {
        if ((s = deposit_launch<T>(ctx, p.scheme, w != nullptr, b, t, (const uint4 *)b.items,
                                   (const unsigned *)b.nitems, b.max_items, half ? meshB : mesh, accumulate, half,
                                   nullptr, nullptr)) != CM_OK)
            return s;
    }
    return CM_OK;
}

constexpr size_t kGroupTiles = 12288;

// Tiled deposit of the whole mesh: one pass when one workgroup's LDS holds a
// counter per tile, else x-slab groups (level-1 binning, one host
// synchronisation for the group sizes, exact record sizing per group).  A
// caller may ask for groups on smaller meshes too (tune.mas_groups): for
// particles in no spatial order the records of one group are then written
// while that group's part of the record buffer stays in the caches.
template <int T>
static cm_status mas_tiled_t(cm_ctx *ctx, const MasK &p, const float *pos, const float *w, float *mesh,
                             float *meshB, int accumulate, unsigned long long *dropped, bool exact,
                             unsigned *overflow) {
    const int ntx = p.N / T, R = mas_radius(p.scheme), Rup = R + (meshB ? 1 : 0);
    const size_t plane_tiles = (size_t)ntx * ntx;
    size_t per_lds = ctx->lds_per_block / sizeof(unsigned);
    if (ctx->tune.mas_hist > 0) per_lds = std::min(per_lds, (size_t)ctx->tune.mas_hist);
    if (ctx->tune.mas_groups > 1 && ntx > 1)  // caller asks for x-slab groups (unordered inputs)
        per_lds = std::min(per_lds, (size_t)((ntx + ctx->tune.mas_groups - 1) / ctx->tune.mas_groups) * plane_tiles);
    if ((size_t)ntx * plane_tiles <= per_lds || plane_tiles > per_lds)
        return mas_tiled_range<T>(ctx, p, pos, w, mesh, meshB, accumulate, dropped, exact, overflow, 0, ntx);
    // once grouped, groups of about kGroupTiles tiles: the record region of a group
    // then stays cache-resident while it is written (1024^3 interlaced mock with
    // 20% randoms: 7 groups 81.9 ms, 11: 79.1, 16: 77.3, 22: 76.3, 32: 76.5)
    per_lds = std::min(per_lds, std::max(plane_tiles, kGroupTiles));
    const int gs = (int)std::min<size_t>((size_t)ntx, per_lds / plane_tiles);  // tile planes per group
    const int G = (ntx + gs - 1) / gs;
    const int nb = ntx * ntx;  // tile rows
    if (nb > 4096) return fail(ctx, CM_ERR_UNSUPPORTED, "mas tiled: %d tile rows", nb);
    const int nrows = std::max(1, ctx->cu_count);
    const size_t chunk = (p.n + nrows - 1) / nrows;
    void *hist, *rtot, *rstart;
    cm_status s;
    if ((s = scratch(ctx, "mas.ghist", sizeof(unsigned) * (size_t)nrows * nb, &hist)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.rtot", sizeof(unsigned long long) * (size_t)nb, &rtot)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.gstart", sizeof(unsigned long long) * (size_t)(nb + 1), &rstart)) != CM_OK) return s;
    if ((s = launch(ctx, "mas_group_count", p.soa ? k_group_count<T, true> : k_group_count<T, false>, dim3(nrows),
                    dim3(1024), sizeof(unsigned) * (size_t)nb, pos, p, R, Rup, ntx, gs, G, chunk, (unsigned *)hist,
                    dropped)) != CM_OK)
        return s;
    if ((s = launch(ctx, "mas_group_scan", k_group_colscan, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0,
                    (unsigned *)hist, nrows, nb, (unsigned long long *)rtot)) != CM_OK)
        return s;
    if ((s = launch(ctx, "mas_group_scan", k_group_scan, dim3(1), dim3(1024), 0, (const unsigned long long *)rtot, nb,
                    (unsigned long long *)rstart)) != CM_OK)
        return s;
    std::vector<unsigned long long> rs_h((size_t)nb + 1);
    CM_HIP(ctx, hipMemcpyAsync(rs_h.data(), rstart, sizeof(unsigned long long) * rs_h.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ++ctx->host_syncs;
    // group g holds the rows of tile planes [g gs, (g+1) gs)
    std::vector<unsigned long long> gs_h((size_t)G + 1);
    for (int g = 0; g <= G; ++g) gs_h[(size_t)g] = rs_h[(size_t)std::min(g * gs, ntx) * ntx];
    const unsigned long long total = gs_h[(size_t)G];
    if (total > 0xffffffffull) return fail(ctx, CM_ERR_UNSUPPORTED, "mas tiled: %llu grouped particles", total);
    void *gpos, *gw = nullptr;
    if ((s = scratch(ctx, "mas.gpos", sizeof(float3) * (size_t)(total ? total : 1), &gpos)) != CM_OK) return s;
    if (w && (s = scratch(ctx, "mas.gw", sizeof(float) * (size_t)(total ? total : 1), &gw)) != CM_OK) return s;
    if ((s = launch(ctx, "mas_group_scatter", p.soa ? k_group_scatter<T, true> : k_group_scatter<T, false>,
                    dim3(nrows), dim3(1024), sizeof(unsigned long long) * (size_t)nb, pos, w, p, R, Rup, ntx, gs, G,
                    chunk, (const unsigned *)hist, (const unsigned long long *)rstart, (float3 *)gpos,
                    (float *)gw)) != CM_OK)
        return s;
    for (int g = 0; g < G; ++g) {
        MasK q = p;
        q.soa = 0;
        q.n = (size_t)(gs_h[(size_t)g + 1] - gs_h[(size_t)g]);
        const float *qp = (const float *)gpos + 3 * (size_t)gs_h[(size_t)g];
        const float *qw = gw ? (const float *)gw + (size_t)gs_h[(size_t)g] : nullptr;
        if ((s = mas_tiled_range<T>(ctx, q, qp, qw, mesh, meshB, accumulate, dropped, true, overflow, g * gs,
                                    std::min(gs, ntx - g * gs))) != CM_OK)
            return s;
    }
    return CM_OK;
}

// Capacity of the sorted-record buffer: when `exact`, the total number of
// records is read back after the scan (one host synchronisation) and the buffer
// sized to it; otherwise a generous estimate is used and the scatter reports
// overflow through `overflow` (checked by the caller).
static cm_status mas_tiled(cm_ctx *ctx, const MasK &p, const float *pos, const float *w, float *mesh,
                           float *meshB, int accumulate, unsigned long long *dropped, bool exact,
                           unsigned *overflow) {
    switch (tile_edge(p.N)) {
    case 16: return mas_tiled_t<16>(ctx, p, pos, w, mesh, meshB, accumulate, dropped, exact, overflow);
    case 8: return mas_tiled_t<8>(ctx, p, pos, w, mesh, meshB, accumulate, dropped, exact, overflow);
    }
    return fail(ctx, CM_ERR_UNSUPPORTED, "mas tiled: N=%d", p.N);
}

// ------------------------------------------------------ packed deposits
// Union of the work lists of several mocks: mock m's items go to
// out[off_m + i] with off_m the sum of the earlier mocks' item counts (device
// values); item_mock records the mock of every item.
struct PackSrc {
    const uint4 *items;
    const unsigned *nitems;
};

__global__ void __launch_bounds__(256) k_pack_items(const PackSrc *__restrict__ src, int nm, uint4 *__restrict__ out,
                                                    unsigned *__restrict__ item_mock, unsigned *__restrict__ total) {
    const int m = blockIdx.y;
    unsigned off = 0;
    for (int k = 0; k < m; ++k) off += *src[k].nitems;
    const unsigned cnt = *src[m].nitems;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        out[off + i] = src[m].items[i];
        item_mock[off + i] = (unsigned)m;
    }
    if (m == nm - 1 && blockIdx.x == 0 && threadIdx.x == 0) *total = off + cnt;
}

bool mas_packable(cm_ctx *ctx, const cm_mas_desc *d) {
    if (d->pos_dtype != CM_F32 || tile_edge(d->N) != 16) return false;
    const size_t ntx = (size_t)d->N / 16;
    return sizeof(unsigned) * ntx * ntx * ntx <= ctx->lds_per_block;  // one binning range
}

cm_status mas_bin_mock(cm_ctx *ctx, const cm_mas_desc *d, const float *pos, const float *w, size_t n,
                       const char *tag, unsigned *overflow, TileBins *b) {
    if (!mas_packable(ctx, d)) return fail(ctx, CM_ERR_UNSUPPORTED, "packed binning: N=%d", d->N);
    MasK p = make_mask(d, n);
    void *dropped;
    cm_status s;
    if ((s = scratch(ctx, "mas.dropped", sizeof(unsigned long long), &dropped)) != CM_OK) return s;
    if ((s = fill(ctx, (unsigned long long *)dropped, 1, 0ull)) != CM_OK) return s;
    TileK t;
    if ((s = mas_bin_range<16>(ctx, p, pos, w, false, (unsigned long long *)dropped, true, overflow, 0, p.N / 16,
                               tag, b, &t)) != CM_OK)
        return s;
    return CM_OK;
}

cm_status mas_deposit_packed(cm_ctx *ctx, const std::vector<TileBins> &bins, const std::vector<float *> &meshes) {
    const int nm = (int)bins.size();
    if (nm == 0) return CM_OK;
    if (meshes.size() != bins.size()) return fail(ctx, CM_ERR_INVALID, "packed deposit: %d meshes", (int)meshes.size());
    const int scheme = bins[0].scheme;
    const bool weighted = bins[0].weighted;
    std::vector<DepMock> dm((size_t)nm);
    std::vector<PackSrc> src((size_t)nm);
    size_t max_items = 0, per_mock = 0;
    cm_status s;
    for (int m = 0; m < nm; ++m) {
        const TileBins &b = bins[(size_t)m];
        if (b.scheme != scheme || b.weighted != weighted || b.T != 16)
            return fail(ctx, CM_ERR_INVALID, "packed deposit: mixed schemes or weights");
        TileK t;
        t.T = 16;
        t.R = mas_radius(scheme);
        t.Rup = t.R;
        t.ntx = b.ntx;
        t.CH = kTileCH;
        t.tx0 = 0;
        t.nxl = b.ntx;
        t.rl = 0;
        if ((s = mas_zero_split<16>(ctx, b, t, meshes[(size_t)m])) != CM_OK) return s;
        dm[(size_t)m] = DepMock{(const uint3 *)b.sorted, (const float *)b.wsorted, meshes[(size_t)m], b.N, b.ntx, 0, 0};
        src[(size_t)m] = PackSrc{(const uint4 *)b.items, (const unsigned *)b.nitems};
        max_items += b.max_items;
        per_mock = std::max(per_mock, b.max_items);
    }
    void *ddm, *dsrc, *uitems, *umock, *utotal;
    if ((s = scratch(ctx, "mas.pk.dm", sizeof(DepMock) * (size_t)nm, &ddm)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.pk.src", sizeof(PackSrc) * (size_t)nm, &dsrc)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.pk.items", sizeof(uint4) * max_items, &uitems)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.pk.mock", sizeof(unsigned) * max_items, &umock)) != CM_OK) return s;
    if ((s = scratch(ctx, "mas.pk.total", sizeof(unsigned), &utotal)) != CM_OK) return s;
    // the descriptor tables are small; a synchronous copy keeps the host
    // vectors out of the stream's lifetime
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CM_HIP(ctx, host_upload(ctx, ddm, dm.data(), sizeof(DepMock) * (size_t)nm));
    CM_HIP(ctx, host_upload(ctx, dsrc, src.data(), sizeof(PackSrc) * (size_t)nm));
    ++ctx->host_syncs;
    const unsigned gx = (unsigned)std::min<size_t>(1024, (per_mock + 255) / 256);
    if ((s = launch(ctx, "mas_pack_items", k_pack_items, dim3(gx, (unsigned)nm), dim3(256), 0, (const PackSrc *)dsrc,
                    nm, (uint4 *)uitems, (unsigned *)umock, (unsigned *)utotal)) != CM_OK)
        return s;
    TileK t0;
    t0.tx0 = 0;
    return deposit_launch<16>(ctx, scheme, weighted, bins[0], t0, (const uint4 *)uitems, (const unsigned *)utotal,
                              max_items, meshes[0], 0, 0, (const DepMock *)ddm, (const unsigned *)umock);
}

cm_status mesh_assign_impl(cm_ctx *ctx, const cm_mas_desc *d, const void *pos, const void *w, size_t n,
                           void *mesh, cm_dtype mdt, int accumulate, bool exact, unsigned **overflow_out,
                           bool reset_overflow, unsigned *overflow_at, void *meshB) {
    if (overflow_out) *overflow_out = nullptr;
    MasK p = make_mask(d, n);
    const size_t n3 = (size_t)d->N * d->N * d->N;
    void *dropped;
    cm_status s;
    if ((s = scratch(ctx, "mas.dropped", sizeof(unsigned long long), &dropped)) != CM_OK) return s;
    bool tiled_ok = d->pos_dtype == CM_F32 && mdt == CM_F32 && tile_edge(d->N) > 0;
    bool use_tiled = d->algo == CM_MAS_TILED || (d->algo == CM_MAS_AUTO && tiled_ok);
    if (use_tiled && !tiled_ok)
        return fail(ctx, CM_ERR_UNSUPPORTED, "tiled mass assignment needs float32 positions and mesh and N%%8==0, N>=16");
    if (use_tiled) {
        void *ovf = overflow_at;
        if (!ovf && (s = scratch(ctx, "mas.overflow", sizeof(unsigned), &ovf)) != CM_OK) return s;
        if ((s = launch(ctx, "fill", k_clear_flags, dim3(1), dim3(64), 0, (unsigned long long *)dropped,
                        reset_overflow ? (unsigned *)ovf : (unsigned *)nullptr)) != CM_OK)
            return s;
        if (overflow_out) *overflow_out = (unsigned *)ovf;
        return mas_tiled(ctx, p, (const float *)pos, (const float *)w, (float *)mesh, (float *)meshB, accumulate,
                         (unsigned long long *)dropped, exact || !overflow_out, (unsigned *)ovf);
    }
    if ((s = fill(ctx, (unsigned long long *)dropped, 1, 0ull)) != CM_OK) return s;
    if (meshB) {  // plain path: the shifted mesh is a second deposit
        cm_mas_desc db = *d;
        db.interlace_shift += 0.5;
        if ((s = mesh_assign_impl(ctx, &db, pos, w, n, meshB, mdt, accumulate, exact, nullptr, false, overflow_at)) !=
            CM_OK)
            return s;
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
