// T-web classification for gfx950 (plain kernels).
//   delta -> r2c -> exp(-k^2R^2/2)(-1/k^2) -> c2r -> phi            (Poisson)
//   phi -> G_a = D_a phi (a = x,y,z)                                 (k_grad3)
//   G -> T_ab = D_b G_a -> FP64 eigenvalues -> label, counts         (k_hess_eig)
// The lower triangle the reference feeds to GSL is A10 = D_x G_y,
// A20 = D_x G_z, A21 = D_y G_z (Miscelanious.cpp:1684-1740).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

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

// Spectral Hessian component (calc_LapPhiv, Source/NumericalMethods.cpp:
// 2605-2679): out = scale W(k) k_a k_b / k^2 delta_k, zero at k = 0, with the
// wavenumbers of kcoord (index N/2 at +N/2 k_f, as the reference's coords).
template <typename T>
__global__ void __launch_bounds__(256) k_kab(const typename std::conditional<sizeof(T) == 4, float2, double2>::type
                                                 *__restrict__ dk,
                                             typename std::conditional<sizeof(T) == 4, float2, double2>::type
                                                 *__restrict__ out,
                                             int N, double kf, double R2h, double scale, int a, int b) {
    const int M1 = N / 2 + 1;
    const size_t total = (size_t)N * N * M1;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(g % M1);
        const size_t r = g / M1;
        const int iy = (int)(r % N), ix = (int)(r / N);
        const double k[3] = {kcoord(ix, N, kf), kcoord(iy, N, kf), kf * (double)iz};
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const double f = k2 > 0.0 ? scale * exp(-k2 * R2h) * (k[a] * k[b]) / k2 : 0.0;
        auto v = dk[g];
        v.x = (T)((double)v.x * f);
        v.y = (T)((double)v.y * f);
        out[g] = v;
    }
}

// Labels, lambdas and counters from the six tensor fields T (planes xx, yy,
// zz, yx, zx, zy of n3 values each): FP64 closed-form eigenvalues per cell,
// as the plain FD path.
template <typename T>
__global__ void __launch_bounds__(256) k_tensor_eig(const T *__restrict__ t6, size_t n3, double lth, TolK tk,
                                                    int8_t *__restrict__ labels, T *__restrict__ lam,
                                                    unsigned long long *__restrict__ stats) {
    __shared__ unsigned cnt[5];
    if (threadIdx.x < 5) cnt[threadIdx.x] = 0;
    __syncthreads();
    const double tol = tol_value(tk);
    unsigned loc[5] = {0, 0, 0, 0, 0};
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n3; c += (size_t)gridDim.x * blockDim.x) {
        double e1, e2, e3;
        eig3(t6[c], t6[n3 + c], t6[2 * n3 + c], t6[3 * n3 + c], t6[4 * n3 + c], t6[5 * n3 + c], e1, e2, e3);
        const int v = (e1 > lth) + (e2 > lth) + (e3 > lth);
        const int lab = 4 - v;
        labels[c] = (int8_t)lab;
        if (lam) {
            lam[c] = (T)e1;
            lam[n3 + c] = (T)e2;
            lam[2 * n3 + c] = (T)e3;
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

// ||T||_F,rms = sqrt(sum over cells of sum_ab T_ab^2 / n3) from the per-
// workgroup sums of squares of the six planes (part[c * np + i]; the three
// off-diagonal planes count twice), as float bits.
__global__ void __launch_bounds__(1024) k_rms_tensor(const double *__restrict__ part, int np, double n,
                                                     unsigned *__restrict__ out) {
    __shared__ double red[1024];
    double s = 0.0;
    for (int c = 0; c < 6; ++c)
        for (int i = threadIdx.x; i < np; i += blockDim.x) s += (c < 3 ? 1.0 : 2.0) * part[(size_t)c * np + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = __float_as_uint((float)sqrt(red[0] / n));
}

// G_a = D_a phi, stored in the potential's type (float32: the stencil in
// double, rounded to float, as the reference's float build; FP64: all double).
template <typename T>
__device__ __forceinline__ T fd4_t(T fll, T fl, T fr, T frr, double ih) {
    if constexpr (sizeof(T) == 4)
        return fd4(fll, fl, fr, frr, ih);
    else
        return fd4d(fll, fl, fr, frr, ih);
}

template <typename T>
__global__ void __launch_bounds__(256) k_grad3(const T *__restrict__ phi, int N, double ih, T *__restrict__ gx,
                                               T *__restrict__ gy, T *__restrict__ gz) {
    const size_t n3 = (size_t)N * N * N;
    const size_t NN = (size_t)N;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n3; c += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(c % N), y = (int)((c / N) % N), x = (int)(c / (NN * N));
        auto at = [&](int i, int j, int k) { return phi[((size_t)i * NN + j) * NN + k]; };
        const int xm2 = (x - 2 + N) % N, xm1 = (x - 1 + N) % N, xp1 = (x + 1) % N, xp2 = (x + 2) % N;
        const int ym2 = (y - 2 + N) % N, ym1 = (y - 1 + N) % N, yp1 = (y + 1) % N, yp2 = (y + 2) % N;
        const int zm2 = (z - 2 + N) % N, zm1 = (z - 1 + N) % N, zp1 = (z + 1) % N, zp2 = (z + 2) % N;
        gx[c] = fd4_t<T>(at(xm2, y, z), at(xm1, y, z), at(xp1, y, z), at(xp2, y, z), ih);
        gy[c] = fd4_t<T>(at(x, ym2, z), at(x, ym1, z), at(x, yp1, z), at(x, yp2, z), ih);
        gz[c] = fd4_t<T>(at(x, y, zm2), at(x, y, zm1), at(x, y, zp1), at(x, y, zp2), ih);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_hess_eig(const T *__restrict__ gx, const T *__restrict__ gy,
                                                  const T *__restrict__ gz, int N, double ih, double lth, TolK tk,
                                                  int8_t *__restrict__ labels, T *__restrict__ lam,
                                                  unsigned long long *__restrict__ stats) {
    __shared__ unsigned cnt[5];
    if (threadIdx.x < 5) cnt[threadIdx.x] = 0;
    __syncthreads();
    const double tol = tol_value(tk);
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
        const T txx = fd4_t<T>(gx[id(xm2, y, z)], gx[id(xm1, y, z)], gx[id(xp1, y, z)], gx[id(xp2, y, z)], ih);
        const T tyx = fd4_t<T>(gy[id(xm2, y, z)], gy[id(xm1, y, z)], gy[id(xp1, y, z)], gy[id(xp2, y, z)], ih);
        const T tzx = fd4_t<T>(gz[id(xm2, y, z)], gz[id(xm1, y, z)], gz[id(xp1, y, z)], gz[id(xp2, y, z)], ih);
        // T_yy = D_y G_y, A21 = D_y G_z
        const T tyy = fd4_t<T>(gy[id(x, ym2, z)], gy[id(x, ym1, z)], gy[id(x, yp1, z)], gy[id(x, yp2, z)], ih);
        const T tzy = fd4_t<T>(gz[id(x, ym2, z)], gz[id(x, ym1, z)], gz[id(x, yp1, z)], gz[id(x, yp2, z)], ih);
        // T_zz = D_z G_z
        const T tzz = fd4_t<T>(gz[id(x, y, zm2)], gz[id(x, y, zm1)], gz[id(x, y, zp1)], gz[id(x, y, zp2)], ih);
        double e1, e2, e3;
        eig3(txx, tyy, tzz, tyx, tzx, tzy, e1, e2, e3);
        const int v = (e1 > lth) + (e2 > lth) + (e3 > lth);
        const int lab = 4 - v;
        labels[c] = (int8_t)lab;
        if (lam) {
            lam[c] = (T)e1;
            lam[n3 + c] = (T)e2;
            lam[2 * n3 + c] = (T)e3;
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

// Per-workgroup FP64 sums of squares of a[0..n) into part[blockIdx.x].
template <typename T>
__global__ void __launch_bounds__(256) k_sumsq(const T *__restrict__ a, size_t n, double *__restrict__ part) {
    __shared__ double red[4];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double v = (double)a[i];
        s += v * v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// *out = bits of the float sqrt(sum(part[0..np)) / n) (one workgroup).
__global__ void __launch_bounds__(1024) k_rms_final(const double *__restrict__ part, int np, double n,
                                                    unsigned *__restrict__ out) {
    __shared__ double red[16];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; ++w) t += red[w];
        *out = __float_as_uint((float)sqrt(t / n));
    }
}

// Fused FD Hessian + labels.  One workgroup owns a TY x TZ column of cells in
// (y,z) and a segment of x planes and marches along x, one barrier per plane:
//   step s: phi plane s enters the LDS ring; barrier; G_a = D_a phi of plane
//   s-2 (halo 2 in y,z for G_y, G_z) from ring planes s-2..s and the centre
//   values of planes s-4, s-3 (kept in registers from the two steps before),
//   G_y and G_z written to the G slot of plane s-2; T of plane c = s-4 from
//   G_x, G_y, G_z of planes c-2..c+2 (x derivative, register histories) and
//   the G slot of plane c (y, z derivatives, written two steps earlier).
// Five phi slots (plane s-4 is no longer read from LDS in step s, so plane
// s+1 may replace it before the next barrier) and three G slots make every
// LDS slot's next write follow the barrier after its last read.  Each thread
// owns two z-adjacent cells (z even, z+1), so every LDS access is one aligned
// 8-byte pair and the z stencils of the two cells share their reads; it keeps
// the pair's last five G_x, G_y, G_z values in registers, and the march is
// unrolled by five so the history slots are compile-time indices.  The next
// phi plane is loaded into registers while the current one is processed.  G
// and T never touch HBM.  Composition order of the reference: G_a = D_a phi,
// then D_x G_x, D_x G_y, D_x G_z, D_y G_y, D_y G_z, D_z G_z
// (Miscelanious.cpp:1684-1700).
// Labels from the inertia of T - lambda_th I: float signs with a certified
// rounding bound, recounted in FP64 in the kernel where the float signs are
// not certain.  The same coefficients, shifted by -+ tol, give the sign
// counts of T - (lambda_th -+ tol) I: the cell is marginal when they differ
// (an eigenvalue lies within tol of lambda_th), decided in FP64 when a shifted
// sign is not certain either.  Class and margin counts are kept per wave from
// the sign masks of the wave's cells (scalar population counts).  The
// marching loop makes no calls except for the lambda output (INL = true),
// which takes the out-of-line FP64 branch for every cell (a call site inside
// the loop costs ~35 VGPRs of saved registers).

// Lambda output of the fused kernel (kept out of line so its FP64 registers do
// not count against the marching loop): closed-form FP64 eigenvalues of T =
// T' ih2 (the kernel works with T' = T/ih^2), descending.
__device__ __attribute__((noinline)) void eig_store(float txx, float tyy, float tzz, float a10, float a20, float a21,
                                                    float ih2, float *lam, size_t cell, size_t n3) {
    double e1, e2, e3;
    eig3(txx * ih2, tyy * ih2, tzz * ih2, a10 * ih2, a20 * ih2, a21 * ih2, e1, e2, e3);
    lam[cell] = (float)e1;
    lam[n3 + cell] = (float)e2;
    lam[2 * n3 + cell] = (float)e3;
}

// FP64 decision of the uncertain float case, out of line (lambda variant).
__device__ __attribute__((noinline)) int positive_eigs_band_d(float b00, float b11, float b22, float a10, float a20,
                                                              float a21, float tol, bool &band) {
    return positive_eigs_band64(b00, b11, b22, a10, a20, a21, tol, band);
}

// LDS of the fused T-web: five phi slots, each padded to a whole number of
// the workgroup's loads ((TY+8)(TZ+8) rounded up to a multiple of NT), and
// three G_y and three G_z planes of (TY+4)(TZ+4).
template <int TY, int TZ, int NT>
constexpr size_t tweb_lds_floats() {
    return 5 * (((size_t)(TY + 8) * (TZ + 8) + NT - 1) / NT * NT) + 6 * (size_t)(TY + 4) * (TZ + 4);
}

// An LDS address the compiler treats as opaque: the reads through it take it
// as their own base register with small immediate offsets.  Without it every
// ring slot's reads share one base, and the compiler pairs reads whose
// offsets do not fit the paired instruction by adding a new base per pair.
__device__ __forceinline__ const lmem<float> *lds_base(const lmem<float> *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(p));
#endif
    return p;
}

// Per lane: the sum of the lane's bits in three lane masks of the wave (a
// select and two adds with carry-in from the masks; written out so that the
// compiler keeps the masks for the scalar counts instead of re-deriving them
// from the sum).
__device__ __forceinline__ unsigned lane_bits3(unsigned long long a, unsigned long long b, unsigned long long c) {
    unsigned r = 0;
#if defined(__HIP_DEVICE_COMPILE__)
    asm("v_cndmask_b32_e64 %0, 0, 1, %1\n\t"
        "v_addc_co_u32_e64 %0, vcc, 0, %0, %2\n\t"
        "v_addc_co_u32_e64 %0, vcc, 0, %0, %3"
        : "=&v"(r)
        : "s"(a), "s"(b), "s"(c)
        : "vcc");
#endif
    return r;
}

__device__ __forceinline__ float2 lds_pair(const lmem<float> *p, int off) {
    return *(const lmem<float2> *)(p + off);
}

template <int TY, int TZ, int NT, bool INL, bool LZ>
__global__ void __launch_bounds__(NT) k_tweb_fused(const float *__restrict__ phi_a, int N_a, int segLen_a,
                                                   float ih2_a, double lth_a, TolK tk_a, int8_t *__restrict__ labels_a,
                                                   float *__restrict__ lam, unsigned long long *__restrict__ stats_a,
                                                   const TwMock *__restrict__ tm, int nm) {
    constexpr int HY = TY + 8, HZ = TZ + 8, HP = HY * HZ;  // haloed phi plane (halo 4)
    constexpr int GZ = TZ + 4, GP = (TY + 4) * GZ;         // G plane (halo 2)
    constexpr int TZ2 = TZ / 2;                            // cell pairs per row
    constexpr int PF = (HP + NT - 1) / NT;                 // phi values per thread per plane
    constexpr int HPP = PF * NT;                           // ring slot stride (padded)
    constexpr int NH1 = 4 * TZ2, NH2 = 2 * TY;             // halo pairs: y rows, z column pairs
    static_assert(TZ % 2 == 0 && TY * TZ == 2 * NT && NH1 + NH2 <= NT, "tile does not match the workgroup");
    extern __shared__ __attribute__((aligned(16))) float tw_sm[];
    const lmem<float> *ring = as_local(tw_sm);  // 5 phi planes
    lmem<float> *sgy = as_local(tw_sm) + 5 * HPP;  // 3 G_y planes
    lmem<float> *sgz = sgy + 3 * GP;              // 3 G_z planes
    __shared__ unsigned scnt[5];
    const int tid = threadIdx.x;
    if (tid < 5) scnt[tid] = 0;
    // the workgroup's potential and work item: (tile column, x segment) of the
    // one potential (tm null), or of the mock that owns it in a packed launch
    const float *phi = phi_a;
    int8_t *labels = labels_a;
    unsigned long long *stats = stats_a;
    int N = N_a, segLen = segLen_a;
    float ih2 = ih2_a;
    double lth = lth_a;
    TolK tk = tk_a;
    // tile columns in XCD order: workgroups are dispatched to the eight XCDs in
    // turn, so tile t runs on XCD bx mod 8; this map gives each XCD a run of
    // gridDim.x / 8 consecutive tiles, whose neighbours (the halo re-reads)
    // then come from that XCD's L2 (512^3: phi reads 2.0 -> 0.87 GB,
    // 0.461 -> 0.450 ms)
    int bx = gridDim.x % 8 == 0 ? (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8 : blockIdx.x, by = blockIdx.y;
    if (tm) {
        int m = 0;
        while (m + 1 < nm && (int)blockIdx.x >= tm[m + 1].first) ++m;
        const TwMock &t = tm[m];
        phi = t.phi;
        labels = t.labels;
        stats = t.stats;
        N = t.N;
        segLen = t.segLen;
        ih2 = t.ih2;
        lth = t.lth;
        tk = t.tk;
        const int local = (int)blockIdx.x - t.first, tiles = (N / TY) * (N / TZ);
        bx = local % tiles;
        by = local / tiles;
    }
    const int ntz = N / TZ;
    const int ty0 = (bx / ntz) * TY, tz0 = (bx % ntz) * TZ;
    const int x0 = by * segLen, x1 = min(N, x0 + segLen);
    const size_t NN = (size_t)N, n3 = NN * NN * NN;
    // T' = T/ih^2: threshold and margin scaled to match
    const double tol = tol_value(tk);
    // (uniform: the tolerance may come through a generic pointer, whose load
    // the compiler takes as per-lane, and a per-lane tolerance would turn the
    // band test's branch and the wave masks into per-lane code)
    // (LZ: lambda_th = 0, whose subtraction is the identity in float)
    const float lthf =
        LZ ? 0.0f : __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)(lth / (double)ih2))));
    const float tols = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)(tol / (double)ih2))));

    // this thread's phi loads: byte offsets in a plane (N^2 floats, below
    // 4 GiB for every supported N) with periodic wrap in y and z; the slot's
    // padding positions take an offset beyond the plane, which the buffer
    // load's range check reads as zero
    unsigned foff[PF];
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const int q = tid + k * NT;
        const int yy = q / HZ, zz = q - (q / HZ) * HZ;
        foff[k] = q < HP ? 4u * (unsigned)(((ty0 - 4 + yy + N) % N) * N + (tz0 - 4 + zz + N) % N) : 0xfffffff0u;
    }
    // the cell pair: row ia, columns ib and ib+1 of the tile.  rb[j] points
    // at (row - 2, column - 2) of the pair in ring slot j: every read of a
    // slot lies within 1.2 KB above it.
    const int ia = tid / TZ2, ib = 2 * (tid - (tid / TZ2) * TZ2);
    constexpr int PB = 2 * HZ + 2, GB = 2 * GZ + 2;  // base offsets below a cell's own index
    const int po = (ia + 4) * HZ + ib + 4, gq = (ia + 2) * GZ + ib + 2;
    const unsigned coff = (unsigned)((ty0 + ia) * N + tz0 + ib);  // output offset in a plane
    const lmem<float> *rb[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) rb[j] = lds_base(ring + j * HPP + (po - PB));
    // at most one halo pair per thread: kind 1 = y-halo row (G_y, G_z),
    // kind 2 = z-halo column pair (G_z)
    int hkind = 0, hq = 0, hpo = 0;
    if (tid < NH1) {
        const int r = tid / TZ2, a = r < 2 ? r : TY + r, b = 2 + 2 * (tid - r * TZ2);
        hkind = 1;
        hq = a * GZ + b;
        hpo = (a + 2) * HZ + b + 2;
    } else if (tid < NH1 + NH2) {
        const int j = tid - NH1, a = 2 + (j >> 1), b = (j & 1) ? TZ + 2 : 0;
        hkind = 2;
        hq = a * GZ + b;
        hpo = (a + 2) * HZ + b + 2;
    }

    // histories of the pair: G_x, G_y, G_z of the last five planes, phi at
    // the cells of the last five planes
    float hx[2][5], hy[2][5], hz[2][5], pc[2][5];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int i = 0; i < 5; ++i) hx[k][i] = hy[k][i] = hz[k][i] = pc[k][i] = 0.f;
    // per-wave counts (uniform): cells with 3 and with at least one positive
    // eigenvalue, with an odd count, marginal cells
    unsigned wn3 = 0, wany = 0, wodd = 0, wmarg = 0;
    float pf[PF];
    // planes are fetched in order x0-4, x0-3, ... (periodic in x): fp is the
    // next one's index, fptr its first value
    const unsigned plane_bytes = (unsigned)(NN * NN * sizeof(float));
    int fp = ((x0 - 4) % N + N) % N;
    const float *fptr = phi + (size_t)fp * NN * NN;
    auto fetch = [&]() {
        const auto rs = buffer_rsrc(fptr, plane_bytes);
#pragma unroll
        for (int k = 0; k < PF; ++k) pf[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, foff[k], 0, 0));
        if (++fp == N) {
            fp = 0;
            fptr = phi;
        } else {
            fptr += NN * NN;
        }
    };
    // G of plane p (p >= x0 - 2) in G slot (p - x0 + 3) mod 3; gk is the slot
    // of plane s-2 in step s (0 in the first step, s = x0-4)
    int gk = 0;
    // labels of plane c, the next output plane
    int8_t *lptr = labels + (size_t)x0 * NN * NN;

    // Step s (J = (s - x0 + 4) mod 5): plane s enters ring slot (J+4)%5 (the
    // slot of plane s-5); phi of plane s-2 at the pair and G at plane s-2 go
    // to history slot J; cells of plane c = s-4 are output, with G(c-2..c+2)
    // in history slots (J+1)%5, (J+2)%5, (J+3)%5, (J+4)%5, J.
    auto step = [&](auto Jc, int s) {
        constexpr int J = decltype(Jc)::value;
        constexpr int S2 = (J + 2) % 5, S1 = (J + 3) % 5, S0 = (J + 4) % 5;
        lmem<float> *cur = as_local(tw_sm) + S0 * HPP;
#pragma unroll
        for (int k = 0; k < PF; ++k) cur[tid + k * NT] = pf[k];
        if (s + 1 < x1 + 4) fetch();  // plane s+1, in flight while plane s is processed
        // G slots of planes s-2 (written) and s-4 (read)
        const int gw = gk * GP, gs = (gk == 2 ? 0 : gk + 1) * GP;
        gk = gk == 2 ? 0 : gk + 1;
        __syncthreads();
        // registers are written unconditionally (before x0 the values come from
        // a partly filled ring and are never read), so no history value has to
        // survive a skipped write
        {
            const lmem<float> *b = rb[S2];
            const float2 y0 = lds_pair(b, 2), y1 = lds_pair(b, HZ + 2), y3 = lds_pair(b, 3 * HZ + 2),
                         y4 = lds_pair(b, 4 * HZ + 2);
            const float2 z0 = lds_pair(b, 2 * HZ), zc = lds_pair(b, 2 * HZ + 2), z2 = lds_pair(b, 2 * HZ + 4);
            const float2 c1 = lds_pair(rb[S1], 2 * HZ + 2), c0 = lds_pair(rb[S0], 2 * HZ + 2);
            // phi of planes s-4 and s-3 at the pair: read in the two steps before
            hx[0][J] = fd4s(pc[0][(J + 3) % 5], pc[0][(J + 4) % 5], c1.x, c0.x);
            hx[1][J] = fd4s(pc[1][(J + 3) % 5], pc[1][(J + 4) % 5], c1.y, c0.y);
            pc[0][J] = zc.x;
            pc[1][J] = zc.y;
            hy[0][J] = fd4s(y0.x, y1.x, y3.x, y4.x);
            hy[1][J] = fd4s(y0.y, y1.y, y3.y, y4.y);
            hz[0][J] = fd4s(z0.x, z0.y, zc.y, z2.x);
            hz[1][J] = fd4s(z0.y, zc.x, z2.x, z2.y);
        }
        if (s >= x0) {
            *(lmem<float2> *)(sgy + gw + gq) = make_float2(hy[0][J], hy[1][J]);
            *(lmem<float2> *)(sgz + gw + gq) = make_float2(hz[0][J], hz[1][J]);
            if (hkind) {
                const lmem<float> *b = lds_base(ring + S2 * HPP + (hpo - PB));
                const float2 z0 = lds_pair(b, 2 * HZ), zc = lds_pair(b, 2 * HZ + 2), z2 = lds_pair(b, 2 * HZ + 4);
                *(lmem<float2> *)(sgz + gw + hq) =
                    make_float2(fd4s(z0.x, z0.y, zc.y, z2.x), fd4s(z0.y, zc.x, z2.x, z2.y));
                if (hkind == 1) {
                    const float2 y0 = lds_pair(b, 2), y1 = lds_pair(b, HZ + 2), y3 = lds_pair(b, 3 * HZ + 2),
                                 y4 = lds_pair(b, 4 * HZ + 2);
                    *(lmem<float2> *)(sgy + gw + hq) =
                        make_float2(fd4s(y0.x, y1.x, y3.x, y4.x), fd4s(y0.y, y1.y, y3.y, y4.y));
                }
            }
        }
        const int c = s - 4;
        if (c < x0) return;
        const lmem<float> *gyb = lds_base(sgy + gs + (gq - GB)), *gzb = lds_base(sgz + gs + (gq - GB));
        constexpr int JC = (J + 3) % 5;  // history slot of plane c
        // T of the pair: txx, tyy, tzz, a10 = D_x G_y, a20 = D_x G_z,
        // a21 = D_y G_z; the z stencil of G_z takes the pair's own values of
        // plane c from the history
        auto tensor = [&](float t[2][6]) {
            const float2 a0 = lds_pair(gyb, 2), a1 = lds_pair(gyb, GZ + 2), a3 = lds_pair(gyb, 3 * GZ + 2),
                         a4 = lds_pair(gyb, 4 * GZ + 2);
            const float2 b0 = lds_pair(gzb, 2), b1 = lds_pair(gzb, GZ + 2), b3 = lds_pair(gzb, 3 * GZ + 2),
                         b4 = lds_pair(gzb, 4 * GZ + 2);
            const float2 w0 = lds_pair(gzb, 2 * GZ), w2 = lds_pair(gzb, 2 * GZ + 4);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                t[k][0] = fd4s(hx[k][(J + 1) % 5], hx[k][(J + 2) % 5], hx[k][(J + 4) % 5], hx[k][J]);
                t[k][3] = fd4s(hy[k][(J + 1) % 5], hy[k][(J + 2) % 5], hy[k][(J + 4) % 5], hy[k][J]);
                t[k][4] = fd4s(hz[k][(J + 1) % 5], hz[k][(J + 2) % 5], hz[k][(J + 4) % 5], hz[k][J]);
            }
            t[0][1] = fd4s(a0.x, a1.x, a3.x, a4.x);
            t[1][1] = fd4s(a0.y, a1.y, a3.y, a4.y);
            t[0][5] = fd4s(b0.x, b1.x, b3.x, b4.x);
            t[1][5] = fd4s(b0.y, b1.y, b3.y, b4.y);
            t[0][2] = fd4s(w0.x, w0.y, hz[1][JC], w2.x);
            t[1][2] = fd4s(w0.y, hz[0][JC], w2.x, w2.y);
        };
        float t[2][6];
        tensor(t);
        // sign count and margin band, certified in float, as lane masks
        // (count = ma + mb + mc per lane)
        unsigned long long ma[2], mb[2], mc[2], mu[2], md[2];
#pragma unroll
        for (int k = 0; k < 2; ++k)
            positive_eigs_band_wave(t[k][0] - lthf, t[k][1] - lthf, t[k][2] - lthf, t[k][3], t[k][4], t[k][5], tols,
                                    ma[k], mb[k], mc[k], mu[k], md[k]);
        const auto lrs = buffer_rsrc(lptr, (unsigned)(NN * NN));
        lptr += NN * NN;
        // uniform branch: a wave with an uncertain cell (or the lambda output)
        // takes the per-lane path, the others count from the masks
        if (__builtin_expect(!(INL && lam) && (mu[0] | mu[1]) == 0ull, 1)) {
            const unsigned v0 = lane_bits3(ma[0], mb[0], mc[0]), v1 = lane_bits3(ma[1], mb[1], mc[1]);
            __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(0x0404u - (v0 | (v1 << 8))), lrs, coff, 0, 0);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                wn3 += (unsigned)__popcll(ma[k] & mb[k] & mc[k]);
                wany += (unsigned)__popcll(ma[k] | mb[k] | mc[k]);
                wodd += (unsigned)__popcll(ma[k] ^ mb[k] ^ mc[k]);
                wmarg += (unsigned)__popcll(md[k]);
            }
        } else {
            // T is recomputed here, so it is not live across the common path
            float u[2][6];
            tensor(u);
            unsigned v[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                v[k] = lane_bits3(ma[k], mb[k], mc[k]);
                bool marg = lane_bits3(md[k], 0ull, 0ull) != 0u;
                if (lane_bits3(mu[k], 0ull, 0ull) != 0u) {
                    if constexpr (INL)
                        v[k] = (unsigned)positive_eigs_band_d(u[k][0] - lthf, u[k][1] - lthf, u[k][2] - lthf, u[k][3],
                                                              u[k][4], u[k][5], tols, marg);
                    else
                        v[k] = (unsigned)positive_eigs_band64(u[k][0] - lthf, u[k][1] - lthf, u[k][2] - lthf,
                                                              u[k][3], u[k][4], u[k][5], tols, marg);
                }
                if constexpr (INL) {
                    if (lam)
                        eig_store(u[k][0], u[k][1], u[k][2], u[k][3], u[k][4], u[k][5], ih2, lam,
                                  (size_t)c * NN * NN + coff + k, n3);
                }
                wn3 += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(v[k] == 3u));
                wany += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(v[k] != 0u));
                wodd += (unsigned)__popcll(__builtin_amdgcn_ballot_w64((v[k] & 1u) != 0u));
                wmarg += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(marg));
            }
            __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(0x0404u - (v[0] | (v[1] << 8))), lrs, coff, 0, 0);
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    const unsigned lanes = (unsigned)__popcll(__builtin_amdgcn_ballot_w64(true));
    fetch();
    const int send = x1 + 4;
    for (int s = x0 - 4; s < send; s += 5) {
        step(I0{}, s);
        if (s + 1 < send) step(I1{}, s + 1);
        if (s + 2 < send) step(I2{}, s + 2);
        if (s + 3 < send) step(I3{}, s + 3);
        if (s + 4 < send) step(I4{}, s + 4);
    }
    // labels 1..4 = 4 - count: count 3 -> scnt[0], 2 -> [1], 1 -> [2], 0 -> [3]
    if ((tid & 63) == 0) {
        const unsigned wall = 2u * lanes * (unsigned)max(0, x1 - x0);
        const unsigned wn0 = wall - wany, wn1 = wodd - wn3, wn2 = wany - wn1 - wn3;
        if (wn3) atomicAdd(&scnt[0], wn3);
        if (wn2) atomicAdd(&scnt[1], wn2);
        if (wn1) atomicAdd(&scnt[2], wn1);
        if (wn0) atomicAdd(&scnt[3], wn0);
        if (wmarg) atomicAdd(&scnt[4], wmarg);
    }
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


// x segments of the fused T-web: a segment re-reads 8 halo planes; the count
// minimises (rounds of resident workgroups) x (planes per segment).  Returns
// the segment length.
static int tweb_seg_len(cm_ctx *ctx, int N, int tiles, int per_cu) {
    const int slots = std::max(1, ctx->cu_count) * std::max(1, per_cu);
    int nseg = 1;
    double best = 1e300;
    for (int g = 1; g <= std::max(1, N / 16); ++g) {
        const int seg = (N + g - 1) / g;
        const double rounds = std::ceil((double)tiles * g / slots);
        const double cost = rounds * (seg + 8);
        if (cost < best * 0.999) {
            best = cost;
            nseg = g;
        }
    }
    if (ctx->tune.tseg > 0 && ctx->tune.tseg <= N / 4) nseg = ctx->tune.tseg;
    return (N + nseg - 1) / nseg;
}

// LZ: a variant for lambda_th = 0 (the common threshold), without the
// subtraction of the threshold from the diagonal.
template <int TY, int TZ, int NT, bool LZ = false>
static cm_status tweb_fused_run(cm_ctx *ctx, int N, double L, const float *phi, double lth, TolK tol,
                                int8_t *labels, float *lam, unsigned long long *stats, bool stats_zeroed) {
    const bool inl = lam != nullptr;
    auto kern = inl ? k_tweb_fused<TY, TZ, NT, true, false>
                    : (LZ && lth == 0.0 ? k_tweb_fused<TY, TZ, NT, false, LZ> : k_tweb_fused<TY, TZ, NT, false, false>);
    const size_t lds = sizeof(float) * tweb_lds_floats<TY, TZ, NT>();
    if (lds > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "tweb: tile %dx%d needs too much LDS", TY, TZ);
    const int tiles = (N / TY) * (N / TZ);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const int segLen = tweb_seg_len(ctx, N, tiles, per_cu);
    const int nseg = (N + segLen - 1) / segLen;
    dim3 grid((unsigned)tiles, (unsigned)nseg);
    const double ih = 1.0 / (12.0 * (L / (double)N));
    const float ih2 = (float)(ih * ih);
    cm_status s;
    if (!stats_zeroed && (s = launch(ctx, "tweb_zero", k_zero_u64, dim3(1), dim3(64), 0, stats, 5)) != CM_OK) return s;
    return launch(ctx, "tweb_fused", kern, grid, dim3(NT), lds, phi, N, segLen, ih2, lth, tol, labels, lam, stats,
                  (const TwMock *)nullptr, 0);
}

// Packed T-web: one launch over the (tile column, x segment) items of every
// mock (16 x 64 tiles, so N a multiple of 64).  mocks[i].first and .segLen
// are set here; each mock's counters are cleared first.
cm_status tweb_fused_packed(cm_ctx *ctx, std::vector<TwMock> &mocks, const std::vector<double> &L) {
    constexpr int TY = 16, TZ = 64, NT = 512;
    const int nm = (int)mocks.size();
    if (nm == 0) return CM_OK;
    bool lz = true;  // every mock thresholds at lambda_th = 0
    for (const TwMock &t : mocks) lz = lz && t.lth == 0.0;
    auto kern = lz ? k_tweb_fused<TY, TZ, NT, false, true> : k_tweb_fused<TY, TZ, NT, false, false>;
    const size_t lds = sizeof(float) * tweb_lds_floats<TY, TZ, NT>();
    if (lds > ctx->lds_per_block) return fail(ctx, CM_ERR_UNSUPPORTED, "tweb: tile %dx%d needs too much LDS", TY, TZ);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    long long blocks = 0;
    cm_status s;
    for (int m = 0; m < nm; ++m) {
        TwMock &t = mocks[(size_t)m];
        if (t.N % TZ != 0 || t.N % TY != 0) return fail(ctx, CM_ERR_UNSUPPORTED, "packed tweb: N=%d", t.N);
        const int tiles = (t.N / TY) * (t.N / TZ);
        {
            TuneScope ts(ctx, t.N);
            t.segLen = tweb_seg_len(ctx, t.N, tiles, per_cu);
        }
        const int nseg = (t.N + t.segLen - 1) / t.segLen;
        const double ih = 1.0 / (12.0 * (L[(size_t)m] / (double)t.N));
        t.ih2 = (float)(ih * ih);
        t.first = (int)blocks;
        blocks += (long long)tiles * nseg;
        if ((s = launch(ctx, "tweb_zero", k_zero_u64, dim3(1), dim3(64), 0, t.stats, 5)) != CM_OK) return s;
    }
    if (blocks > 0x7fffffffLL) return fail(ctx, CM_ERR_UNSUPPORTED, "packed tweb: %lld workgroups", blocks);
    void *dtm;
    if ((s = scratch(ctx, "tweb.pk.mocks", sizeof(TwMock) * (size_t)nm, &dtm)) != CM_OK) return s;
    CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CM_HIP(ctx, host_upload(ctx, dtm, mocks.data(), sizeof(TwMock) * (size_t)nm));
    ++ctx->host_syncs;
    return launch(ctx, "tweb_fused", kern, dim3((unsigned)blocks), dim3(NT), lds, (const float *)nullptr, 0, 0, 0.0f,
                  0.0, TolK{0.0, 0.0, nullptr}, (int8_t *)nullptr, (float *)nullptr, (unsigned long long *)nullptr,
                  (const TwMock *)dtm, nm);
}

// Fused T-web from phi: labels (and optional lambdas) plus the five counters.
// Tile (TY x TZ cells, threads; two cells per thread): 16x32/256 when 32 | N
// (four workgroups per CU; 0.462 against 0.473 ms for 16x64/512 at 512^3,
// which the tuning knob tz = 64 selects), else 8x16/64 or 8x8/32 by the
// largest of 16, 8 dividing N.
// stats_zeroed: the counters were cleared by an earlier launch on the stream.
cm_status tweb_fused_launch(cm_ctx *ctx, int N, double L, const float *phi, double lth, TolK tol, int8_t *labels,
                            float *lam, unsigned long long *stats, bool stats_zeroed) {
    int TZ = N % 32 == 0 ? 32 : (N % 16 == 0 ? 16 : (N % 8 == 0 ? 8 : 0));
    if (ctx->tune.tz > 0 && N % ctx->tune.tz == 0 &&
        (ctx->tune.tz == 64 || ctx->tune.tz == 32 || ctx->tune.tz == 16 || ctx->tune.tz == 8))
        TZ = ctx->tune.tz;
    switch (TZ) {
    case 64: return tweb_fused_run<16, 64, 512, true>(ctx, N, L, phi, lth, tol, labels, lam, stats, stats_zeroed);
    case 32: return tweb_fused_run<16, 32, 256, true>(ctx, N, L, phi, lth, tol, labels, lam, stats, stats_zeroed);
    case 16: return tweb_fused_run<8, 16, 64>(ctx, N, L, phi, lth, tol, labels, lam, stats, stats_zeroed);
    case 8: return tweb_fused_run<8, 8, 32>(ctx, N, L, phi, lth, tol, labels, lam, stats, stats_zeroed);
    }
    return fail(ctx, CM_ERR_UNSUPPORTED, "tweb: N=%d has no tile decomposition", N);
}

// Coefficient K(N) of the rounding tolerance K u ||phi||_rms / d^2: how far
// rounding of the potential's arithmetic (unit roundoff u) may move an
// eigenvalue of the FD tensor, from a random-rounding model.
//   * The FFT chain (r2c, multiplier, c2r of N^3 points; phi rounded on store)
//     leaves phi with a white error of rms g u ||phi||_rms, g = sqrt(log2 N^3):
//     each of the log2 N^3 radix-2-equivalent stages adds an independent
//     rounding of relative size at most u.
//   * To first order an eigenvalue moves by v^T dT v = (v.D)^2 dphi, the
//     second difference of the error along the eigenvector v.  The stencil of
//     (v.D)^2 (4th-order gradfindif composed with itself) has an L2 norm of at
//     most 1.5 / d^2 over all directions (1.12 / d^2 along an axis, 1.43 / d^2
//     along a diagonal), so white noise of rms e moves it by rms <= 1.5 e / d^2.
//   * The largest of the 3 N^3 eigenvalue errors of a Gaussian model lies
//     below z = sqrt(2 ln(3 N^3)) standard deviations.
// K(N) = 1.5 sqrt(log2 N^3) sqrt(2 ln(3 N^3)): 49.1 at N = 512.  The stencil
// sums of the float32 T-web round in proportion to differences of nearby phi
// values, far below the rounding of phi itself.
double tweb_tol_coeff(int N) {
    const double n3 = (double)N * N * N;
    return 1.5 * std::sqrt(std::log2(n3)) * std::sqrt(2.0 * std::log(3.0 * n3));
}

TolK tweb_tol(int N, double L, double margin_tol, bool fp64, const unsigned *norm) {
    TolK t{margin_tol, 0.0, nullptr};
    if (margin_tol < 0.0) {
        const double d = L / (double)N, u = fp64 ? std::ldexp(1.0, -53) : std::ldexp(1.0, -24);
        t.tol = 0.0;
        t.per_norm = -margin_tol * tweb_tol_coeff(N) * u / (d * d);
        t.norm = norm;
    }
    return t;
}

// ||phi||_rms of n values from per-workgroup sums of squares, as float bits.
cm_status phi_rms_final(cm_ctx *ctx, const double *part, int np, size_t n, unsigned *out) {
    return launch(ctx, "tweb_rms", k_rms_final, dim3(1), dim3(1024), 0, part, np, (double)n, out);
}

cm_status tweb_zero_words(cm_ctx *ctx, unsigned long long *p, int n) {
    return launch(ctx, "tweb_zero", k_zero_u64, dim3(1), dim3(64), 0, p, n);
}

// phi = c2r(W(k)(-1/k^2) r2c(delta)*scale) in the potential's type.
template <typename T>
static cm_status poisson_t(cm_ctx *ctx, const cm_tweb_opts *o, const T *delta, T *phi, double subtract) {
    using C2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    constexpr cm_dtype DT = sizeof(T) == 4 ? CM_F32 : CM_F64;
    const int N = o->N;
    const size_t nh = (size_t)N * N * (N / 2 + 1);
    void *dk;
    cm_status s;
    if ((s = scratch(ctx, DT == CM_F32 ? "tweb.dk" : "tweb.dk64", sizeof(C2) * nh, &dk)) != CM_OK) return s;
    if ((s = fft_r2c_impl(ctx, N, DT, delta, dk, subtract)) != CM_OK) return s;
    const double kf = 2.0 * M_PI / o->L;
    const double R2h = o->R_smooth > 0 ? 0.5 * o->R_smooth * o->R_smooth : 0.0;
    const double scale = o->delta_scale == 0.0 ? 1.0 : o->delta_scale;
    if ((s = launch(ctx, "tweb_kmul", k_kmul<T>, dim3(grid_for(nh, 256, 16384)), dim3(256), 0, (C2 *)dk, N, kf, R2h,
                    scale)) != CM_OK)
        return s;
    const double n3 = (double)N * N * N;
    return fft_c2r_impl(ctx, N, DT, dk, phi, 1.0 / n3);
}

cm_status poisson_impl(cm_ctx *ctx, const cm_tweb_opts *o, const float *delta, float *phi, double subtract) {
    return poisson_t<float>(ctx, o, delta, phi, subtract);
}

// Labels, optional lambdas and the counters from phi.  plain: separate
// gradient and Hessian+eigen kernels (float32 or FP64); otherwise the fused
// marching kernel (float32).  A negative margin_tol takes max|phi| from one
// reduction over phi first.
template <typename T>
static cm_status tweb_from_phi_t(cm_ctx *ctx, const cm_tweb_opts *o, const T *phi, int8_t *labels, T *lam,
                                 cm_tweb_stats *st, bool plain) {
    const int N = o->N;
    const size_t n3 = (size_t)N * N * N;
    constexpr bool F64 = sizeof(T) == 8;
    void *stats;
    cm_status s;
    // counters[5] and ||phi||_rms (one word each)
    if ((s = scratch(ctx, "tweb.stats", sizeof(unsigned long long) * 6, &stats)) != CM_OK) return s;
    auto *w = (unsigned long long *)stats;
    auto *norm = (unsigned *)(w + 5);
    if ((s = tweb_zero_words(ctx, w, 6)) != CM_OK) return s;
    const unsigned gr = grid_for(n3, 256, 65536);
    if (o->margin_tol < 0.0) {
        constexpr unsigned NP = 2048;
        void *part;
        if ((s = scratch(ctx, "tweb.sq", sizeof(double) * NP, &part)) != CM_OK) return s;
        const unsigned g = std::min(gr, NP);
        if ((s = launch(ctx, "tweb_sumsq", k_sumsq<T>, dim3(g), dim3(256), 0, phi, n3, (double *)part)) != CM_OK)
            return s;
        if ((s = phi_rms_final(ctx, (const double *)part, (int)g, n3, norm)) != CM_OK) return s;
    }
    const TolK tk = tweb_tol(N, o->L, o->margin_tol, F64, norm);
    const double ih = 1.0 / (12.0 * (o->L / (double)N));
    if (plain || F64) {
        void *g;
        if ((s = scratch(ctx, F64 ? "tweb.grad64" : "tweb.grad", sizeof(T) * 3 * n3, &g)) != CM_OK) return s;
        T *gx = (T *)g, *gy = gx + n3, *gz = gy + n3;
        if ((s = launch(ctx, "tweb_grad", k_grad3<T>, dim3(gr), dim3(256), 0, phi, N, ih, gx, gy, gz)) != CM_OK)
            return s;
        if ((s = launch(ctx, "tweb_hess_eig", k_hess_eig<T>, dim3(gr), dim3(256), 0, (const T *)gx, (const T *)gy,
                        (const T *)gz, N, ih, o->lambda_th, tk, labels, lam, w)) != CM_OK)
            return s;
    } else {
        // the counters were cleared above
        if ((s = tweb_fused_launch(ctx, N, o->L, (const float *)phi, o->lambda_th, tk, labels, (float *)lam, w,
                                   true)) != CM_OK)
            return s;
    }
    if (st) {
        unsigned long long h[6];
        CM_HIP(ctx, hipMemcpyAsync(h, stats, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
        CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < 4; ++i) st->counts[i] = (int64_t)h[i];
        st->n_marginal = (int64_t)h[4];
        float pn;
        const unsigned bits = (unsigned)h[5];
        std::memcpy(&pn, &bits, sizeof pn);
        st->margin_tol = tk.norm ? tk.per_norm * (double)pn : tk.tol;
    }
    return CM_OK;
}

cm_status tweb_from_phi_impl(cm_ctx *ctx, const cm_tweb_opts *o, const float *phi, int8_t *labels, float *lam,
                             cm_tweb_stats *st, bool plain) {
    return tweb_from_phi_t<float>(ctx, o, phi, labels, lam, st, plain);
}

// Spectral T-web (the reference's _USE_GFFT_EIGENV_ branch of EigenValuesTweb,
// Source/Miscelanious.cpp:1703-1718): T_ab = c2r(scale W(k) k_a k_b / k^2
// r2c(delta)) for the six independent components, then labels from the FP64
// eigenvalues of each cell's tensor.  A negative margin_tol asks for
// |margin_tol| K_s u ||T||_F,rms with K_s = sqrt(log2 N^3) sqrt(2 ln(3 N^3)):
// each component carries the white rounding of its FFT pair, rms
// sqrt(log2 N^3) u times its own rms, an eigenvalue moves by at most the
// Frobenius norm of that error, and the last factor bounds the largest of the
// 3 N^3 eigenvalue errors of a Gaussian model.
template <typename T>
static cm_status tweb_spectral_t(cm_ctx *ctx, const cm_tweb_opts *o, const T *delta, int8_t *labels, T *lam,
                                 cm_tweb_stats *st) {
    using C2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    constexpr cm_dtype DT = sizeof(T) == 4 ? CM_F32 : CM_F64;
    constexpr bool F64 = sizeof(T) == 8;
    const int N = o->N;
    const size_t n3 = (size_t)N * N * N, nh = (size_t)N * N * (N / 2 + 1);
    void *dk, *dab, *t6, *stats;
    cm_status s;
    if ((s = scratch(ctx, F64 ? "tweb.dk64" : "tweb.dk", sizeof(C2) * nh, &dk)) != CM_OK) return s;
    if ((s = scratch(ctx, F64 ? "tweb.dkab64" : "tweb.dkab", sizeof(C2) * nh, &dab)) != CM_OK) return s;
    if ((s = scratch(ctx, F64 ? "tweb.t6_64" : "tweb.t6", sizeof(T) * 6 * n3, &t6)) != CM_OK) return s;
    if ((s = scratch(ctx, "tweb.stats", sizeof(unsigned long long) * 6, &stats)) != CM_OK) return s;
    auto *w = (unsigned long long *)stats;
    auto *norm = (unsigned *)(w + 5);
    if ((s = tweb_zero_words(ctx, w, 6)) != CM_OK) return s;
    if ((s = fft_r2c_impl(ctx, N, DT, delta, dk, 0.0)) != CM_OK) return s;
    const double kf = 2.0 * M_PI / o->L;
    const double R2h = o->R_smooth > 0 ? 0.5 * o->R_smooth * o->R_smooth : 0.0;
    const double scale = o->delta_scale == 0.0 ? 1.0 : o->delta_scale;
    static const int ab[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 0}, {2, 0}, {2, 1}};
    T *tp = (T *)t6;
    for (int c = 0; c < 6; ++c) {
        if ((s = launch(ctx, "tweb_kab", k_kab<T>, dim3(grid_for(nh, 256, 16384)), dim3(256), 0, (const C2 *)dk,
                        (C2 *)dab, N, kf, R2h, scale, ab[c][0], ab[c][1])) != CM_OK)
            return s;
        if ((s = fft_c2r_impl(ctx, N, DT, dab, tp + (size_t)c * n3, 1.0 / (double)n3)) != CM_OK) return s;
    }
    const unsigned gr = grid_for(n3, 256, 65536);
    TolK tk{o->margin_tol, 0.0, nullptr};
    if (o->margin_tol < 0.0) {
        constexpr unsigned NP = 2048;
        void *part;
        if ((s = scratch(ctx, "tweb.sq6", sizeof(double) * 6 * NP, &part)) != CM_OK) return s;
        const unsigned g = std::min(gr, NP);
        for (int c = 0; c < 6; ++c)
            if ((s = launch(ctx, "tweb_sumsq", k_sumsq<T>, dim3(g), dim3(256), 0, (const T *)(tp + (size_t)c * n3),
                            n3, (double *)part + (size_t)c * g)) != CM_OK)
                return s;
        if ((s = launch(ctx, "tweb_rms", k_rms_tensor, dim3(1), dim3(1024), 0, (const double *)part, (int)g,
                        (double)n3, norm)) != CM_OK)
            return s;
        const double n3d = (double)n3, u = F64 ? std::ldexp(1.0, -53) : std::ldexp(1.0, -24);
        tk.tol = 0.0;
        tk.per_norm = -o->margin_tol * std::sqrt(std::log2(n3d)) * std::sqrt(2.0 * std::log(3.0 * n3d)) * u;
        tk.norm = norm;
    }
    if ((s = launch(ctx, "tweb_tensor_eig", k_tensor_eig<T>, dim3(gr), dim3(256), 0, (const T *)tp, n3,
                    o->lambda_th, tk, labels, lam, w)) != CM_OK)
        return s;
    if (st) {
        unsigned long long h[6];
        CM_HIP(ctx, hipMemcpyAsync(h, stats, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
        CM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < 4; ++i) st->counts[i] = (int64_t)h[i];
        st->n_marginal = (int64_t)h[4];
        float pn;
        const unsigned bits = (unsigned)h[5];
        std::memcpy(&pn, &bits, sizeof pn);
        st->margin_tol = tk.norm ? tk.per_norm * (double)pn : tk.tol;
    }
    return CM_OK;
}

static cm_status check_opts(cm_ctx *ctx, const cm_tweb_opts *o) {
    if (!o) return fail(ctx, CM_ERR_INVALID, "tweb: null options");
    if (o->hessian != CM_HESSIAN_FD && o->hessian != CM_HESSIAN_SPECTRAL)
        return fail(ctx, CM_ERR_INVALID, "tweb: hessian %d (0: FD, 1: spectral)", o->hessian);
    if (!fft_size_ok(o->N)) return fail(ctx, CM_ERR_INVALID, "tweb: N=%d not supported by the FFT", o->N);
    if (!(o->L > 0)) return fail(ctx, CM_ERR_INVALID, "tweb: L must be > 0");
    if (!std::isfinite(o->margin_tol)) return fail(ctx, CM_ERR_INVALID, "tweb: margin_tol must be finite");
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
    if (o->hessian != CM_HESSIAN_FD)
        return fail(ctx, CM_ERR_UNSUPPORTED, "cm_tweb_from_phi: the spectral Hessian is formed from delta (cm_tweb)");
    return tweb_from_phi_impl(ctx, o, phi, labels, lambdas, stats, false);
}

cm_status cm_tweb(cm_ctx *ctx, const cm_tweb_opts *o, const float *delta, int8_t *labels, float *lambdas,
                  cm_tweb_stats *stats) {
    if (!ctx) return CM_ERR_INVALID;
    cm_status s = check_opts(ctx, o);
    if (s != CM_OK) return s;
    if (!delta || !labels) return fail(ctx, CM_ERR_INVALID, "cm_tweb: null pointer");
    if (o->hessian == CM_HESSIAN_SPECTRAL) return tweb_spectral_t<float>(ctx, o, delta, labels, lambdas, stats);
    void *phi;
    const size_t n3 = (size_t)o->N * o->N * o->N;
    if ((s = scratch(ctx, "tweb.phi", sizeof(float) * n3, &phi)) != CM_OK) return s;
    if ((s = poisson_impl(ctx, o, delta, (float *)phi, 0.0)) != CM_OK) return s;
    return tweb_from_phi_impl(ctx, o, (const float *)phi, labels, lambdas, stats, false);
}

cm_status cm_tweb_f64(cm_ctx *ctx, const cm_tweb_opts *o, const double *delta, int8_t *labels, double *lambdas,
                      double *phi_out, cm_tweb_stats *stats) {
    if (!ctx) return CM_ERR_INVALID;
    cm_status s = check_opts(ctx, o);
    if (s != CM_OK) return s;
    if (!delta || !labels) return fail(ctx, CM_ERR_INVALID, "cm_tweb_f64: null pointer");
    if (o->hessian == CM_HESSIAN_SPECTRAL) {
        if (phi_out) return fail(ctx, CM_ERR_INVALID, "cm_tweb_f64: the spectral Hessian forms no potential");
        return tweb_spectral_t<double>(ctx, o, delta, labels, lambdas, stats);
    }
    void *phi = phi_out;
    const size_t n3 = (size_t)o->N * o->N * o->N;
    if (!phi && (s = scratch(ctx, "tweb.phi64", sizeof(double) * n3, &phi)) != CM_OK) return s;
    if ((s = poisson_t<double>(ctx, o, delta, (double *)phi, 0.0)) != CM_OK) return s;
    return tweb_from_phi_t<double>(ctx, o, (const double *)phi, labels, lambdas, stats, true);
}

}  // extern "C"

// ------------------------------------------------------------ label combos
namespace cm {

// 4-bit class mask of a combo code (bit c-1 for class c); 0: matches nothing
static unsigned combo_mask(int code) {
    switch (code) {
    case 1: return 1u;
    case 2: return 2u;
    case 3: return 4u;
    case 4: return 8u;
    case 12: return 3u;
    case 23: return 6u;
    case 34: return 12u;
    case 123: return 7u;
    case 234: return 14u;
    }
    return 0u;
}

struct ComboTab {
    unsigned mask[16];
};

__global__ void __launch_bounds__(256) k_tclass(const int8_t *__restrict__ labels, size_t n, ComboTab tab, int ncwt,
                                                int single, int32_t *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (single) {
            out[i] = single;
            continue;
        }
        const int lab = labels[i];
        const unsigned bit = (lab >= 1 && lab <= 4) ? (1u << (lab - 1)) : 0u;
        int ans = 0;
        for (int c = 0; c < ncwt; ++c)
            if (tab.mask[c] & bit) ans = c;
        out[i] = ans;
    }
}

}  // namespace cm

extern "C" cm_status cm_tweb_tclass(cm_ctx *ctx, const int8_t *labels, size_t n, const int *cwt_used, int ncwt,
                                    int32_t *out) {
    if (!ctx) return CM_ERR_INVALID;
    if (!cwt_used || ncwt < 1 || ncwt > 16) return cm::fail(ctx, CM_ERR_INVALID, "cm_tweb_tclass: 1..16 codes");
    if (n && (!labels || !out)) return cm::fail(ctx, CM_ERR_INVALID, "cm_tweb_tclass: null pointer");
    if (!n) return CM_OK;
    cm::ComboTab tab{};
    for (int c = 0; c < ncwt; ++c) tab.mask[c] = cm::combo_mask(cwt_used[c]);
    // one code: every cell gets the code itself (the reference returns
    // cwt_used[0]; for code 0 the loop below also gives 0)
    const int single = ncwt == 1 ? cwt_used[0] : 0;
    size_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return cm::launch(ctx, "tweb_tclass", cm::k_tclass, dim3((unsigned)g), dim3(256), 0, labels, n, tab, ncwt, single,
                      out);
}
