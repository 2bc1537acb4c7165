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
//   s-2 (halo 2 in y,z for G_y, G_z) from ring planes s-3..s and the thread's
//   own plane s-4 values (kept in registers from the step before), G_y and
//   G_z written to the G slot of plane s-2; T of plane c = s-4 from G_x, G_y,
//   G_z of planes c-2..c+2 (x derivative, register histories) and the G slot
//   of plane c (y, z derivatives, written two steps earlier).
// Five phi slots (plane s-4 is no longer read from LDS in step s, so plane
// s+1 may replace it before the next barrier) and three G slots make every
// LDS slot's next write follow the barrier after its last read.  Each thread owns IP interior cells and keeps
// their last five G_x, G_y, G_z values in registers; the march is unrolled by
// five so the history slots are compile-time indices.  The next phi plane is
// loaded into registers while the current one is processed.  G and T never
// touch HBM.  Composition order of the reference: G_a = D_a phi, then D_x G_x,
// D_x G_y, D_x G_z, D_y G_y, D_y G_z, D_z G_z (Miscelanious.cpp:1684-1700).
// Labels from the inertia of T - lambda_th I: float signs with a certified
// rounding bound, recounted in FP64 in the kernel where the float signs are
// not certain.  The same coefficients, shifted by -+ tol, give the sign
// counts of T - (lambda_th -+ tol) I: the cell is marginal when they differ
// (an eigenvalue lies within tol of lambda_th), decided in FP64 when a shifted
// sign is not certain either.  The marching loop makes no calls except for the
// lambda output (INL = true), which takes the out-of-line FP64 branch for
// every cell (a call site inside the loop costs ~35 VGPRs of saved registers).

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

template <int TY, int TZ>
constexpr size_t tweb_lds_floats() {
    return 5 * (size_t)(TY + 8) * (TZ + 8) + 6 * (size_t)(TY + 4) * (TZ + 4);
}

template <int TY, int TZ, int NT, bool INL>
__global__ void __launch_bounds__(NT) k_tweb_fused(const float *__restrict__ phi_a, int N_a, int segLen_a,
                                                   float ih2_a, double lth_a, TolK tk_a, int8_t *__restrict__ labels_a,
                                                   float *__restrict__ lam, unsigned long long *__restrict__ stats_a,
                                                   const TwMock *__restrict__ tm, int nm) {
    constexpr int HY = TY + 8, HZ = TZ + 8, HP = HY * HZ;  // haloed phi plane (halo 4)
    constexpr int GZ = TZ + 4, GP = (TY + 4) * GZ;         // G plane (halo 2)
    constexpr int IP = TY * TZ / NT;                       // interior cells per thread
    constexpr int PF = (HP + NT - 1) / NT;                 // phi values per thread per plane
    constexpr int NH1 = 4 * TZ, NH2 = 4 * TY;              // halo positions: y rows, z columns
    static_assert(IP * NT == TY * TZ && NH1 + NH2 <= NT, "tile does not match the workgroup");
    extern __shared__ float sm[];
    float *ring = sm;             // 5 phi planes
    float *sgy = ring + 5 * HP;   // 3 G_y planes
    float *sgz = sgy + 3 * GP;    // 3 G_z planes
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
    int bx = blockIdx.x, by = blockIdx.y;
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
    const float lthf = (float)(lth / (double)ih2);
    const float tols = (float)(tol / (double)ih2);

    // this thread's phi loads: in-plane offsets with periodic wrap in y and z
    int foff[PF];
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const int q = tid + k * NT;
        const int yy = q / HZ, zz = q - (q / HZ) * HZ;
        foff[k] = q < HP ? ((ty0 - 4 + yy + N) % N) * N + (tz0 - 4 + zz + N) % N : 0;
    }
    // interior cells: phi-plane index, G-plane index, in-plane output offset.
    // The stencil reads of a cell address the LDS from one base per cell
    // (pb = its phi index - 2 HZ - 2, gb = its G index - 2 GZ - 2) plus
    // non-negative compile-time offsets, which the LDS instructions encode as
    // immediates: no per-read address arithmetic.
    int po[IP], gq[IP], coff[IP];
#pragma unroll
    for (int k = 0; k < IP; ++k) {
        const int i = tid + k * NT, ia = i / TZ, ib = i % TZ;
        po[k] = (ia + 4) * HZ + ib + 4;
        gq[k] = (ia + 2) * GZ + ib + 2;
        coff[k] = (ty0 + ia) * N + tz0 + ib;
    }
    constexpr int PB = 2 * HZ + 2, GB = 2 * GZ + 2;  // base offsets below a cell's own index
    // at most one halo position per thread: kind 1 = y-halo row (G_y, G_z),
    // kind 2 = z-halo column (G_z)
    int hkind = 0, hq = 0, hpo = 0;
    if (tid < NH1) {
        const int r = tid / TZ, a = r < 2 ? r : TY + r, b = 2 + tid % TZ;
        hkind = 1;
        hq = a * GZ + b;
        hpo = (a + 2) * HZ + b + 2;
    } else if (tid < NH1 + NH2) {
        const int j = tid - NH1, r = j & 3, a = 2 + (j >> 2), b = r < 2 ? r : TZ + r;
        hkind = 2;
        hq = a * GZ + b;
        hpo = (a + 2) * HZ + b + 2;
    }

    float hx[IP][5], hy[IP][5], hz[IP][5], p3[IP][5];
#pragma unroll
    for (int k = 0; k < IP; ++k)
#pragma unroll
        for (int i = 0; i < 5; ++i) hx[k][i] = hy[k][i] = hz[k][i] = p3[k][i] = 0.f;
    unsigned long long cls = 0;  // class counts, 16 bits each (at most 2048 cells per thread)
    unsigned nmarg = 0;
    float pf[PF];
    auto fetch = [&](int s) {
        const int sp = (s % N + N) % N;
        const auto plane = as_global(phi) + (size_t)sp * NN * NN;
#pragma unroll
        for (int k = 0; k < PF; ++k)
            if (tid + k * NT < HP) pf[k] = plane[foff[k]];
    };
    // G of plane p (p >= x0 - 2) in G slot (p - x0 + 3) % 3
    auto gslot = [&](int p) { return (p - x0 + 3) % 3 * GP; };

    // Step s (J = (s - x0 + 4) mod 5): plane s enters ring slot (J+4)%5 (the
    // slot of plane s-5); G at plane s-2 goes to history slot J; cells of
    // plane c = s-4 are output, with G(c-2..c+2) in history slots (J+1)%5,
    // (J+2)%5, (J+3)%5, (J+4)%5, J.  p3[k][J] keeps plane s-3 at the thread's
    // cell for step s+1 (where it is plane s-4).
    auto step = [&](auto Jc, int s) {
        constexpr int J = decltype(Jc)::value;
        float *cur = ring + ((J + 4) % 5) * HP;
#pragma unroll
        for (int k = 0; k < PF; ++k)
            if (tid + k * NT < HP) cur[tid + k * NT] = pf[k];
        if (s + 1 < x1 + 4) fetch(s + 1);  // in flight while plane s is processed
        __syncthreads();
        constexpr int S3 = ((J + 1) % 5) * HP, S2 = ((J + 2) % 5) * HP, S1 = ((J + 3) % 5) * HP,
                      S0 = ((J + 4) % 5) * HP;
        const float *pm2 = ring + S2;
        // registers are written unconditionally (before x0 the values come from
        // a partly filled ring and are never read), so no history value has to
        // survive a skipped write
#pragma unroll
        for (int k = 0; k < IP; ++k) {
            const float *b = ring + (po[k] - PB);  // loop-invariant per cell
            p3[k][J] = b[S3 + PB];
            hx[k][J] = fd4s(p3[k][(J + 4) % 5], p3[k][J], b[S1 + PB], b[S0 + PB]);
            hy[k][J] = fd4s(b[S2 + 2], b[S2 + HZ + 2], b[S2 + 3 * HZ + 2], b[S2 + 4 * HZ + 2]);
            hz[k][J] = fd4s(b[S2 + 2 * HZ], b[S2 + 2 * HZ + 1], b[S2 + 2 * HZ + 3], b[S2 + 2 * HZ + 4]);
        }
        if (s >= x0) {
            float *gyw = sgy + gslot(s - 2), *gzw = sgz + gslot(s - 2);
#pragma unroll
            for (int k = 0; k < IP; ++k) {
                gyw[gq[k]] = hy[k][J];
                gzw[gq[k]] = hz[k][J];
            }
            if (hkind) {
                const float *b = pm2 + (hpo - PB);
                if (hkind == 1) gyw[hq] = fd4s(b[2], b[HZ + 2], b[3 * HZ + 2], b[4 * HZ + 2]);
                gzw[hq] = fd4s(b[2 * HZ], b[2 * HZ + 1], b[2 * HZ + 3], b[2 * HZ + 4]);
            }
        }
        const int c = s - 4;
        if (c < x0) return;
        const int gs = gslot(c);  // uniform
        const auto lplane = as_global(labels) + (size_t)c * NN * NN;
        // T of cell k: txx, tyy, tzz, a10 = D_x G_y, a20 = D_x G_z, a21 = D_y G_z;
        // G_y at gy[gs + i], G_z at gy[gs + 3 GP + i]
        constexpr int ZO = 3 * GP;
        auto tensor = [&](int k, float t[6]) {
            const float *g = sgy + gs + (gq[k] - GB);
            t[0] = fd4s(hx[k][(J + 1) % 5], hx[k][(J + 2) % 5], hx[k][(J + 4) % 5], hx[k][J]);
            t[1] = fd4s(g[2], g[GZ + 2], g[3 * GZ + 2], g[4 * GZ + 2]);
            t[2] = fd4s(g[ZO + 2 * GZ], g[ZO + 2 * GZ + 1], g[ZO + 2 * GZ + 3], g[ZO + 2 * GZ + 4]);
            t[3] = fd4s(hy[k][(J + 1) % 5], hy[k][(J + 2) % 5], hy[k][(J + 4) % 5], hy[k][J]);
            t[4] = fd4s(hz[k][(J + 1) % 5], hz[k][(J + 2) % 5], hz[k][(J + 4) % 5], hz[k][J]);
            t[5] = fd4s(g[ZO + 2], g[ZO + GZ + 2], g[ZO + 3 * GZ + 2], g[ZO + 4 * GZ + 2]);
        };
        int v[IP];
        bool unc[IP];
        bool any = INL && lam != nullptr;
#pragma unroll
        for (int k = 0; k < IP; ++k) {
            float t[6];
            tensor(k, t);
            // sign count and margin band, certified in float
            bool bd;
            v[k] = positive_eigs_band(t[0] - lthf, t[1] - lthf, t[2] - lthf, t[3], t[4], t[5], tols, unc[k], bd);
            nmarg += (!unc[k] && bd) ? 1u : 0u;
            any = any || unc[k];
        }
        if (__builtin_expect(any, 0)) {
#pragma unroll
            for (int k = 0; k < IP; ++k) {
                if (!(unc[k] || (INL && lam))) continue;
                // T is recomputed here, so it is not live across the common path
                float t[6];
                tensor(k, t);
                if (unc[k]) {
                    bool bd;
                    if constexpr (INL)
                        v[k] = positive_eigs_band_d(t[0] - lthf, t[1] - lthf, t[2] - lthf, t[3], t[4], t[5], tols, bd);
                    else
                        v[k] = positive_eigs_band64(t[0] - lthf, t[1] - lthf, t[2] - lthf, t[3], t[4], t[5], tols, bd);
                    nmarg += bd ? 1u : 0u;
                }
                if constexpr (INL) {
                    if (lam) eig_store(t[0], t[1], t[2], t[3], t[4], t[5], ih2, lam, (size_t)c * NN * NN + coff[k], n3);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < IP; ++k) {
            const int lab = 4 - v[k];
            lplane[coff[k]] = (int8_t)lab;
            cls += 1ull << (16 * (lab - 1));
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    fetch(x0 - 4);
    const int send = x1 + 4;
    for (int s = x0 - 4; s < send; s += 5) {
        step(I0{}, s);
        if (s + 1 < send) step(I1{}, s + 1);
        if (s + 2 < send) step(I2{}, s + 2);
        if (s + 3 < send) step(I3{}, s + 3);
        if (s + 4 < send) step(I4{}, s + 4);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const unsigned ci = (unsigned)((cls >> (16 * i)) & 0xffffull);
        if (ci) atomicAdd(&scnt[i], ci);
    }
    if (nmarg) atomicAdd(&scnt[4], nmarg);
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

template <int TY, int TZ, int NT>
static cm_status tweb_fused_run(cm_ctx *ctx, int N, double L, const float *phi, double lth, TolK tol,
                                int8_t *labels, float *lam, unsigned long long *stats, bool stats_zeroed) {
    const bool inl = lam != nullptr;
    auto kern = inl ? k_tweb_fused<TY, TZ, NT, true> : k_tweb_fused<TY, TZ, NT, false>;
    const size_t lds = sizeof(float) * tweb_lds_floats<TY, TZ>();
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
    auto kern = k_tweb_fused<TY, TZ, NT, false>;
    const size_t lds = sizeof(float) * tweb_lds_floats<TY, TZ>();
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
    CM_HIP(ctx, hipMemcpy(dtm, mocks.data(), sizeof(TwMock) * (size_t)nm, hipMemcpyHostToDevice));
    ++ctx->host_syncs;
    return launch(ctx, "tweb_fused", kern, dim3((unsigned)blocks), dim3(NT), lds, (const float *)nullptr, 0, 0, 0.0f,
                  0.0, TolK{0.0, 0.0, nullptr}, (int8_t *)nullptr, (float *)nullptr, (unsigned long long *)nullptr,
                  (const TwMock *)dtm, nm);
}

// Fused T-web from phi: labels (and optional lambdas) plus the five counters.
// Tile (TY x TZ cells, threads): 16x64/512 when 64 | N, else 8x32/256,
// 8x16/128 or 8x8/64 by the largest of 32, 16, 8 dividing N.
// stats_zeroed: the counters were cleared by an earlier launch on the stream.
cm_status tweb_fused_launch(cm_ctx *ctx, int N, double L, const float *phi, double lth, TolK tol, int8_t *labels,
                            float *lam, unsigned long long *stats, bool stats_zeroed) {
    int TZ = N % 64 == 0 ? 64 : (N % 32 == 0 ? 32 : (N % 16 == 0 ? 16 : (N % 8 == 0 ? 8 : 0)));
    if (ctx->tune.tz > 0 && N % ctx->tune.tz == 0 && ctx->tune.tz <= TZ &&
        (ctx->tune.tz == 64 || ctx->tune.tz == 32 || ctx->tune.tz == 16 || ctx->tune.tz == 8))
        TZ = ctx->tune.tz;
    switch (TZ) {
    case 64: return tweb_fused_run<16, 64, 512>(ctx, N, L, phi, lth, tol, labels, lam, stats, stats_zeroed);
    case 32: return tweb_fused_run<8, 32, 256>(ctx, N, L, phi, lth, tol, labels, lam, stats, stats_zeroed);
    case 16: return tweb_fused_run<8, 16, 128>(ctx, N, L, phi, lth, tol, labels, lam, stats, stats_zeroed);
    case 8: return tweb_fused_run<8, 8, 64>(ctx, N, L, phi, lth, tol, labels, lam, stats, stats_zeroed);
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
