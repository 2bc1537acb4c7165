// Mode enumeration and shell index of the P(k) estimator, shared by the shell
// kernel and the fused FFT/shell pass.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace cm {

template <typename T> struct v2pk_t;
template <> struct v2pk_t<float> { using type = float2; };
template <> struct v2pk_t<double> { using type = double2; };
template <typename T> using v2pk = typename v2pk_t<T>::type;

struct PkK {
    int N;
    int nshells;
    int int_rule;    // Dmin and ndel are integers: exact integer edges
    int float_rule;  // the reference's float32 shell index
    int Dmin_i, ndel_i;
    double Dmin, ndel;  // kmin/k_f, dk/k_f
    float dkx_f, kmin_f, rdk_f;
    float rndel_f;  // 1/ndel for the integer rule's estimate
    // logarithmic shells: shell s holds q = |k/k_f|^2 with qedge[s] <= q <
    // qedge[s+1] (device table of nshells+1 thresholds); nullptr: linear
    const int *qedge;
    // window mode set of the random branch (FftwFunctions.cpp:1248-1262):
    // the first octant (both folded x and y below N/2) is left out
    int window;
};

// Shell of q under the log rule: the s with qedge[s] <= q < qedge[s+1], else -1.
__device__ __forceinline__ int log_shell(const int *__restrict__ qedge, int ns, int q) {
    if (q < qedge[0] || q >= qedge[ns]) return -1;
    int lo = 0, hi = ns;  // qedge[lo] <= q < qedge[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (q >= qedge[mid]) lo = mid; else hi = mid;
    }
    return lo;
}

// Which stored half-spectrum modes power_spectrum_fkp visits
// (FftwFunctions.cpp:1061-1084,1180,1273,1373,1474): the loop runs i,j < N/2,
// k <= N/2 and adds the copies (i,j,k) [not the zero mode], (i,N-j,k) if
// j>0 && k>0, (N-i,j,k) if i>0 && (j>0 || k>0), (N-i,N-j,k) if i,j,k>0.  So the
// x- and y-Nyquist planes are never visited.  Returns the folded |i|,|j|.
__device__ __forceinline__ bool mode_counted(int N, int ix, int iy, int iz, int &fx, int &fy) {
    const int h = N / 2;
    if (ix == h || iy == h) return false;
    const bool xl = ix < h, yl = iy < h;
    fx = xl ? ix : N - ix;
    fy = yl ? iy : N - iy;
    if (xl && yl) return (ix | iy | iz) != 0;
    if (xl) return iz > 0;           // (i, N-j, k): j>0 && k>0
    if (yl) return iy > 0 || iz > 0;  // (N-i, j, k): i>0 && (j>0 || k>0)
    return iz > 0;                    // (N-i, N-j, k): i,j,k>0
}

// Shell of the folded mode (i,j,k); -1 outside [0, nshells).
__device__ __forceinline__ int shell_of(const PkK &p, int i, int j, int k) {
    if (p.float_rule) {
#pragma clang fp contract(off)
        // real_prec=float arithmetic of FftwFunctions.cpp:1068-1097.  Each float
        // operation is evaluated in double and rounded to float, which gives the
        // correctly rounded float result (double has more than 2*24+2 bits); FP
        // contraction is off in this block so no multiply-add becomes an FMA.
        const double dk = (double)p.dkx_f;
        float ti = (float)((double)(float)(i * i) * dk);
        ti = (float)((double)ti * dk);
        float tj = (float)((double)(float)(j * j) * dk);
        tj = (float)((double)tj * dk);
        float tk = (float)((double)(float)(k * k) * dk);
        tk = (float)((double)tk * dk);
        const float s1 = (float)((double)ti + (double)tj);
        const float s2 = (float)((double)s1 + (double)tk);
        const float kv = (float)sqrt((double)s2);
        const float d = (float)((double)kv - (double)p.kmin_f);
        const float t = (float)((double)d * (double)p.rdk_f);
        int s = (int)floorf(t);
        return (s < 0 || s >= p.nshells) ? -1 : s;
    }
    const int q = i * i + j * j + k * k;
    if (p.qedge) return log_shell(p.qedge, p.nshells, q);
    if (p.int_rule) {
        // integer DeltaKmin and ndel: float estimate (hardware sqrt and a
        // reciprocal are enough), then exact integer edges
        // (s*ndel + Dmin)^2 <= q < ((s+1)*ndel + Dmin)^2 in 32-bit arithmetic
        // (q <= 3 (N/2)^2 and the edges stay below 2^31 for N <= 1024)
        int s = (int)floorf((__builtin_sqrtf((float)q) - (float)p.Dmin_i) * p.rndel_f);
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const int lo = s * p.ndel_i + p.Dmin_i, hi = (s + 1) * p.ndel_i + p.Dmin_i;
            if (lo > 0 && q < lo * lo) --s;
            else if (hi >= 0 && q >= hi * hi) ++s;
        }
        if (q < p.Dmin_i * p.Dmin_i) s = -1;
        return (s < 0 || s >= p.nshells) ? -1 : s;
    }
    const double t = (sqrt((double)q) - p.Dmin) / p.ndel;
    int s = (int)floor(t);
    if (p.int_rule) {
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            long lo = (long)s * p.ndel_i + p.Dmin_i, hi = (long)(s + 1) * p.ndel_i + p.Dmin_i;
            if (lo > 0 && (long)q < lo * lo) --s;
            else if (hi >= 0 && (long)q >= hi * hi) ++s;
        }
    }
    if (t < 0.0 && s >= 0) s = -1;
    return (s < 0 || s >= p.nshells) ? -1 : s;
}

}  // namespace cm
