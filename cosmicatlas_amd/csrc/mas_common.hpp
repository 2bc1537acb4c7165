// Per-particle geometry shared by every mass-assignment kernel: shift by the
// box corner, periodic wrap in the position dtype, cell index by FP64 floor,
// and the 1-D MAS weights of the nodes around the particle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cm {

struct MasK {
    int N;
    int scheme;   // 0 NGP, 1 CIC, 2 TSC, 3 PCS
    int soa;
    int pad;
    double L;
    double d;      // L/N
    double rd;     // 1/d rounded to double
    double xmin[3];
    double ishift; // interlace shift in length units
    size_t n;
};

// Stencil radius of the scheme in cells (nodes c-R .. c+R).
__host__ __device__ inline int mas_radius(int scheme) { return scheme == 0 ? 0 : (scheme == 3 ? 2 : 1); }

// x -= xmin in double then rounded to the position dtype; wrap with the
// reference's while-loops in that dtype (far images are first reduced by fmod).
template <typename PT>
__device__ __forceinline__ bool wrap_coord(PT v, double xmin, double L, PT &out) {
    PT s = (PT)((double)v - xmin);
    if (!isfinite((double)s)) return false;
    const PT Lp = (PT)L;
    if (fabs((double)s) > 4.0 * L) {
        double t = fmod((double)s, L);
        if (t < 0) t += L;
        s = (PT)t;
    }
    int it = 0;
    while (s < (PT)0 && it < 8) { s += Lp; ++it; }
    while (s >= Lp && it < 16) { s -= Lp; ++it; }
    if (s < (PT)0 || s >= Lp) return false;
    out = s;
    return true;
}

// Shifted coordinate (interlacing) in double from the wrapped value.
__device__ __forceinline__ double shifted(double s, double ishift, double L) {
    if (ishift == 0.0) return s;
    double x = s + ishift;
    if (x >= L) x -= L;
    return x;
}

// get_bin: floor(double(x)/double(d)), clamped to [0, N-1].
__device__ __forceinline__ int cell_of(double x, double d, int N) {
    int c = (int)floor(x / d);
    c = c >= N ? N - 1 : c;
    return c < 0 ? 0 : c;
}

// get_bin's cell and the in-cell offset q - c - 1/2 without an FP64 division
// in the common case.  For 0 <= x < L the product x*rd is within 2^-42 of
// x/d and the rounded quotient within 2^-44 (x/d < N <= 1024), so when no
// integer lies within 2^-40 of the product, both floor to the same cell;
// otherwise the quotient is taken by division, as in cell_of.
__device__ __forceinline__ int cell_fast(double x, double d, double rd, int N, double &q) {
    q = x * rd;
    if (__builtin_expect(floor(q - 0x1p-40) != floor(q + 0x1p-40), 0)) {
        q = x / d;
        __asm__ volatile("" ::: "memory");  // keeps the division in the branch (not evaluated speculatively)
    }
    int c = (int)floor(q);
    c = c >= N ? N - 1 : c;
    return c < 0 ? 0 : c;
}

// Weights of nodes c+m, m = -2..2 (stored at m+2), for the offset
// f = x/d - c - 1/2 in [-1/2, 1/2): MAS_{NGP,CIC,TSC,PCS}(node - x).
template <typename MT>
__device__ __forceinline__ void axis_weights(int scheme, MT f, MT w[5]) {
    const MT h = (MT)0.5;
    w[0] = w[1] = w[2] = w[3] = w[4] = (MT)0;
    switch (scheme) {
    case 0: w[2] = (MT)1; break;
    case 1:  // CIC: 1-|s|
        w[2] = (MT)1 - fabs(f);
        if (f < (MT)0) w[1] = -f; else w[3] = f;
        break;
    case 2: {  // TSC
        MT a = h - f, b = h + f;
        w[1] = h * a * a;
        w[2] = (MT)0.75 - f * f;
        w[3] = h * b * b;
        break;
    }
    case 3: {  // PCS: cubic B-spline at s = m - f
        const MT s6 = (MT)(1.0 / 6.0);
#pragma unroll
        for (int m = -2; m <= 2; ++m) {
            MT y = fabs((MT)m - f);
            MT r = (MT)0;
            if (y < (MT)1) r = s6 * ((MT)4 - (MT)6 * y * y + (MT)3 * y * y * y);
            else if (y < (MT)2) { MT t = (MT)2 - y; r = s6 * t * t * t; }
            w[m + 2] = r;
        }
        break;
    }
    }
}

template <typename PT>
__device__ __forceinline__ PT load_coord(const PT *pos, size_t i, int a, int soa, size_t n) {
    return soa ? pos[(size_t)a * n + i] : pos[3 * i + (size_t)a];
}

}  // namespace cm
