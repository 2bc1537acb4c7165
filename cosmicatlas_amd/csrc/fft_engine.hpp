// Workgroup-level mixed-radix Stockham FFT engine for gfx950.
//
// A workgroup transforms `nb` independent sequences of length n.  Each stage
// of radix R reads R inputs strided by n/R into registers, applies the twiddles
// exp(-+2 pi i r k / (Ns R)) (k = j mod Ns) from an FP64-generated table, runs
// the radix-R butterfly and writes the outputs to (j/Ns) Ns R + k + q Ns.  The
// first stage reads through a Src functor (global memory, coalesced) and the
// last stage writes through a Dst functor, so only intermediate stages touch
// LDS.  Each thread owns E elements per stage (E divisible by every radix).
#pragma once

#include <hip/hip_runtime.h>

#include "memspace.hpp"

namespace cm {

template <typename T> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <typename T> using v2 = typename vec2<T>::type;

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return C{a.x, -a.y}; }
// multiply by s*i (s = -1 forward, +1 inverse)
template <bool FWD, typename C> __device__ __forceinline__ C mul_si(C a) {
    return FWD ? C{a.y, -a.x} : C{-a.y, a.x};
}

template <int R, bool FWD, typename C> struct Bfly;

template <bool FWD, typename C> struct Bfly<2, FWD, C> {
    __device__ __forceinline__ static void run(C *v) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool FWD, typename C> struct Bfly<4, FWD, C> {
    __device__ __forceinline__ static void run(C *v) {
        C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        C s13 = cadd(v[1], v[3]), d13 = mul_si<FWD>(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    }
};

template <bool FWD, typename C> struct Bfly<8, FWD, C> {
    __device__ __forceinline__ static void run(C *v) {
        using T = decltype(v[0].x);
        const T h = (T)0.70710678118654752440;
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        Bfly<4, FWD, C>::run(e);
        Bfly<4, FWD, C>::run(o);
        // o[q] *= w8^q, w8 = exp(s 2 pi i/8)
        C o1 = o[1], o3 = o[3];
        if (FWD) {
            o[1] = C{h * (o1.x + o1.y), h * (o1.y - o1.x)};
            o[3] = C{h * (-o3.x + o3.y), h * (-o3.y - o3.x)};
        } else {
            o[1] = C{h * (o1.x - o1.y), h * (o1.y + o1.x)};
            o[3] = C{h * (-o3.x - o3.y), h * (-o3.y + o3.x)};
        }
        o[2] = mul_si<FWD>(o[2]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = cadd(e[q], o[q]);
            v[q + 4] = csub(e[q], o[q]);
        }
    }
};

template <bool FWD, typename C> struct Bfly<3, FWD, C> {
    __device__ __forceinline__ static void run(C *v) {
        using T = decltype(v[0].x);
        const T c3 = (T)0.86602540378443864676;
        C t1 = cadd(v[1], v[2]);
        C d = csub(v[1], v[2]);
        // t2 = s*i*(sqrt3/2)*(a1-a2)
        C t2 = mul_si<FWD>(C{c3 * d.x, c3 * d.y});
        C m = C{v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, t2);
        v[2] = csub(m, t2);
    }
};

struct DevPlan {
    int n;
    int nst;
    int r[8];
};

// Index decomposition of a work item into (sequence b, butterfly j).
// COLS: sequences are adjacent in memory (b fastest); ROWS: elements are.
template <bool COLS> struct Items;
template <> struct Items<true> {
    __device__ __forceinline__ static void split(int it, int nb, int nR, int &b, int &j) {
        b = it % nb;
        j = it / nb;
    }
    __device__ __forceinline__ static int lds(int b, int e, int nb, int stride) { return e * nb + b; }
};
// Row layout: element e of row b at b*stride + rpad(e); one padding slot per 8
// elements breaks the stride-8 stage writes across LDS banks.
__device__ __forceinline__ int rpad(int e) { return e + (e >> 3); }
__host__ __device__ inline int row_stride(int M) { return M + 1 + ((M + 1) >> 3) + 1; }
template <> struct Items<false> {
    __device__ __forceinline__ static void split(int it, int nb, int nR, int &b, int &j) {
        j = it % nR;
        b = it / nR;
    }
    __device__ __forceinline__ static int lds(int b, int e, int nb, int stride) { return b * stride + rpad(e); }
};

template <typename T, int E, int R, bool COLS, bool FWD, class Src, class Dst>
__device__ __forceinline__ void fft_stage(v2<T> *lds_p, int ldsStride, int n, int nb, int Ns,
                                          bool first, bool last, const v2<T> *__restrict__ tw_p,
                                          const Src &src, const Dst &dst) {
    using C = v2<T>;
    const auto lds = as_local(lds_p);
    const auto tw = as_global(tw_p);
    constexpr int IT = E / R;
    const int NT = blockDim.x;
    const int tid = threadIdx.x;
    const int nR = n / R;
    const int tws = n / (Ns * R);
    C v[E];
    int bb[IT], jj[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        int it = tid + i * NT;
        Items<COLS>::split(it, nb, nR, bb[i], jj[i]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int e = jj[i] + r * nR;
            v[i * R + r] = first ? src(bb[i], e) : lds[Items<COLS>::lds(bb[i], e, nb, ldsStride)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        int k = jj[i] % Ns;
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C w = tw[r * k * tws];
                if (!FWD) w.y = -w.y;
                v[i * R + r] = cmul(v[i * R + r], w);
            }
        }
        Bfly<R, FWD, C>::run(&v[i * R]);
        int idxD = (jj[i] / Ns) * Ns * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            int e = idxD + q * Ns;
            if (last)
                dst(bb[i], e, v[i * R + q]);
            else
                lds[Items<COLS>::lds(bb[i], e, nb, ldsStride)] = v[i * R + q];
        }
    }
    if (!last) __syncthreads();
}

// Runs all stages of `pl` on nb sequences of length pl.n.
template <typename T, int E, bool COLS, bool FWD, class Src, class Dst>
__device__ __forceinline__ void fft_run(v2<T> *lds, int ldsStride, const DevPlan &pl, int nb,
                                        const v2<T> *__restrict__ tw, const Src &src,
                                        const Dst &dst) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.r[s];
        const bool first = (s == 0), last = (s == pl.nst - 1);
        switch (R) {
        case 8:
            if constexpr (E % 8 == 0)
                fft_stage<T, E, 8, COLS, FWD>(lds, ldsStride, pl.n, nb, Ns, first, last, tw, src, dst);
            break;
        case 4:
            if constexpr (E % 4 == 0)
                fft_stage<T, E, 4, COLS, FWD>(lds, ldsStride, pl.n, nb, Ns, first, last, tw, src, dst);
            break;
        case 2:
            if constexpr (E % 2 == 0)
                fft_stage<T, E, 2, COLS, FWD>(lds, ldsStride, pl.n, nb, Ns, first, last, tw, src, dst);
            break;
        case 3:
            if constexpr (E % 3 == 0)
                fft_stage<T, E, 3, COLS, FWD>(lds, ldsStride, pl.n, nb, Ns, first, last, tw, src, dst);
            break;
        default:
            break;
        }
        Ns *= R;
    }
}

}  // namespace cm
