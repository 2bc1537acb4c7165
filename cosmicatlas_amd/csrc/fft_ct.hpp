// Compile-time variant of the Stockham engine (fft_engine.hpp) for float32.
//
// The transform length n, the number of sequences per workgroup NB and the LDS
// row stride are template parameters, so every stage's radix, butterfly span
// Ns, twiddle stride and index split are constants: the integer divisions and
// the runtime radix switch of the generic engine disappear.  Stage order and
// arithmetic are those of the generic engine (radix 8s, then 4/2, then 3).
#pragma once

#include "fft_engine.hpp"

namespace cm {

struct CtPlan {
    int nst;
    int r[8];
};

constexpr CtPlan ct_plan(int n) {
    CtPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0}};
    int m = n, n3 = 0, k2 = 0;
    while (m % 3 == 0) {
        m /= 3;
        ++n3;
    }
    while (m > 1 && m % 2 == 0) {
        m /= 2;
        ++k2;
    }
    int n8 = k2 / 3, rem = k2 % 3, n4 = 0, n2 = 0;
    if (rem == 2) n4 = 1;
    if (rem == 1) {
        if (n8 > 0) {
            --n8;
            n4 = 2;
        } else {
            n2 = 1;
        }
    }
    int s = 0;
    for (int i = 0; i < n8; ++i) p.r[s++] = 8;
    for (int i = 0; i < n4; ++i) p.r[s++] = 4;
    for (int i = 0; i < n2; ++i) p.r[s++] = 2;
    for (int i = 0; i < n3; ++i) p.r[s++] = 3;
    p.nst = s;
    return p;
}

constexpr int ct_nst(int n) { return ct_plan(n).nst; }
constexpr int ct_radix(int n, int s) { return ct_plan(n).r[s]; }

// elements per thread: divisible by every radix of the plan
constexpr int ct_E(int n) {
    bool has3 = false, has8 = false;
    CtPlan p = ct_plan(n);
    for (int i = 0; i < p.nst; ++i) {
        has3 = has3 || p.r[i] == 3;
        has8 = has8 || p.r[i] == 8;
    }
    return has3 ? 24 : ((has8 && n >= 64) ? 16 : 8);
}

constexpr int ct_row_stride(int M) { return M + 1 + ((M + 1) >> 3) + 1; }

// Stage S of the plan for length n (Ns = product of the earlier radices).
// COLS: element e of sequence b at LDS e*STRIDE + b (STRIDE >= NB; NB+1 keeps
// per-element rows on distinct banks); rows: b*STRIDE + rpad(e).
// The first stage reads through src(b, e), the last writes through
// dst(b, e, v); the caller owns the barrier before the first stage.
// TWL: the twiddle table is in LDS (else global memory).
template <typename T, int n, int NB, int STRIDE, bool COLS, bool FWD, int S, int Ns, int E, bool TWL, class Src,
          class Dst>
__device__ __forceinline__ void ct_stage(v2<T> *lds_p, const v2<T> *__restrict__ tw_p, const Src &src,
                                         const Dst &dst) {
    using C = v2<T>;
    // LDS and global views: the pointers may arrive through a non-inlined
    // function, where they would otherwise be accessed with flat instructions
    const auto lds = as_local(lds_p);
    auto twv = [&](int i) -> C {
        if constexpr (TWL) {
            const C x = as_local(tw_p)[i];
            return x;
        } else {
            const C x = as_global(tw_p)[i];
            return x;
        }
    };
    constexpr int R = ct_radix(n, S), NST = ct_nst(n);
    constexpr bool FIRST = S == 0, LAST = S == NST - 1;
    constexpr int IT = E / R, nR = n / R, NT = NB * n / E, tws = n / (Ns * R);
    static_assert(E % R == 0 && NT * E == NB * n, "plan does not tile the workgroup");
    const int tid = threadIdx.x;
    C v[E];
    int bb[IT], jj[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int it = tid + i * NT;
        if constexpr (COLS) {
            bb[i] = it % NB;
            jj[i] = it / NB;
        } else {
            jj[i] = it % nR;
            bb[i] = it / nR;
        }
    }
    // twiddles of this thread's butterflies, all loaded before the stage's
    // barrier so that their latency overlaps it (loaded one at a time next to
    // their use, each load waited for the previous one)
    C w[IT][R];
    if constexpr (Ns > 1) {
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int k = jj[i] % Ns;
#pragma unroll
            for (int r = 1; r < R; ++r) w[i][r] = twv(r * k * tws);
        }
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int e = jj[i] + r * nR;
            if constexpr (FIRST)
                v[i * R + r] = src(bb[i], e);
            else
                v[i * R + r] = lds[COLS ? e * STRIDE + bb[i] : bb[i] * STRIDE + rpad(e)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int k = jj[i] % Ns;
        if constexpr (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C wr = w[i][r];
                if (!FWD) wr.y = -wr.y;
                v[i * R + r] = cmul(v[i * R + r], wr);
            }
        }
        Bfly<R, FWD, C>::run(&v[i * R]);
        const int idxD = (jj[i] / Ns) * Ns * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int e = idxD + q * Ns;
            if constexpr (LAST)
                dst(bb[i], e, v[i * R + q]);
            else
                lds[COLS ? e * STRIDE + bb[i] : bb[i] * STRIDE + rpad(e)] = v[i * R + q];
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
        ct_stage<T, n, NB, STRIDE, COLS, FWD, S + 1, Ns * R, E, TWL>(lds_p, tw_p, src, dst);
    }
}

// E: elements per thread (ct_E(n) by default; any multiple of every radix of
// the plan), so the workgroup has NB*n/E threads.  TWL: tw points into LDS.
template <typename T, int n, int NB, int STRIDE, bool COLS, bool FWD, int E = ct_E(n), bool TWL = false, class Src,
          class Dst>
__device__ __forceinline__ void ct_fft(v2<T> *lds, const v2<T> *__restrict__ tw, const Src &src, const Dst &dst) {
    ct_stage<T, n, NB, STRIDE, COLS, FWD, 0, 1, E, TWL>(lds, tw, src, dst);
}

// Workgroup shapes of the three passes for a float32 N^3 transform.
// z rows (length M = N/2): about 4096 complex values per workgroup.
template <int N> struct RowShape {
    static constexpr int M = N / 2;
    static constexpr int E = ct_E(M);
    static constexpr int pick() {
        const int rows = N * N, rows_pow2 = rows & -rows;
        int nb = 1;
        while (nb * 2 * M <= 4096 && nb * 2 <= rows_pow2) nb *= 2;
        while ((nb * M) % E) nb *= 2;
        while (nb * M / E > 512 && nb > 1) nb /= 2;
        return nb;
    }
    static constexpr int NB = pick();
    static constexpr int NT = NB * M / E;
    static constexpr int STRIDE = ct_row_stride(M);
    static constexpr size_t LDS = sizeof(float2) * (size_t)NB * STRIDE;
};

// y / x columns (length N): NB adjacent kz columns per workgroup.
template <int N, int NBC, int EC = ct_E(N)> struct ColShape {
    static constexpr int E = EC;
    static constexpr int NB = NBC;
    static constexpr int NT = NB * N / E;
    static constexpr size_t LDS = sizeof(float2) * (size_t)NB * N;
    static_assert(NT <= 1024, "column workgroup too large");
};

// kz row pitch (complex values) of the internal half spectrum: N/2+1 rounded
// up to 16 so that 16-column segments are whole 128-byte lines.
constexpr int padded_pitch(int N) { return ((N / 2 + 1) + 15) & ~15; }

}  // namespace cm

// Supported transform sizes (N = 2^a or 3*2^a, 8 <= N <= 1024).
#define CM_FFT_SIZES(X) X(8) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)
