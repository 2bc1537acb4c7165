// Device helpers of the T-web: k-space multiplier, FD stencil, 3x3 eigenvalues.
#pragma once

#include <hip/hip_runtime.h>

namespace cm {

// Wavenumber of index i on an N grid: i<=N/2 -> +i k_f, else -(N-i) k_f
// (PoissonSolver, Miscelanious.cpp:1236-1241).
__device__ __forceinline__ double kcoord(int i, int N, double kf) {
    return kf * (i <= N / 2 ? (double)i : -(double)(N - i));
}

// scale * exp(-k^2 R^2/2) * (-1/k^2), zero at k=0 (kernelcomp filtertype 1,
// NumericalMethods.cpp:2263-2264, times PoissonSolver's -1/k^2).
__device__ __forceinline__ double smooth_poisson_factor(int ix, int iy, int iz, int N, double kf, double R2h,
                                                        double scale) {
    const double kx = kcoord(ix, N, kf), ky = kcoord(iy, N, kf), kz = kf * (double)iz;
    const double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 <= 0.0) return 0.0;
    double f = -scale / k2;
    if (R2h > 0.0) f *= exp(-k2 * R2h);
    return f;
}

// gradfindif stencil (NumericalMethods.cpp:1123-1217); the reference's float
// build evaluates it in double (8.0 is a double literal) and stores to float.
__device__ __forceinline__ float fd4(float fll, float fl, float fr, float frr, double ih) {
    return (float)((-(double)frr + 8.0 * (double)fr - 8.0 * (double)fl + (double)fll) * ih);
}

// Eigenvalues (descending) of the symmetric matrix with diagonal a00,a11,a22 and
// lower triangle a10,a20,a21: trigonometric closed form in FP64.
__device__ __forceinline__ void eig3(double a00, double a11, double a22, double a10, double a20, double a21,
                                     double &e1, double &e2, double &e3) {
    const double p1 = a10 * a10 + a20 * a20 + a21 * a21;
    const double q = (a00 + a11 + a22) * (1.0 / 3.0);
    if (p1 == 0.0) {
        double x = a00, y = a11, z = a22, t;
        if (x < y) { t = x; x = y; y = t; }
        if (y < z) { t = y; y = z; z = t; }
        if (x < y) { t = x; x = y; y = t; }
        e1 = x; e2 = y; e3 = z;
        return;
    }
    const double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
    const double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * p1;
    const double p = sqrt(p2 * (1.0 / 6.0));
    const double ip = 1.0 / p;
    const double c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip;
    const double c10 = a10 * ip, c20 = a20 * ip, c21 = a21 * ip;
    const double det = c00 * (c11 * c22 - c21 * c21) - c10 * (c10 * c22 - c21 * c20) + c20 * (c10 * c21 - c11 * c20);
    double r = 0.5 * det;
    r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
    const double phi = acos(r) * (1.0 / 3.0);
    e1 = q + 2.0 * p * cos(phi);
    e3 = q + 2.0 * p * cos(phi + 2.0943951023931954923);
    e2 = 3.0 * q - e1 - e3;
}

// The gradfindif stencil in the FP64 T-web (the reference's DOUBLE_PREC build).
__device__ __forceinline__ double fd4d(double fll, double fl, double fr, double frr, double ih) {
    return (-frr + 8.0 * fr - 8.0 * fl + fll) * ih;
}

// Margin tolerance of a T-web call.  tol >= 0 is absolute.  A negative tol
// asks for the rounding tolerance |tol| K(N) u ||phi||_rms / d^2 of the
// potential's arithmetic (u = 2^-24 for float32 phi, 2^-53 for FP64), taken on
// the device from ||phi||_rms (a float's bits in *norm, written by a reduction
// over phi), so no host synchronisation is needed; per_norm = |tol| K u / d^2.
struct TolK {
    double tol;
    double per_norm;
    const unsigned *norm;
};
__device__ __forceinline__ double tol_value(const TolK &t) {
    return t.norm ? t.per_norm * (double)__uint_as_float(*t.norm) : t.tol;
}

// One mock of a packed T-web launch (several potentials, one launch over the
// union of their (tile column, x segment) work items): its potential, labels,
// counters, tolerance and geometry; first = its first workgroup.
struct TwMock {
    const float *phi;
    int8_t *labels;
    unsigned long long *stats;
    TolK tk;
    double lth;
    float ih2;
    int N, segLen, first;
};

// The same stencil in float (fused kernel): 3 FMAs and a multiply.
__device__ __forceinline__ float fd4f(float fll, float fl, float fr, float frr, float ih) {
    return ((fll - frr) + 8.0f * (fr - fl)) * ih;
}

// The stencil without the 1/(12 d) factor: the fused T-web works with
// T' = T (12 d)^2, whose inertia relative to lambda_th (12 d)^2 is that of T.
__device__ __forceinline__ float fd4s(float fll, float fl, float fr, float frr) {
    return (fll - frr) + 8.0f * (fr - fl);
}

}  // namespace cm

namespace cm {

// Number of eigenvalues of the symmetric matrix B (diagonal b00,b11,b22, lower
// triangle a10,a20,a21) that are > 0, from the signs of the coefficients of its
// characteristic polynomial t^3 - e1 t^2 + e2 t - e3 (Descartes' rule of signs,
// exact for a polynomial whose roots are all real; zero roots are factored
// out).  FP64.  Also returns |e3| and the squared Frobenius norm, which bound
// the smallest |eigenvalue| from below: min|mu| >= |e3| / ||B||_F^2.
__device__ __forceinline__ int positive_eigs(double b00, double b11, double b22, double a10, double a20,
                                             double a21, double &abs_det, double &fro2) {
    const double e1 = b00 + b11 + b22;
    const double e2 = b00 * b11 + b00 * b22 + b11 * b22 - a10 * a10 - a20 * a20 - a21 * a21;
    const double e3 = b00 * (b11 * b22 - a21 * a21) - a10 * (a10 * b22 - a21 * a20) + a20 * (a10 * a21 - b11 * a20);
    abs_det = fabs(e3);
    fro2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * (a10 * a10 + a20 * a20 + a21 * a21);
    // coefficient sequence 1, -e1, e2, -e3 with trailing zero roots removed
    double c[4] = {1.0, -e1, e2, -e3};
    int deg = 3;
    while (deg > 0 && c[deg] == 0.0) --deg;
    int changes = 0;
    double prev = 1.0;
    for (int i = 1; i <= deg; ++i) {
        if (c[i] == 0.0) continue;
        if ((c[i] > 0.0) != (prev > 0.0)) ++changes;
        prev = c[i];
    }
    return changes;
}


// Same count from float T: the coefficients are evaluated in float together
// with a bound on their rounding error (a few ulps of the sum of absolute
// values of the terms).  When every coefficient is farther from zero than its
// bound its sign is certain and the float count equals the FP64 count;
// otherwise the FP64 evaluation decides.
__device__ __forceinline__ int positive_eigs_f(float b00, float b11, float b22, float a10, float a20, float a21,
                                               float &det, float &det_err, float &fro2) {
    const float e1 = b00 + b11 + b22;
    const float m1 = fabsf(b00) + fabsf(b11) + fabsf(b22);
    const float p01 = b00 * b11, p02 = b00 * b22, p12 = b11 * b22;
    const float q10 = a10 * a10, q20 = a20 * a20, q21 = a21 * a21;
    const float e2 = (p01 + p02 + p12) - (q10 + q20 + q21);
    const float m2 = fabsf(p01) + fabsf(p02) + fabsf(p12) + q10 + q20 + q21;
    const float c0 = b11 * b22 - q21, c1 = a10 * b22 - a21 * a20, c2 = a10 * a21 - b11 * a20;
    const float e3 = b00 * c0 - a10 * c1 + a20 * c2;
    const float m3 = fabsf(b00) * (fabsf(p12) + q21) + fabsf(a10) * (fabsf(a10 * b22) + fabsf(a21 * a20)) +
                     fabsf(a20) * (fabsf(a10 * a21) + fabsf(b11 * a20));
    const float eps = 8.0f * 5.9604645e-8f;  // 8 units of float roundoff
    det = e3;
    det_err = 3.0f * eps * m3;
    fro2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0f * (q10 + q20 + q21);
    if (fabsf(e1) <= eps * m1 || fabsf(e2) <= 2.0f * eps * m2 || fabsf(e3) <= 3.0f * eps * m3 || !(m3 < 1e30f)) {
        double ad, f2;
        return positive_eigs(b00, b11, b22, a10, a20, a21, ad, f2);
    }
    // no zero roots: sign changes of 1, -e1, e2, -e3
    int ch = 0;
    ch += (e1 > 0.f);            // 1 -> -e1
    ch += ((-e1 > 0.f) != (e2 > 0.f));
    ch += ((e2 > 0.f) != (-e3 > 0.f));
    return ch;
}


// Same count with one scale for all rounding bounds: s = sum of |entries|
// bounds the sums of absolute values of the terms of e1, e2, e3 by s, s^2, s^3,
// so the float signs are certain when |e1| > 8u s, |e2| > 16u s^2 and
// |e3| > 24u s^3 (u = 2^-24; the factors cover every rounding of the
// evaluation with room to spare).  Otherwise FP64 decides.  Also returns |e3|
// minus its bound (a lower bound of |det B|) and ||B||_F^2.
// The float count of positive_eigs_s without its FP64 branch: unc is set when
// a coefficient lies within its rounding bound (the count is then not
// certain and the FP64 count must replace it).  marg is set when the
// eigenvalue of B closest to zero may lie within tol of it, from a bound on
// the roots of the characteristic polynomial: the reciprocals w = 1/mu solve
// w^3 - (e2/e3) w^2 + (e1/e3) w - 1/e3 = 0, so by Fujiwara's bound
// |w| <= 2 max(|e2/e3|, |e1/e3|^(1/2), |1/(2 e3)|^(1/3)), and min |mu| >= tol
// whenever |e3| >= max(2 tol |e2|, 4 tol^2 |e1|, 4 tol^3) (each coefficient
// taken at the unfavourable end of its rounding bound).  Unlike
// |det|/||B||_F^2 this stays within a factor ~2 of min |mu| when the other two
// eigenvalues differ widely.  Below tol = 1e-10 (tol^3 near float underflow)
// the Frobenius bound min |mu| >= 2 |det| / ||B||_F^2 is used instead.
__device__ __forceinline__ int positive_eigs_nb(float b00, float b11, float b22, float a10, float a20, float a21,
                                                bool &unc, float tol, bool &marg) {
    const float e1 = b00 + b11 + b22;
    const float q10 = a10 * a10, q20 = a20 * a20, q21 = a21 * a21;
    const float e2 = (b00 * b11 + b00 * b22 + b11 * b22) - (q10 + q20 + q21);
    const float c0 = b11 * b22 - q21, c1 = a10 * b22 - a21 * a20, c2 = a10 * a21 - b11 * a20;
    const float e3 = b00 * c0 - a10 * c1 + a20 * c2;
    const float sc = (fabsf(b00) + fabsf(b11) + fabsf(b22)) + (fabsf(a10) + fabsf(a20) + fabsf(a21));
    const float u = 5.9604645e-8f;
    const float t1 = 8.0f * u * sc, t2 = 2.0f * t1 * sc, t3 = 1.5f * t2 * sc;
    const float ae3 = fabsf(e3), det_lo = ae3 - t3;
    unc = (fabsf(e1) <= t1) | (fabsf(e2) <= t2) | (ae3 <= t3) | !(sc < 1e12f);
    if (tol >= 1e-10f) {
        const float thr = fmaxf(fmaxf(2.0f * tol * (fabsf(e2) + t2), 4.0f * tol * tol * (fabsf(e1) + t1)),
                                4.0f * tol * tol * tol);
        marg = tol > 0.f && det_lo < thr * 1.0001f;
    } else {
        const float fro2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0f * (q10 + q20 + q21);
        marg = tol > 0.f && 2.0f * det_lo < tol * fro2 * 1.0001f;
    }
    return (e1 > 0.f) + ((e1 < 0.f) != (e2 > 0.f)) + ((e2 > 0.f) != (e3 < 0.f));
}

// FP64 count for a cell whose float signs are not certain, with the margin
// pre-filter of positive_eigs_nb evaluated on the FP64 coefficients (bounds
// with u = 2^-53): cells flagged only because of float rounding drop out.
__device__ __forceinline__ int positive_eigs_nb64(float b00, float b11, float b22, float a10, float a20, float a21,
                                                  float tol, bool &marg) {
    double ad, f2;
    const int v = positive_eigs(b00, b11, b22, a10, a20, a21, ad, f2);
    const double d00 = b00, d11 = b11, d22 = b22, d10 = a10, d20 = a20, d21 = a21;
    const double e1 = d00 + d11 + d22;
    const double e2 = d00 * d11 + d00 * d22 + d11 * d22 - d10 * d10 - d20 * d20 - d21 * d21;
    const double sc = (fabs(d00) + fabs(d11) + fabs(d22)) + (fabs(d10) + fabs(d20) + fabs(d21));
    const double u = 1.1102230246251565e-16;
    const double t1 = 8.0 * u * sc, t2 = 2.0 * t1 * sc, t3 = 1.5 * t2 * sc;
    const double det_lo = ad - t3, tl = (double)tol;
    const double thr = fmax(fmax(2.0 * tl * (fabs(e2) + t2), 4.0 * tl * tl * (fabs(e1) + t1)), 4.0 * tl * tl * tl);
    marg = tol > 0.f && det_lo < thr * 1.0001;
    return v;
}

// FP64 count of positive eigenvalues of B from float entries (inline).
__device__ __forceinline__ int positive_eigs_count64(float b00, float b11, float b22, float a10, float a20,
                                                     float a21) {
    double ad, f2;
    return positive_eigs(b00, b11, b22, a10, a20, a21, ad, f2);
}

// Sign count of B and its margin band from one set of coefficients.  With
// e1, e2, e3 the coefficients of B's characteristic polynomial
// t^3 - e1 t^2 + e2 t - e3, det(B - sI) = e3 - s (e2 - s (e1 - s)).  An odd
// number of eigenvalues lies in (-tol, tol) exactly when det(B - sI) has
// opposite signs at s = -tol and s = +tol.  Two eigenvalues mu1, mu2 in the
// band would force both |e3| = |mu1 mu2 mu3| < tol^2 m and
// |e2| = |mu1 mu2 + mu3 (mu1 + mu2)| < tol^2 + 2 tol m with
// m = |e1| + 2 tol > |mu3| (e1 = mu1 + mu2 + mu3), so when either exceeds its
// bound (widened by the rounding bounds of e1..e3) the cell is marginal (an
// eigenvalue within tol of zero) exactly when the sign changes; otherwise
// (3e-5 of the cells of the 512^3 mock) unc is set and FP64 decides.  The two
// determinants are taken in float as they come: a cell whose eigenvalue lies
// within their rounding of -+ tol may be counted either way, the float
// rounding of the boundary itself.  tol = 0: no band test.
// Wave form: called by every active lane together, it returns lane masks of
// the wave (bit l for lane l): the count's three sign changes ma (1 -> -e1),
// mb (-e1 -> e2) and mc (e2 -> -e3), so count = ma + mb + mc per lane; munc
// (the float signs are not all certain) and mband (an eigenvalue within tol of
// zero).  Each mask is taken straight from a compare and combined in scalar
// registers.
__device__ __forceinline__ void positive_eigs_band_wave(float b00, float b11, float b22, float a10, float a20,
                                                        float a21, float tol, unsigned long long &ma,
                                                        unsigned long long &mb, unsigned long long &mc,
                                                        unsigned long long &munc, unsigned long long &mband) {
    const float e1 = b00 + b11 + b22;
    const float q10 = a10 * a10, q20 = a20 * a20, q21 = a21 * a21;
    const float e2 = (b00 * b11 + b00 * b22 + b11 * b22) - (q10 + q20 + q21);
    const float c0 = b11 * b22 - q21, c1 = a10 * b22 - a21 * a20, c2 = a10 * a21 - b11 * a20;
    const float e3 = b00 * c0 - a10 * c1 + a20 * c2;
    const float sc = (fabsf(b00) + fabsf(b11) + fabsf(b22)) + (fabsf(a10) + fabsf(a20) + fabsf(a21));
    const float u = 5.9604645e-8f;
    const float t1 = 8.0f * u * sc, t2 = 2.0f * t1 * sc, t3 = 1.5f * t2 * sc;
    const unsigned long long p1 = __builtin_amdgcn_ballot_w64(e1 > 0.f), n1 = __builtin_amdgcn_ballot_w64(e1 < 0.f),
                             p2 = __builtin_amdgcn_ballot_w64(e2 > 0.f), n3 = __builtin_amdgcn_ballot_w64(e3 < 0.f);
    ma = p1;
    mb = n1 ^ p2;
    mc = p2 ^ n3;
    munc = __builtin_amdgcn_ballot_w64(fabsf(e1) <= t1) | __builtin_amdgcn_ballot_w64(fabsf(e2) <= t2) |
           __builtin_amdgcn_ballot_w64(fabsf(e3) <= t3) | __builtin_amdgcn_ballot_w64(!(sc < 1e12f));
    mband = 0ull;
    if (tol > 0.f) {  // uniform
        const float dm = e3 + tol * (e2 + tol * (e1 + tol));  // det(B + tol I)
        const float dp = e3 - tol * (e2 - tol * (e1 - tol));  // det(B - tol I)
        const float m = fabsf(e1) + 2.0f * tol + t1;
        munc |= __builtin_amdgcn_ballot_w64(fabsf(e3) <= tol * tol * m + t3) &
                __builtin_amdgcn_ballot_w64(fabsf(e2) <= tol * (tol + 2.0f * m) + t2);
        mband = __builtin_amdgcn_ballot_w64(dm > 0.f) ^ __builtin_amdgcn_ballot_w64(dp > 0.f);
    }
}

// FP64 decision of positive_eigs_band_wave for a cell whose float signs are not
// all certain: the count of B and the band test from B -+ tol I in FP64.
__device__ __forceinline__ int positive_eigs_band64(float b00, float b11, float b22, float a10, float a20,
                                                    float a21, float tol, bool &band) {
    double ad, f2;
    const int v = positive_eigs(b00, b11, b22, a10, a20, a21, ad, f2);
    band = false;
    if (tol > 0.f) {
        const double t = tol;
        const int lo = positive_eigs((double)b00 + t, (double)b11 + t, (double)b22 + t, a10, a20, a21, ad, f2);
        const int hi = positive_eigs((double)b00 - t, (double)b11 - t, (double)b22 - t, a10, a20, a21, ad, f2);
        band = lo != hi;
    }
    return v;
}

__device__ __forceinline__ int positive_eigs_s(float b00, float b11, float b22, float a10, float a20, float a21,
                                               float &det_lo, float &fro2) {
    const float e1 = b00 + b11 + b22;
    const float q10 = a10 * a10, q20 = a20 * a20, q21 = a21 * a21;
    const float e2 = (b00 * b11 + b00 * b22 + b11 * b22) - (q10 + q20 + q21);
    const float c0 = b11 * b22 - q21, c1 = a10 * b22 - a21 * a20, c2 = a10 * a21 - b11 * a20;
    const float e3 = b00 * c0 - a10 * c1 + a20 * c2;
    const float sc = (fabsf(b00) + fabsf(b11) + fabsf(b22)) + (fabsf(a10) + fabsf(a20) + fabsf(a21));
    const float u = 5.9604645e-8f;
    const float t1 = 8.0f * u * sc, t2 = 2.0f * t1 * sc, t3 = 1.5f * t2 * sc;
    const float ae3 = fabsf(e3);
    det_lo = ae3 - t3;
    fro2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0f * (q10 + q20 + q21);
    if (fabsf(e1) <= t1 || fabsf(e2) <= t2 || ae3 <= t3 || !(sc < 1e12f)) {
        double ad, f2;
        return positive_eigs(b00, b11, b22, a10, a20, a21, ad, f2);
    }
    // no zero roots: sign changes of 1, -e1, e2, -e3
    return (e1 > 0.f) + ((e1 < 0.f) != (e2 > 0.f)) + ((e2 > 0.f) != (e3 < 0.f));
}

}  // namespace cm
