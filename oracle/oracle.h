/*
 * oracle.h -- FP64 CPU restatement of the COSMICATLAS mesh-measurement path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in cosmicatlas_amd/ links or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates
 * (paths relative to the COSMICATLAS source tree).
 */
#ifndef COSMO_ORACLE_H
#define COSMO_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mass-assignment schemes, numbered like Params::mass_assignment
 * (Source/Params.cpp:3531-3542: NGP 0, CIC 1, TSC 2, PCS 3). */
enum { OR_NGP = 0, OR_CIC = 1, OR_TSC = 2, OR_PCS = 3 };

/* Particle -> mesh assignment, the per-particle grid_assignment_{NGP,CIC,TSC,PCS}
 * (Source/Miscelanious.cpp:2597-2868) applied to every particle in turn, serially
 * (the reference's OpenMP loop over tracers races, Source/Catalog.cpp:14117-14211).
 *   pos      : positions, AoS xyz (layout 0) or SoA x[n],y[n],z[n] (layout 1)
 *   pos_f32  : 1 if pos holds float32 (the periodic wrap then happens in float32,
 *              like real_prec=float in the reference), 0 for float64
 *   w        : per-particle weights in the position dtype, or NULL for 1
 *   xmin     : box corner (Xoffset-L/2 in the reference), subtracted first
 *   interlace_shift : added to every coordinate after the wrap, in cell units
 *              (0 for the plain mesh, 0.5 for the interlaced mesh)
 *   mesh     : N^3 float64, C order z fastest; ACCUMULATED into (caller zeroes)
 * PCS uses Yff+2 for the y index (the reference's Xff+2 at Miscelanious.cpp:2828
 * is a typo).  Returns the number of particles deposited. */
size_t or_mas(int N, double L, const double xmin[3], const void *pos, int pos_f32,
              int layout_soa, size_t n, const void *w, int scheme,
              double interlace_shift, double *mesh);

/* 1-D MAS kernels W(s), s in cell units (Source/Miscelanious.cpp:1277-1321). */
double or_mas_kernel(int scheme, double s);

/* 3-D real-to-complex FFT, unnormalised forward e^{-ik.x}, FFTW r2c layout
 * N x N x (N/2+1) interleaved (re,im) (Source/NumericalMethods.cpp:1718-1753).
 * Own mixed-radix FFT (radix 2,3,5 and a direct DFT for other factors). */
void or_fft_r2c(int N, const double *in, double *out);

/* 3-D complex-to-real inverse FFT divided by N^3, the do_fftw_c2r convention
 * (Source/NumericalMethods.cpp:1818-1856).  The input half spectrum is taken as
 * Hermitian: the full spectrum is rebuilt from it, transformed, and its real
 * part returned. */
void or_fft_c2r(int N, const double *in, double *out);

/* 1-D complex DFT of length n (sign -1 forward, +1 backward, unnormalised),
 * out-of-place, interleaved.  Exposed for the known-answer tests. */
void or_fft_1d(int n, const double *in, double *out, int sign);

/* Estimator options, the members power_spectrum_fkp reads
 * (Source/FftwFunctions.cpp:958-1657; binning constants Source/Params.cpp:4079-4115). */
typedef struct {
    int    N;             /* Nft */
    double L;             /* Lbox */
    int    nshells;       /* Nnp_data = Nft/ndel_data/2 for linear binning */
    double kmin;          /* DeltaKmin * deltak_0 */
    double dk;            /* ndel_data * deltak_0 */
    int    scheme;        /* mass_assignment 0..3 */
    int    mas_correction;/* MAS_correction */
    double shot_noise;    /* SN term V/N_gal (0 when SN_correction is off) */
    double normal_power;  /* N_gal^2/V */
    int    float_shell_rule; /* 1: the reference's float32 floor (FftwFunctions.cpp:1097);
                                0: exact (FP64) rule */
    int    log_binning;   /* type_of_binning = log: shell floor(log10(|k|/kmin)/dlog) + 1
                             (FftwFunctions.cpp:1103-1108), exact rule */
    double dlog;          /* Deltal (Params.cpp:4106) */
} or_pk_opts;

/* Window shells of the random branch (FftwFunctions.cpp:1248-1262, 1354-1362,
 * 1454-1462, 1557-1568): the zero mode and the three non-first octant copies. */
void or_window_shells(const double *dk, const or_pk_opts *o, double *pw, int64_t *nmodes);

/* power_spectrum_fkp box branch (Source/FftwFunctions.cpp:958-1657): enumerates
 * the four octant copies of every mode, bins |k| into shells, subtracts the
 * aliased shot noise (Jing 2005, Miscelanious.cpp:2532-2554), deconvolves the
 * MAS window (Miscelanious.cpp:2557-2562) and normalises.  pk2/pk4 are the
 * Legendre multipoles with the line of sight along z (def.h:1714), ell=2,4
 * without shot-noise subtraction (FftwFunctions.cpp:1177,1207-1211,1620-1628).
 *   dk      : N x N x (N/2+1) interleaved complex spectrum
 *   raw_sum : if non-NULL, receives the per-shell sums before normalisation
 */
void or_pk_shells(const double *dk, const or_pk_opts *o, double *pk0, double *pk2,
                  double *pk4, int64_t *nmodes);

/* Shell index of the integer mode (i,j,k), both rules; -1 when outside. */
int or_shell_index(const or_pk_opts *o, int i, int j, int k);

/* Gaussian kernel exp(-k^2 R^2/2) on the r2c grid (kernelcomp filtertype 1,
 * Source/NumericalMethods.cpp:2197-2306) times the Poisson multiplier -1/k^2
 * with phi(0)=0 (PoissonSolver, Source/Miscelanious.cpp:1189-1274), applied in
 * place to an N x N x (N/2+1) spectrum.  R <= 0 skips the smoothing. */
void or_smooth_poisson(int N, double L, double R, double *dk);

/* Periodic 4th-order central difference along dim 1 (x, slowest), 2 or 3
 * (Source/NumericalMethods.cpp:1123-1217). */
void or_gradfindif(int N, double L, const double *in, double *out, int dim);

/* Eigenvalues of a symmetric 3x3 matrix, descending (gsl_eigen_symmv + sort
 * VAL_DESC, Source/NumericalMethods.cpp:572-592; GSL reads the lower triangle
 * m[i+3j], i>=j).  Cyclic Jacobi in FP64. */
void or_eigen3(const double m[9], double ev[3]);

/* T-web (Cwclass::get_CWC, Source/Cwclass.cpp:235-894, default macros):
 * delta (N^3 real) -> r2c -> Gaussian(R) * (-1/k^2) -> c2r -> phi;
 * EigenValuesTweb FD branch (Source/Miscelanious.cpp:1668-1769): T_ij from
 * gradfindif o gradfindif; eigenvalues (desc); label = 4 - #{lambda_j > lambda_th}
 * (Cwclass.cpp:874-894).  lambdas (3 N^3, l1|l2|l3 planes) and phi may be NULL.
 * counts[4] receives the number of cells with label 1..4. */
void or_tweb(int N, double L, const double *delta, double R, double lambda_th,
             int8_t *labels, double *lambdas, double *phi, int64_t counts[4]);

/* T-web with the spectral tidal tensor (EigenValuesTweb _USE_GFFT_EIGENV_
 * branch, Source/Miscelanious.cpp:1703-1718; calc_LapPhiv,
 * Source/NumericalMethods.cpp:2605-2679): T_ab = c2r(W(k) k_a k_b / k^2
 * delta_k); eigenvalues, labels and counts as or_tweb. */
void or_tweb_spectral(int N, double L, const double *delta, double R, double lambda_th,
                      int8_t *labels, double *lambdas, int64_t counts[4]);

/* T-web from a given potential (EigenValuesTweb FD branch + labels only). */
void or_tweb_from_phi(int N, double L, const double *phi, double lambda_th,
                      int8_t *labels, double *lambdas, int64_t counts[4]);

/* Direct (brute-force) bispectrum of a box, FftwFunctions::bispectrum_fkp
 * (Source/FftwFunctions.cpp:3105-3297): every ordered pair (k1, k2) of the
 * (N+1)^3 grid q in [-N/2, N/2]^3 (both Nyquist planes, :3146-3147; the
 * half-spectrum lookup of remap, Source/Miscelanious.cpp:43-62) whose
 * k3 = -k1-k2 lies in the same grid (:3262), none of the three zero (:3279),
 * accumulated in the bin triple sorted ascending (:3284-3287):
 *   bsum[(a nb + b) nb + c] += Re(d1 d2 d3), snsum += |d1|^2+|d2|^2+|d3|^2,
 *   cnt += 1   (a <= b <= c < nb);
 *   psum[a] += |d1|^2, pcnt[a] += 1 over k1 (:3194-3199).
 * dk: complex128 half spectrum N x N x (N/2+1).  d = dk / W with
 * W = prod_c sinc(pi q_c/N)^(scheme+1) when mas_corr (correction_MAS,
 * Miscelanious.cpp:2565-2577, applied to all three vectors).  Bins: shell
 * s of |q| with Dmin + s ndel <= |q| < Dmin + (s+1) ndel (|q| in units of k_f):
 * integer comparisons of |q|^2 when Dmin and ndel are integers, else FP64;
 * modes below Dmin or at s >= nb are left out.  Outputs are accumulated into
 * (caller zeroes). */
void or_bk_direct(int N, const double *dk, double Dmin, double ndel, int nb, int scheme, int mas_corr,
                  double *bsum, double *snsum, int64_t *cnt, double *psum, int64_t *pcnt);

#ifdef __cplusplus
}
#endif
#endif
