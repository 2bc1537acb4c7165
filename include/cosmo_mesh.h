/*
 * cosmo_mesh.h -- C ABI of the MI355X (gfx950) mesh-measurement library
 * libcosmo_mesh.so: particle->mesh mass assignment, 3-D r2c/c2r FFT, the P(k)
 * shell estimator and the T-web classification.
 *
 * Plain C: pointers, sizes and PODs only.  Every compute entry point takes
 * DEVICE pointers (allocate them with cm_dev_alloc or any HIP allocator) and
 * runs on the context's stream; small results (shell sums, counts) come back to
 * host memory and those calls synchronise that stream.  Errors are reported as
 * cm_status codes with a message in cm_last_error(); nothing exits the process.
 *
 * Each entry point names the reference interface it replaces (paths relative to
 * the COSMICATLAS source tree).
 */
#ifndef COSMO_MESH_H
#define COSMO_MESH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CM_OK = 0,
    CM_ERR_INVALID = 1,     /* bad argument (null pointer, N not supported, ...) */
    CM_ERR_UNSUPPORTED = 2, /* valid request that this build does not implement */
    CM_ERR_NOMEM = 3,       /* device allocation failed */
    CM_ERR_DEVICE = 4,      /* HIP runtime error */
    CM_ERR_BUDGET = 5,      /* request larger than the context's HBM budget */
    CM_ERR_TIMEOUT = 6,     /* serving: request not finished before its deadline */
    CM_ERR_CANCELLED = 7    /* serving: request cancelled */
} cm_status;

typedef enum { CM_F32 = 0, CM_F64 = 1 } cm_dtype;

/* Mass-assignment schemes numbered like Params::mass_assignment
 * (Source/Params.cpp:3531-3542). */
typedef enum { CM_NGP = 0, CM_CIC = 1, CM_TSC = 2, CM_PCS = 3 } cm_mas;

typedef enum { CM_AOS = 0, CM_SOA = 1 } cm_layout;

/* Deposit algorithm: plain = one thread per particle with global atomics;
 * tiled = particles binned to 16^3-cell tiles (with ghost copies for the
 * stencil halo) and accumulated in LDS, each tile written once. */
typedef enum { CM_MAS_AUTO = 0, CM_MAS_PLAIN = 1, CM_MAS_TILED = 2 } cm_mas_algo;

typedef struct cm_ctx cm_ctx;

/* ---------------------------------------------------------------- context */
/* Opens a context on HIP device `device`.  hbm_budget (bytes, 0 = 90% of the
 * device) bounds the scratch the context may keep resident.  No reference
 * counterpart: the reference holds state in Params/FftwFunctions objects
 * (Headers/FftwFunctions.h:1083, Source/FftwFunctions.cpp:28-156). */
cm_status cm_ctx_create(int device, size_t hbm_budget, cm_ctx **out);
void cm_ctx_destroy(cm_ctx *ctx);
const char *cm_last_error(const cm_ctx *ctx);
const char *cm_status_string(cm_status s);
/* Run subsequent work on `stream` (a hipStream_t; NULL = the context's own). */
cm_status cm_set_stream(cm_ctx *ctx, void *stream);
void *cm_get_stream(cm_ctx *ctx);
cm_status cm_synchronize(cm_ctx *ctx);

/* hipGraph replay of repeated fused measurements (cm_measure, fused = 1): the
 * first call with given options, pointers, particle count, stream and launch
 * shapes runs eagerly, the second is captured into a graph, later ones launch
 * the graph (one submission instead of ~20 kernel launches).  A graph is
 * recaptured when one of the context's work buffers has been reallocated.  On
 * by default; 0 turns it off for this context.  No reference counterpart (the
 * reference has no GPU path). */
cm_status cm_set_graphs(cm_ctx *ctx, int on);
/* Graphs captured and graph launches made by this context so far. */
cm_status cm_graph_counts(cm_ctx *ctx, int64_t *captured, int64_t *launched);

/* Device memory and copies for hosts that do not include HIP headers. */
cm_status cm_dev_alloc(cm_ctx *ctx, size_t bytes, void **ptr);
cm_status cm_dev_free(cm_ctx *ctx, void *ptr);
cm_status cm_memset(cm_ctx *ctx, void *dptr, int value, size_t bytes);
cm_status cm_copy_h2d(cm_ctx *ctx, void *dst, const void *src, size_t bytes);
cm_status cm_copy_d2h(cm_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Per-kernel timing with HIP events recorded on the stream each kernel is
 * launched on.  While enabled, every launch of the named kernels is bracketed
 * by events; cm_kernel_time returns the mean duration (ms) and count since the
 * last reset. */
cm_status cm_profile_enable(cm_ctx *ctx, int on);
cm_status cm_kernel_time(cm_ctx *ctx, const char *kernel, double *mean_ms, int64_t *count);
cm_status cm_profile_reset(cm_ctx *ctx);

/* ------------------------------------------------------------ mass assignment */
typedef struct {
    int N;                  /* Nft: cells per side */
    double L;               /* Lbox */
    double xmin[3];         /* box corner, Xoffset - Lbox/2 (Source/Params.cpp:3612) */
    cm_layout layout;       /* AoS xyz or SoA x[n],y[n],z[n] */
    cm_dtype pos_dtype;     /* float32 / float64 positions and weights */
    cm_mas scheme;
    double interlace_shift; /* cells added to every coordinate (0 or 0.5) */
    cm_mas_algo algo;
} cm_mas_desc;

/* Replaces grid_assignment_{NGP,CIC,TSC,PCS} over a catalogue
 * (Source/Miscelanious.cpp:2597-2868, driven by
 * Catalog::get_interpolated_density_field, Source/Catalog.cpp:14186-14211) and
 * getDensity_{NGP,CIC,TSC} (Source/Miscelanious.cpp:65-147,454-557,693-815).
 * pos/w/mesh are device pointers; w may be NULL (unit weights).  The mesh
 * (N^3, C order, z fastest) is overwritten unless `accumulate`.  Positions are
 * wrapped periodically in their own dtype; the cell index is
 * floor(double(x)/double(L/N)) as in get_bin (Miscelanious.cpp:1447-1457).
 * NGP with unit weights gives exact integer counts. */
cm_status cm_mesh_assign(cm_ctx *ctx, const cm_mas_desc *desc, const void *pos,
                         const void *w, size_t n, void *mesh, cm_dtype mesh_dtype,
                         int accumulate);

/* FP64 sum of a mesh: replaces get_nobjects (Source/Miscelanious.cpp:1421-1426),
 * which sums in float. */
cm_status cm_mesh_sum(cm_ctx *ctx, const void *mesh, cm_dtype dtype, size_t n, double *sum);

/* mesh[i] = mesh[i]*scale + shift: FftwFunctions::get_fluctuation box branch
 * (Source/FftwFunctions.cpp:293-294, shift = -N_gal/NGRID) and delta=rho/mean-1. */
cm_status cm_mesh_affine(cm_ctx *ctx, void *mesh, cm_dtype dtype, size_t n, double scale,
                         double shift);

/* ---------------------------------------------------------------------- FFT */
/* Sizes supported: N = 2^a * 3^b with b <= 1, N a multiple of 8, 8 <= N <= 1024. */
int cm_fft_supported(int N);

/* Replaces do_fftw_r2c (Source/NumericalMethods.cpp:1718-1753): unnormalised
 * forward transform e^{-ik.x} of a real N^3 mesh into the FFTW half spectrum
 * N x N x (N/2+1) interleaved (re,im).  `in` is not modified.  `subtract` is
 * removed from every input value on load (0 for a plain transform). */
cm_status cm_fft_r2c(cm_ctx *ctx, int N, cm_dtype dtype, const void *in, void *out,
                     double subtract);

/* Replaces do_fftw_c2r (Source/NumericalMethods.cpp:1818-1856): inverse of a
 * Hermitian half spectrum, divided by N^3.  `in` is used as scratch and is
 * overwritten (FFTW's c2r also destroys its input). */
cm_status cm_fft_c2r(cm_ctx *ctx, int N, cm_dtype dtype, void *in, void *out);

/* ---------------------------------------------------------------- P(k) shells */
typedef struct {
    int N;                /* Nft */
    double L;             /* Lbox */
    int nshells;          /* Nnp_data (Source/Params.cpp:4104) */
    double kmin;          /* DeltaKmin*deltak_0 (Source/Params.cpp:4099) */
    double dk;            /* ndel_data*deltak_0 (Source/Params.cpp:4101) */
    cm_mas scheme;        /* mass_assignment */
    int mas_correction;   /* MAS_correction */
    double shot_noise;    /* SN = V/N_gal, 0 to switch off (SN_correction) */
    double normal_power;  /* N_gal^2/V (FftwFunctions::raw_sampling, :159-174) */
    int float_shell_rule; /* 1 = the reference's float32 shell index
                             (FftwFunctions.cpp:1097); 0 = exact integer rule
                             (linear binning only) */
    int log_binning;      /* type_of_binning = log (Source/Params.cpp:4104-4106):
                             shell s = floor(log10(|k|/kmin)/dlog) + 1 for
                             s = 0..nshells-1 (FftwFunctions.cpp:1103-1108),
                             evaluated exactly (integer |k|^2 thresholds); dk
                             is unused; kmin must be > 0 */
    double dlog;          /* Deltal = log10(kmax/kmin)/N_log_bins */
    int window_set;       /* 1: the window W(k) of the random branch
                             (FftwFunctions.cpp:1248-1262, 1354-1568): the
                             zero mode (in the shell of k = 0) and the stored
                             modes outside the first octant (kx, ky both
                             below N/2 left out); pass the randoms' raw
                             spectrum, shot_noise 0 and normal_window =
                             normal_power/alpha^2 */
} cm_pk_opts;

/* Replaces FftwFunctions::power_spectrum_fkp box branch
 * (Source/FftwFunctions.cpp:958-1657): octant enumeration, |k| shells, aliased
 * shot noise and MAS deconvolution, normalisation; monopole pk0 and the z-LOS
 * multipoles pk2, pk4.  dk is a device half spectrum (cm_dtype); outputs are
 * host arrays of nshells (any may be NULL).  Shell sums are FP64; mode counts
 * exact. */
cm_status cm_pk_shells(cm_ctx *ctx, const cm_pk_opts *o, cm_dtype dtype, const void *dk,
                       double *pk0, double *pk2, double *pk4, int64_t *nmodes);

/* cm_pk_shells without any device allocation, for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results as cm_pk_shells; its work memory (tables, shell sums and counts,
 * log thresholds, results: 78,880 bytes) is one fixed array per device that
 * the library holds until the process ends.  Limits: N <= 1024 and
 * nshells <= 1024 (CM_ERR_UNSUPPORTED beyond).  Every call on a device uses
 * that device's one array, so calls that overlap on one device race; calls
 * in sequence (any contexts and streams) are safe: each starts by clearing
 * its sums.  Synchronises the context's stream like cm_pk_shells. */
cm_status cm_pk_shells_noalloc(cm_ctx *ctx, const cm_pk_opts *o, cm_dtype dtype, const void *dk,
                               double *pk0, double *pk2, double *pk4, int64_t *nmodes);

/* ---------------------------------------------------------------------- T-web */
typedef struct {
    int N;
    double L;
    double R_smooth;     /* Gaussian smoothing radius (smscale, Params.cpp:572); <=0: none */
    double lambda_th;    /* lambdath (Params.cpp:552) */
    double margin_tol;   /* cells with min_j |lambda_j - lambda_th| < tol are counted:
                            tol = margin_tol when >= 0; a negative value -s asks
                            for s times the rounding tolerance
                            K(N) u ||phi||_rms / d^2 (see CM_TWEB_TOL_ROUNDING) */
    double delta_scale;  /* input multiplied by this before the Poisson solve (1/mean for rho-mean) */
    int hessian;         /* CM_HESSIAN_FD (0): gradfindif applied twice to phi, the reference's default
                            branch (Source/Miscelanious.cpp:1684-1700); CM_HESSIAN_SPECTRAL (1): T_ab =
                            c2r(k_a k_b / k^2 W(k) delta_k), its _USE_GFFT_EIGENV_ branch (calc_LapPhiv,
                            Source/NumericalMethods.cpp:2605-2679; Source/Miscelanious.cpp:1703-1718), with
                            FP64 closed-form eigenvalues per cell and, for a negative margin_tol, the
                            rounding tolerance s K_s(N) u ||T||_F,rms, K_s = sqrt(log2 N^3)
                            sqrt(2 ln(3 N^3)).  cm_tweb and cm_tweb_f64 take either; the fused
                            measurement and cm_tweb_from_phi the FD Hessian only. */
} cm_tweb_opts;
#define CM_HESSIAN_FD 0
#define CM_HESSIAN_SPECTRAL 1

/* margin_tol value that counts the cells whose label the rounding of the
 * potential's arithmetic can change: tol = K(N) u ||phi||_rms / d^2 with
 * d = L/N, u the unit roundoff of phi (2^-24 float32, 2^-53 FP64) and
 * K(N) = 1.5 sqrt(log2 N^3) sqrt(2 ln(3 N^3)) (49.1 at N = 512), from a
 * random-rounding model: the FFT chain leaves phi a white error of rms
 * sqrt(log2 N^3) u ||phi||_rms; an eigenvalue moves by the second difference
 * of that error along its eigenvector, a stencil of L2 norm <= 1.5/d^2; and
 * sqrt(2 ln(3 N^3)) standard deviations bound the largest of the 3 N^3
 * eigenvalue errors.  ||phi||_rms is reduced on the device (no host
 * synchronisation); cm_tweb_stats.margin_tol returns the value applied. */
#define CM_TWEB_TOL_ROUNDING (-1.0)

typedef struct {
    int64_t counts[4];   /* cells labelled knot, filament, sheet, void (1..4) */
    int64_t n_marginal;  /* cells whose eigenvalue margin is below the tolerance */
    double margin_tol;   /* the absolute tolerance applied */
} cm_tweb_stats;

/* Replaces Cwclass::get_CWC (Source/Cwclass.cpp:235-894) with the Gaussian
 * pre-smoothing of kernelcomp (Source/NumericalMethods.cpp:2197-2306):
 * delta (device N^3 float32) -> r2c -> W(k)(-1/k^2) (PoissonSolver,
 * Source/Miscelanious.cpp:1189-1274) -> c2r -> 4th-order FD Hessian
 * (gradfindif o gradfindif, EigenValuesTweb FD branch,
 * Source/Miscelanious.cpp:1668-1769) -> FP64 3x3 eigenvalues -> label
 * 4 - #{lambda_j > lambda_th}.  labels: device int8 N^3; lambdas: device float32
 * 3*N^3 (planes l1|l2|l3, descending) or NULL; stats: host, may be NULL. */
cm_status cm_tweb(cm_ctx *ctx, const cm_tweb_opts *o, const float *delta, int8_t *labels,
                  float *lambdas, cm_tweb_stats *stats);

/* FP64 T-web: Cwclass::get_CWC of the reference's DOUBLE_PREC build
 * (real_prec = double, Headers/def.h:594-595; Source/Cwclass.cpp:235-894), the
 * same steps as cm_tweb with every step in double: FP64 r2c, multiplier, c2r,
 * FD stencils (NumericalMethods.cpp:1123-1217) and closed-form eigenvalues.
 * delta: device double N^3; labels: device int8 N^3; lambdas: device double
 * 3*N^3 or NULL; phi_out: device double N^3 receiving the potential, or NULL. */
cm_status cm_tweb_f64(cm_ctx *ctx, const cm_tweb_opts *o, const double *delta, int8_t *labels,
                      double *lambdas, double *phi_out, cm_tweb_stats *stats);

/* Replaces PoissonSolver (Source/Miscelanious.cpp:1189-1274) with optional
 * Gaussian smoothing: phi = c2r(W(k) * (-1/k^2) * r2c(delta*delta_scale)). */
cm_status cm_poisson(cm_ctx *ctx, const cm_tweb_opts *o, const float *delta, float *phi);

/* Replaces EigenValuesTweb FD branch + the label loop
 * (Source/Miscelanious.cpp:1668-1769, Source/Cwclass.cpp:874-894) for a given
 * potential phi (device float32 N^3). */
cm_status cm_tweb_from_phi(cm_ctx *ctx, const cm_tweb_opts *o, const float *phi,
                           int8_t *labels, float *lambdas, cm_tweb_stats *stats);

/* Per-cell index of the classification combos requested by the parameter file
 * (Cwclass::get_Tclassification, Source/Cwclass.cpp:1983-2030): cwt_used lists
 * codes 1..4 (one class: 1 knot, 2 filament, 3 sheet, 4 void) or the combos
 * 12, 23, 34, 123, 234.  With one code, every cell gets that code; with
 * several, a cell gets the index of the last code whose classes contain its
 * label, else 0 (codes outside this list, 0 included, match no cell).  labels:
 * device int8[n]; out: device int32[n]. */
cm_status cm_tweb_tclass(cm_ctx *ctx, const int8_t *labels, size_t n, const int *cwt_used, int ncwt, int32_t *out);

/* ------------------------------------------------- whole-mock measurement */
typedef struct {
    cm_mas_desc mas;     /* positions -> mesh */
    int nshells;
    double kmin, dk;     /* binning, as in cm_pk_opts */
    int mas_correction;
    int sn_correction;
    int float_shell_rule;
    int do_tweb;         /* 0: P(k) only */
    double R_smooth, lambda_th, margin_tol; /* margin_tol as in cm_tweb_opts */
    int fused;           /* 1: fused pipeline; 0: the plain step-by-step GPU path */
    int interlace;       /* 1 (fused only): second mesh of the particles shifted by
                            half a cell on every axis, spectra averaged as
                            (d1 + exp(i k.(d/2)(1,1,1)) d2)/2 and the aliased shot
                            noise (prod C1 + prod A)/2 with A the alternating
                            aliasing sum; no counterpart in the reference */
    int log_binning;     /* logarithmic shells, as in cm_pk_opts */
    double dlog;
} cm_measure_opts;

typedef struct {
    double n_obj;        /* FP64 mesh sum (N_gal) */
    double shot_noise, normal_power;
    cm_tweb_stats tweb;
} cm_measure_info;

/* One mock: the -p workflow for Input_type=catalog (PowerSpectrumF::
 * compute_power_spectrum, Source/PowerSpectrumF.cpp:741-1363) followed by the
 * T-web of delta = rho/mean - 1 (Bam's cwclass.do_CWC call,
 * Source/Bam.cpp:2210-2283).  pos (and w) are device arrays of n particles;
 * outputs: host pk0/pk2/pk4/nmodes[nshells] (may be NULL), device int8 labels
 * N^3 (NULL unless do_tweb).  The context keeps the meshes resident between
 * calls of the same N. */
cm_status cm_measure(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, const void *w,
                     size_t n, double *pk0, double *pk2, double *pk4, int64_t *nmodes,
                     int8_t *labels, cm_measure_info *info);

/* One mock of a packed measurement: the arguments of cm_measure. */
typedef struct {
    cm_measure_opts opts;
    const void *pos;     /* device particles (float32) */
    const void *w;       /* device weights or NULL */
    size_t n;
    double *pk0, *pk2, *pk4; /* host, opts.nshells each (pk2/pk4 may be NULL) */
    int64_t *nmodes;     /* host, opts.nshells (may be NULL) */
    int8_t *labels;      /* device int8 N^3 when opts.do_tweb */
    cm_measure_info *info; /* host, may be NULL */
} cm_packed_mock;

/* Packed measurement of a batch of uneven mocks (no reference counterpart;
 * the batched configuration).  Each mock gets exactly what cm_measure gives it,
 * but the tile deposits of all mocks run as one persistent launch per MAS
 * scheme over the union of their (mock, tile) work lists, and the T-web of all
 * mocks as one launch over their (mock, tile column, x segment) items; the
 * binning (exact record sizing, one host synchronisation per mock), the FFT
 * passes and the fused shell pass run per mock in between.  Mocks are taken
 * in groups that hold at most half of the context's free budget.  Mocks the
 * packed path does not take (plain or interlaced, N above 512, N not a
 * multiple of 64 with the T-web) run through cm_measure.  statuses (nullable):
 * per mock; returns the first failure. */
cm_status cm_measure_packed(cm_ctx *ctx, const cm_packed_mock *mocks, int nm, cm_status *statuses);

/* ------------------------------------------------------------- autotuner */
/* Launch shapes of the fused path; 0 = built-in choice. */
typedef struct {
    int xnb;     /* kz columns per workgroup of the fused x pass: 8 or 16 */
    int xe;      /* elements per thread of the fused x pass: 8 or 16 */
    int tz;      /* T-web tile width along z: 64, 32, 16 or 8 */
    int tseg;    /* T-web x segments per tile column (0: occupancy model) */
    int dep_wg;  /* deposit workgroups per CU (0: 8) */
} cm_tune;

/* Tile-shape autotuner (no reference counterpart).  Times the fused
 * measurement of the given request (o->fused is taken as 1) under candidate
 * shapes, one knob at a time in the order of the struct (coordinate descent on
 * the summed kernel time per mock, best of two runs per candidate, a change
 * kept only when it gains more than 1%), and stores the best shapes for meshes
 * of this N on this device in a process-wide table that every context's
 * cm_measure and cm_measure_fkp consult (server slots included).  best
 * (nullable) receives the shapes, ms_before / ms_after (nullable) the kernel
 * time per mock with the built-in and with the chosen shapes. */
cm_status cm_autotune(cm_ctx *ctx, const cm_measure_opts *o, const void *pos, const void *w, size_t n,
                      int8_t *labels, cm_tune *best, double *ms_before, double *ms_after);
/* Set (t NULL: clear) or read the process-wide shapes for meshes of N on the
 * context's device; cm_get_tune returns zeros when none are set. */
cm_status cm_set_tune(cm_ctx *ctx, int N, const cm_tune *t);
cm_status cm_get_tune(cm_ctx *ctx, int N, cm_tune *t);

/* ------------------------------------------------ FKP with a random catalogue */
typedef struct {
    double nbar;      /* mean number density n(r) of the sample (constant in the box);
                         <= 0: W_g / L^3 */
    double P_est;     /* FKP weight 1/(1 + nbar P_est) when fkp_weight (Params _Pest) */
    int fkp_weight;   /* Params _FKP_weight */
    int window;       /* also the window: shells of the randoms' power */
    const void *nbar_g; /* device float32 n(r) of each object (Halo.mean_density,
                           Source/Catalog.cpp:14150-14152); NULL: the constant nbar */
    const void *nbar_r; /* the same for each random; NULL: the constant nbar */
} cm_fkp_opts;

/* The random-catalogue branch of the estimator
 * (FftwFunctions::get_parameters_estimator, Source/FftwFunctions.cpp:177-200;
 * get_fluctuation, :275-281; per-object sums of Catalog::
 * get_interpolated_density_field, Source/Catalog.cpp:14153-14171): weights
 * w = w_user * w_fkp, w_fkp = 1/(1 + nbar P_est), W = sum w, S = sum w^2,
 * normal = sum nbar w^2 per catalogue, nbar per object when f->nbar_g / nbar_r
 * are given; alpha = W_g/W_r, delta = rho_g - alpha rho_r (one mesh: the randoms
 * are deposited with weight -alpha w), normal_power = alpha normal_r,
 * shot noise = (alpha + 1) S_r / normal_r.  pk_window (NULL unless
 * f->window) receives the randoms' shells over normal_power/alpha^2
 * (power_spectrum_fkp window branch, :1250-1262, :1635-1638).  Positions are
 * device float32 AoS; weights may be NULL.  The T-web and interlacing options
 * must be off. */
cm_status cm_measure_fkp(cm_ctx *ctx, const cm_measure_opts *o, const cm_fkp_opts *f, const void *pos,
                         const void *w, size_t n, const void *rpos, const void *rw, size_t nr, double *pk0,
                         double *pk2, double *pk4, int64_t *nmodes, double *pk_window, cm_measure_info *info);

/* ------------------------------------------------------------- bispectrum */
typedef struct {
    int N;
    double L;
    int nshells;      /* Nshells_bk: shells [kmin + s dk, kmin + (s+1) dk) */
    double kmin, dk;  /* d_kmin, DeltaK_Bis; the last edge must lie below N/3 k_f */
    int shell_min;    /* first shell of the triangle loop (Nshell_min) */
    cm_mas scheme;
    int mas_correction, sn_correction;
    double n_obj;     /* objects in the box: normal_b = n^3/V^2, nbar = n/V */
} cm_bk_opts;

/* Bispectrum of a box (FftwFunctions::get_bispectrum_fkp_fast and
 * loop_shells_bispectrum, Source/FftwFunctions.cpp:3358-3718, 4878) from the
 * device half spectrum dk (complex64, FFTW layout) of the fluctuation: for each
 * triangle bin (i >= j >= k, the reference's loop) ijk[3b..3b+2], the
 * bispectrum bk[b] with the Poisson shot noise sn[b] = <P_i+P_j+P_k>/nbar +
 * 1/nbar^2 subtracted (when sn_correction), and the number of closed ordered
 * triangles ntri[b].  *nbins receives the bin count; the arrays hold max_bins. */
cm_status cm_bispectrum(cm_ctx *ctx, const cm_bk_opts *o, const void *dk, int max_bins, int32_t *ijk,
                        double *bk, double *sn, double *ntri, int *nbins);

typedef struct {
    int N;            /* mesh side of the half spectrum (even, at most 1024) */
    double L;
    int nbins;        /* k bins s = 0 .. nbins-1, at most 64 */
    double kmin, dk;  /* bin s: kmin + s dk <= |k| < kmin + (s+1) dk (d_kmin, DeltaK_data) */
    cm_mas scheme;
    int mas_correction, sn_correction;
    double n_obj;     /* objects in the box: normal_b = n^3/V^2, normal_p = n^2/V */
} cm_bkd_opts;

/* Direct (brute-force) bispectrum of a box (FftwFunctions::get_bispectrum_fkp
 * and bispectrum_fkp, Source/FftwFunctions.cpp:3105-3350) from the device
 * half spectrum dk (complex64, FFTW layout) of the fluctuation: every ordered
 * pair (k1, k2) of the grid q in [-N/2, N/2]^3 (both Nyquist planes, like the
 * reference's (N+1)^3 loops) with k3 = -k1-k2 in the same grid and no vector
 * zero, binned by the sorted bin triple (a <= b <= c).  Host outputs of
 * nbins^3 entries at (a nbins + b) nbins + c (zero where a <= b <= c fails):
 * ntri (ordered pairs), bk = sum Re(d1 d2 d3)/(ntri normal_b) - sn with the
 * Poisson shot noise sn = <P1+P2+P3>/nbar + 1/nbar^2 (P_i = |d_i|^2/normal_p -
 * V/n) when sn_correction; optional raw sums bsum (sum Re(d1 d2 d3)) and snsum
 * (sum |d1|^2+|d2|^2+|d3|^2), which may be NULL.  Per bin (nbins entries): the
 * power pk = <|d|^2>/normal_p - V/n (sn_correction) over k1 and its mode count
 * nk (:3194-3199, 3336).  d = delta/W with the MAS window W when mas_correction. */
cm_status cm_bispectrum_direct(cm_ctx *ctx, const cm_bkd_opts *o, const void *dk, double *bk, double *sn,
                               int64_t *ntri, double *pk, int64_t *nk, double *bsum, double *snsum);

/* ----------------------------------------------------------- synthetic mocks */
typedef struct {
    int NL;          /* Lagrangian grid per side (an FFT size) */
    double L;        /* box side */
    double A, ns;    /* P(k) = A k^ns */
    uint64_t seed;
} cm_zeldovich_opts;

/* GPU replacement of create_GARFIELD_FIXED_AMP (Source/NumericalMethods.cpp:
 * 2147-2189) followed by the first-order LPT displacement
 * (LPT::get_displacement, Source/LPT.cpp:285): fixed-amplitude Gaussian field
 * with Philox4x32-10 phases keyed by (seed, mode index), Psi(k) = i k delta/k^2,
 * particle (i,j,k) at row (i NL + j) NL + k of pos (device, NL^3 x 3 float32,
 * AoS), positions wrapped into [0, L). */
cm_status cm_zeldovich(cm_ctx *ctx, const cm_zeldovich_opts *o, float *pos);

/* n uniform positions in [0, L)^3 (device, n x 3 float32 AoS), Philox4x32-10
 * of counter (i, 0, 1, 0) and key seed: the Poisson randoms of an FKP
 * catalogue (no reference generator; Catalog reads them from files). */
cm_status cm_uniform(cm_ctx *ctx, size_t n, double L, uint64_t seed, float *pos);

/* ------------------------------------------------- batching and serving */
/* One measurement request: the options of cm_measure, the input (device
 * particles, or pos == NULL and gen.NL > 0 for a seeded Zel'dovich mock made on
 * the GPU), and the outputs (host pk0/pk2/pk4/nmodes[nshells] and info, device
 * labels when opts.do_tweb).  The buffers must stay valid until the request
 * completes.  No counterpart in the reference, whose drivers
 * (PowerSpectrumF::compute_power_spectrum, Source/PowerSpectrumF.cpp:741-1363;
 * Bam, Source/Bam.cpp:2210-2283) measure one catalogue per call. */
typedef struct {
    cm_measure_opts opts;
    const void *pos;
    const void *w;
    size_t n;
    cm_zeldovich_opts gen;
    double *pk0, *pk2, *pk4;
    int64_t *nmodes;
    int8_t *labels;
    cm_measure_info *info;
} cm_request;

typedef struct cm_server cm_server;

/* Device bytes a request holds while it runs (the admission footprint). */
size_t cm_request_footprint(const cm_request *r);

/* Continuous-batching server on `device`: `slots` (1..16) concurrent requests,
 * each on its own stream with its own resident meshes.  All slots share one
 * HBM budget (hbm_budget, 0: 90% of the device).  A queued request starts when
 * a slot is free and its reservation (footprint + 25% + 256 MB, or what the
 * slot already holds when more) fits beside what the other slots hold:
 * running requests' reservations and idle slots' resident buffers, which are
 * freed, largest first, when the request needs the room.  A running slot may
 * not grow past its reservation; a request that would is queued again to run
 * alone with the whole budget. */
cm_status cm_server_create(int device, int slots, size_t hbm_budget, cm_server **out);

/* Queue a request (thread-safe, returns at once).  CM_ERR_BUDGET if the
 * request alone exceeds the budget. */
cm_status cm_server_submit(cm_server *s, const cm_request *r, int64_t *ticket);

/* Wait for a ticket (timeout_s <= 0: no limit; CM_ERR_TIMEOUT leaves it
 * pending).  Returns the request's status; latency_ms = completion - submission;
 * err receives the request's error message.  A ticket is waited for once. */
cm_status cm_server_wait(cm_server *s, int64_t ticket, double timeout_s, double *latency_ms, char *err,
                         size_t errlen);

/* Cancel queued requests (their waits return CM_ERR_CANCELLED), finish the
 * running ones, free the slots. */
cm_status cm_server_destroy(cm_server *s);

/* Admission counters: idle-slot evictions and the bytes they freed, the most
 * the slots held together (reservations of running requests plus idle slots'
 * resident buffers, tables aside), the budget, and the requests that outgrew
 * their reservation and ran again alone.  Any pointer may be NULL. */
cm_status cm_server_stats(cm_server *s, int64_t *evictions, size_t *evicted_bytes, size_t *peak_held,
                          size_t *budget, int64_t *reruns);

/* Message of the most recent request that failed on this server ("" if none),
 * for callers of cm_measure_batch, which returns status codes only. */
cm_status cm_server_last_error(cm_server *s, char *buf, size_t len);

/* Submit n requests and wait for all: per-request status and latency (either
 * array may be NULL); returns the first failure or CM_OK. */
cm_status cm_measure_batch(cm_server *s, const cm_request *reqs, int n, cm_status *statuses,
                           double *latency_ms);

#ifdef __cplusplus
}
#endif
#endif
